"""nnx_cg_solve_async_noalloc, the allocation-free variant of
nnx_cg_solve_async (its working set in the library's fixed per-GPU array),
called through the C ABI: same iterations, same solution bits and same
result record as nnx_cg_solve_async on two calls in a row, the second on a
smaller grid (the array keeps the first call's contents), and a grid beyond
the array's size refused before anything is enqueued."""
import ctypes as C

import numpy as np
import pytest

from neuronephys_clean_amd import _lib, host
from neuronephys_clean_amd.devmem import DeviceArray as DA

pytestmark = pytest.mark.gpu


def _grid(ctx, nx, ny, nz, c, variable, seed):
    keep = []
    if variable:
        rng = np.random.default_rng(seed)
        cf = (c * np.exp(rng.normal(0.0, 0.5, nx * ny * nz))).astype(np.float32)
        keep = [DA.from_numpy(ctx, k) for k in host.face_coefficients(cf.reshape(nz, ny, nx))]
        return _lib.Grid(nx, ny, nz, 7, 0.0, 0.0, keep[0].ptr, keep[1].ptr, keep[2].ptr), keep
    return _lib.Grid(nx, ny, nz, 7, c, 0.0, None, None, None), keep


def _solve(ctx, name, g, b, rtol, maxit):
    db = DA.from_numpy(ctx, b)
    dx = DA.zeros(ctx, b.size, np.float32)
    res = DA.zeros(ctx, 4, np.float64)
    o = _lib.CgOpts(rtol=rtol, maxit=maxit, precond=_lib.PRECOND_JACOBI, check_every=1)
    ctx.call(name, C.byref(g), db.ptr, dx.ptr, C.byref(o), res.ptr)
    return dx.numpy(), res.numpy()


@pytest.mark.parametrize("order", ["noalloc_first", "async_first"])
def test_noalloc_equals_async_on_two_calls(ctx, order):
    cases = [(96, 96, 96, 1.0, False, 1e-7, 200),    # constant coefficients: the fused stencil kernel
             (40, 36, 30, 4.0, True, 1e-6, 400)]     # smaller, variable coefficients: Jacobi diagonal in the array
    for nx, ny, nz, c, var, rtol, maxit in cases:
        print(f"noalloc CG {nx}x{ny}x{nz} c={c} variable={var} ({order})", flush=True)
        g, keep = _grid(ctx, nx, ny, nz, c, var, seed=nx)
        b = np.random.default_rng(nx + 1).random(nx * ny * nz).astype(np.float32)
        names = ["nnx_cg_solve_async_noalloc", "nnx_cg_solve_async"]
        if order == "async_first":
            names.reverse()
        out = {nm: _solve(ctx, nm, g, b, rtol, maxit) for nm in names}
        (xn, rn), (xa, ra) = out["nnx_cg_solve_async_noalloc"], out["nnx_cg_solve_async"]
        print(f"  iterations {rn[0]:.0f}, relres {rn[1]:.3e}, converged {rn[2]:.0f}", flush=True)
        assert rn[2] == 1.0 and 0 < rn[0] < maxit
        assert np.array_equal(rn, ra)
        assert np.array_equal(xn, xa)
        del keep


def test_noalloc_refuses_a_grid_beyond_its_array(ctx):
    g = _lib.Grid(256, 256, 256, 7, 1.0, 0.0, None, None, None)
    b, x, res = DA.zeros(ctx, 1, np.float32), DA.zeros(ctx, 1, np.float32), DA.zeros(ctx, 4, np.float64)
    o = _lib.CgOpts(rtol=1e-6, maxit=10, precond=_lib.PRECOND_JACOBI, check_every=1)
    with pytest.raises(_lib.NNXError) as e:
        ctx.call("nnx_cg_solve_async_noalloc", C.byref(g), b.ptr, x.ptr, C.byref(o), res.ptr)
    assert e.value.status == 1 and "working memory" in str(e.value)
    ctx.synchronize()
