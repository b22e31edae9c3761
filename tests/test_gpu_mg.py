"""Multigrid V-cycle (nnx_mg_apply) and flexible CG preconditioned by it, on
the variable-diffusivity operator: the cycle contracts the error in the
energy norm, its iteration count does not grow with the grid (the property
that separates a working multigrid from a broken one), and solves reach the
FP64 tolerance with far fewer iterations than Jacobi."""
import ctypes as C

import numpy as np
import pytest

import mg_ref
from neuronephys_clean_amd import _lib, configs
from neuronephys_clean_amd.devmem import DeviceArray as DA
from neuronephys_clean_amd.precond import UNet, device_field

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def apply_A(ctx, g, x):
    dx = DA.from_numpy(ctx, x.astype(np.float32))
    dy = DA.zeros(ctx, x.size, np.float32)
    ctx.call("nnx_stencil_apply", C.byref(g), dx.ptr, dy.ptr)
    return dy.numpy().astype(np.float64)


def const_field(n, k):
    kx = np.full((n, n, n + 1), k, np.float32).ravel()  # (nx+1) faces fastest: same count for all three
    return kx


@pytest.mark.parametrize("dims", [(32, 32, 32), (32, 16, 24), (12, 10, 8), (64, 64, 64)])
@pytest.mark.parametrize("nu", [1, 2, 4])
@pytest.mark.parametrize("smoother", ["chebyshev", "jacobi"])
@pytest.mark.parametrize("gamma", [1, 2])
def test_vcycle_matches_numpy_restatement(ctx, dims, nu, smoother, gamma):
    """nnx_mg_apply against tests/mg_ref.py (FP64) on log-normal faces,
    including grids whose coarsening stops early (odd or small dims)."""
    nx, ny, nz = dims
    rng = np.random.default_rng(nx * 7 + ny + nz + nu)
    kx = rng.lognormal(2.0, 1.0, (nz, ny, nx + 1)).astype(np.float32)
    ky = rng.lognormal(2.0, 1.0, (nz, ny + 1, nx)).astype(np.float32)
    kz = rng.lognormal(2.0, 1.0, (nz + 1, ny, nx)).astype(np.float32)
    r = rng.normal(size=(nz, ny, nx)).astype(np.float32)
    d = [DA.from_numpy(ctx, a.ravel()) for a in (kx, ky, kz, r)]
    z = DA.zeros(ctx, r.size, np.float32)
    g = _lib.Grid(nx, ny, nz, 7, 0.0, 0.0, d[0].ptr, d[1].ptr, d[2].ptr)
    sm = _lib.MG_CHEBYSHEV if smoother == "chebyshev" else _lib.MG_JACOBI
    ctx.call("nnx_mg_apply", C.byref(g), nu, sm, gamma, d[3].ptr, z.ptr)
    ref = mg_ref.vcycle(r.astype(np.float64), kx.astype(np.float64), ky.astype(np.float64),
                        kz.astype(np.float64), nu=nu, smoother=smoother, gamma=gamma)
    got = z.numpy().reshape(nz, ny, nx)
    err = np.abs(got - ref).max() / np.abs(ref).max()
    say("dims", dims, "nu", nu, smoother, "gamma", gamma, "max rel diff", err)
    assert err < 1e-4


@pytest.mark.parametrize("kind", ["smooth", "rough"])
@pytest.mark.parametrize("smoother", [0, 1])
def test_vcycle_contracts_energy_norm(ctx, kind, smoother):
    n = 64
    f = configs.c5_field(seed=8, n=n)
    g, keep = device_field(ctx, f)
    rng = np.random.default_rng(3)
    if kind == "rough":
        e = rng.normal(0, 1, n ** 3)
    else:
        ax = np.arange(n) + 0.5
        X, Y, Z = np.meshgrid(ax, ax, ax, indexing="ij")
        e = (np.sin(np.pi * X / n) * np.sin(2 * np.pi * Y / n) * np.sin(np.pi * Z / n)).transpose(2, 1, 0).ravel()
        e = e + 0.05 * rng.normal(0, 1, n ** 3)
    e = e.astype(np.float32).astype(np.float64)
    Ae = apply_A(ctx, g, e)
    r, z = DA.from_numpy(ctx, Ae.astype(np.float32)), DA.zeros(ctx, n ** 3, np.float32)
    for nu in (1, 2):
        ctx.call("nnx_mg_apply", C.byref(g), nu, smoother, 1, r.ptr, z.ptr)
        e1 = e - z.numpy().astype(np.float64)
        ratio = np.sqrt(e1 @ apply_A(ctx, g, e1) / (e @ Ae))
        say(kind, "smoother", smoother, "nu", nu, "energy-norm contraction", ratio)
        assert ratio < (0.6 if nu == 1 else 0.45)


def _fcg(ctx, g, b, pc, net=None, feat=None, rtol=1e-6, nu=0):
    x = DA.zeros(ctx, b.size, np.float32)
    db = DA.from_numpy(ctx, b.astype(np.float32))
    o = _lib.CgOpts(rtol=rtol, maxit=5000, precond=pc, check_every=1, stall_iters=10, mg_nu=nu)
    r = _lib.CgResult()
    ctx.call("nnx_fcg_solve", C.byref(g), net.h if net else None, feat.ptr if feat else None, db.ptr, x.ptr,
             C.byref(o), C.byref(r))
    return r


def test_fcg_mg_far_fewer_iterations_than_jacobi(ctx):
    n = 64
    f = configs.c5_field(seed=9, n=n)
    g, keep = device_field(ctx, f)
    b = configs.point_sources(21, n, 300)
    rj = _fcg(ctx, g, b, _lib.PRECOND_JACOBI)
    rm = _fcg(ctx, g, b, _lib.PRECOND_MG)
    say("jacobi", rj.as_dict(), "mg", rm.as_dict())
    assert rm.converged and rm.precond_used == _lib.PRECOND_MG
    assert rm.true_relres < 3 * max(rj.true_relres, 1e-6)
    assert 5 * rm.iters < rj.iters


def test_fcg_mg_iterations_do_not_grow_with_the_grid(ctx):
    its = {}
    for n in (32, 64, 128):
        k = DA.from_numpy(ctx, const_field(n, 50.0))
        g = _lib.Grid(n, n, n, 7, 0.0, 0.0, k.ptr, k.ptr, k.ptr)
        b = configs.point_sources(22, n, 50)
        r = _fcg(ctx, g, b, _lib.PRECOND_MG, rtol=1e-8)
        its[n] = r.iters
        say("n", n, r.as_dict())
        assert r.converged
    assert its[128] <= its[32] + 4


def test_fcg_mg_refined_and_mg_learned(ctx):
    n = 64
    f = configs.c5_field(seed=10, n=n)
    g, keep = device_field(ctx, f)
    b = configs.point_sources(23, n, 300).astype(np.float64)
    db, dx = DA.from_numpy(ctx, b), DA.zeros(ctx, n ** 3, np.float64)
    net = UNet(ctx, seed=5)
    for pc in (_lib.PRECOND_MG, _lib.PRECOND_MG_LEARNED):
        o = _lib.CgOpts(rtol=1e-10, maxit=3000, precond=pc, check_every=1, stall_iters=10, smooth=1)
        r = _lib.CgResult()
        ctx.call("nnx_fcg_solve_refine", C.byref(g), net.h, keep["feat"].ptr, db.ptr, dx.ptr, C.byref(o),
                 1e-6, 20, C.byref(r))
        say("refine", pc, r.as_dict())
        assert r.converged and r.relres <= 1e-10
        # the true residual, recomputed on the host in FP64
        x = dx.numpy().reshape(n, n, n)
        res = b.reshape(n, n, n) - mg_ref.apply(x, *(a.astype(np.float64) for a in (f.kx, f.ky, f.kz)))
        rel = np.linalg.norm(res) / np.linalg.norm(b)
        say("host FP64 residual", rel)
        assert rel < 2e-10


def test_multigrid_argument_errors(ctx):
    """Multigrid needs the variable-coefficient operator and a known
    smoother; the errors come back as status codes, not faults."""
    n = 16
    b, x = DA.from_numpy(ctx, np.ones(n ** 3, np.float32)), DA.zeros(ctx, n ** 3, np.float32)
    g = _lib.Grid(n, n, n, 7, 1.0, 0.0, None, None, None)
    r = _lib.CgResult()
    o = _lib.CgOpts(rtol=1e-6, maxit=100, precond=_lib.PRECOND_MG, check_every=1)
    st = ctx._l.nnx_fcg_solve(ctx.h, C.byref(g), None, None, b.ptr, x.ptr, C.byref(o), C.byref(r))
    assert st == 1 and "variable-coefficient" in ctx._l.nnx_last_error(ctx.h).decode()
    assert ctx._l.nnx_mg_apply(ctx.h, C.byref(g), 2, 0, 0, b.ptr, x.ptr) == 1
    k = DA.from_numpy(ctx, const_field(n, 5.0))
    gv = _lib.Grid(n, n, n, 7, 0.0, 0.0, k.ptr, k.ptr, k.ptr)
    assert ctx._l.nnx_mg_apply(ctx.h, C.byref(gv), 2, 7, 0, b.ptr, x.ptr) == 1
    assert ctx._l.nnx_mg_apply(ctx.h, C.byref(gv), 17, 0, 0, b.ptr, x.ptr) == 1
    assert ctx._l.nnx_mg_apply(ctx.h, C.byref(gv), 2, 0, 3, b.ptr, x.ptr) == 1
    o2 = _lib.CgOpts(rtol=1e-6, maxit=100, precond=_lib.PRECOND_MG, check_every=1, mg_smoother=5)
    assert ctx._l.nnx_fcg_solve(ctx.h, C.byref(gv), None, None, b.ptr, x.ptr, C.byref(o2), C.byref(r)) == 1
