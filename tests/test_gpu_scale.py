"""Full-size solves checked through size-independent properties.

The precision-split solve (512^3 grid, dt*D/h^2 = 20, 50,000 unit point
releases): fp32 inner CG with FP64 dots, FP64 outer refinement to a 1e-10
relative residual.  The residual of the returned FP64 solution is
recomputed on the host by the oracle's FP64 operator.
"""
import ctypes as C

import numpy as np
import pytest

import oracle
from neuronephys_clean_amd import _lib, configs
from neuronephys_clean_amd.devmem import DeviceArray as DA

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_c3_refinement_512_to_1e10(ctx):
    n, c = configs.C3_GRID, configs.C3_C
    b = configs.c3_sources(n=n)
    say("c3 refine", n, c, "sources", int(b.sum()))
    g = _lib.Grid(n, n, n, 7, c, 0.0, None, None, None)
    db, dx = DA.from_numpy(ctx, b), DA.zeros(ctx, n ** 3, np.float64)
    o = _lib.CgOpts(rtol=1e-10, maxit=5000, precond=_lib.PRECOND_JACOBI, check_every=8)
    r = _lib.CgResult()
    ctx.call("nnx_cg_solve_refine", C.byref(g), db.ptr, dx.ptr, C.byref(o), 1e-6, 30, C.byref(r))
    say("c3 refine result", r.as_dict())
    assert r.converged and r.relres <= 1e-10 and r.outer >= 2
    x = dx.numpy()
    del db, dx
    res = b - oracle.apply(oracle.grid(n, n, n, 7, c), x)
    rel = np.linalg.norm(res) / np.linalg.norm(b)
    say("c3 host FP64 residual", rel)
    assert rel <= 1.05e-10
    # the solution of a non-negative source under an M-matrix is non-negative
    assert x.min() > -1e-12 * x.max()
