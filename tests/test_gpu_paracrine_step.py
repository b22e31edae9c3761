"""The paracrine solve inside the timestep (nnx_step_plan.paracrine): the
north-star loop HH + coupling + paracrine CG closed on the device, replayed
from hipGraphs with no host read.

Checked against the oracle on the same seeded inputs: rasters and the
release counts each solve consumes bit for bit; each solve's field by its
FP64 relative residual with the oracle's operator (<= 1e-8, the stated
tolerance) and against the exact DST-I solution of the same right-hand side;
and the capturable CG (nnx_cg_solve_async) against the host-checked one."""
import ctypes as C

import numpy as np
import pytest
import scipy.fft

import oracle
from neuronephys_clean_amd import _lib, configs
from neuronephys_clean_amd.devmem import DeviceArray as DA
from neuronephys_clean_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def _lam7(nx, ny, nz, c):
    th = [2 * np.cos(np.pi * np.arange(1, m + 1) / (m + 1)) for m in (nz, ny, nx)]
    tz, ty, tx = np.meshgrid(*th, indexing="ij")
    return 1 + c * ((2 - tx) + (2 - ty) + (2 - tz))


def _dst_solve(b, lam):
    return scipy.fft.dstn(scipy.fft.dstn(b, type=1, norm="ortho") / lam, type=1, norm="ortho")


def _oracle_windows(net, steps, every):
    """Oracle fp32 run of `steps` steps: its raster and the per-voxel
    release counts of each window of `every` steps (from the raster)."""
    p = oracle.params(_lib.params_squid().as_dict())
    T = np.float32
    V, m, h, ng = (np.asarray(a, T).copy() for a in (net.V0, net.m0, net.h0, net.n0))
    s = np.zeros(net.n, T)
    iext = np.asarray(net.iext, T)
    nvox = int(np.prod(net.grid))
    vox = net.ensure_voxel()
    cap = 1 << 21
    ras, ln, vc, _ = oracle.run_rk4(p, net.dt, steps, net.off, net.tgt, net.w, V, m, h, ng, s, iext,
                                    cap=cap, voxel_of=vox, nvox=nvox)
    assert ln <= cap
    counts = []
    for w in range(steps // every):
        sel = (ras[:, 0] >= w * every) & (ras[:, 0] < (w + 1) * every)
        counts.append(np.bincount(vox[ras[sel, 1]], minlength=nvox).astype(np.int32))
    assert np.array_equal(np.sum(counts, axis=0), vc)
    return ras.copy(), counts


@pytest.mark.parametrize("couple,graph", [("event", False), ("event", True), ("fused", True)])
def test_paracrine_in_step_matches_oracle(ctx, couple, graph):
    net = configs.c1(seed=0)
    every, windows = 50, 4
    say(f"paracrine in step: c1 n={net.n} couple={couple} graph={graph} every={every}")
    ref_ras, ref_counts = _oracle_windows(net, every * windows, every)
    sim = Simulation(net, np.float32, couple=couple, record_cap=1 << 20, deposit=True, ctx=ctx)
    sim.set_paracrine(every, rtol=1e-8, max_inner=48, max_outer=3, at_neurons=True)
    nx, ny, nz = net.grid
    og = oracle.grid(nx, ny, nz, 7, net.c)
    lam = _lam7(nx, ny, nz, net.c)
    for w in range(windows + 1):
        conc_prev = sim.a["conc"].numpy()
        counts = sim.a["voxel_counts"].numpy()
        # the solve at step w*every consumes the releases of the window before it
        if w > 0:
            assert np.array_equal(counts, ref_counts[w - 1]), f"window {w - 1} release counts differ"
        else:
            assert not counts.any()
        if w == windows:
            sim.run(1, use_graph=False)  # just the solve of the last window (and one step)
        else:
            sim.run(every, use_graph=graph, chunk=every if graph else None)
        conc = sim.a["conc"].numpy()
        b = conc_prev + net.amount * counts.astype(np.float64)
        if not b.any():
            assert not conc.any()
            continue
        rel = np.linalg.norm(b - oracle.apply(og, conc)) / np.linalg.norm(b)
        xs = _dst_solve(b.reshape(nz, ny, nx), lam).ravel()
        err = np.linalg.norm(conc - xs) / np.linalg.norm(xs)
        say(f"  solve {w}: releases {int(counts.sum())}, FP64 relres {rel:.2e}, error vs DST {err:.2e}")
        assert rel <= 1e-8
        assert err <= 1e-7
        at = sim.a["at_neurons"].numpy()
        assert np.array_equal(at, conc[net.ensure_voxel()].astype(np.float32))
    st = sim.paracrine_stats()
    say("  stats", st)
    assert st["solves"] == windows + 1 and st["unconverged"] == 0 and st["max_relres"] <= 1e-8
    ras, ln = sim.raster()
    ras = ras[ras[:, 0] < every * windows]
    assert len(ras) == len(ref_ras) and np.array_equal(ras, ref_ras)


def test_paracrine_graph_replays_equal_eager(ctx):
    """Graph replays of a 100-step chunk with a solve at its first step give
    the eager run's field and rasters bit for bit, replay after replay."""
    net = configs.c1(seed=1)
    sims = []
    for graph in (False, True):
        sim = Simulation(net, np.float32, couple="event", record_cap=1 << 20, deposit=True, ctx=ctx)
        sim.set_paracrine(100, max_inner=40, max_outer=2)
        for _ in range(4):
            sim.run(100, use_graph=graph, chunk=100 if graph else None)
        sim.synchronize()
        sims.append(sim)
    a, b = sims
    assert np.array_equal(a.a["conc"].numpy(), b.a["conc"].numpy())
    assert np.array_equal(a.raster()[0], b.raster()[0])
    sa, sb = a.paracrine_stats(), b.paracrine_stats()
    assert sa == sb and sa["solves"] == 4 and sa["unconverged"] == 0


def test_cg_solve_async_equals_host_checked(ctx):
    """The capturable CG (fixed iteration budget, device stop test) performs
    the same iterations as nnx_cg_solve and leaves the same bits."""
    for nx, ny, nz, c, var in ((64, 48, 40, 1.0, False), (30, 22, 18, 4.0, True)):
        n = nx * ny * nz
        rng = np.random.default_rng(5)
        b = rng.random(n).astype(np.float32)
        keep = []
        if var:
            cf = (4.0 * np.exp(rng.normal(0, 0.5, n))).astype(np.float32)
            from neuronephys_clean_amd import host
            kx, ky, kz = host.face_coefficients(cf.reshape(nz, ny, nx))
            keep = [DA.from_numpy(ctx, k) for k in (kx, ky, kz)]
            g = _lib.Grid(nx, ny, nz, 7, 0.0, 0.0, keep[0].ptr, keep[1].ptr, keep[2].ptr)
        else:
            g = _lib.Grid(nx, ny, nz, 7, c, 0.0, None, None, None)
        db = DA.from_numpy(ctx, b)
        x1, x2 = DA.zeros(ctx, n, np.float32), DA.zeros(ctx, n, np.float32)
        o = _lib.CgOpts(rtol=1e-6, maxit=400, precond=_lib.PRECOND_JACOBI, check_every=1)
        r = _lib.CgResult()
        ctx.call("nnx_cg_solve", C.byref(g), db.ptr, x1.ptr, C.byref(o), C.byref(r))
        res = DA.zeros(ctx, 4, np.float64)
        ctx.call("nnx_cg_solve_async", C.byref(g), db.ptr, x2.ptr, C.byref(o), res.ptr)
        out = res.numpy()
        say(f"async CG {nx}x{ny}x{nz} var={var}: host-checked {r.iters} it, async {out}")
        assert int(out[0]) == r.iters and out[2] == 1.0
        assert abs(out[1] - r.relres) <= 1e-12 * max(1.0, r.relres)
        assert np.array_equal(x1.numpy(), x2.numpy())
