"""The batched ensemble (nnx_ensemble_run / nnx_ensemble_autotune) on the
GPU: every network's trajectory equals the oracle's run of that network
alone, bit for bit; results do not depend on the launch configuration or on
graph replay; the autotuner leaves the run untouched."""
import ctypes as C

import numpy as np
import pytest
import scipy.fft

import oracle
from neuronephys_clean_amd import _lib, configs
from neuronephys_clean_amd.sim import EnsembleSim

pytestmark = pytest.mark.gpu

SIZES = [1000, 3000, 1500, 2048, 1024, 4000, 700, 2500]


def say(*a):
    print(*a, flush=True)


@pytest.fixture(scope="module")
def ens():
    return configs.c4_ensemble(seed=3, count=len(SIZES), sizes=SIZES, keep_members=True)


def oracle_member(net, steps):
    T = np.float32
    V, m, h, ng = (np.asarray(a, T).copy() for a in (net.V0, net.m0, net.h0, net.n0))
    s = np.zeros(net.n, T)
    p = oracle.params(_lib.params_squid().as_dict())
    ras, ln, _, _ = oracle.run_rk4(p, net.dt, steps, net.off, net.tgt, net.w, V, m, h, ng, s,
                                   np.asarray(net.iext, T), cap=1 << 20)
    return ras, ln, {"v": V, "m": m, "h": h, "ng": ng, "s": s}


def test_ensemble_matches_oracle_per_network(ctx, ens):
    steps = 400
    say("ensemble", len(SIZES), "networks", ens.n, "neurons", ens.nnz, "edges", steps, "steps")
    sim = EnsembleSim(ens, ctx=ctx, record_cap=1 << 20, deposit=True)
    sim.run(steps)
    sim.synchronize()
    got = sim.state()
    ras, ln = sim.raster()
    assert ln == len(ras)
    total = 0
    for i, net in enumerate(ens.members):
        ref, rln, st = oracle_member(net, steps)
        b, s = int(ens.begin[i]), int(ens.size[i])
        for k in ("v", "m", "h", "ng", "s"):
            assert np.array_equal(got[k][b:b + s], st[k]), (i, k)
        mine = ras[(ras[:, 1] >= b) & (ras[:, 1] < b + s)].copy()
        mine[:, 1] -= b
        assert np.array_equal(mine, ref), i
        total += rln
    say("spikes", total)
    assert total == ln and total > 0
    assert np.array_equal(sim.a["net_spikes"].numpy(),
                          [np.count_nonzero((ras[:, 1] >= b) & (ras[:, 1] < b + s))
                           for b, s in zip(ens.begin, ens.size)])
    vc = sim.a["voxel_counts"].numpy()
    ref_vc = np.bincount(ens.voxel_of[ras[:, 1]], minlength=vc.size)
    assert np.array_equal(vc, ref_vc)


@pytest.mark.parametrize("cfg", [(64, 64, 7, 1), (512, 8192, 3, -1), (128, 512, 2048, 0), (256, 256, 512, 50)])
def test_ensemble_result_independent_of_config_and_graph(ctx, ens, cfg):
    """Any launch shape, with events pushed with target flags (bulk -1),
    without them (bulk 1) or switching (bulk 50, and 0: the library's
    rule), gives the default run's bits."""
    steps = 200
    a = EnsembleSim(ens, ctx=ctx, record_cap=1 << 20)
    a.run(steps)
    b = EnsembleSim(ens, ctx=ctx, record_cap=1 << 20)
    c = _lib.EnsConfig(*cfg)
    say("config", cfg)
    b.run(steps, cfg=c, use_graph=True, chunk=100)
    a.synchronize()
    b.synchronize()
    sa, sb = a.state(), b.state()
    for k in sa:
        assert np.array_equal(sa[k], sb[k]), k
    ra, rb = a.raster(), b.raster()
    assert ra[1] == rb[1] and np.array_equal(ra[0], rb[0])


def test_ensemble_autotune_leaves_the_run_untouched(ctx, ens):
    """The search times every candidate on the same window (the first 150
    steps: the initial volley) and restores the state after it: the run
    that follows equals one that never tuned, outputs included."""
    t = EnsembleSim(ens, ctx=ctx, record_cap=1 << 20, deposit=True)
    t.run(37)
    best, us, ran = t.autotune(steps_per_candidate=150)
    say("autotune", best.as_dict(), us, "us/step over", ran, "steps")
    assert ran == 0 and us > 0
    assert best.block in (64, 128, 256, 512) and best.tile % 64 == 0 and best.grid > 0
    t.run(163, use_graph=True, chunk=50)
    ref = EnsembleSim(ens, ctx=ctx, record_cap=1 << 20, deposit=True)
    ref.run(200)
    t.synchronize()
    ref.synchronize()
    assert t.step == ref.step == 200
    for k, v in t.state().items():
        assert np.array_equal(v, ref.state()[k]), k
    assert t.raster()[1] == ref.raster()[1] > 0
    assert np.array_equal(t.raster()[0], ref.raster()[0])
    for k in ("net_spikes", "voxel_counts"):
        assert np.array_equal(t.a[k].numpy(), ref.a[k].numpy()), k


def test_full_ensemble_matches_oracle_per_network(ctx):
    """The full batched configuration (256 networks of 1,024 to 65,536
    neurons, 4.56M neurons, 2.3e8 edges, the benchmark's seed) for its first
    5 ms (500 steps: the initial volley, in which every neuron fires) with
    the library's default launch shape in graphs of 100: every network's
    raster and state equal the oracle's run of that network alone."""
    full = configs.c4_ensemble(seed=0, count=256, keep_members=True)
    steps = 500
    say("full ensemble", len(full.members), "networks", full.n, "neurons", full.nnz, "edges", steps, "steps")
    sim = EnsembleSim(full, ctx=ctx, record_cap=8 << 20)
    sim.run(steps, use_graph=True, chunk=100)
    sim.synchronize()
    got = sim.state()
    ras, ln = sim.raster()
    del sim
    net_of = np.searchsorted(np.asarray(full.begin), ras[:, 1], side="right") - 1
    order = np.argsort(net_of, kind="stable")
    ras, net_of = ras[order], net_of[order]
    cuts = np.searchsorted(net_of, np.arange(len(full.members) + 1))
    total = 0
    for i, net in enumerate(full.members):
        ref, rln, st = oracle_member(net, steps)
        b, s = int(full.begin[i]), int(full.size[i])
        for k in ("v", "m", "h", "ng", "s"):
            assert np.array_equal(got[k][b:b + s], st[k]), (i, k)
        mine = ras[cuts[i]:cuts[i + 1]].copy()
        mine[:, 1] -= b
        assert np.array_equal(mine, ref), i
        total += rln
    say("spikes", total)
    assert total == ln and total > len(full.members) * 1000


def _dst_solve64(b, c):
    th = 2 * np.cos(np.pi * np.arange(1, 65) / 65)
    tz, ty, tx = np.meshgrid(th, th, th, indexing="ij")
    lam = 1 + c * ((2 - tx) + (2 - ty) + (2 - tz))
    return scipy.fft.dstn(scipy.fft.dstn(b, type=1, norm="ortho") / lam, type=1, norm="ortho")


@pytest.mark.parametrize("graph", [False, True])
def test_ensemble_per_network_paracrine_solves(ctx, ens, graph):
    """Every network's own 64^3 grid solved inside the ensemble's steps
    every 100 steps (one solve of the box with Dirichlet faces between the
    grids): after each solve every network's block equals the exact DST-I
    solution of its own system, b = its previous field + 0.01 x the
    releases of its neurons in the window, and the box's FP64 relative
    residual is <= 1e-8."""
    every, windows = 100, 6
    sim = EnsembleSim(ens, ctx=ctx, record_cap=1 << 20, deposit=True)
    sim.set_paracrine(every, rtol=1e-8, max_inner=48, max_outer=3)
    nx, ny, nz = ens.grid
    say("ensemble paracrine: box", ens.grid, "lattice", ens.lattice, "graph", graph)
    og = oracle.grid(64, 64, 64, 7, ens.c)
    solved = 0
    for w in range(windows):
        prev = sim.a["conc"].numpy().reshape(nz, ny, nx)
        counts = sim.a["voxel_counts"].numpy().reshape(nz, ny, nx)
        sim.run(every, use_graph=graph, chunk=every)
        conc = sim.a["conc"].numpy().reshape(nz, ny, nx)
        if w == 0:
            assert not counts.any() and not conc.any()
            continue
        worst_rel, worst_err = 0.0, 0.0
        for i in range(len(SIZES)):
            ox, oy, oz = ens.block_of(i)
            sl = np.s_[oz:oz + 64, oy:oy + 64, ox:ox + 64]
            b = prev[sl] + 0.01 * counts[sl].astype(np.float64)
            x = conc[sl]
            if not b.any():  # no release in this network yet
                assert not x.any(), i
                continue
            solved += 1
            xs = _dst_solve64(b, ens.c)
            worst_err = max(worst_err, np.linalg.norm(x - xs) / np.linalg.norm(xs))
            r = b.ravel() - oracle.apply(og, np.ascontiguousarray(x).ravel())
            worst_rel = max(worst_rel, np.linalg.norm(r) / np.linalg.norm(b))
        say(f"  solve {w}: releases {int(counts.sum())}, worst network relres {worst_rel:.2e}, "
            f"worst error vs DST {worst_err:.2e}")
        assert worst_err <= 1e-7
        assert worst_rel <= 1e-7
    st = sim.paracrine_stats()
    say("  stats", st, "network solves checked", solved)
    assert solved >= len(SIZES)
    assert st["solves"] == windows and st["unconverged"] == 0 and st["max_relres"] <= 1e-8
