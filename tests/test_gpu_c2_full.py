"""C2 at full scale (1,048,576 RK4 HH neurons, ~2.1e8 distance-dependent
edges, release deposit on 256^3): the first 100 ms, 10,000 steps at dt
0.01 ms, of the fused fp32 step (one persistent launch per step, replayed
from hipGraphs) against the fp32 oracle on the same seeded inputs.  Every
spike (step, neuron) of the raster, the final state and the per-voxel
release counts are identical.  About 75 s of oracle time on 16 host cores.
The 10,000-step oracle runs go in calls of 1,000 steps (the last step's
spikes carried over), each reported as it ends.
"""
import numpy as np
import pytest

import oracle
from neuronephys_clean_amd import _lib, configs
from neuronephys_clean_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def oracle_rk4(capfd, *args, **kw):
    """The oracle's RK4 run in calls of 1,000 steps, each reported past
    pytest's output capture as it ends."""
    def done(k, n, t):
        with capfd.disabled():
            print(f"  oracle: {k} steps, {n} spikes, {t:.0f} s", flush=True)
    return oracle.run_rk4_chunked(*args, chunk=1000, done=done, **kw)


def test_c2_full_scale_100ms_bitexact(ctx, capfd):
    net = configs.c2()
    steps, cap = 10000, 16_000_000
    say("c2 full scale", net.n, net.nnz, steps)
    sim = Simulation(net, np.float32, couple="fused", record_cap=cap, deposit=True, ctx=ctx)
    sim.run(steps, use_graph=True, chunk=100)
    sim.synchronize()
    ras, ln = sim.raster()
    st = sim.state()
    vc = sim.a["voxel_counts"].numpy()
    del sim
    say("device spikes", ln, "; oracle run")
    T = np.float32
    V, m, h, ng = (np.asarray(a, T).copy() for a in (net.V0, net.m0, net.h0, net.n0))
    s = np.zeros(net.n, T)
    p = oracle.params(_lib.params_squid().as_dict())
    ref, rln, rvc, sec = oracle_rk4(capfd, p, net.dt, steps, net.off, net.tgt, net.w, V, m, h, ng, s,
                                    np.asarray(net.iext, T), cap=cap, voxel_of=net.ensure_voxel(),
                                    nvox=int(np.prod(net.grid)))
    say("oracle spikes", rln, f"in {sec:.1f} s")
    assert ln == rln and ln > 100_000
    assert np.array_equal(ras, ref)
    for k, a in dict(v=V, m=m, h=h, ng=ng, s=s).items():
        assert np.array_equal(st[k], a), k
    assert np.array_equal(vc, rvc)


def test_c2_full_scale_pull_and_event_1000_steps(ctx):
    """The same network's first 10 ms with the dense pull coupling (spike
    bitmap in LDS at exactly 2^20 neurons) and with the unfused event
    coupling: rasters and states identical to the oracle."""
    net = configs.c2()
    steps, cap = 1000, 4_000_000
    T = np.float32
    V, m, h, ng = (np.asarray(a, T).copy() for a in (net.V0, net.m0, net.h0, net.n0))
    s = np.zeros(net.n, T)
    p = oracle.params(_lib.params_squid().as_dict())
    ref, rln, rvc, sec = oracle.run_rk4(p, net.dt, steps, net.off, net.tgt, net.w, V, m, h, ng, s,
                                        np.asarray(net.iext, T), cap=cap, voxel_of=net.ensure_voxel(),
                                        nvox=int(np.prod(net.grid)))
    say("oracle spikes", rln, f"in {sec:.1f} s")
    for couple in ("pull", "event"):
        say("c2 full scale", couple, steps)
        sim = Simulation(net, T, couple=couple, record_cap=cap, deposit=True, ctx=ctx)
        sim.run(steps, use_graph=True, chunk=100)
        sim.synchronize()
        ras, ln = sim.raster()
        st = sim.state()
        assert ln == rln and ln > 100_000, couple
        assert np.array_equal(ras, ref), couple
        for k, a in dict(v=V, m=m, h=h, ng=ng, s=s).items():
            assert np.array_equal(st[k], a), (couple, k)
        assert np.array_equal(sim.a["voxel_counts"].numpy(), rvc), couple
        del sim


def test_c2_full_scale_fp64_equals_oracle_and_fp32_tolerance(ctx):
    """Precision split at full scale: the fp64 device run of the first 100
    steps equals the FP64 oracle bit for bit (raster and state), and the fp32
    run stays within the stated tolerance of it: |dV| <= 1e-4 mV after one
    step and <= 1e-3 mV after 100, gating variables within 1e-5."""
    net = configs.c2()
    steps, cap = 100, 4_000_000
    dev = {}
    for T in (np.float64, np.float32):
        sim = Simulation(net, T, couple="event", record_cap=cap, ctx=ctx)
        sim.run(1)
        sim.synchronize()
        v1 = sim.state()["v"].astype(np.float64)
        sim.run(steps - 1, use_graph=True, chunk=99)
        sim.synchronize()
        dev[T] = (v1, {k: a.astype(np.float64) for k, a in sim.state().items() if k in ("v", "m", "h", "ng")},
                  sim.raster())
        del sim
    T = np.float64
    V, m, h, ng = (np.asarray(a, T).copy() for a in (net.V0, net.m0, net.h0, net.n0))
    s = np.zeros(net.n, T)
    p = oracle.params(_lib.params_squid().as_dict())
    ref, rln, _, sec = oracle.run_rk4(p, net.dt, steps, net.off, net.tgt, net.w, V, m, h, ng, s,
                                      np.asarray(net.iext, T), cap=cap)
    ras, ln = dev[np.float64][2]
    assert ln == rln and np.array_equal(ras, ref)
    for k, a in dict(v=V, m=m, h=h, ng=ng).items():
        assert np.array_equal(dev[np.float64][1][k], a), k
    d1 = np.abs(dev[np.float32][0] - dev[np.float64][0]).max()
    d = {k: np.abs(dev[np.float32][1][k] - dev[np.float64][1][k]).max() for k in ("v", "m", "h", "ng")}
    say("fp32 vs fp64: max|dV| after 1 step", d1, "after 100:", d)
    assert d1 <= 1e-4 and d["v"] <= 1e-3
    assert max(d["m"], d["h"], d["ng"]) <= 1e-5


def test_c2_full_scale_fp64_100ms_equals_fp64_oracle_and_fp32_agreement(ctx, capfd):
    """The configuration's rule at full scale: the fp64 device run of C2's
    first 100 ms (10,000 steps, unfused event coupling, graphs of 100)
    produces exactly the FP64 oracle's raster (every step and neuron) and
    state.  Reported beside it: the fp32 run's (fused step) agreement with
    that FP64 raster - the fraction of (step, neuron) pairs they share, the
    fraction of FP64 spikes the fp32 run has within one step, the first
    divergent pair - and |dV| after 1, 100 and 10,000 steps.  About 2 min of
    oracle time on 16 host cores."""
    net = configs.c2()
    steps, cap = 10000, 16_000_000
    res = {}
    for T, couple in ((np.float64, "event"), (np.float32, "fused")):
        sim = Simulation(net, T, couple=couple, record_cap=cap, ctx=ctx, device_csc=True)
        vs = {}
        done = 0
        for k in (1, 100, steps):
            sim.run(k - done, use_graph=True, chunk=100 if (k - done) % 100 == 0 else k - done)
            done = k
            sim.synchronize()
            vs[k] = sim.state()["v"].astype(np.float64)
        res[T] = (vs, sim.state(), sim.raster())
        say(f"device {np.dtype(T).name} ({couple}): {res[T][2][1]} spikes")
        del sim
    T = np.float64
    V, m, h, ng = (np.asarray(a, T).copy() for a in (net.V0, net.m0, net.h0, net.n0))
    s = np.zeros(net.n, T)
    p = oracle.params(_lib.params_squid().as_dict())
    ref, rln, _, sec = oracle_rk4(capfd, p, net.dt, steps, net.off, net.tgt, net.w, V, m, h, ng, s,
                                  np.asarray(net.iext, T), cap=cap)
    say(f"FP64 oracle: {rln} spikes in {sec:.1f} s")
    ras64, ln64 = res[np.float64][2]
    assert ln64 == rln and rln > 1_000_000
    assert np.array_equal(ras64, ref)
    for k, a in dict(v=V, m=m, h=h, ng=ng, s=s).items():
        assert np.array_equal(res[np.float64][1][k], a), k
    # fp32 against FP64
    ras32, ln32 = res[np.float32][2]
    n = net.n
    k32 = ras32[:, 0].astype(np.int64) * n + ras32[:, 1]
    k64 = ref[:, 0].astype(np.int64) * n + ref[:, 1]
    match = np.intersect1d(k32, k64, assume_unique=True).size / max(k32.size, k64.size)
    near = float((np.isin(k64, k32) | np.isin(k64 - n, k32) | np.isin(k64 + n, k32)).mean())
    diff = np.setxor1d(k32, k64, assume_unique=True)
    first = (int(diff.min()) // n, int(diff.min()) % n) if diff.size else None
    dv = {k: float(np.abs(res[np.float32][0][k] - res[np.float64][0][k]).max()) for k in (1, 100, steps)}
    say(f"fp32 vs FP64: spikes {ln32} / {ln64}, shared pairs {match:.6f}, within one step {near:.6f}, "
        f"first divergence (step, neuron) {first}, max|dV| {dv}")
    assert dv[1] <= 1e-4 and dv[100] <= 1e-3
    assert match > 0.5 and near > 0.9
