"""C2 at full scale (1,048,576 RK4 HH neurons, ~2.1e8 distance-dependent
edges, release deposit on 256^3): the first 100 ms, 10,000 steps at dt
0.01 ms, of the fused fp32 step (one persistent launch per step, replayed
from hipGraphs) against the fp32 oracle on the same seeded inputs.  Every
spike (step, neuron) of the raster, the final state and the per-voxel
release counts are identical.  About 75 s of oracle time on 16 host cores.
"""
import numpy as np
import pytest

import oracle
from neuronephys_clean_amd import _lib, configs
from neuronephys_clean_amd.sim import Simulation

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def test_c2_full_scale_100ms_bitexact(ctx):
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
    ref, rln, rvc, sec = oracle.run_rk4(p, net.dt, steps, net.off, net.tgt, net.w, V, m, h, ng, s,
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
