"""Event-driven Hebbian learning on the device against the oracle's dense
restatement of the reference's rule (Hebbian::UpdateWeights,
src_wholebrain/hebbian.cu:3-69): weights bit-exact and the set of running
countdowns identical after every call; and inside nnx_run_steps."""
import ctypes as C

import numpy as np
import pytest

import oracle
from neuronephys_clean_amd import _lib, configs, hebbian, host
from neuronephys_clean_amd.devmem import DeviceArray as DA
from neuronephys_clean_amd.sim import Simulation, upload_graph

pytestmark = pytest.mark.gpu


def say(*a):
    print(*a, flush=True)


def pack(flags):
    n = len(flags)
    b = np.zeros(((n + 63) // 64) * 64, np.uint8)
    b[:n] = flags
    return np.packbits(b, bitorder="little").view(np.uint64)


def random_net(seed, n, deg):
    """Targets drawn with replacement (duplicate synapses) and self-loops allowed."""
    rng = np.random.default_rng(seed)
    d = rng.poisson(deg, n).astype(np.int32)
    off = np.concatenate([[0], np.cumsum(d)]).astype(np.int32)
    tgt = rng.integers(0, n, off[-1]).astype(np.int32)
    tgt[off[:-1][d > 0][:50]] = np.nonzero(d > 0)[0][:50]  # some self-loops
    w = rng.uniform(0, 2, off[-1]).astype(np.float32)
    net = configs.Network(n, off, tgt, w, *(np.zeros(n),) * 5, np.zeros((n, 3), np.float32))
    return net


@pytest.mark.parametrize("window", [1, 2, 5, 40])
def test_hebbian_update_matches_dense_rule(ctx, window):
    say("hebbian", window)
    net = random_net(window, 3000 + 17, 12)
    g, keep = upload_graph(ctx, net)
    rng = np.random.default_rng(100 + window)
    w0 = rng.uniform(0, 2, net.nnz).astype(np.float32)
    st = hebbian.HebbianState(ctx, g, window, 0.05, w0)
    timers = np.zeros(net.nnz, np.int32)
    w = w0.copy()
    bits = DA(ctx, (net.n + 63) // 64, np.uint64)
    ups = downs = 0
    for step in range(150):
        p = 0.4 if step % 37 == 5 else 0.03
        flags = (rng.random(net.n) < p).astype(np.uint8)
        bits.upload(pack(flags))
        st.update(step, bits)
        before = w.copy()
        oracle.hebbian_update(net.off, net.tgt, flags, window, 0.05, timers, w)
        ups += int((w > before).sum())
        downs += int((w < before).sum())
        got = st.weights()
        assert np.array_equal(got, w), (step, int(np.sum(got != w)))
        assert np.array_equal(st.armed() != 0, timers != 0), step
    say("ups", ups, "downs", downs)
    assert downs > 100 and (ups > 100 if window > 1 else ups == 0)  # window 1: only decreases


def test_hebbian_errors(ctx):
    net = random_net(1, 100, 3)
    g, keep = upload_graph(ctx, net)
    with pytest.raises(_lib.NNXError):
        hebbian.HebbianState(ctx, g, 0, 0.1, net.w)


@pytest.mark.parametrize("couple,use_graph", [("event", True), ("fused", True), ("pull", False)])
def test_hebbian_in_step_plan(ctx, couple, use_graph):
    """HebbianLearningExample's order (src_wholebrain/main.cu:158-168): after
    each step, the rule on the spikes that step's coupling read."""
    say("hebbian plan", couple, use_graph)
    net = configs.c1(seed=3)
    sim = Simulation(net, np.float32, couple=couple, record_cap=1 << 20)
    w0 = np.random.default_rng(5).uniform(0, 2, net.nnz).astype(np.float32)
    st = hebbian.attach(sim, 25, 0.02, w0)
    steps = 400
    for k in (30, 100, 100, 170):
        sim.run(k, use_graph=use_graph, chunk=50 if use_graph else None)
    ras, ln = sim.raster()
    assert ln > 0
    timers = np.zeros(net.nnz, np.int32)
    w = w0.copy()
    for s in range(steps):
        flags = np.zeros(net.n, np.uint8)
        flags[ras[ras[:, 0] == s - 1, 1]] = 1  # the spikes step s's coupling reads
        oracle.hebbian_update(net.off, net.tgt, flags, 25, 0.02, timers, w)
    got = st.weights()
    say("changed", int((w != w0).sum()), "of", net.nnz)
    assert (w != w0).sum() > 1000
    assert np.array_equal(got, w)
    assert np.array_equal(st.armed() != 0, timers != 0)
