"""Electrode readout on the device against the oracle: region means and
firing counts (bit-exact: the sum order is the reference's), the network
mean and the synchrony sums (bit-exact: the oracle restates the device's
fixed order), and the print-step readout inside nnx_run_steps (bit-exact
against measurements taken between single steps)."""
import ctypes as C

import numpy as np
import pytest

import oracle
from neuronephys_clean_amd import _lib, configs, host, readout
from neuronephys_clean_amd.devmem import DeviceArray as DA
from neuronephys_clean_amd.sim import SOA, ReferenceSimulation, Simulation

pytestmark = pytest.mark.gpu

DT = {np.float32: _lib.F32, np.float64: _lib.F64}


def say(*a):
    print(*a, flush=True)


def flags_of(words, n):
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:n].astype(np.int32)


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("shape", ["small", "large"])
def test_electrode_measure_bitexact(ctx, T, shape):
    """Small regions take the thread-per-region kernel, large ones the
    wavefront kernel; both give the reference's sequential sums."""
    say("electrode_measure", T, shape)
    rng = np.random.default_rng(3)
    n = 100_003
    V = rng.normal(-60, 20, n).astype(T)
    V[:8] = [1e8, 1.0, -1e8, 1.0, 2.0 ** 53, 1.0, -(2.0 ** 53), 1.0]
    words = rng.integers(0, 2 ** 63, (n + 63) // 64, dtype=np.int64).view(np.uint64)
    if shape == "small":
        sizes = rng.integers(0, 9, 20_000)
    else:
        sizes = rng.integers(0, 3000, 300)
        sizes[:3] = [64, 65, 1]
    sizes[5] = 0  # an empty region
    sizes[1] = sizes[2] = 4
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    idx = rng.integers(0, n, off[-1]).astype(np.int32)
    idx[off[1]:off[3]] = np.arange(8)  # the cancelling sums: order-dependent means
    idx[-1] = n + 5  # out of range: NaN mean, not firing
    ro = readout.ElectrodeNet(ctx, idx, off)
    dv, db = DA.from_numpy(ctx, V), DA.from_numpy(ctx, words)
    avg, cnt = ro.measure(dv, db)
    ref, rcnt = oracle.electrode_measure(off, idx, V, flags_of(words, n))
    assert np.array_equal(avg, ref, equal_nan=True)
    assert np.array_equal(cnt, rcnt)
    assert np.isnan(avg[5]) and np.isnan(avg[-1])
    avg2, none = ro.measure(dv)  # without a bitmap
    assert none is None and np.array_equal(avg2, ref, equal_nan=True)


def test_electrode_measure_clamps_offsets(ctx):
    V = np.arange(10, dtype=np.float32)
    idx = np.arange(10, dtype=np.int32)
    off = np.array([-5, 3, 2, 50], np.int32)  # [0,3), empty (end < start), [2,10)
    ro = readout.ElectrodeNet(ctx, idx, off)
    avg, _ = ro.measure(DA.from_numpy(ctx, V))
    assert avg[0] == 1.0 and np.isnan(avg[1]) and avg[2] == np.float32(5.5)


def test_electrode_measure_errors(ctx):
    r = _lib.Regions(n_regions=-1, nnz=0)
    v = DA.zeros(ctx, 4, np.float32)
    with pytest.raises(_lib.NNXError):
        ctx.call("nnx_electrode_measure", C.byref(r), _lib.F32, 4, v.ptr, None, v.ptr, None)
    r = _lib.Regions(n_regions=2, nnz=4, offsets=None, idx=None)
    with pytest.raises(_lib.NNXError):
        ctx.call("nnx_electrode_measure", C.byref(r), _lib.F32, 4, v.ptr, None, v.ptr, None)


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 1000, 262_145, 1 << 20])
def test_network_mean_bitexact(ctx, T, n):
    say("network_mean", T, n)
    V = np.random.default_rng(n).normal(-60, 20, n).astype(T)
    got = readout.network_mean(ctx, DA.from_numpy(ctx, V))
    assert got == oracle.network_mean(V)


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_synchrony_bitexact(ctx, T):
    say("synchrony", T)
    rng = np.random.default_rng(7)
    n, n_idx = 50_000, 5000
    idx = rng.choice(n, n_idx, replace=False).astype(np.int32)
    counts = np.array([1, 2, 100, 1025, 5000, 0, 5001], np.int32)  # the last two are invalid
    sn = readout.SynchronyNet(ctx, idx, counts)
    sums = [np.zeros(n_idx), np.zeros(n_idx), np.zeros(len(counts)), np.zeros(len(counts))]
    base = rng.normal(-60, 10, n)
    for k in range(6):
        V = (base + 5 * np.sin(0.7 * k) + rng.normal(0, 1, n)).astype(T)
        sn.measure_variances(DA.from_numpy(ctx, V))
        oracle.sync_accumulate(V, idx, counts, sums)
    for d, h in zip(sn.sums, sums):
        assert np.array_equal(d.numpy()[:len(h)], h, equal_nan=True)
    chi = sn.measures()
    ref = oracle.sync_measures(n_idx, counts, sums, 6)
    assert np.array_equal(chi, ref, equal_nan=True)
    assert np.isnan(chi[5:]).all() and np.isfinite(chi[:5]).all()


def _manual_rows(make_sim, steps, every, off, idx, ctx, firing=True):
    """Reference rows: single steps, measuring before each print step."""
    sim = make_sim()
    ro = readout.ElectrodeNet(ctx, idx, off)
    rows_avg, rows_fire, rows_mean = [], [], []
    v = sim.a["v"] if "v" in sim.a else sim.sv["V"]
    for s in range(steps):
        if s % every == 0:
            bits = sim.a["bits0"] if s % 2 == 0 else sim.a["bits1"]
            avg, cnt = ro.measure(v, bits if firing else None)
            rows_avg.append(avg)
            rows_fire.append(cnt)
            rows_mean.append(readout.network_mean(ctx, v))
        sim.run(1)
    return np.array(rows_avg), np.array(rows_fire), np.array(rows_mean)


@pytest.mark.parametrize("couple,use_graph", [("event", False), ("event", True), ("pull", True),
                                              ("fused", True), ("fused", False)])
def test_step_readout_matches_single_steps(ctx, couple, use_graph):
    """The print-step readout in nnx_run_steps (graph replays, chunks that
    start off the print phase) equals measurements taken between steps."""
    say("step readout", couple, use_graph)
    net = configs.c1(seed=2)
    off, idx = host.spherical_regions(net.pos, 37, center_idx=np.arange(0, net.n, 97))
    steps, every, cap = 230, 7, 30  # 33 print steps: the last 3 rows are dropped
    ref_avg, ref_fire, ref_mean = _manual_rows(lambda: Simulation(net, np.float32, couple="event"),
                                               steps, every, off, idx, ctx)
    assert ref_fire.sum() > 0
    sim = Simulation(net, np.float32, couple=couple)
    sim.set_readout(off, idx, every, cap)
    done = 0
    for k in (5, 100, 100, 25):  # chunk starts at phases 0, 5, 0 (step 105 = 15*7), 2
        sim.run(k, use_graph=use_graph)
        done += k
    assert done == steps
    got = sim.readout()
    assert np.array_equal(got["avg"], ref_avg[:cap])
    assert np.array_equal(got["firing"], ref_fire[:cap])
    assert np.array_equal(got["mean"], ref_mean[:cap])


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_step_readout_reference_leapfrog(ctx, T):
    """The reference's own loop (src/main.cu:133-141): mean V of each region
    every print step, leapfrog network with COUNT coupling."""
    say("leapfrog readout", T)
    net = configs.c1(seed=5)
    in_deg = net.ensure_csc()[4]
    rng = np.random.default_rng(14)
    sv0 = {k: np.zeros(net.n) for k in SOA}
    sv0["V"] = rng.normal(-55, 25, net.n)
    sv0["LOCAL_COUNT"] = in_deg.astype(np.float64)
    sv0["APPCUR"] = rng.normal(0, 2, net.n)
    off, idx = host.spherical_regions(net.pos, 50, center_idx=(net.n // 16) * np.arange(16))
    steps, every = 60, 10

    def make():
        return ReferenceSimulation(net, sv0, T, couple="count")

    ref_avg, ref_fire, ref_mean = _manual_rows(make, steps, every, off, idx, ctx)
    sim = make()
    sim.set_readout(off, idx, every, steps // every)
    sim.run(steps, use_graph=True)
    got = sim.readout()
    assert np.array_equal(got["avg"], ref_avg) and np.array_equal(got["firing"], ref_fire)
    assert np.array_equal(got["mean"], ref_mean)
    # and against the oracle's restatement of the reference's loop
    sv = [np.asarray(sv0[k], T).copy() for k in SOA]
    q = _lib.params_reference()
    q.noise_std = 0.0
    oracle.run_leapfrog(oracle.params(q.as_dict()), 0.4, every * 3, net.off, net.tgt, 2, sv, cap=1)
    oavg, _ = oracle.electrode_measure(off, idx, sv[0])
    assert np.array_equal(got["avg"][3], oavg)
