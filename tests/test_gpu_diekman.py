"""The Diekman SCN neuron step on the device against the oracle, bit for bit
in fp32 (the reference's float/double mix) and fp64, with the level spike
bitmap on the output."""
import ctypes as C

import numpy as np
import pytest

import oracle
from neuronephys_clean_amd import _lib
from neuronephys_clean_amd.devmem import DeviceArray as DA

pytestmark = pytest.mark.gpu


def flags_of(words, n):
    return np.unpackbits(words.view(np.uint8), bitorder="little")[:n]


def start(seed, n, T):
    rng = np.random.default_rng(seed)
    V = rng.uniform(-70, -40, n)
    e = np.exp
    x = {"v": V, "m": 1 / (1 + e(-(V + 35.2) / 8.1)), "h": 1 / (1 + e((V + 62) / 2)),
         "ng": (1 + e((V - 14) / -17.0)) ** -0.25, "rl": 1 / (1 + e(-(V + 36) / 5.1)),
         "rnl": 1 / (1 + e(-(V + 21.6) / 6.7)), "fnl": 1 / (1 + e((V + 260) / 65)), "s": rng.uniform(0, 0.1, n),
         "p": (1 + e(-(V + 25) / 7.4)) ** -1.5, "y": rng.uniform(0, 0.2, n), "cas": rng.uniform(0, 3e-4, n),
         "cac": rng.uniform(0, 3e-4, n)}
    cnt = rng.integers(0, 4, n).astype(np.float64)
    inp = {"iin": rng.uniform(0, 1, n) * cnt, "icount": cnt, "egaba": rng.uniform(-80, -30, n),
           "gkca": rng.normal(100, 0.5, n), "gkleak": rng.normal(0.12, 0.01, n)}
    return ({k: np.ascontiguousarray(v, T) for k, v in x.items()}, {k: np.ascontiguousarray(v, T) for k, v in inp.items()})


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_diekman_step_bitexact(ctx, T):
    print("diekman", T, flush=True)
    n = 100_003
    x, inp = start(11, n, T)
    d = {k: DA.from_numpy(ctx, v) for k, v in {**x, **inp}.items()}
    d["out"] = DA.zeros(ctx, n, T)
    bits = DA.zeros(ctx, (n + 63) // 64, np.uint64)
    st = _lib.DiekmanState(dtype=_lib.F32 if T == np.float32 else _lib.F64, n=n,
                           **{k: d[k].ptr for k in _lib.DIEKMAN_STATE}, iin=d["iin"].ptr, icount=d["icount"].ptr,
                           ioutput=d["out"].ptr, egaba=d["egaba"].ptr, gkca=d["gkca"].ptr, gkleak=d["gkleak"].ptr)
    p = _lib.params_diekman()
    fired = 0
    for step in range(25):
        ctx.call("nnx_diekman_step", C.byref(st), C.byref(p), 0.4, bits.ptr)
        out, spk = oracle.diekman_step(p.consts(), 0.4, x, inp)
        for k in _lib.DIEKMAN_STATE:
            assert np.array_equal(d[k].numpy(), x[k], equal_nan=True), (step, k)
        assert np.array_equal(d["out"].numpy(), out, equal_nan=True), step
        assert np.array_equal(flags_of(bits.numpy(), n), spk), step
        fired += int(spk.sum())
    assert fired > 1000 and np.isfinite(x["v"]).all()


def test_diekman_errors(ctx):
    st = _lib.DiekmanState(dtype=_lib.F32, n=10)
    p = _lib.params_diekman()
    with pytest.raises(_lib.NNXError):
        ctx.call("nnx_diekman_step", C.byref(st), C.byref(p), 0.4, None)
