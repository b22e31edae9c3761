"""The SAN neuron step on the device against the oracle, bit for bit in fp32
(the reference's float/double mix) and fp64, with the level spike bitmap."""
import ctypes as C

import numpy as np
import pytest

import oracle
from neuronephys_clean_amd import _lib
from neuronephys_clean_amd.devmem import DeviceArray as DA

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_san_step_bitexact(ctx, T):
    print("san", T, flush=True)
    n = 70_001
    rng = np.random.default_rng(5)
    x = {"v": rng.uniform(-75, -40, n), "nk": rng.uniform(0, 0.3, n), "ca": rng.uniform(1, 60, n),
         "ye": rng.uniform(0, 0.5, n), "yi": rng.uniform(0, 0.5, n)}
    x["v"][:64] = np.linspace(-34.5, -33.5, 64)
    x["v"][7] = -34.0  # alpha_n's removable singularity
    ec, ic = rng.integers(0, 4, n).astype(float), rng.integers(0, 4, n).astype(float)
    inp = {"ein": rng.uniform(0, 1, n) * ec, "ecount": ec, "iin": rng.uniform(0, 1, n) * ic, "icount": ic}
    x = {k: np.ascontiguousarray(v, T) for k, v in x.items()}
    inp = {k: np.ascontiguousarray(v, T) for k, v in inp.items()}
    d = {k: DA.from_numpy(ctx, v) for k, v in {**x, **inp}.items()}
    d["eout"], d["iout"] = DA.zeros(ctx, n, T), DA.zeros(ctx, n, T)
    bits = DA.zeros(ctx, (n + 63) // 64, np.uint64)
    st = _lib.SanState(dtype=_lib.F32 if T == np.float32 else _lib.F64, n=n,
                       **{k: d[k].ptr for k in ("v", "nk", "ca", "ye", "yi", "ein", "ecount", "iin", "icount")},
                       eoutput=d["eout"].ptr, ioutput=d["iout"].ptr)
    p = _lib.params_san()
    fired = 0
    for step in range(30):
        ctx.call("nnx_san_step", C.byref(st), C.byref(p), 0.04, bits.ptr)
        out, spk = oracle.san_step(p.consts(), 0.04, x, inp)
        for k in x:
            assert np.array_equal(d[k].numpy(), x[k], equal_nan=True), (step, k)
        assert np.array_equal(d["iout"].numpy(), out) and np.array_equal(d["eout"].numpy(), out), step
        got = np.unpackbits(bits.numpy().view(np.uint8), bitorder="little")[:n]
        assert np.array_equal(got, spk), step
        fired += int(spk.sum())
    assert np.isfinite(x["v"]).all()
    print("fired", fired, flush=True)


def test_san_errors(ctx):
    st = _lib.SanState(dtype=_lib.F32, n=10)
    p = _lib.params_san()
    with pytest.raises(_lib.NNXError):
        ctx.call("nnx_san_step", C.byref(st), C.byref(p), 0.4, None)
