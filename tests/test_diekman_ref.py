"""CPU pins of the oracle's Diekman step (Leapfrog_diekman, src/ephys.cu:
250-347) by the equations each update solves: every gate, the two Ca pools,
the Ca-activated K gate, the GABA synapse and V follow the trapezoid rule of
their ODE (the submembrane Ca through the root of its quadratic), and the
output is the reference's sigmoid.  Checked on the fp64 path (double
throughout) with residuals near round-off; the fp32 path (the reference's
float/double mix) against fp64."""
import numpy as np
import pytest

import oracle
from neuronephys_clean_amd import _lib

P = _lib.params_diekman()
DT = 0.4


def state(seed, n, dtype):
    rng = np.random.default_rng(seed)
    x = {"v": rng.uniform(-80, 0, n), "m": rng.uniform(0.05, 0.95, n), "h": rng.uniform(0.05, 0.95, n),
         "ng": rng.uniform(0.05, 0.95, n), "rl": rng.uniform(0, 1, n), "rnl": rng.uniform(0, 1, n),
         "fnl": rng.uniform(0, 1, n), "s": rng.uniform(0, 1, n), "p": rng.uniform(0, 1, n), "y": rng.uniform(0, 1, n),
         "cas": rng.uniform(0, 4e-4, n), "cac": rng.uniform(0, 4e-4, n)}
    cnt = rng.integers(0, 5, n).astype(np.float64)
    inp = {"iin": rng.uniform(0, 3, n) * cnt, "icount": cnt, "egaba": rng.uniform(-80, -30, n),
           "gkca": rng.normal(100, 0.5, n), "gkleak": rng.normal(0.12, 0.01, n)}
    return ({k: np.ascontiguousarray(v, dtype) for k, v in x.items()},
            {k: np.ascontiguousarray(v, dtype) for k, v in inp.items()})


def test_fp64_step_solves_the_trapezoid_equations():
    x0, inp = state(1, 4000, np.float64)
    x = {k: v.copy() for k, v in x0.items()}
    out, spk = oracle.diekman_step(P.consts(), DT, x, inp)
    V, h = x0["v"], DT / 2
    e = np.exp
    inf = {"m": 1 / (1 + e(-(V + 35.2) / 8.1)), "h": 1 / (1 + e((V + 62) / 2)),
           "ng": (1 + e((V - 14) / -17.0)) ** -0.25, "rl": 1 / (1 + e(-(V + 36) / 5.1)),
           "rnl": 1 / (1 + e(-(V + 21.6) / 6.7)), "fnl": 1 / (1 + e((V + 260) / 65)),
           "p": (1 + e(-(V + 25) / 7.4)) ** -1.5}
    tau = {"m": e(-(V + 286) / 160), "h": 0.51 + e(-(V + 26.6) / 7.1), "ng": e(-(V - 67) / 68),
           "rl": np.full_like(V, P.tau_rl), "rnl": np.full_like(V, P.tau_rnl), "fnl": e(-(V - 444) / 220),
           "p": np.full_like(V, P.tau_p)}
    for k in inf:  # x1 - x0 = h ((xinf - x0)/tau + (xinf - x1)/tau)
        r = x[k] - x0[k] - h * ((inf[k] - x0[k]) / tau[k] + (inf[k] - x[k]) / tau[k])
        assert np.abs(r).max() < 1e-12, k
    dV = V - P.eca

    def f_cas(ca, rl, rnl, fnl):
        return -P.kcas * (P.gcal * rl * P.k1 / (P.k2 + ca) + P.gcanl * rnl * fnl) * dV - ca / P.taucas + P.bcas
    f0 = f_cas(x0["cas"], x0["rl"], x0["rnl"], x0["fnl"])
    f1 = f_cas(x["cas"], x["rl"], x["rnl"], x["fnl"])
    r = x["cas"] - x0["cas"] - h * (f0 + f1)
    assert np.abs(r).max() < 1e-12 * max(1.0, np.abs(h * f0).max() * 1e6)
    assert (x["cas"] >= 0).all()

    def s_inf_tau(ca):
        c2 = 1e7 * ca * ca
        return c2 / (c2 + 5.6), 500 / (c2 + 5.6)
    si1, t1 = s_inf_tau(x0["cas"])
    si2, t2 = s_inf_tau(x["cas"])
    r = x["s"] * (1 + DT / (2 * t2)) - x0["s"] * (1 - DT / (2 * t1)) - h * (si1 / t1 + si2 / t2)
    assert np.abs(r).max() < 1e-13
    # cytosolic Ca: the reference's new-time current uses Cac and new_Rl for the non-L channel
    i_old = P.gcal * x0["rl"] * (P.k1 / (P.k2 + x0["cas"])) + P.gcanl * x0["rnl"] * x0["fnl"]
    i_new = P.gcal * x["rl"] * (P.k1 / (P.k2 + x0["cac"])) + P.gcanl * x["rl"] * x["fnl"]
    r = (x["cac"] * (1 + DT / (2 * P.taucac)) - x0["cac"] * (1 - DT / (2 * P.taucac)) - P.bcac * DT
         + DT * P.kcac / 2 * (i_new + i_old) * dV)
    assert np.abs(r).max() < 1e-17
    tavg = np.where(inp["icount"] == 0, 0.0, inp["iin"] / np.where(inp["icount"] == 0, 1, inp["icount"]))
    a = P.ar * tavg + P.ad
    r = x["y"] * (1 + h * a) - x0["y"] * (1 - h * a) - P.ar * tavg * DT
    assert np.abs(r).max() < 1e-13
    n4 = x["ng"] ** 4
    fl = P.k1 / (P.k2 + x["cas"])
    gna = P.gna * x["m"] ** 3 * x["h"]
    G = (gna + P.gk * n4 + P.gcal * x["rl"] * fl + P.gcanl * x["rnl"] * x["fnl"] + inp["gkca"] * x["s"] ** 2
         + inp["gkleak"] + P.gnaleak + P.gnap * x["p"] + x["y"])
    E = (P.ena * (gna + P.gnaleak + P.gnap * x["p"]) + P.ek * (P.gk * n4 + inp["gkca"] * x["s"] ** 2 + inp["gkleak"])
         + P.eca * (P.gcal * x["rl"] * fl + P.gcanl * x["rnl"] * x["fnl"]) + P.g_gaba * x["y"] * inp["egaba"])
    r = x["v"] * (1 + h * G) - V * (1 - h * G) - (P.appcur + E) * DT  # dV/dt = appcur + E - G V
    assert np.abs(r).max() < 1e-9 * np.abs(V * h * G).max()
    o = P.tmax / (1 + e(-(x["v"] - P.vt) / P.kp))
    assert np.abs(out - o).max() < 1e-15
    assert np.array_equal(spk, (out > 0.1).astype(np.uint8))


def test_fp32_follows_fp64():
    x32, in32 = state(2, 4000, np.float32)
    x64 = {k: v.astype(np.float64) for k, v in x32.items()}
    in64 = {k: v.astype(np.float64) for k, v in in32.items()}
    o32, _s = oracle.diekman_step(P.consts(), DT, x32, in32)
    o64, _s = oracle.diekman_step(P.consts(), DT, x64, in64)
    for k in x32:
        scale = np.abs(x64[k]).max() + 1e-30
        assert np.abs(x32[k] - x64[k]).max() <= 1e-4 * scale, k
    assert np.abs(o32 - o64).max() < 1e-4


def test_long_run_from_rest_fires_tonically():
    """5,000 steps (2 s) at the reference's dt 0.4 ms from V = -60 mV with
    every gate at its steady state: the state stays finite and bounded and
    every neuron fires repeatedly (appcur 3.5 drives tonic firing).  At this
    dt the trapezoid gates overshoot during a spike (m reaches ~1.13): that is
    the reference's integrator, not checked against [0, 1]."""
    n = 64
    V = np.full(n, -60.0)
    e = np.exp
    x = {"v": V.copy(), "m": 1 / (1 + e(-(V + 35.2) / 8.1)), "h": 1 / (1 + e((V + 62) / 2)),
         "ng": (1 + e((V - 14) / -17.0)) ** -0.25, "rl": 1 / (1 + e(-(V + 36) / 5.1)),
         "rnl": 1 / (1 + e(-(V + 21.6) / 6.7)), "fnl": 1 / (1 + e((V + 260) / 65)), "s": np.zeros(n),
         "p": (1 + e(-(V + 25) / 7.4)) ** -1.5, "y": np.zeros(n), "cas": np.full(n, 1e-4), "cac": np.full(n, 1e-4)}
    x = {k: np.ascontiguousarray(v, np.float64) for k, v in x.items()}
    rng = np.random.default_rng(4)
    inp = {"iin": np.zeros(n), "icount": np.zeros(n), "egaba": np.full(n, -60.0),
           "gkca": rng.normal(100, 0.5, n), "gkleak": rng.normal(0.12, 0.01, n)}
    prev = np.zeros(n, bool)
    onsets = np.zeros(n, np.int64)
    vmin, vmax = 0.0, -200.0
    for _ in range(5000):
        _o, spk = oracle.diekman_step(P.consts(), DT, x, inp)
        s = spk.astype(bool)
        onsets += s & ~prev
        prev = s
        vmin, vmax = min(vmin, x["v"].min()), max(vmax, x["v"].max())
    assert all(np.isfinite(v).all() for v in x.values())
    assert -100 < vmin and vmax < 100, (vmin, vmax)
    assert onsets.min() >= 10  # at least 5 Hz in every neuron
