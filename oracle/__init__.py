"""CPU oracle (test infrastructure only).

ctypes front end of liboracle.so, the from-scratch C++ restatement of the
reference's per-timestep algorithm (oracle/oracle.hpp).  Only tests/, the
smoke check and bench.py's CPU-baseline leg import this package; the product
never does.
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(_HERE, "liboracle.so")
_lib = None

PARAM_FIELDS = ("C", "gna", "gk", "gl", "ena", "ek", "el", "v_shift", "tau_syn", "e_syn",
                "tmax", "vt", "kp", "ar", "ad", "g_local", "e_local", "g_i", "e_i", "g_e",
                "e_e", "noise_mean", "noise_std", "noise_seed", "v_thresh", "out_thresh")


class Params(C.Structure):
    _fields_ = [(f, C.c_uint64 if f == "noise_seed" else C.c_double) for f in PARAM_FIELDS]


class Grid(C.Structure):
    _fields_ = [("nx", C.c_int32), ("ny", C.c_int32), ("nz", C.c_int32), ("stencil", C.c_int32),
                ("c", C.c_double), ("decay_dt", C.c_double), ("kx", C.c_void_p),
                ("ky", C.c_void_p), ("kz", C.c_void_p)]


def build():
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        _lib = C.CDLL(LIB)
        _lib.orc_run_rk4_f32.restype = C.c_int64
        _lib.orc_run_rk4_f64.restype = C.c_int64
        _lib.orc_run_leapfrog_f32.restype = C.c_int64
        _lib.orc_run_leapfrog_f64.restype = C.c_int64
        for f in ("orc_cg_f32", "orc_cg_f64", "orc_refine", "orc_update_density_reference"):
            getattr(_lib, f).restype = C.c_int
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def set_threads(n: int) -> int:
    """Sets the OpenMP thread count of the oracle's loops (n <= 0: all
    processors); returns the count now in effect."""
    lib().orc_set_threads(int(n))
    return int(lib().orc_threads())


def params(d) -> Params:
    """Params from a dict or any object with the PARAM_FIELDS attributes."""
    p = Params()
    for f in PARAM_FIELDS:
        setattr(p, f, d[f] if isinstance(d, dict) else getattr(d, f))
    return p


def grid(nx, ny, nz, stencil=7, c=1.0, decay_dt=0.0, kx=None, ky=None, kz=None) -> Grid:
    ks = [None if a is None else np.ascontiguousarray(a, np.float32) for a in (kx, ky, kz)]
    g = Grid(nx, ny, nz, stencil, c, decay_dt, *(_p(a) for a in ks))
    g._keep = ks  # the structure holds raw pointers: keep the arrays alive
    return g


def _sfx(dtype):
    return "f32" if np.dtype(dtype) == np.float32 else "f64"


def exp(x):
    x = np.ascontiguousarray(x)
    y = np.empty_like(x)
    getattr(lib(), "orc_exp_" + _sfx(x.dtype))(_p(x), _p(y), C.c_int64(x.size))
    return y


def rates(u):
    u = np.ascontiguousarray(u)
    out = np.empty((u.size, 6), u.dtype)
    getattr(lib(), "orc_rates_" + _sfx(u.dtype))(_p(u), _p(out), C.c_int64(u.size))
    return out  # am, bm, ah, bh, an, bn


def hh_rk4(p, dt, V, m, h, ng, s, iext, gin=None):
    """One RK4 step in place; returns the 0/1 crossing flags."""
    n = V.size
    spk = np.zeros(n, np.uint8)
    getattr(lib(), "orc_hh_rk4_" + _sfx(V.dtype))(C.c_int64(n), C.byref(p), C.c_double(dt),
                                                   _p(V), _p(m), _p(h), _p(ng), _p(s), _p(iext),
                                                   _p(gin), _p(spk))
    return spk


def hh_leapfrog(p, dt, sv, step=0):
    """One reference leapfrog step over the 17 SoA arrays (list, in place);
    `step` keys the white noise when p.noise_std != 0."""
    n = sv[0].size
    spk = np.zeros(n, np.uint8)
    arr = (C.c_void_p * 17)(*[a.ctypes.data for a in sv])
    getattr(lib(), "orc_hh_leapfrog_" + _sfx(sv[0].dtype))(C.c_int64(n), C.byref(p),
                                                            C.c_double(dt), arr, _p(spk),
                                                            C.c_int64(step))
    return spk


def philox4x64(counter, key):
    """The four 64-bit words of Philox4x64-10 at counter (4 words), key (2)."""
    c = np.ascontiguousarray(counter, np.uint64)
    k = np.ascontiguousarray(key, np.uint64)
    out = np.zeros(4, np.uint64)
    lib().orc_philox4x64_10(_p(c), _p(k), _p(out))
    return out


def noise_log(u):
    u = np.ascontiguousarray(u, np.float64)
    out = np.empty_like(u)
    lib().orc_noise_log(_p(u), _p(out), C.c_int64(u.size))
    return out


def noise_cos2pi(u):
    u = np.ascontiguousarray(u, np.float64)
    out = np.empty_like(u)
    lib().orc_noise_cos2pi(_p(u), _p(out), C.c_int64(u.size))
    return out


def noise_normal(seed, step, n, first=0):
    """Standard normals of neurons first .. first+n-1 at `step` (the white
    noise of the leapfrog step)."""
    out = np.empty(n, np.float64)
    lib().orc_noise_normal(C.c_uint64(seed), C.c_uint64(step), C.c_int64(first), C.c_int64(n), _p(out))
    return out


def couple(off, tgt, w, fire, mode, out=None, dtype=np.float32):
    n = len(off) - 1
    g = np.empty(n, dtype)
    out = None if out is None else np.ascontiguousarray(out, dtype)
    getattr(lib(), "orc_couple_" + _sfx(dtype))(C.c_int64(n), _p(off), _p(tgt), _p(w),
                                                _p(np.ascontiguousarray(fire, np.uint8)), mode,
                                                _p(out), _p(g))
    return g


def apply(g: Grid, x):
    x = np.ascontiguousarray(x)
    y = np.empty_like(x)
    getattr(lib(), "orc_apply_" + _sfx(x.dtype))(C.byref(g), _p(x), _p(y))
    return y


def cg(g: Grid, b, rtol, maxit=10000, jacobi=True, x0=None):
    b = np.ascontiguousarray(b)
    x = np.zeros_like(b) if x0 is None else np.array(x0, b.dtype)
    rel = C.c_double()
    it = getattr(lib(), "orc_cg_" + _sfx(b.dtype))(C.byref(g), _p(b), _p(x), C.c_double(rtol),
                                                   maxit, int(jacobi), int(x0 is not None),
                                                   C.byref(rel))
    return x, it, rel.value


def refine(g: Grid, b, rtol, inner_rtol, maxit=10000, max_outer=50):
    b = np.ascontiguousarray(b, np.float64)
    x = np.zeros_like(b)
    rel, inner = C.c_double(), C.c_int()
    outer = lib().orc_refine(C.byref(g), _p(b), _p(x), C.c_double(rtol), C.c_double(inner_rtol),
                             maxit, max_outer, C.byref(rel), C.byref(inner))
    return x, outer, inner.value, rel.value


def update_density_reference(data, h, D, k, dt, eb=1e-6, max_iter=20):
    data = np.array(data, np.float32, order="C")
    nz, ny, nx = data.shape
    rr = C.c_double()
    it = lib().orc_update_density_reference(nx, ny, nz, C.c_double(h), C.c_double(D),
                                            C.c_double(k), C.c_double(dt), C.c_double(eb),
                                            max_iter, _p(data), C.byref(rr))
    return data, it, rr.value


def run_rk4(p, dt, steps, off, tgt, w, V, m, h, ng, s, iext, cap=1 << 20, voxel_of=None,
            nvox=0, fire=None):
    """Full RK4 run (state updated in place); returns raster pairs, total
    count, voxel counts and the wall seconds of the step loop.  fire
    (optional uint8 array of n): in, the spikes of the step before the first;
    out, those of the last step (a run continued in the next call)."""
    T = V.dtype
    raster = np.zeros(2 * cap, np.int32)
    vc = np.zeros(max(nvox, 1), np.int32)
    sec = C.c_double()
    args = [C.c_int64(V.size), C.byref(p), C.c_double(dt), C.c_int64(steps), _p(off), _p(tgt), _p(w),
            _p(V), _p(m), _p(h), _p(ng), _p(s), _p(iext), _p(raster), C.c_int64(cap), _p(voxel_of),
            _p(vc), C.byref(sec)]
    if fire is None:
        ln = getattr(lib(), "orc_run_rk4_" + _sfx(T))(*args)
    else:
        assert fire.dtype == np.uint8 and fire.size == V.size and fire.flags.c_contiguous
        ln = getattr(lib(), "orc_run_rk4_cont_" + _sfx(T))(*args, _p(fire))
    return raster.reshape(-1, 2)[:min(ln, cap)], ln, vc, sec.value


def run_rk4_chunked(p, dt, steps, off, tgt, w, V, m, h, ng, s, iext, cap=1 << 20, voxel_of=None,
                    nvox=0, chunk=1000, done=None):
    """run_rk4 over `steps` in calls of `chunk` steps (the spikes of each
    call's last step carried into the next): one call's raster, count, voxel
    counts and seconds; done(steps_so_far, spikes_so_far, seconds) after
    each call."""
    fire = np.zeros(V.size, np.uint8)
    parts, total, sec = [], 0, 0.0
    vc = np.zeros(max(nvox, 1), np.int32)
    for k0 in range(0, steps, chunk):
        k = min(chunk, steps - k0)
        r, ln, c, t = run_rk4(p, dt, k, off, tgt, w, V, m, h, ng, s, iext, cap=max(cap - total, 1),
                              voxel_of=voxel_of, nvox=nvox, fire=fire)
        r = r.copy()
        r[:, 0] += k0
        parts.append(r)
        vc += c
        total += ln
        sec += t
        if done:
            done(k0 + k, total, sec)
    ras = np.concatenate(parts)[:min(total, cap)] if parts else np.zeros((0, 2), np.int32)
    return ras, total, vc, sec


def run_leapfrog(p, dt, steps, off, tgt, mode, sv, cap=1 << 20, first_step=0):
    raster = np.zeros(2 * cap, np.int32)
    arr = (C.c_void_p * 17)(*[a.ctypes.data for a in sv])
    ln = getattr(lib(), "orc_run_leapfrog_" + _sfx(sv[0].dtype))(
        C.c_int64(sv[0].size), C.byref(p), C.c_double(dt), C.c_int64(steps), _p(off), _p(tgt),
        mode, arr, _p(raster), C.c_int64(cap), C.c_int64(first_step))
    return raster.reshape(-1, 2)[:min(ln, cap)], ln


def electrode_measure(off, idx, V, fire=None):
    """Per-region (mean V, firing count) as ElectrodeNet::DoElectrodeMeasurements
    (src_wholebrain/electrode_net.cu:3-30); fire: 0/1 per neuron or None."""
    off = np.ascontiguousarray(off, np.int32)
    idx = np.ascontiguousarray(idx, np.int32)
    V = np.ascontiguousarray(V)
    nreg = len(off) - 1
    avg = np.empty(nreg, V.dtype)
    cnt = np.zeros(nreg, np.int32)
    f = None if fire is None else np.ascontiguousarray(fire, np.int32)
    getattr(lib(), "orc_electrode_measure_" + _sfx(V.dtype))(
        C.c_int64(nreg), C.c_int64(len(idx)), _p(off), _p(idx), C.c_int64(len(V)), _p(V), _p(f), _p(avg),
        _p(cnt))
    return avg, (cnt if fire is not None else None)


def network_mean(V):
    """Mean of V in the device's fixed FP64 summation order, rounded to V's dtype."""
    V = np.ascontiguousarray(V)
    fn = getattr(lib(), "orc_network_mean_" + _sfx(V.dtype))
    fn.restype = C.c_float if V.dtype == np.float32 else C.c_double
    return V.dtype.type(fn(C.c_int64(len(V)), _p(V)))


def sync_accumulate(V, idx, counts, sums):
    """One SynchronyNet::MeasureVariances step; sums = (vi, vi2, vavg, vavg2) FP64 arrays, updated."""
    V = np.ascontiguousarray(V)
    idx = np.ascontiguousarray(idx, np.int32)
    counts = np.ascontiguousarray(counts, np.int32)
    getattr(lib(), "orc_sync_accumulate_" + _sfx(V.dtype))(
        C.c_int64(len(V)), _p(V), C.c_int64(len(idx)), _p(idx), C.c_int(len(counts)), _p(counts),
        *(_p(a) for a in sums))


def sync_measures(n_idx, counts, sums, steps):
    """chi^2 per sub-population (SynchronyNet::GetSynchronyMeasures)."""
    counts = np.ascontiguousarray(counts, np.int32)
    chi = np.empty(len(counts), np.float64)
    lib().orc_sync_measures(C.c_int64(n_idx), C.c_int(len(counts)), _p(counts), *(_p(a) for a in sums),
                            C.c_int64(steps), _p(chi))
    return chi


def hebbian_update(off, tgt, fire, window, heb_inc, timers, w):
    """One Hebbian::UpdateWeights call; timers (int32) and w (float32) updated in place."""
    off = np.ascontiguousarray(off, np.int32)
    tgt = np.ascontiguousarray(tgt, np.int32)
    f = np.ascontiguousarray(fire, np.uint8)
    assert timers.dtype == np.int32 and w.dtype == np.float32
    lib().orc_hebbian_update(C.c_int64(len(off) - 1), _p(off), _p(tgt), _p(f), C.c_int32(window),
                             C.c_float(heb_inc), _p(timers), _p(w))


DIEKMAN_STATE = ("v", "m", "h", "ng", "rl", "rnl", "fnl", "s", "p", "y", "cas", "cac")
DIEKMAN_IN = ("iin", "icount", "egaba", "gkca", "gkleak")


def diekman_step(consts, dt, x, inputs, thr=0.1):
    """One Leapfrog_diekman step; x: dict of the 12 state arrays (updated in
    place), inputs: dict of the 5 input arrays; consts: the 27 constants in
    nnx_diekman_params order.  Returns (IOUTPUT, spike flags)."""
    a = [x[k] for k in DIEKMAN_STATE]
    dt_ = a[0].dtype
    assert all(v.dtype == dt_ and v.flags.c_contiguous for v in a)
    ins = [np.ascontiguousarray(inputs[k], dt_) for k in DIEKMAN_IN]
    n = len(a[0])
    out = np.empty(n, dt_)
    sp = np.empty(n, np.uint8)
    c = np.ascontiguousarray(consts, np.float64)
    xs = (C.c_void_p * 12)(*[v.ctypes.data for v in a])
    ii = (C.c_void_p * 5)(*[v.ctypes.data for v in ins])
    getattr(lib(), "orc_diekman_" + _sfx(dt_))(C.c_int64(n), _p(c), C.c_double(dt), C.c_double(thr), xs, ii,
                                               _p(out), _p(sp))
    return out, sp


SAN_STATE = ("v", "nk", "ca", "ye", "yi")
SAN_IN = ("ein", "ecount", "iin", "icount")


def san_step(consts, dt, x, inputs, thr=0.1):
    """One Euler_an_ie_model step (noise = its mean); x: dict of the 5 state
    arrays (updated in place), inputs: dict of the 4 input arrays; consts:
    the 24 constants in nnx_san_params order.  Returns (output, spike flags)."""
    a = [x[k] for k in SAN_STATE]
    dt_ = a[0].dtype
    assert all(v.dtype == dt_ and v.flags.c_contiguous for v in a)
    ins = [np.ascontiguousarray(inputs[k], dt_) for k in SAN_IN]
    n = len(a[0])
    out = np.empty(n, dt_)
    sp = np.empty(n, np.uint8)
    c = np.ascontiguousarray(consts, np.float64)
    xs = (C.c_void_p * 5)(*[v.ctypes.data for v in a])
    ii = (C.c_void_p * 4)(*[v.ctypes.data for v in ins])
    getattr(lib(), "orc_san_" + _sfx(dt_))(C.c_int64(n), _p(c), C.c_double(dt), C.c_double(thr), xs, ii,
                                           _p(out), _p(sp))
    return out, sp
