"""Benchmark: neuron-steps/s of the 1M-neuron HH + coupling timestep, and
paracrine CG milliseconds to a 1e-8 relative residual, on one MI355X.

Workload (the single-GPU configuration of the headline metric): 1,048,576
squid HH neurons, RK4 at dt = 0.01 ms in fp32, ~210M-edge CSR with
distance-dependent targets (exp(-d/10) in a 128^3 box), log-normal weights,
event-driven deterministic coupling, release deposit onto a 256^3 grid.
A step = coupling on the previous step's spikes + release deposit + HH RK4
update + ascending spike list + raster append, enqueued as a captured
hipGraph (chunks of --chunk steps).  Inputs are resident in HBM before the
timed region starts.

Prints ONE JSON line (rank 0).  Extra objects:
  roofline     - the dominant kernel's algorithmic bytes per launch over its
                 average launch time (HIP events on the library's stream),
                 against 8 TB/s
  cpu_baseline - the CPU oracle (fp32 restatement of the same step) on the
                 host cores for a bounded sample of the same workload
  cg           - plain Jacobi-CG to 1e-8 on the 256^3 paracrine grid (and the
                 FP64-refined solve), milliseconds and iterations
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import re
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK = 8.0e12  # B/s, MI355X HBM3E
METRIC = "neuron-steps/s at 1M HH; CG ms to 1e-8 resid (learned-PC vs plain)"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def timed_events(ctx, fn, reps):
    """Average ms of fn() over reps launches, HIP events on ctx's stream."""
    e0, e1 = C.c_void_p(), C.c_void_p()
    ctx.call("nnx_event_create", C.byref(e0))
    ctx.call("nnx_event_create", C.byref(e1))
    fn()
    ctx.call("nnx_event_record", e0)
    for _ in range(reps):
        fn()
    ctx.call("nnx_event_record", e1)
    ms = C.c_float()
    ctx.call("nnx_event_elapsed", e0, e1, C.byref(ms))
    ctx.call("nnx_event_destroy", e0)
    ctx.call("nnx_event_destroy", e1)
    return ms.value / reps


def cpu_baseline(net, steps):
    import oracle
    from neuronephys_clean_amd import _lib
    T = np.float32
    V, m, h, ng = (np.asarray(a, T).copy() for a in (net.V0, net.m0, net.h0, net.n0))
    s = np.zeros(net.n, T)
    p = oracle.params(_lib.params_squid().as_dict())
    _, ln, _, sec = oracle.run_rk4(p, net.dt, steps, net.off, net.tgt, net.w, V, m, h, ng, s,
                                   np.asarray(net.iext, T), cap=1,
                                   voxel_of=net.ensure_voxel(), nvox=int(np.prod(net.grid)))
    threads = int(os.environ.get("OMP_NUM_THREADS", os.cpu_count() or 1))
    return {"value": net.n * steps / sec, "unit": "neuron-steps/s", "cores": threads,
            "kind": "port",
            "sample": f"first {steps} steps of the same network and initial state, fp32 oracle "
                      f"(RK4 + push coupling in ascending source order + release counting), "
                      f"OpenMP over neurons, {sec:.2f} s"}


def cg_bench(ctx, grid_n, b_host, rtol=1e-8, c=1.0):
    """The configuration's paracrine solve on its own: 256^3, dt D / h^2 = 1,
    b = the fixed release field; Jacobi-CG with FP64 refinement to a true
    (FP64) relative residual of 1e-8 - the headline CG milliseconds - median
    of 3; beside it the fp32 CG to a recursive 1e-8 and its true residual."""
    from neuronephys_clean_amd import _lib
    from neuronephys_clean_amd.devmem import DeviceArray as DA
    nvox = grid_n ** 3
    g = _lib.Grid(grid_n, grid_n, grid_n, 7, c, 0.0, None, None, None)
    o = _lib.CgOpts(rtol=rtol, maxit=4000, precond=_lib.PRECOND_JACOBI, check_every=8)
    db64 = DA.from_numpy(ctx, np.asarray(b_host, np.float64))
    dx64 = DA.zeros(ctx, nvox, np.float64)
    runs = []
    for _ in range(4):
        r2 = _lib.CgResult()
        ctx.call("nnx_cg_solve_refine", C.byref(g), db64.ptr, dx64.ptr, C.byref(o), 1e-6, 20, C.byref(r2))
        runs.append((r2.ms, r2.iters, r2.outer, r2.relres, r2.converged))
    runs = sorted(runs[1:])
    ms, it, outer, rel, conv = runs[1]
    db = DA.from_numpy(ctx, np.asarray(b_host, np.float32))
    dx = DA.zeros(ctx, nvox, np.float32)
    r = _lib.CgResult()
    st = ctx._l.nnx_cg_solve(ctx.h, C.byref(g), db.ptr, dx.ptr, C.byref(o), C.byref(r))
    return {"grid": f"{grid_n}^3", "c": c, "rtol": rtol, "precond": "jacobi",
            "rhs": f"0.01 x the per-voxel spike counts of the first {FIXED_RELEASE_STEPS} C2 steps",
            "plain_ms": ms, "plain_iters": it, "plain_outer": outer, "plain_true_relres_fp64": rel,
            "plain_converged": bool(conv),
            "method": "Jacobi-CG, fp32 vectors, FP64 dots, FP64 iterative refinement (inner 1e-6)",
            "fp32_only": {"ms": r.ms, "iters": r.iters, "relres": r.relres, "true_relres_fp64": r.true_relres,
                          "status": int(st)},
            "learned_pc_ms": None,
            "bytes_per_iter_alg": 40 * nvox}


def c3_bench(ctx):
    """Precision split, paracrine half: 512^3 grid, dt*D/h^2 = 20, 50,000
    unit point releases; fp32 inner CG with FP64 dots, FP64 outer
    refinement to a 1e-10 relative residual (inputs resident in HBM)."""
    from neuronephys_clean_amd import _lib, configs
    from neuronephys_clean_amd.devmem import DeviceArray as DA
    n, c = configs.C3_GRID, configs.C3_C
    b = configs.c3_sources(n=n)
    g = _lib.Grid(n, n, n, 7, c, 0.0, None, None, None)
    db, dx = DA.from_numpy(ctx, b), DA.zeros(ctx, n ** 3, np.float64)
    del b
    o = _lib.CgOpts(rtol=1e-10, maxit=5000, precond=_lib.PRECOND_JACOBI, check_every=8)
    runs = []
    for _ in range(2):
        r = _lib.CgResult()
        ctx.call("nnx_cg_solve_refine", C.byref(g), db.ptr, dx.ptr, C.byref(o), 1e-6, 30,
                 C.byref(r))
        runs.append(r)
    r = min(runs, key=lambda q: q.ms)
    return {"grid": f"{n}^3", "c": c, "sources": 50000, "rtol": 1e-10, "inner_rtol": 1e-6,
            "refined_ms": r.ms, "outer": r.outer, "inner_iters": r.iters,
            "relres_fp64": r.relres, "converged": bool(r.converged),
            "bytes_per_inner_iter_alg": 44 * n ** 3}


def raster_agreement(r32, r64, n):
    """fp32 raster against the FP64 one: the fraction of (step, neuron)
    pairs they share, the fraction of fp64 spikes the fp32 run has within
    one step, and the first step where they differ."""
    k32 = r32[:, 0].astype(np.int64) * n + r32[:, 1]
    k64 = r64[:, 0].astype(np.int64) * n + r64[:, 1]
    both = np.intersect1d(k32, k64, assume_unique=True)
    near = np.isin(k64, k32) | np.isin(k64 - n, k32) | np.isin(k64 + n, k32)
    diff = np.setxor1d(k32, k64, assume_unique=True)
    first = None
    if diff.size:
        d0 = int(diff.min())
        first = {"step": d0 // n, "neuron": d0 % n, "in_fp32": bool(np.isin(d0, k32)),
                 "in_fp64": bool(np.isin(d0, k64))}
    return {"spikes_fp32": int(k32.size), "spikes_fp64": int(k64.size),
            "match_fraction": both.size / max(k32.size, k64.size, 1),
            "within_one_step_fraction": float(near.mean()) if k64.size else 1.0,
            "first_divergence": first}


def precision_split(ctx, net, steps=(1, 100, 10000)):
    """max |V_fp32 - V_fp64| of the same network after 1, 100 and 10,000
    steps, and the fp32 raster against the fp64 one over the whole run.
    The fp64 device run is checked bit for bit against the FP64 oracle over
    the first 100 steps; later steps use it as the FP64 trajectory."""
    import oracle
    from neuronephys_clean_amd import _lib
    from neuronephys_clean_amd.sim import Simulation
    cap = 1 << 23
    s32 = Simulation(net, np.float32, couple="event", ctx=ctx, device_csc=True, record_cap=cap)
    s64 = Simulation(net, np.float64, couple="event", ctx=ctx, device_csc=True, record_cap=cap)
    out, done = {}, 0
    v64_100 = None
    for k in steps:
        for sim in (s32, s64):
            sim.run(k - done, use_graph=True, chunk=100 if (k - done) % 100 == 0 else k - done)
        done = k
        v32, v64 = s32.state()["v"], s64.state()["v"]
        out[str(k)] = float(np.max(np.abs(v32.astype(np.float64) - v64)))
        if k == 100:
            v64_100 = v64
    res = {"max_abs_dV_mV": out}
    (r32, l32), (r64, l64) = s32.raster(), s64.raster()
    if l32 <= cap and l64 <= cap:
        res["raster_fp32_vs_fp64"] = raster_agreement(r32, r64, net.n)
    if v64_100 is not None:
        T = np.float64
        V, m, h, ng = (np.asarray(a, T).copy() for a in (net.V0, net.m0, net.h0, net.n0))
        s = np.zeros(net.n, T)
        p = oracle.params(_lib.params_squid().as_dict())
        oracle.run_rk4(p, net.dt, 100, net.off, net.tgt, net.w, V, m, h, ng, s,
                       np.asarray(net.iext, T), cap=1)
        res["fp64_device_equals_oracle_at_100"] = bool(np.array_equal(V, v64_100))
    del s32, s64
    return res


def c4_bench(ctx, steps=5000, chunk=100, tune_steps=1000, warmup=100, tune=True):
    """Batched ensemble: 256 networks (sizes 2^U(10,16), Poisson(50)
    out-degree), one fused persistent launch per step, 100 steps per
    hipGraph, in the launch shape the autotuner picks: every candidate is
    timed on the same window of `tune_steps` steps after `warmup` steps (the
    initial volley and the quiet steps after it), which the search restores.
    The library's default shape is timed on the same steps from the same
    state in a second simulation (the trajectory does not depend on the
    shape), so the two times compare like for like."""
    from neuronephys_clean_amd import _lib, configs
    from neuronephys_clean_amd.sim import EnsembleSim
    t0 = time.time()
    ens = configs.c4_ensemble(seed=0, count=256)
    live = int(ens.size.sum())
    log(f"bench: C4 ensemble {len(ens.size)} networks, {live} neurons, {ens.nnz} edges, "
        f"built in {time.time() - t0:.1f}s")
    default = _lib.EnsConfig(block=256, tile=256, grid=8192)
    e0, e1 = C.c_void_p(), C.c_void_p()
    ctx.call("nnx_event_create", C.byref(e0))
    ctx.call("nnx_event_create", C.byref(e1))

    def timed(cfg, autotune, para=0):
        sim = EnsembleSim(ens, ctx=ctx, deposit=True)
        sim.cfg = cfg
        if para:
            sim.set_paracrine(para, rtol=1e-8, inner_rtol=1e-6, max_inner=40, max_outer=2)
        sim.run(warmup, use_graph=True, chunk=chunk)
        info = None
        if autotune:
            t1 = time.time()
            best, tune_us, ran = sim.autotune(tune_steps)
            info = {"config": best.as_dict(), "us_per_step": tune_us, "window": [sim.step, sim.step + tune_steps],
                    "search_s": time.time() - t1}
            log(f"bench: C4 autotune {info}")
            cfg = sim.cfg = best
        sim.run(chunk - sim.step % chunk, use_graph=False)  # align the graph chunks
        sim.run(chunk, use_graph=True, chunk=chunk)         # capture
        sim.synchronize()
        first = sim.step
        spk0 = int(sim.a["net_spikes"].numpy().sum())
        ctx.call("nnx_event_record", e0)
        sim.run(steps, use_graph=True, chunk=chunk)
        ctx.call("nnx_event_record", e1)
        sim.synchronize()
        ms = C.c_float()
        ctx.call("nnx_event_elapsed", e0, e1, C.byref(ms))
        spikes = int(sim.a["net_spikes"].numpy().sum()) - spk0
        v = sim.a["v"].numpy()
        if para:
            info = sim.paracrine_stats()
        del sim
        return cfg, ms.value, spikes, first, info, v

    best, ms, spikes, first, info, v_best = timed(default, tune)
    out = {"networks": len(ens.size), "neurons": live, "edges": ens.nnz, "steps": steps,
           "timed_steps": [first, first + steps], "graph_chunk": chunk}
    sec = ms * 1e-3
    bytes_step = 48.125 * live
    out.update({"ms": ms, "us_per_step": 1e6 * sec / steps, "neuron_steps_per_s": live * steps / sec,
                "spikes_per_step": spikes / steps, "config": best.as_dict(), "autotuned": tune,
                "autotune": info, "warmup": warmup,
                "hbm_GBps_state": bytes_step * steps / sec / 1e9,
                "frac_hbm_state": bytes_step * steps / sec / HBM_PEAK})
    if tune:
        if best.as_dict() == default.as_dict():
            out["default"] = {"config": default.as_dict(), "us_per_step": out["us_per_step"], "same_shape": True}
        else:
            _, ms_d, spikes_d, first_d, _, v_d = timed(default, False)
            out["default"] = {"config": default.as_dict(), "us_per_step": 1e3 * ms_d / steps,
                              "same_steps": first_d == first and spikes_d == spikes,
                              "same_final_v": bool(np.array_equal(v_d, v_best))}
        out["autotuned_over_default"] = out["default"]["us_per_step"] / out["us_per_step"]
    # every network's own 64^3 paracrine grid solved every 100 steps (one
    # solve of the box of grids with Dirichlet faces between them)
    _, ms_p, spikes_p, first_p, pst, v_p = timed(best, False, para=100)
    n_timed = sum(1 for s_ in range(first_p, first_p + steps) if s_ % 100 == 0)
    nx, ny, nz = ens.grid
    out["with_paracrine"] = {"every": 100, "grids": f"{len(ens.size)} x 64^3 in a {nx}x{ny}x{nz} box", "c": ens.c,
                             "rtol_true_fp64": 1e-8, "us_per_step": 1e3 * ms_p / steps,
                             "neuron_steps_per_s": live * steps / (ms_p * 1e-3), "solves": pst,
                             "solves_timed": n_timed, "ms_per_solve": (ms_p - ms) / max(n_timed, 1),
                             "same_trajectory": bool(np.array_equal(v_p, v_best))}
    ctx.call("nnx_event_destroy", e0)
    ctx.call("nnx_event_destroy", e1)
    return out


UNET_FLOP_PER_VOXEL = (2 * 27 * (2 * 16 + 16 * 16)          # level 0: e0, d0
                       + 2 * 27 * (16 * 32 + 32 * 16) / 8.0    # level 1: e1, d1
                       + 2 * 27 * (32 * 32) / 64.0)            # level 2: e2
BF16_PEAK = 2.5e15  # dense BF16 MFMA FLOP/s, MI355X


def c5_bench(ctx, counts, smooth=2):
    """Variable-diffusivity 256^3 solve (log-normal D, log-sigma 1,
    correlation length 8 voxels, mean dt D / h^2 = 50, harmonic-mean faces)
    with the C2 network's release counts as source, every solver inside FP64
    refinement to a true 1e-8 relative residual: Jacobi-CG (the plain
    baseline), flexible CG with the U-Net on Jacobi smoothing, with the
    multigrid V-cycle, and with the U-Net correcting the V-cycle."""
    from neuronephys_clean_amd import _lib, configs
    from neuronephys_clean_amd.devmem import DeviceArray as DA
    from neuronephys_clean_amd.precond import MG_WEIGHTS, UNet, device_field
    n = 256
    t0 = time.time()
    f = configs.c5_field(seed=0, n=n)
    g, keep = device_field(ctx, f)
    log(f"bench: C5 field built in {time.time() - t0:.1f}s")
    b = 0.01 * counts.astype(np.float64)
    db, dx = DA.from_numpy(ctx, b), DA.zeros(ctx, n ** 3, np.float64)
    net = UNet.load(ctx)
    net_mg = UNet.load(ctx, MG_WEIGHTS) if os.path.exists(MG_WEIGHTS) else None
    out = {"grid": f"{n}^3", "field": "log-normal D, log-sigma 1, corr 8, mean dt D/h^2 50",
           "rtol_true_fp64": 1e-8, "smooth_sweeps": smooth}
    solvers = [("jacobi_cg", None, None), ("learned_fcg", _lib.PRECOND_LEARNED, net),
               ("mg_fcg", _lib.PRECOND_MG, None)]
    if net_mg is not None:
        solvers.append(("mg_learned_fcg", _lib.PRECOND_MG_LEARNED, net_mg))
    for name, pc, nw in solvers:
        best = None
        for _ in range(2):
            r = _lib.CgResult()
            o = _lib.CgOpts(rtol=1e-8, maxit=20000, precond=_lib.PRECOND_JACOBI if pc is None else pc,
                            check_every=8, stall_iters=10, smooth=smooth if pc == _lib.PRECOND_LEARNED else 1)
            if pc is None:
                ctx.call("nnx_cg_solve_refine", C.byref(g), db.ptr, dx.ptr, C.byref(o), 1e-6, 30, C.byref(r))
            else:
                ctx.call("nnx_fcg_solve_refine", C.byref(g), nw.h if nw else None, keep["feat"].ptr, db.ptr,
                         dx.ptr, C.byref(o), 1e-6, 30, C.byref(r))
            if best is None or r.ms < best["ms"]:
                best = r.as_dict()
        if pc in (_lib.PRECOND_LEARNED, _lib.PRECOND_MG_LEARNED):
            best["fell_back"] = best["precond_used"] != pc
        out[name] = best
        log(f"bench: C5 {name} {best}")
    for name, _, _ in solvers[1:]:
        out["speedup_" + name] = out["jacobi_cg"]["ms"] / out[name]["ms"]
    learned = [k for k in ("learned_fcg", "mg_learned_fcg") if k in out]
    best_l = max(learned, key=lambda k: out["jacobi_cg"]["ms"] / out[k]["ms"])
    out["speedup_learned_vs_plain"] = out["jacobi_cg"]["ms"] / out[best_l]["ms"]
    # what the network itself adds: the learned solver against its own base
    # (the same cycle or smoother without the network) and the mean weight
    # the energy criterion gave the network's correction
    out["learned_best"] = {"solver": best_l, "fell_back": out[best_l]["fell_back"],
                           "vs_its_base": (out["mg_fcg"]["ms"] if best_l == "mg_learned_fcg"
                                           else out["jacobi_cg"]["ms"]) / out[best_l]["ms"],
                           "network_weight_mean": out[best_l]["learned_weight"]}
    # one network application (the preconditioner's dense part) on the MFMA roofline
    r32, z32 = DA.from_numpy(ctx, b.astype(np.float32)), DA.zeros(ctx, n ** 3, np.float32)

    def f_apply():
        net.apply((n, n, n), 1, r32, keep["feat"], z32)
    ms = timed_events(ctx, f_apply, 10)
    flop = UNET_FLOP_PER_VOXEL * n ** 3
    out["unet_apply"] = {"ms": ms, "flop": flop, "achieved_TFLOPs": flop / (ms * 1e-3) / 1e12,
                         "peak_TFLOPs": BF16_PEAK / 1e12, "mfma_frac": flop / (ms * 1e-3) / BF16_PEAK}

    def f_mg():
        ctx.call("nnx_mg_apply", C.byref(g), 0, 0, 0, r32.ptr, z32.ptr)
    out["mg_vcycle_ms"] = timed_events(ctx, f_mg, 10)
    net.close()
    if net_mg is not None:
        net_mg.close()
    return out


def stencil_roofline(ctx, n=256, reps=50):
    """y = A x on an n^3 grid (constant 7-point): 8 algorithmic bytes per
    voxel (read x, write y), HIP events on the library stream."""
    from neuronephys_clean_amd import _lib
    from neuronephys_clean_amd.devmem import DeviceArray as DA
    g = _lib.Grid(n, n, n, 7, 1.0, 0.0, None, None, None)
    x = DA.from_numpy(ctx, np.random.default_rng(0).random(n ** 3).astype(np.float32))
    y = DA.zeros(ctx, n ** 3, np.float32)

    def f():
        ctx.call("nnx_stencil_apply", C.byref(g), x.ptr, y.ptr)
    ms = timed_events(ctx, f, reps)
    b = 8.0 * n ** 3
    return {"kernel": "k_stencil7x4<false> (y = A x, 7-point, fp32)", "grid": f"{n}^3",
            "bound": "hbm", "achieved": b / (ms * 1e-3) / 1e9, "peak": HBM_PEAK / 1e9,
            "unit": "GB/s", "frac": b / (ms * 1e-3) / HBM_PEAK, "avg_launch_us": ms * 1e3,
            "alg_bytes_per_launch": b}


def profile_pass(ctx, sim, net):
    from neuronephys_clean_amd import _lib, configs
    from neuronephys_clean_amd.devmem import DeviceArray as DA
    from neuronephys_clean_amd.precond import UNet
    from neuronephys_clean_amd.sim import EnsembleSim
    # 1,000 steps, then one step at step 1,000 with the in-step paracrine
    # solve on the releases of those steps (eager dispatches throughout)
    every = sim.plan.paracrine.every
    sim.plan.paracrine.every = 0
    sim.run(1000)
    sim.plan.paracrine.every = every
    sim.run(1)
    sim.synchronize()
    log(f"profile: in-step solve {sim.paracrine_stats()}")
    sim.plan.paracrine.every = 0
    sim.run(20)  # eager: one dispatch per kernel per step
    sim.synchronize()
    if sim.plan.couple_mode == _lib.COUPLE_WEIGHTED_FUSED:
        # and 20 steps of the unfused event path (push, HH + gather, compaction)
        sim.plan.couple_mode = _lib.COUPLE_WEIGHTED_EVENT
        sim.run(20)
        sim.synchronize()
        sim.plan.couple_mode = _lib.COUPLE_WEIGHTED_FUSED
    # the HH kernel alone (the roofline measurement of the benchmark)
    st, p = sim.plan.state, sim.params
    for _ in range(20):
        ctx.call("nnx_hh_step", C.byref(st), C.byref(p), _lib.INT_RK4, _lib.SPIKE_CROSSING, net.dt, 0,
                 sim.a["g_in"].ptr, sim.a["bits0"].ptr)
    ctx.synchronize()
    # 20 dense pulls (WEIGHTED_PULL) on a seeded bitmap of 1% of the neurons
    fire = np.random.default_rng(11).random(net.n) < 0.01
    pb = DA.from_numpy(ctx, np.packbits(np.concatenate([fire, np.zeros(-net.n % 64, bool)]),
                                        bitorder="little").view(np.uint64))
    for _ in range(20):
        ctx.call("nnx_couple", C.byref(sim.graph), _lib.COUPLE_WEIGHTED_PULL, _lib.F32, pb.ptr, None, None, None,
                 sim.a["g_in"].ptr)
    ctx.synchronize()
    log(f"profile: dense pull spiking in-edges E_s={int(np.diff(net.off.astype(np.int64))[fire].sum())}")
    # 20 electrode readouts (1,000 electrodes x 50 nearest neurons) and network means
    from neuronephys_clean_amd import host, readout
    off, idx = host.spherical_regions(net.pos, 50, center_idx=(net.n // 1000) * np.arange(1000))
    enet = readout.ElectrodeNet(ctx, idx, off)
    avg, mean = DA(ctx, 1000, np.float32), DA(ctx, 1, np.float32)
    for _ in range(20):
        ctx.call("nnx_electrode_measure", C.byref(enet.r), _lib.F32, net.n, sim.a["v"].ptr, sim.a["bits0"].ptr,
                 avg.ptr, enet.d_firing.ptr)
        ctx.call("nnx_network_mean", _lib.F32, net.n, sim.a["v"].ptr, mean.ptr)
    ctx.synchronize()
    # 21 launches each of the Diekman and SAN steps at the same size
    model_bench(ctx, net.n, reps=20)
    # 20 Hebbian updates (window 300) on the firing of 20 more fused steps
    from neuronephys_clean_amd import hebbian
    hs = hebbian.HebbianState(ctx, sim.graph, 300, 0.1, net.w)
    for _ in range(20):
        sim.run(1)
        hs.update(sim.step, sim.a["bits0"] if sim.step % 2 else sim.a["bits1"])  # the step just run
    ctx.synchronize()
    del hs
    # 20 steps of the 256-network ensemble (default launch shape)
    ens = EnsembleSim(configs.c4_ensemble(seed=0, count=256), ctx=ctx)
    ens.run(20)
    ens.synchronize()
    del ens
    # 5 applications of the preconditioner network at 256^3
    u = UNet.load(ctx)
    rr = DA.from_numpy(ctx, np.random.default_rng(0).normal(0, 1, 256 ** 3).astype(np.float32))
    zz = DA.zeros(ctx, 256 ** 3, np.float32)
    for _ in range(5):
        u.apply((256, 256, 256), 1, rr, None, zz)
    ctx.synchronize()
    u.close()
    stencil_roofline(ctx, 256, 20)
    # 3 multigrid V-cycles (default sweeps) on the 256^3 variable-diffusivity operator
    from neuronephys_clean_amd.precond import device_field
    fg, keep = device_field(ctx, configs.c5_field(seed=0, n=256))
    for _ in range(3):
        ctx.call("nnx_mg_apply", C.byref(fg), 0, 0, 0, rr.ptr, zz.ptr)
    ctx.synchronize()
    # 60 Jacobi-CG iterations on the same operator (the tiled variable-coefficient iteration)
    zz.upload(np.zeros(256 ** 3, np.float32))
    ov = _lib.CgOpts(rtol=0.0, maxit=60, precond=_lib.PRECOND_JACOBI, check_every=20)
    ctx._l.nnx_cg_solve(ctx.h, C.byref(fg), rr.ptr, zz.ptr, C.byref(ov), C.byref(_lib.CgResult()))
    ctx.synchronize()
    del keep
    vc = sim.a["voxel_counts"].numpy().astype(np.float32)
    b = 0.01 * vc
    if not b.any():
        b[len(b) // 2] = 1.0
    g = _lib.Grid(256, 256, 256, 7, 1.0, 0.0, None, None, None)
    db, dx = DA.from_numpy(ctx, b), DA.zeros(ctx, b.size, np.float32)
    o = _lib.CgOpts(rtol=1e-8, maxit=2000, precond=_lib.PRECOND_JACOBI, check_every=8)
    r = _lib.CgResult()
    ctx._l.nnx_cg_solve(ctx.h, C.byref(g), db.ptr, dx.ptr, C.byref(o), C.byref(r))
    src, dst = DA(ctx, 1 << 28, np.uint8), DA(ctx, 1 << 28, np.uint8)
    ctx.call("nnx_memcpy_d2d", dst.ptr, src.ptr, C.c_size_t(1 << 28))
    ctx.synchronize()
    log("profile pass done: 1,000 steps, 1 step with the in-step solve, 20 fused + 20 unfused steps, 20 HH launches, 20 dense pulls, 20 electrode readouts, "
        "21 Diekman and SAN "
        "steps, 20 Hebbian updates, 20 ensemble steps, "
        "5 network applications, "
        f"20 stencil applies, 3 V-cycles, 1 CG solve ({r.iters} iterations), 256 MiB copy")


def profiled_traffic(label):
    """HBM bytes per launch of a kernel as measured by the FETCH_SIZE and
    WRITE_SIZE passes summarised in profiles/hbm_traffic_profile_mode.md
    (None when the summary is absent)."""
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), "profiles", "hbm_traffic_profile_mode.md")
    try:
        with open(path) as fh:
            m = re.search(re.escape(label) + r" traffic per launch: ([0-9.eE+-]+) bytes", fh.read())
    except OSError:
        return None
    return float(m.group(1)) if m else None


SNAP_KEYS = ("v", "m", "h", "ng", "s", "bits0", "bits1", "list", "count", "voxel_counts", "raster_len")


def snapshot(ctx, sim, extra=()):
    """Device copy of everything a fused C2 run reads between calls (and
    the `extra` arrays, e.g. the paracrine field and its statistics)."""
    from neuronephys_clean_amd.devmem import DeviceArray
    snap = {}
    for kk in SNAP_KEYS + tuple(extra):
        snap[kk] = DeviceArray(ctx, sim.a[kk].shape, sim.a[kk].dtype)
        snap[kk].copy_from(sim.a[kk])
    return snap, sim.step


def restore_snapshot(sim, snap):
    snapd, step0 = snap
    for kk, arr in snapd.items():
        sim.a[kk].copy_from(arr)
    sim.step = step0


def cg_iteration_roofline(ctx, n, b32, c, k1=10, k2=40):
    """One Jacobi-CG iteration on the n^3 constant 7-point grid (its two
    tiled launches): the difference of two capturable solves of k2 and k1
    iterations with rtol 0 (every iteration works), HIP events on the
    library stream.  Algorithmic bytes per iteration: 32 per voxel - first
    half: read r, p_old and x, write p and x (the previous iteration's
    deferred update), q = A p formed in LDS and not stored; second half:
    read p and r, write r (q recomputed in LDS)."""
    from neuronephys_clean_amd import _lib
    from neuronephys_clean_amd.devmem import DeviceArray as DA
    g = _lib.Grid(n, n, n, 7, c, 0.0, None, None, None)
    db, dx = DA.from_numpy(ctx, b32), DA.zeros(ctx, n ** 3, np.float32)
    res = DA.zeros(ctx, 4, np.float64)

    def solve(k):
        o = _lib.CgOpts(rtol=0.0, maxit=k, precond=_lib.PRECOND_JACOBI, check_every=1)
        ctx.call("nnx_cg_solve_async", C.byref(g), db.ptr, dx.ptr, C.byref(o), res.ptr)
    t1 = timed_events(ctx, lambda: solve(k1), 5)
    t2 = timed_events(ctx, lambda: solve(k2), 5)
    us = (t2 - t1) * 1e3 / (k2 - k1)
    b = 32.0 * n ** 3
    return {"kernel": f"k_stencil7x4<2> + k_stencil7x4<3> (one Jacobi-CG iteration, {n}^3 constant 7-point)",
            "bound": "hbm", "achieved": b / (us * 1e-6) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
            "frac": b / (us * 1e-6) / HBM_PEAK, "traffic": profiled_traffic("cg-iteration"),
            "avg_iteration_us": us, "alg_bytes_per_launch": b,
            "alg_bytes_rule": "32 B per voxel per iteration (20 + 12, two launches)",
            "time_source": f"HIP events: ({k2}-iteration solve - {k1}-iteration solve) / {k2 - k1}"}


def model_bench(ctx, n, reps=100):
    """The reference's other neuron models at n neurons, fp32: the Diekman
    SCN step (the reference's default program, src/main.cu:570-577) and the
    SAN step, each timed over `reps` back-to-back launches from a
    physiological start, with its HBM roofline (algorithmic bytes: every
    array read and written once, plus the spike bit)."""
    from neuronephys_clean_amd import _lib
    from neuronephys_clean_amd.devmem import DeviceArray as DA

    rng = np.random.default_rng(0)
    V = rng.uniform(-70, -40, n)
    e = np.exp
    dk = {"v": V, "m": 1 / (1 + e(-(V + 35.2) / 8.1)), "h": 1 / (1 + e((V + 62) / 2)),
          "ng": (1 + e((V - 14) / -17.0)) ** -0.25, "rl": 1 / (1 + e(-(V + 36) / 5.1)),
          "rnl": 1 / (1 + e(-(V + 21.6) / 6.7)), "fnl": 1 / (1 + e((V + 260) / 65)), "s": np.zeros(n),
          "p": (1 + e(-(V + 25) / 7.4)) ** -1.5, "y": np.zeros(n), "cas": np.full(n, 1e-4), "cac": np.full(n, 1e-4),
          "iin": np.zeros(n), "icount": np.zeros(n), "ioutput": np.zeros(n), "egaba": np.full(n, -60.0),
          "gkca": rng.normal(100, 0.5, n), "gkleak": rng.normal(0.12, 0.01, n)}
    d = {k: DA.from_numpy(ctx, np.asarray(v, np.float32)) for k, v in dk.items()}
    bits = DA.zeros(ctx, (n + 63) // 64, np.uint64)
    st = _lib.DiekmanState(dtype=_lib.F32, n=n, **{k: d[k].ptr for k in dk})
    pd = _lib.params_diekman()

    def f_dk():
        ctx.call("nnx_diekman_step", C.byref(st), C.byref(pd), 0.4, bits.ptr)
    us_dk = timed_events(ctx, f_dk, reps) * 1e3
    sn = {"v": rng.uniform(-70, -40, n), "nk": rng.uniform(0, 0.3, n), "ca": rng.uniform(1, 60, n),
          "ye": np.zeros(n), "yi": np.zeros(n), "ein": np.zeros(n), "ecount": np.zeros(n), "iin": np.zeros(n),
          "icount": np.zeros(n), "eoutput": np.zeros(n), "ioutput": np.zeros(n)}
    d2 = {k: DA.from_numpy(ctx, np.asarray(v, np.float32)) for k, v in sn.items()}
    st2 = _lib.SanState(dtype=_lib.F32, n=n, **{k: d2[k].ptr for k in sn})
    ps = _lib.params_san()

    def f_san():
        ctx.call("nnx_san_step", C.byref(st2), C.byref(ps), 0.4, bits.ptr)
    us_san = timed_events(ctx, f_san, reps) * 1e3
    out = {}
    for name, us, b, kern in (("diekman", us_dk, 120.125, "k_diekman<float>"), ("san", us_san, 64.125, "k_san<float>")):
        ach = b * n / (us * 1e-6)
        out[name] = {"kernel": kern, "neurons": n, "avg_launch_us": us, "neuron_steps_per_s": n / (us * 1e-6),
                     "alg_bytes_per_neuron": b, "bound": "hbm", "achieved_GBs": ach / 1e9, "frac": ach / HBM_PEAK}
    return out


def replay_window(ctx, sim, snap, chunk, steps):
    """Restores the saved state, captures the plan's graphs on 2 chunks,
    restores again and times `steps` steps (HIP events).  Returns the time,
    (sum of V, spikes) after the window and a sample of V's bits."""
    snapd, step0 = snap

    def restore():
        for kk in SNAP_KEYS:
            sim.a[kk].copy_from(snapd[kk])
        sim.step = step0

    restore()
    sim.run(2 * chunk, use_graph=True, chunk=chunk)
    restore()
    sim.synchronize()
    e0, e1 = C.c_void_p(), C.c_void_p()
    ctx.call("nnx_event_create", C.byref(e0))
    ctx.call("nnx_event_create", C.byref(e1))
    ctx.call("nnx_event_record", e0)
    sim.run(steps, use_graph=True, chunk=chunk)
    ctx.call("nnx_event_record", e1)
    sim.synchronize()
    msv = C.c_float()
    ctx.call("nnx_event_elapsed", e0, e1, C.byref(msv))
    ctx.call("nnx_event_destroy", e0)
    ctx.call("nnx_event_destroy", e1)
    v = sim.a["v"].numpy()
    spk = int(sim.a["raster_len"].numpy()[0]) - int(snapd["raster_len"].numpy()[0])
    return msv.value, (float(np.sum(v.astype(np.float64))), spk), v.view(np.uint32)[::4099].copy()


def hebbian_bench(ctx, sim, net, chunk, snap, steps=2000, window=300, inc=0.1):
    """Hebbian learning on every synapse of C2 during the timed window, with
    the reference's HebbianLearningExample settings (window 300 steps,
    heb_inc 0.1, src_wholebrain/main.cu:150): the window replayed from its
    saved start without and with the rule."""
    from neuronephys_clean_amd import _lib, hebbian

    t_none, s0, p0 = replay_window(ctx, sim, snap, chunk, steps)
    st = hebbian.attach(sim, window, inc, net.w)
    t_h, s1, p1 = replay_window(ctx, sim, snap, chunk, steps)
    changed = int((st.weights() != net.w).sum())
    armed = int(st.armed().astype(np.int64).sum())
    sim.plan.hebbian = _lib.Hebbian()
    del sim.hebbian, st
    return {"workload": f"Hebbian rule (window {window}, heb_inc {inc}) on all {net.nnz} synapses, "
                        f"the {steps}-step timed C2 window replayed from its saved start",
            "window_ms": {"no_hebbian": t_none, "hebbian": t_h},
            "per_step_us": (t_h - t_none) * 1e3 / steps,
            "weights_changed": changed, "armed_after": armed,
            "dense_rule_bytes_per_step": 8 * net.nnz,
            "simulation_unchanged": bool(s0 == s1 and np.array_equal(p0, p1))}


def readout_bench(ctx, sim, net, chunk, snap, steps=2000, every=40, n_elec=1000, k=50):
    """Electrode readout on the C2 run: 1,000 electrodes, each the mean of
    its 50 nearest neurons, plus the network mean, read every 40 steps (the
    reference's 2.5 kHz sampling, src/main.cu:547,574-575, at dt 0.01 ms).
    The benchmark's timed window is run three more times from the state it
    started from (the run is deterministic): without readout, with it every
    40 steps and every step; the time differences are the readout's cost
    inside the step graph.  Graphs are captured before each timed window."""
    from neuronephys_clean_amd import host

    centers = (net.n // n_elec) * np.arange(n_elec)
    off, idx = host.spherical_regions(net.pos, k, center_idx=centers)

    def window():
        return replay_window(ctx, sim, snap, chunk, steps)

    t_none, sum0, probe0 = window()
    sim.set_readout(off, idx, every, steps // every + 1)
    t_40, sum40, probe40 = window()
    rows = sim.readout()
    sim.set_readout(off, idx, 1, steps + 1)
    t_1, sum1, probe1 = window()
    sim.plan.readout_every = 0
    same = sum0 == sum40 == sum1 and np.array_equal(probe0, probe40) and np.array_equal(probe0, probe1)
    per_read_us = (t_1 - t_none) * 1e3 / steps
    # the kernels alone, eager launches (an upper bound: includes launch gaps)
    r = sim.plan.regions
    fire = sim.readout()["firing"]

    def f_measure():
        ctx.call("nnx_electrode_measure", C.byref(r), _lib_f32(), net.n, sim.a["v"].ptr, sim.a["bits0"].ptr,
                 sim.a["ro_avg"].ptr, sim.a["ro_firing"].ptr)

    def f_mean():
        ctx.call("nnx_network_mean", _lib_f32(), net.n, sim.a["v"].ptr, sim.a["ro_mean"].ptr)
    m_us = timed_events(ctx, f_measure, 200) * 1e3
    mean_us = timed_events(ctx, f_mean, 200) * 1e3
    alg = (8.125 * len(idx) + 12 * n_elec)  # idx + V + spike bit per index; offset, mean, count per region
    return {"workload": f"{n_elec} electrodes x {k} nearest neurons + network mean, every {every} steps, "
                        f"the {steps}-step timed C2 window replayed from its saved start",
            "window_ms": {"no_readout": t_none, f"every_{every}": t_40, "every_step": t_1},
            "overhead_frac_every_%d" % every: (t_40 - t_none) / t_none,
            "per_readout_us_in_graph": per_read_us,
            "eager_launch_us": {"electrode_measure": m_us, "network_mean": mean_us},
            "electrode_measure_alg_bytes": alg,
            "rows": int(len(rows["avg"]) - 1), "firing_in_rows": int(rows["firing"].sum()),
            "firing_every_step_rows": int(fire.sum()),
            "spikes_in_window": sum0[1],
            "simulation_unchanged": bool(same)}


def _lib_f32():
    from neuronephys_clean_amd import _lib
    return _lib.F32


def tune_fused(ctx, sim, chunk):
    """Picks the fused step's launch shape by timing whole graph-replayed
    chunks of real steps (untimed by the benchmark; results do not depend on
    the shape)."""
    e0, e1 = C.c_void_p(), C.c_void_p()
    ctx.call("nnx_event_create", C.byref(e0))
    ctx.call("nnx_event_create", C.byref(e1))
    best, steps = None, 0
    for block in (256, 512):
        for tile in (256, 512, 1024, 2048):
            if tile < block:
                continue
            for grid in (1024, 2048, 4096):
                sim.set_fused(block, tile, grid)
                sim.run(chunk, use_graph=True, chunk=chunk)  # capture
                ctx.call("nnx_event_record", e0)
                sim.run(2 * chunk, use_graph=True, chunk=chunk)
                ctx.call("nnx_event_record", e1)
                sim.synchronize()
                ms = C.c_float()
                ctx.call("nnx_event_elapsed", e0, e1, C.byref(ms))
                us = ms.value * 1e3 / (2 * chunk)
                steps += 3 * chunk
                if best is None or us < best["us_per_step"]:
                    best = {"block": block, "tile": tile, "grid": grid, "us_per_step": us}
    sim.set_fused(best["block"], best["tile"], best["grid"])
    best["tune_steps"] = steps
    return best


C2_STEPS = 10000      # the configuration's 100 ms at dt 0.01 ms
FIXED_RELEASE_STEPS = 1000
# CG iterations enqueued for the first refinement step of each in-step solve
# (half as many for the second): C2's solves take at most 25 in the first
# and 34 in all (measured over the configuration's 100 solves); iterations
# past convergence are launches that return at once
MAX_INNER = 32


def cpu_model():
    try:
        with open("/proc/cpuinfo") as fh:
            for line in fh:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline_full(net, every, b_fixed, all_steps=300, one_steps=20, reps=3):
    """The oracle (fp32 restatement of the same step: RK4, push coupling in
    ascending source order, release counting) on the host cores, on a
    bounded sample of the same workload: the first `all_steps` steps on all
    threads and the first `one_steps` on one thread, median of `reps` runs
    each, plus one FP64-refined Jacobi-CG solve of the 256^3 grid to 1e-8 on
    the fixed release field (all threads).  The loop rate is
    N / (t_step + t_solve / every)."""
    import oracle
    from neuronephys_clean_amd import _lib
    T = np.float32
    p = oracle.params(_lib.params_squid().as_dict())
    nthreads = oracle.set_threads(0)

    def step_time(steps):
        ts = []
        for _ in range(reps):
            V, m, h, ng = (np.asarray(a, T).copy() for a in (net.V0, net.m0, net.h0, net.n0))
            s = np.zeros(net.n, T)
            _, _, _, sec = oracle.run_rk4(p, net.dt, steps, net.off, net.tgt, net.w, V, m, h, ng, s,
                                          np.asarray(net.iext, T), cap=1, voxel_of=net.ensure_voxel(),
                                          nvox=int(np.prod(net.grid)))
            ts.append(sec / steps)
        return float(np.median(ts))

    t_all = step_time(all_steps)
    oracle.set_threads(1)
    t_one = step_time(one_steps)
    oracle.set_threads(0)
    g = oracle.grid(*net.grid, 7, net.c)
    t0 = time.perf_counter()
    x, outer, inner, rel = oracle.refine(g, b_fixed, 1e-8, 1e-6)
    t_solve = time.perf_counter() - t0
    per_step = t_all + (t_solve / every if every > 0 else 0.0)
    return {"value": net.n / per_step, "unit": "neuron-steps/s", "cores": nthreads, "kind": "port",
            "sample": f"oracle fp32 step (RK4 + push coupling in ascending source order + release counting) over "
                      f"the first {all_steps} steps of the same network and initial state, median of {reps} runs "
                      f"on {nthreads} threads ({t_all * 1e3:.2f} ms/step), plus one FP64-refined Jacobi-CG solve of "
                      f"the 256^3 grid to 1e-8 on the fixed release field ({t_solve:.2f} s, {outer} outer steps, "
                      f"{inner} inner iterations, relres {rel:.2e}) "
                      f"every {every} steps",
            "one_core": {"value": net.n / (t_one + (t_solve * nthreads / every if every > 0 else 0.0)),
                         "step_ms": t_one * 1e3, "steps": one_steps,
                         "note": "first steps on 1 thread, median of 3; the solve share scaled from the "
                                 "all-thread time by the thread count"},
            "all_threads_step_ms": t_all * 1e3, "solve_s": t_solve, "cpu_model": cpu_model(),
            "nproc": os.cpu_count()}


SNAP_PARA = ("conc", "para_stats")


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=C2_STEPS,
                    help="requested timed steps; the timed window is always at least the configuration's "
                         "10,000 steps from step 0")
    ap.add_argument("--warmup", type=int, default=200,
                    help="untimed steps from step 0 before the timed window (at least two graph chunks); the "
                         "state is restored to step 0 afterwards")
    ap.add_argument("--chunk", type=int, default=100)
    ap.add_argument("--para-every", type=int, default=100,
                    help="steps between paracrine solves inside the step (the reference's timescale)")
    ap.add_argument("--n", type=int, default=1 << 20)
    ap.add_argument("--couple", default="fused", choices=["fused", "event", "pull"])
    ap.add_argument("--tune", action="store_true",
                    help="pick the fused step's launch shape by timing 21 shapes on real steps before the timed "
                         "region")
    ap.add_argument("--mode", default="bench", choices=["bench", "profile", "loop"],
                    help="profile: a short eager sequence for rocprofv3 --pmc passes; loop: only the timed C2 "
                         "windows (for a kernel trace); neither prints the JSON line")
    ap.add_argument("--eager", action="store_true",
                    help="loop mode: dispatch every kernel eagerly instead of replaying step graphs (for kernel "
                         "traces; the kernels are the same)")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-cg", action="store_true")
    ap.add_argument("--no-c3", action="store_true")
    ap.add_argument("--no-c4", action="store_true")
    ap.add_argument("--no-c5", action="store_true")
    ap.add_argument("--no-readout", action="store_true")
    args = ap.parse_args()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus > 1 is not supported "
                                   "(multi-GPU sharding is design only)"}))
        sys.exit(2)
    chunk = max(1, args.chunk)
    steps = max(args.steps, C2_STEPS)
    steps = (steps + chunk - 1) // chunk * chunk
    warm = max(args.warmup, 2 * chunk)
    warm = (warm + chunk - 1) // chunk * chunk

    from neuronephys_clean_amd import _lib, configs
    from neuronephys_clean_amd.sim import Simulation

    t0 = time.time()
    net = configs.c2(n=args.n)
    log(f"bench: network n={net.n} nnz={net.nnz} grid={net.grid} built in {time.time() - t0:.1f}s")
    ctx = _lib.context(0)
    t0 = time.time()
    sim = Simulation(net, np.float32, couple=args.couple, record_cap=1 << 24, deposit=True,
                     ctx=ctx, device_csc=True)
    every = args.para_every
    sim.set_paracrine(max(every, 1), rtol=1e-8, inner_rtol=1e-6, max_inner=MAX_INNER, max_outer=2)
    sim.synchronize()
    log(f"bench: uploaded + device CSC in {time.time() - t0:.1f}s")

    if args.mode == "profile":
        profile_pass(ctx, sim, net)
        return

    tune = None
    if args.couple == "fused":
        if args.tune:
            tune = tune_fused(ctx, sim, chunk)
            sim.step = 0
        else:  # the library's default shape (one 256-neuron tile per workgroup, 4096 workgroups)
            tune = {"block": 256, "tile": 256, "grid": 4096, "tune_steps": 0}
        sim.set_fused(tune["block"], tune["tile"], tune["grid"])
    snap = snapshot(ctx, sim, extra=SNAP_PARA)

    def restore():
        restore_snapshot(sim, snap)

    def para(on):
        sim.plan.paracrine.every = every if (on and every > 0) else 0

    # a fixed release field: the per-voxel spike counts of the first 1,000
    # steps (the right-hand side of the standalone CG and C5 solves)
    para(False)
    sim.run(FIXED_RELEASE_STEPS, use_graph=not args.eager, chunk=chunk)
    counts_fixed = sim.a["voxel_counts"].numpy().astype(np.float64)
    b_fixed = net.amount * counts_fixed
    log(f"bench: fixed release field: {int(counts_fixed.sum())} releases in {FIXED_RELEASE_STEPS} steps")

    def window(on):
        """The timed window from step 0: `steps` steps in graphs of `chunk`
        (a solve at the first step of every chunk when on), warmed up on
        the same graphs first; returns wall and event milliseconds."""
        para(on)
        restore()
        sim.run(warm, use_graph=True, chunk=chunk)
        restore()
        e0, e1 = C.c_void_p(), C.c_void_p()
        ctx.call("nnx_event_create", C.byref(e0))
        ctx.call("nnx_event_create", C.byref(e1))
        sim.synchronize()
        w0 = time.perf_counter()
        ctx.call("nnx_event_record", e0)
        sim.run(steps, use_graph=True, chunk=chunk)
        ctx.call("nnx_event_record", e1)
        sim.synchronize()
        wall = time.perf_counter() - w0
        ev = C.c_float()
        ctx.call("nnx_event_elapsed", e0, e1, C.byref(ev))
        ctx.call("nnx_event_destroy", e0)
        ctx.call("nnx_event_destroy", e1)
        return wall * 1e3, ev.value

    if args.mode == "loop":
        for on in (True, False):
            if args.eager:
                para(on)
                restore()
                t0 = time.perf_counter()
                sim.run(steps)
                sim.synchronize()
                wall_ms = (time.perf_counter() - t0) * 1e3
            else:
                wall_ms, ev_ms = window(on)
            log(f"bench loop: paracrine={'on' if on else 'off'} {steps} steps {wall_ms:.1f} ms "
                f"({'eager' if args.eager else 'graphs'}), solves {sim.paracrine_stats()}")
        return

    # the headline: HH + coupling + release + paracrine solve every `every` steps
    wall_ms, ev_ms = window(True)
    spikes = sim.raster()[1]
    pst = sim.paracrine_stats()
    ras_full = sim.raster()[0]
    conc_sum = float(sim.a["conc"].numpy().sum())
    value = net.n * steps / (wall_ms * 1e-3)
    ms_step = wall_ms / steps
    log(f"bench: {steps} steps with paracrine every {every}: {wall_ms:.1f} ms (events {ev_ms:.1f}), {spikes} spikes, "
        f"{value:.3e} neuron-steps/s, solves {pst}")
    # the same window without the solves
    wall0_ms, ev0_ms = window(False)
    same_raster = bool(np.array_equal(sim.raster()[0], ras_full))
    steps_only = net.n * steps / (wall0_ms * 1e-3)
    log(f"bench: steps only {wall0_ms:.1f} ms (events {ev0_ms:.1f}), {steps_only:.3e} neuron-steps/s")
    n_solves = int(pst["solves"])
    paracrine = {"every": every, "solves": n_solves, "grid": f"{net.grid[0]}^3", "c": net.c,
                 "method": "Jacobi-CG (fp32 inner, FP64 dots) with FP64 iterative refinement, warm start from the "
                           f"previous field, captured in the step graphs (2 refinement steps of at most {MAX_INNER} "
                           f"and {MAX_INNER // 2} inner iterations enqueued per solve, kernels after convergence "
                           "return on the device's done flag; "
                           "constant 7-point grid: tiled FP64 residual, two tiled launches per inner iteration)",
                 "rtol_true_fp64": 1e-8, "max_relres": pst["max_relres"], "unconverged": int(pst["unconverged"]),
                 "inner_iters_per_solve": pst["inner_iters"] / max(n_solves, 1),
                 "outer_per_solve": pst["outer_steps"] / max(n_solves, 1),
                 "ms_per_solve": (ev_ms - ev0_ms) / max(n_solves, 1),
                 "share_of_window": (ev_ms - ev0_ms) / ev_ms, "field_sum": conc_sum,
                 "rasters_unchanged_by_solves": same_raster}

    # kernel rooflines: the fused step (the steps-only window over its steps:
    # the step kernel plus its share of one push and one compaction per chunk)
    fused_roof = None
    if args.couple == "fused":
        ev_per_step = spikes / steps * net.nnz / net.n
        fb = 48.125 * net.n + 32.0 * ev_per_step
        us = ev0_ms * 1e3 / steps
        fused_roof = {"kernel": "k_ens_step<%d> (fused gather + RK4 + push)" % tune["block"], "bound": "hbm",
                      "achieved": fb / (us * 1e-6) / 1e9, "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                      "frac": fb / (us * 1e-6) / HBM_PEAK, "traffic": profiled_traffic("fused-step"),
                      "traffic_source": "profiles/hbm_traffic_profile_mode.md (FETCH_SIZE + WRITE_SIZE, "
                                        "calibrated on a 256 MiB copy, quiet-step median)",
                      "avg_step_us": us,
                      "time_source": "HIP events around the steps-only window on the library stream / steps "
                                     "(the step kernel plus its share of one push and one compaction per "
                                     "100-step graph)",
                      "alg_bytes_per_launch": fb,
                      "alg_bytes_rule": "48.125 B per neuron + 32 B per delivered event",
                      "note": "bound by VALU issue, not HBM, on quiet steps: 599 VALU instructions per neuron "
                              "at ~0.8-0.9 of the non-packed issue rate (profiles/valu_rk4_kernels_c2.md), and by "
                              "random event traffic on the ~290 steps of synchronous firing"}
    cg_iter = cg_iteration_roofline(ctx, net.grid[0], b_fixed.astype(np.float32), net.c)
    log(f"bench: CG iteration {cg_iter}")
    t_cg_iters = cg_iter["avg_iteration_us"] * 1e-3 * pst["inner_iters"]
    t_steps = ev0_ms
    use_fused = bool(fused_roof) and t_steps >= t_cg_iters
    roof = dict(fused_roof if use_fused else cg_iter)
    roof["share_of_timed_window"] = (t_steps if use_fused else t_cg_iters) / ev_ms

    # the HH RK4 kernel of the unfused path, timed alone on the library
    # stream over the same state
    st = sim.plan.state
    bits = sim.a["bits0"].ptr
    gin = sim.a["g_in"].ptr
    p = sim.params

    def hh():
        ctx.call("nnx_hh_step", C.byref(st), C.byref(p), _lib.INT_RK4, _lib.SPIKE_CROSSING,
                 net.dt, 0, gin, bits)
    hh_ms = timed_events(ctx, hh, 200)
    hh_bytes = 48.125 * net.n
    roof_hh = {"kernel": "k_hh_rk4<float>", "bound": "hbm", "achieved": hh_bytes / (hh_ms * 1e-3) / 1e9,
               "peak": HBM_PEAK / 1e9, "unit": "GB/s", "frac": hh_bytes / (hh_ms * 1e-3) / HBM_PEAK,
               "traffic": None, "avg_launch_us": hh_ms * 1e3, "alg_bytes_per_launch": hh_bytes}
    models = model_bench(ctx, net.n)
    log(f"bench: models {models}")

    # the dense pull coupling (WEIGHTED_PULL: every target reads all its
    # in-edges, the spike bitmap and the weights of its spiking in-edges),
    # timed alone on seeded spike bitmaps of 0.1%, 1% and 3% of the neurons.
    # Algorithmic bytes: 4(N+1) + 4E + 4E_s + N/8 + 4N (offsets, sources,
    # weights of the E_s spiking in-edges, bitmap, g_in); with the HH kernel
    # (48.125 B per neuron) beside it, the pair is an unfused HH + coupling
    # step whose bytes are counted against the HBM peak.
    from neuronephys_clean_amd.devmem import DeviceArray as DA
    outdeg = np.diff(net.off.astype(np.int64))
    prng = np.random.default_rng(11)
    pbits = DA.zeros(ctx, (net.n + 63) // 64, np.uint64)
    pull_cases = []
    for frac in (0.001, 0.01, 0.03):
        fire = prng.random(net.n) < frac
        pbits.upload(np.packbits(np.concatenate([fire, np.zeros(-net.n % 64, bool)]),
                                 bitorder="little").view(np.uint64))
        e_s = int(outdeg[fire].sum())

        def pull():
            ctx.call("nnx_couple", C.byref(sim.graph), _lib.COUPLE_WEIGHTED_PULL, _lib.F32,
                     pbits.ptr, None, None, None, gin)
        ms = timed_events(ctx, pull, 20)
        b = 4.0 * (net.n + 1) + 4.0 * net.nnz + 4.0 * e_s + net.n / 8.0 + 4.0 * net.n
        pull_cases.append({"firing": frac, "spiking_in_edges": e_s, "avg_launch_us": ms * 1e3,
                           "alg_bytes_per_launch": b, "achieved": b / (ms * 1e-3) / 1e9,
                           "frac": b / (ms * 1e-3) / HBM_PEAK,
                           "hh_plus_pull_frac": (b + hh_bytes) / ((ms + hh_ms) * 1e-3) / HBM_PEAK,
                           "hh_plus_pull_neuron_steps_per_s": net.n / ((ms + hh_ms) * 1e-3)})
    roof_pull = dict(pull_cases[1])
    roof_pull.update({"kernel": "k_pull_lds<float,0> (dense pull over the CSC, spike bitmap in LDS, "
                                "ascending-source sums)", "bound": "hbm", "peak": HBM_PEAK / 1e9, "unit": "GB/s",
                      "alg_bytes_rule": "4(N+1) + 4E + 4E_s + N/8 + 4N", "cases": pull_cases})
    log(f"bench: pull coupling {roof_pull}")

    ro = heb = None
    if args.couple == "fused" and not args.no_readout:
        para(False)
        rsnap = (snap[0], 0)
        ro = readout_bench(ctx, sim, net, chunk, rsnap, steps=2000)
        log(f"bench: readout {ro}")
        heb = hebbian_bench(ctx, sim, net, chunk, rsnap, steps=2000)
        log(f"bench: hebbian {heb}")

    roof_stencil = stencil_roofline(ctx)
    log(f"bench: stencil {roof_stencil}")

    cg = None
    if not args.no_cg:
        cg = cg_bench(ctx, net.grid[0], b_fixed)
        log(f"bench: CG {cg}")

    c5 = None
    if not args.no_c5:
        c5 = c5_bench(ctx, counts_fixed)
        log(f"bench: C5 {c5}")

    c4 = None
    if not args.no_c4:
        c4 = c4_bench(ctx)
        log(f"bench: C4 {c4}")

    c3 = None
    if not args.no_c3:
        c3 = c3_bench(ctx)
        log(f"bench: C3 {c3}")
        c3["precision_split"] = precision_split(ctx, net)
        log(f"bench: precision split {c3['precision_split']}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline_full(net, every, b_fixed)
        log(f"bench: cpu {cpu}")
    n_graphs = ctx.graph_cache()
    out = {"metric": METRIC, "value": value, "unit": "neuron-steps/s", "n_gpus": 1,
           "steps": steps, "warmup": warm, "steps_requested": args.steps, "warmup_requested": args.warmup,
           "ms_per_step": ms_step,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (Philox4x64-10, seed 0)",
           "config": {"workload": "c2: 1,048,576 squid HH neurons, RK4 dt 0.01 ms, ~210M-edge "
                                  "exp(-d/10) CSR, log-normal weights, event-driven deterministic "
                                  "coupling, release deposit on 256^3 with the paracrine Jacobi-CG solve to a "
                                  "true 1e-8 every 100 steps inside the step graphs; the configuration's "
                                  "10,000 steps from step 0",
                      "n": net.n, "nnz": net.nnz, "dt": net.dt, "couple": args.couple, "fused_shape": tune,
                      "graph_chunk": chunk, "spikes_timed": int(spikes), "spikes_per_step": spikes / steps,
                      "paracrine_every": every},
           "roofline": roof, "cpu_baseline": cpu,
           "steps_only": {"neuron_steps_per_s": steps_only, "ms": wall0_ms, "ms_per_step": wall0_ms / steps},
           "paracrine": paracrine, "cg": cg, "roofline_fused_step": fused_roof,
           "roofline_cg_iteration": cg_iter, "roofline_stencil": roof_stencil,
           "roofline_hh_unfused": roof_hh, "roofline_pull_coupling": roof_pull,
           "c3": c3, "c4": c4, "c5": c5, "readout": ro, "hebbian": heb, "models": models,
           "graphs": {"cached": n_graphs[0], "instantiated": n_graphs[1]},
           "event_ms": ev_ms}
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()
