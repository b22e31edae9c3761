"""Trains the U-Net preconditioner of the variable-diffusivity paracrine
solve and writes its parameters (raw float32, nnx_unet parameter order).

  python -m neuronephys_clean_amd.train_unet --out path [--fields F --sources S --epochs E]

Pairs (r_k, e_k) come from Jacobi-CG iterates on fresh 64^3 log-normal
diffusivity fields (log-sigma 1, correlation length 8 voxels, mean
dt D / h^2 = 50) with random point-source right-hand sides; forward,
backward and Adam run in libnnx.  Afterwards flexible CG with the network
is compared with Jacobi-CG on held-out problems (FP64-refined to a 1e-8
relative residual).
"""
from __future__ import annotations

import argparse
import ctypes as C
import json
import os
import sys
import time

import numpy as np

from . import _lib, configs
from .devmem import DeviceArray as DA
from .precond import UNet, device_field, harvest, train


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def compare(ctx, net, n, seed, sources=400, rtol=1e-8, smooth=1, base="jacobi"):
    """Jacobi-CG vs learned FCG (and, for base "mg", multigrid FCG without
    the network), all inside FP64 refinement, on one field."""
    f = configs.c5_field(seed=seed, n=n)
    g, keep = device_field(ctx, f)
    b = configs.point_sources(seed + 7, n, sources).astype(np.float64)
    db, dx = DA.from_numpy(ctx, b), DA.zeros(ctx, n ** 3, np.float64)
    out = {}
    solvers = [("jacobi_cg", None), ("learned_fcg", _lib.PRECOND_LEARNED)]
    if base == "mg":
        solvers = [("jacobi_cg", None), ("mg_fcg", _lib.PRECOND_MG), ("learned_fcg", _lib.PRECOND_MG_LEARNED)]
    for name, pc in solvers:
        best = None
        for _ in range(2):
            r = _lib.CgResult()
            o = _lib.CgOpts(rtol=rtol, maxit=20000, precond=_lib.PRECOND_JACOBI if pc is None else pc,
                            check_every=8, stall_iters=10, smooth=smooth)
            if pc is None:
                ctx.call("nnx_cg_solve_refine", C.byref(g), db.ptr, dx.ptr, C.byref(o), 1e-6, 30, C.byref(r))
            else:
                ctx.call("nnx_fcg_solve_refine", C.byref(g), net.h, keep["feat"].ptr, db.ptr, dx.ptr,
                         C.byref(o), 1e-6, 30, C.byref(r))
            if best is None or r.ms < best["ms"]:
                best = r.as_dict()
        out[name] = best
    out["speedup"] = out["jacobi_cg"]["ms"] / out["learned_fcg"]["ms"]
    if base == "mg":
        out["speedup_vs_mg"] = out["mg_fcg"]["ms"] / out["learned_fcg"]["ms"]
    # cost of one network application at this size
    r32, z32 = DA.from_numpy(ctx, b.astype(np.float32)), DA.zeros(ctx, n ** 3, np.float32)
    e0, e1 = C.c_void_p(), C.c_void_p()
    ctx.call("nnx_event_create", C.byref(e0))
    ctx.call("nnx_event_create", C.byref(e1))
    net.apply((n, n, n), 1, r32, keep["feat"], z32)
    ctx.call("nnx_event_record", e0)
    for _ in range(10):
        net.apply((n, n, n), 1, r32, keep["feat"], z32)
    ctx.call("nnx_event_record", e1)
    ms = C.c_float()
    ctx.call("nnx_event_elapsed", e0, e1, C.byref(ms))
    out["unet_apply_ms"] = ms.value / 10
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--out", required=True)
    ap.add_argument("--fields", type=int, default=160)
    ap.add_argument("--sources", type=int, default=16)
    ap.add_argument("--epochs", type=int, default=3)
    ap.add_argument("--batch", type=int, default=8)
    ap.add_argument("--lr", type=float, default=1.5e-4)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--init", default=None, help="start from these parameters")
    ap.add_argument("--eval-n", type=int, nargs="*", default=[64, 256])
    ap.add_argument("--n", type=int, default=64, help="training grid size")
    ap.add_argument("--coarse", action="store_true", help="two-grid mode")
    ap.add_argument("--smooth", type=int, default=1, help="post-smoothing sweeps in the evaluation")
    ap.add_argument("--loss", choices=["mse", "anorm", "smoothed"], default="smoothed",
                    help="smoothed: A-norm of the error left after --smooth post-smoothing sweeps")
    ap.add_argument("--base", choices=["jacobi", "mg"], default="jacobi",
                    help="what the network corrects: damped-Jacobi smoothing or a multigrid V-cycle")
    args = ap.parse_args()
    ctx = _lib.context(0)
    t0 = time.time()
    sched = (0, 1, 2, 3, 4, 6, 8, 12) if args.base == "mg" else (0, 1, 2, 4, 8, 16, 32, 64)
    ps = harvest(ctx, n=args.n, fields=args.fields, sources=args.sources, seed=args.seed, log=log,
                 base=args.base, schedule=sched)
    ctx.synchronize()
    t_h = time.time() - t0
    log(f"harvested {ps.count} pairs in {t_h:.1f}s")
    net = UNet(ctx, np.fromfile(args.init, np.float32) if args.init else None, seed=args.seed,
               coarse=args.coarse)
    t0 = time.time()
    losses = train(ctx, net, ps, batch=args.batch, epochs=args.epochs, lr=args.lr, anorm=args.loss != "mse",
                   smooth=args.smooth if args.loss == "smoothed" else 0, log=log)
    ctx.synchronize()
    t_t = time.time() - t0
    steps = args.epochs * (ps.count // args.batch)
    log(f"trained {steps} steps in {t_t:.1f}s ({1e3 * t_t / max(steps, 1):.2f} ms/step), "
        f"loss {losses[0][1]:.4f} -> {losses[-1][1]:.4f}")
    net.save(args.out)
    del ps
    report = {"pairs": args.fields * args.sources * 8, "n": args.n, "coarse": args.coarse, "base": args.base,
              "smooth": args.smooth, "loss": args.loss, "harvest_s": t_h, "train_s": t_t,
              "steps": steps, "batch": args.batch, "epochs": args.epochs,
              "loss_first": losses[0][1], "loss_last": losses[-1][1], "losses": losses, "eval": {}}
    for n in args.eval_n:
        report["eval"][str(n)] = compare(ctx, net, n, seed=424242 + n, smooth=args.smooth, base=args.base)
        log(f"eval {n}^3: {report['eval'][str(n)]}")
    with open(os.path.splitext(args.out)[0] + ".json", "w") as fh:
        json.dump(report, fh, indent=1)


if __name__ == "__main__":
    main()
