"""The learned preconditioner of the paracrine solve: a handle on libnnx's
U-Net (nnx_unet_*), the harvesting of (residual, error) pairs from
Jacobi-CG solves on random variable-diffusivity fields, and the training
loop (forward, hand-written backward and Adam all run in libnnx).

No counterpart in the reference: its paracrine solve is unpreconditioned CG
(src/paracrine.cu:106-129).
"""
from __future__ import annotations

import ctypes as C
import json
import math
import os
import time

import numpy as np

from . import _lib, configs
from .devmem import DeviceArray as DA

SMOOTH_W = np.float32(2.0 / 3.0)  # damping of the Jacobi smoothings around the network
DEFAULT_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_c5.f32")
# weights of the network that corrects the multigrid V-cycle (PRECOND_MG_LEARNED)
MG_WEIGHTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "data", "unet_c5_mg.f32")


def param_count() -> int:
    return int(_lib.lib().nnx_unet_param_count())


class UNet:
    """coarse: two-grid mode (the network on the 2x coarser grid between
    block-mean restriction and trilinear prolongation)."""

    def __init__(self, ctx: _lib.Context, params: np.ndarray | None = None, seed: int = 0,
                 coarse: bool = False):
        self.ctx = ctx
        self.h = C.c_void_p()
        arr = None
        if params is not None:
            arr = np.ascontiguousarray(params, np.float32)
            assert arr.size == param_count(), (arr.size, param_count())
        ctx.call("nnx_unet_create", None if arr is None else arr.ctypes.data_as(C.c_void_p), seed,
                 C.byref(self.h))
        self.coarse = bool(coarse)
        ctx.call("nnx_unet_set_coarse", self.h, int(self.coarse))

    @classmethod
    def load(cls, ctx, path: str = DEFAULT_WEIGHTS) -> "UNet":
        """Parameters from `path`; the mode from the report beside it."""
        coarse = False
        meta = os.path.splitext(path)[0] + ".json"
        if os.path.exists(meta):
            with open(meta) as fh:
                coarse = bool(json.load(fh).get("coarse", False))
        return cls(ctx, np.fromfile(path, np.float32), coarse=coarse)

    def save(self, path: str):
        os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
        self.params().tofile(path)

    def params(self) -> np.ndarray:
        out = np.empty(param_count(), np.float32)
        self.ctx.call("nnx_unet_get_params", self.h, out.ctypes.data_as(C.c_void_p))
        return out

    def grad(self) -> np.ndarray:
        out = np.empty(param_count(), np.float32)
        self.ctx.call("nnx_unet_get_grad", self.h, out.ctypes.data_as(C.c_void_p))
        return out

    def apply(self, dims, batch, r: DA, feat: DA | None, z: DA):
        nx, ny, nz = dims
        self.ctx.call("nnx_unet_apply", self.h, nx, ny, nz, batch, r.ptr,
                      None if feat is None else feat.ptr, z.ptr)

    def train_step(self, dims, batch, r: DA, feat: DA | None, e: DA, lr: float, want_loss=True,
                   faces=None, smooth=0):
        """faces: (kx, ky, kz) device arrays of the batch's operators for the
        A-norm loss (smooth > 0: after that many Jacobi post-smoothing
        sweeps), or None for the plain relative MSE."""
        nx, ny, nz = dims
        loss = C.c_double()
        kx, ky, kz = (None, None, None) if faces is None else (f.ptr for f in faces)
        self.ctx.call("nnx_unet_train_step", self.h, nx, ny, nz, batch, r.ptr,
                      None if feat is None else feat.ptr, e.ptr, kx, ky, kz, int(smooth), lr,
                      C.byref(loss) if want_loss else None)
        return loss.value if want_loss else None

    def close(self):
        if self.h:
            self.ctx.call("nnx_unet_destroy", self.h)
            self.h = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def jacobi_inverse(f: configs.DiffusionField) -> np.ndarray:
    """1 / diagonal of the variable 7-point operator, flattened (x fastest)."""
    d = (1.0 + f.kx[:, :, :-1].astype(np.float64) + f.kx[:, :, 1:] + f.ky[:, :-1, :] + f.ky[:, 1:, :]
         + f.kz[:-1] + f.kz[1:])
    return (1.0 / d).astype(np.float32).ravel()


def device_field(ctx, f: configs.DiffusionField):
    """Grid struct and device arrays (kept alive by the returned dict)."""
    d = {"kx": DA.from_numpy(ctx, f.kx), "ky": DA.from_numpy(ctx, f.ky),
         "kz": DA.from_numpy(ctx, f.kz), "feat": DA.from_numpy(ctx, f.feat)}
    g = _lib.Grid(f.n, f.n, f.n, 7, 0.0, 0.0, d["kx"].ptr, d["ky"].ptr, d["kz"].ptr)
    return g, d


class PairSet:
    """(residual, error) pairs and the feature field of each pair, resident
    on the device: r[i], e[i] are n^3 fp32 grids, feature of pair i is
    feats[field_of[i]]."""

    def __init__(self, ctx, n, npairs, nfields):
        self.ctx, self.n, self.count = ctx, n, 0
        vox = n ** 3
        self.r = DA(ctx, (npairs, vox), np.float32)
        self.e = DA(ctx, (npairs, vox), np.float32)
        self.feats = DA(ctx, (nfields, vox), np.float32)
        nf = (n + 1) * n * n
        self.faces = [DA(ctx, (nfields, nf), np.float32) for _ in range(3)]
        self.face_bytes = nf * 4
        self.field_of = np.zeros(npairs, np.int64)
        self.bytes_per = vox * 4

    def _put(self, dst: DA, row: int, a: np.ndarray):
        a = np.ascontiguousarray(a, np.float32)
        self.ctx.call("nnx_memcpy_h2d", C.c_void_p(dst.ptr + row * self.bytes_per),
                      a.ctypes.data_as(C.c_void_p), C.c_size_t(self.bytes_per))

    def add_field(self, idx, feat, faces=None):
        self._put(self.feats, idx, feat)
        if faces is not None:
            for dst, a in zip(self.faces, faces):
                a = np.ascontiguousarray(a, np.float32)
                self.ctx.call("nnx_memcpy_h2d", C.c_void_p(dst.ptr + idx * self.face_bytes),
                              a.ctypes.data_as(C.c_void_p), C.c_size_t(self.face_bytes))

    def add_pair(self, r, e, field):
        i = self.count
        self._put(self.r, i, r)
        self._put(self.e, i, e)
        self.field_of[i] = field
        self.count += 1

    def gather(self, idx, r_out: DA, f_out: DA, e_out: DA, faces_out=None):
        for j, i in enumerate(idx):
            off = j * self.bytes_per
            for dst, src, row in ((r_out, self.r, i), (e_out, self.e, i),
                                  (f_out, self.feats, self.field_of[i])):
                self.ctx.call("nnx_memcpy_d2d", C.c_void_p(dst.ptr + off),
                              C.c_void_p(src.ptr + int(row) * self.bytes_per), C.c_size_t(self.bytes_per))
            if faces_out is not None:
                fo = int(self.field_of[i]) * self.face_bytes
                for dst, src in zip(faces_out, self.faces):
                    self.ctx.call("nnx_memcpy_d2d", C.c_void_p(dst.ptr + j * self.face_bytes),
                                  C.c_void_p(src.ptr + fo), C.c_size_t(self.face_bytes))


def _cg(ctx, g, b: DA, x: DA, rtol, maxit, use_x0=0, mg=False):
    """maxit iterations of Jacobi-CG (or of multigrid-preconditioned
    flexible CG) from x = 0: the iterate snapshots of the pair harvest."""
    o = _lib.CgOpts(rtol=rtol, maxit=maxit, precond=_lib.PRECOND_MG if mg else _lib.PRECOND_JACOBI,
                    check_every=8, use_x0=use_x0)
    r = _lib.CgResult()
    if mg:
        st = ctx._l.nnx_fcg_solve(ctx.h, C.byref(g), None, None, b.ptr, x.ptr, C.byref(o), C.byref(r))
    else:
        st = ctx._l.nnx_cg_solve(ctx.h, C.byref(g), b.ptr, x.ptr, C.byref(o), C.byref(r))
    if st not in (0, 4):  # 4: not converged within maxit (the snapshots)
        ctx.check(st)
    return r


def harvest(ctx, n=64, fields=160, sources=16, schedule=(0, 1, 2, 4, 8, 16, 32, 64), seed=1,
            log=print, base: str = "jacobi") -> PairSet:
    """Pairs (r_k, e_k) from Jacobi-CG iterates x_k (x_0 = 0) on `fields`
    fresh log-normal diffusivity fields (C5 statistics at n^3) and `sources`
    random point-source right-hand sides per field: e_k = x* - x_k with x*
    the FP64-refined solution (relative residual 1e-11), and r_k = A e_k
    (= b - A x_k up to the residual of x*), so that every pair satisfies
    its equation to fp32 rounding.  The network is the coarse correction of
    a two-level preconditioner after damped-Jacobi pre-smoothing
    (z0 = w D^-1 r, w = 2/3), so the stored pair is what the smoother
    leaves: e1 = e_k - w D^-1 r_k and r1 = A e1.
    base "mg": the iterates are those of multigrid-preconditioned flexible
    CG and the network corrects one V-cycle: e1 = e_k - V(r_k)."""
    ps = PairSet(ctx, n, fields * sources * len(schedule), fields)
    vox = n ** 3
    x64, xk, ek, rk, zk = (DA(ctx, vox, np.float64), DA(ctx, vox, np.float32), DA(ctx, vox, np.float32),
                           DA(ctx, vox, np.float32), DA(ctx, vox, np.float32))
    mg = base == "mg"
    t0 = time.time()
    for f in range(fields):
        fld = configs.c5_field(seed=seed * 100003 + f, n=n)
        g, keep = device_field(ctx, fld)
        ps.add_field(f, fld.feat, (fld.kx, fld.ky, fld.kz))
        dinv = jacobi_inverse(fld)
        for s in range(sources):
            cnt = int(20 + (s * 37 + f * 11) % 380)
            b = configs.point_sources(seed * 100003 + f * 1009 + s, n, cnt)
            db, db64 = DA.from_numpy(ctx, b), DA.from_numpy(ctx, b.astype(np.float64))
            o = _lib.CgOpts(rtol=1e-11, maxit=5000, precond=_lib.PRECOND_JACOBI, check_every=8)
            ctx.call("nnx_cg_solve_refine", C.byref(g), db64.ptr, x64.ptr, C.byref(o), 1e-6, 30,
                     C.byref(_lib.CgResult()))
            xstar = x64.numpy()
            for k in schedule:
                if k == 0:
                    e = xstar.astype(np.float32)
                else:
                    _cg(ctx, g, db, xk, 0.0, k, mg=mg)
                    e = (xstar - xk.numpy()).astype(np.float32)
                ek.upload(e)
                ctx.call("nnx_stencil_apply", C.byref(g), ek.ptr, rk.ptr)
                if mg:
                    ctx.call("nnx_mg_apply", C.byref(g), 0, _lib.MG_CHEBYSHEV, 0, rk.ptr, zk.ptr)
                    e1 = (e - zk.numpy()).astype(np.float32)
                else:
                    e1 = (e - SMOOTH_W * dinv * rk.numpy()).astype(np.float32)
                ek.upload(e1)
                ctx.call("nnx_stencil_apply", C.byref(g), ek.ptr, rk.ptr)
                ps.add_pair(rk.numpy(), e1, f)
        if (f + 1) % 20 == 0:
            log(f"harvest: {f + 1}/{fields} fields, {ps.count} pairs, {time.time() - t0:.1f}s")
        del keep
    return ps


def train(ctx, net: UNet, ps: PairSet, batch=8, epochs=3, lr=3e-4, lr_min=1e-5, warmup=500, seed=0,
          anorm=True, smooth=0, log=print):
    """Adam, linear warm-up then a cosine learning-rate schedule, over
    shuffled pairs; the A-norm loss by default (errors the operator
    amplifies - rough ones - weigh most, as they do for the solver)."""
    n, vox = ps.n, ps.n ** 3
    rb, fb, eb = (DA(ctx, (batch, vox), np.float32) for _ in range(3))
    fcs = [DA(ctx, (batch, (n + 1) * n * n), np.float32) for _ in range(3)] if anorm else None
    steps = epochs * (ps.count // batch)
    rng = np.random.default_rng(seed)
    losses, t0, step = [], time.time(), 0
    for ep in range(epochs):
        perm = rng.permutation(ps.count)
        for j in range(ps.count // batch):
            idx = perm[j * batch:(j + 1) * batch]
            ps.gather(idx, rb, fb, eb, fcs)
            cur = lr_min + 0.5 * (lr - lr_min) * (1 + math.cos(math.pi * step / max(steps, 1)))
            if step < warmup:
                cur = lr * (step + 1) / warmup
            want = (step % 50 == 0) or step == steps - 1
            loss = net.train_step((n, n, n), batch, rb, fb, eb, cur, want_loss=want, faces=fcs, smooth=smooth)
            if want:
                losses.append((step, loss))
                if step % 500 == 0:
                    log(f"train: epoch {ep} step {step}/{steps} lr {cur:.2e} loss {loss:.4f} "
                        f"{time.time() - t0:.1f}s")
            step += 1
    return losses
