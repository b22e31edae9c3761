"""numpy restatement of the multigrid V-cycle for the variable-coefficient
7-point backward-Euler operator (the checker of nnx_mg_apply), written
from the definitions: arrays are indexed [k, j, i] (x fastest in memory),
face arrays kx [nz, ny, nx+1], ky [nz, ny+1, nx], kz [nz+1, ny, nx]."""
import numpy as np


def apply(x, kx, ky, kz):
    """(A x)_v = x_v + sum over the 6 faces of k_f (x_v - x_nbr), x = 0
    beyond the boundary."""
    p = np.pad(x, 1)
    c = p[1:-1, 1:-1, 1:-1]
    d = 1.0 + kx[:, :, :-1] + kx[:, :, 1:] + ky[:, :-1, :] + ky[:, 1:, :] + kz[:-1] + kz[1:]
    s = (kx[:, :, :-1] * p[1:-1, 1:-1, :-2] + kx[:, :, 1:] * p[1:-1, 1:-1, 2:]
         + ky[:, :-1, :] * p[1:-1, :-2, 1:-1] + ky[:, 1:, :] * p[1:-1, 2:, 1:-1]
         + kz[:-1] * p[:-2, 1:-1, 1:-1] + kz[1:] * p[2:, 1:-1, 1:-1])
    return d * c - s


def diag(kx, ky, kz):
    return 1.0 + kx[:, :, :-1] + kx[:, :, 1:] + ky[:, :-1, :] + ky[:, 1:, :] + kz[:-1] + kz[1:]


def coarsen(kx, ky, kz):
    """Coarse face = 1/16 of the sum of the 4 fine faces it covers."""
    cx = (kx[0::2, 0::2, 0::2] + kx[0::2, 1::2, 0::2] + kx[1::2, 0::2, 0::2] + kx[1::2, 1::2, 0::2]) / 16
    cy = (ky[0::2, 0::2, 0::2] + ky[0::2, 0::2, 1::2] + ky[1::2, 0::2, 0::2] + ky[1::2, 0::2, 1::2]) / 16
    cz = (kz[0::2, 0::2, 0::2] + kz[0::2, 0::2, 1::2] + kz[0::2, 1::2, 0::2] + kz[0::2, 1::2, 1::2]) / 16
    return cx, cy, cz


def restrict(r):
    """Block mean over 2 x 2 x 2 children."""
    nz, ny, nx = r.shape
    return r.reshape(nz // 2, 2, ny // 2, 2, nx // 2, 2).mean(axis=(1, 3, 5))


def _prolong_axis(e, axis):
    """Cell-centred linear interpolation by 2 along one axis, zero beyond
    the ends: fine 2I gets 3/4 e_I + 1/4 e_{I-1}, fine 2I+1 gets
    3/4 e_I + 1/4 e_{I+1}."""
    e = np.moveaxis(e, axis, 0)
    z = np.zeros_like(e[:1])
    lo = np.concatenate([z, e[:-1]])
    hi = np.concatenate([e[1:], z])
    out = np.empty((2 * e.shape[0],) + e.shape[1:], e.dtype)
    out[0::2] = 0.75 * e + 0.25 * lo
    out[1::2] = 0.75 * e + 0.25 * hi
    return np.moveaxis(out, 0, axis)


def prolong(e):
    return _prolong_axis(_prolong_axis(_prolong_axis(e, 2), 1), 0)


def chebyshev_coefficients(m, lmax=2.0, lmin=2.0 / 16):
    """Three-term recurrence of the Chebyshev polynomial on [lmin, lmax]:
    sweep s sets d = c1 d + c2 D^-1 r and x = x + d."""
    th, de = (lmax + lmin) / 2, (lmax - lmin) / 2
    sig = th / de
    rho = 1 / sig
    out = [(0.0, 1 / th)]
    for _ in range(1, m):
        r1 = 1 / (2 * sig - rho)
        out.append((r1 * rho, 2 * r1 / de))
        rho = r1
    return out


def _smooth(x, r, kx, ky, kz, dinv, nu, w, smoother):
    """nu sweeps from x (None: from zero)."""
    if smoother == "jacobi":
        if x is None:
            x = w * dinv * r
            nu -= 1
        for _ in range(nu):
            x = x + w * dinv * (r - apply(x, kx, ky, kz))
        return x
    d = None
    for c1, c2 in chebyshev_coefficients(nu):
        res = r if x is None else r - apply(x, kx, ky, kz)
        d = c2 * dinv * res if d is None else c1 * d + c2 * dinv * res
        x = d if x is None else x + d
    return x


def vcycle(r, kx, ky, kz, nu=2, w=0.8, coarse_sweeps=16, smoother="jacobi", gamma=1, v_cells=4096):
    """One cycle from zero: nu sweeps of the smoother (damped Jacobi or
    Chebyshev) before and after each level, coarsening while every
    dimension is even and >= 8, coarse_sweeps damped-Jacobi sweeps on the
    coarsest level; each coarser level with more than v_cells cells visited
    gamma times (1: V-cycle, 2: W-cycle, the later visits on the residual
    the earlier ones leave), smaller ones once."""
    nz, ny, nx = r.shape
    dinv = 1.0 / diag(kx, ky, kz)
    can = nx % 2 == 0 and ny % 2 == 0 and nz % 2 == 0 and min(nx, ny, nz) >= 8
    if not can:
        return _smooth(None, r, kx, ky, kz, dinv, coarse_sweeps, w, "jacobi")
    x = _smooth(None, r, kx, ky, kz, dinv, nu, w, smoother)
    rc = restrict(r - apply(x, kx, ky, kz))
    kc = coarsen(kx, ky, kz)
    kw = dict(nu=nu, w=w, coarse_sweeps=coarse_sweeps, smoother=smoother, gamma=gamma, v_cells=v_cells)
    ec = vcycle(rc, *kc, **kw)
    for _ in range(gamma - 1 if rc.size > v_cells else 0):
        ec = ec + vcycle(rc - apply(ec, *kc), *kc, **kw)
    x = x + prolong(ec)
    return _smooth(x, r, kx, ky, kz, dinv, nu, w, smoother)
