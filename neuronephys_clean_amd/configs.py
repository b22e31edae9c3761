"""Synthetic inputs of the benchmark configurations (host side, numpy).

Every draw comes from a Philox4x64-10 stream keyed by (seed, purpose), so
the same seed gives the same network and initial state on any machine and
thread count.  Network construction stays on the host, as in the reference
(Network::Generate, src/network_factory.cuh:23-38).

  c1 : 1,024 squid HH neurons, Erdos-Renyi with exactly 100 distinct
       out-edges per neuron (self-loops excluded), w ~ U[0, 0.1], positions
       U[0,32)^3 on a 32^3 grid with h = 1, dt*D/h^2 = 1.
  c2 : 1,048,576 neurons in a 128^3 box, out-degree LogNormal(ln 200 - 1/8,
       1/2) (mean 200), targets drawn without replacement with probability
       proportional to exp(-d/10) up to d = 60, w ~ LogNormal(-3, 1), a 256^3
       grid (h = 0.5).  `scale` shrinks N and the box at constant density.
  c4 : ensemble members: N ~ round(2^U(10,16)), out-degree Poisson(50),
       I_ext ~ N(mu_i, 2) with mu_i ~ U(6, 14).
Initial state: V ~ U(-70, -60) mV, gates at their steady state for V,
s = 0; I_ext ~ N(10, 2) uA/cm^2; dt = 0.01 ms.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field

import numpy as np

from . import host

# Philox purposes (one stream per kind of draw)
P_V0, P_IEXT, P_W, P_POS, P_MU, P_NSIZE, P_GRIDSRC = 101, 102, 103, 104, 105, 106, 107
P_C3SRC = 108
P_DFIELD, P_PCSRC = 109, 110

SQUID_VSHIFT = 5.0


@dataclass
class Network:
    n: int
    off: np.ndarray
    tgt: np.ndarray
    w: np.ndarray
    V0: np.ndarray
    m0: np.ndarray
    h0: np.ndarray
    n0: np.ndarray
    iext: np.ndarray
    pos: np.ndarray
    grid: tuple = (32, 32, 32)
    h_grid: float = 1.0
    c: float = 1.0           # dt*D/h^2 of the implicit paracrine step
    amount: float = 0.01     # release per spike
    dt: float = 0.01
    name: str = ""
    csc: tuple = field(default=None, repr=False)   # in_off, in_src, in_w, out_to_in, in_degree
    voxel: np.ndarray = field(default=None, repr=False)

    @property
    def nnz(self) -> int:
        return int(self.off[-1])

    def ensure_csc(self):
        if self.csc is None:
            self.csc = host.csr_to_csc(self.n, self.off, self.tgt, self.w)
        return self.csc

    def ensure_voxel(self):
        if self.voxel is None:
            self.voxel = host.voxel_of(self.pos, self.h_grid, *self.grid)
        return self.voxel


def initial_state(seed, n, v_shift=SQUID_VSHIFT, iext_mu=10.0, iext_sigma=2.0):
    V0 = host.uniform(seed, P_V0, n, -70.0, -60.0)
    m0, h0, n0 = host.steady_state(V0, v_shift)
    iext = host.normal(seed, P_IEXT, n, iext_mu, iext_sigma)
    return V0, m0, h0, n0, iext


def c1(seed: int = 0, n: int = 1024, k: int = 100) -> Network:
    off, tgt = host.er_fixed(seed, n, k, allow_self=False)
    w = host.uniform(seed, P_W, len(tgt), 0.0, 0.1).astype(np.float32)
    V0, m0, h0, n0, iext = initial_state(seed, n)
    pos = host.uniform(seed, P_POS, 3 * n, 0.0, 32.0).astype(np.float32).reshape(n, 3)
    return Network(n, off, tgt, w, V0, m0, h0, n0, iext, pos, grid=(32, 32, 32), h_grid=1.0,
                   c=1.0, name=f"c1_n{n}_k{k}")


def c2(seed: int = 0, n: int = 1 << 20, lam: float = 10.0, cutoff: float = 60.0,
       grid: int | None = None) -> Network:
    scale = (n / float(1 << 20)) ** (1.0 / 3.0)
    box = 128.0 * scale
    pos = host.uniform(seed, P_POS, 3 * n, 0.0, box).astype(np.float32).reshape(n, 3)
    deg = host.outdeg_lognormal(seed, n, math.log(200.0) - 0.125, 0.5, min(n - 1, 20000))
    off = host.offsets_of(deg)
    tgt = host.spatial_targets(seed, pos, box, lam, cutoff, off)
    w = host.lognormal(seed, P_W, len(tgt), -3.0, 1.0).astype(np.float32)
    V0, m0, h0, n0, iext = initial_state(seed, n)
    g = grid if grid is not None else max(8, int(round(256 * scale)))
    return Network(n, off, tgt, w, V0, m0, h0, n0, iext, pos, grid=(g, g, g), h_grid=box / g,
                   c=1.0, name=f"c2_n{n}")


def c3_sources(seed: int = 0, n: int = 512, count: int = 50_000) -> np.ndarray:
    """The precision-split solve's right-hand side: `count` unit point
    releases at uniformly random voxels of an n^3 grid (coincident releases
    add), FP64, x fastest."""
    u = host.uniform(seed, P_C3SRC, count, 0.0, 1.0)
    idx = np.minimum((u * float(n) ** 3).astype(np.int64), n ** 3 - 1)
    return np.bincount(idx, minlength=n ** 3).astype(np.float64)


C3_GRID, C3_C = 512, 20.0


@dataclass
class DiffusionField:
    """A variable-diffusivity backward-Euler operator (I - dt div D grad) on
    an n^3 grid: per-voxel c = dt D / h^2, face coefficients and the
    network's feature channel (normalised log c)."""
    n: int
    c: np.ndarray
    kx: np.ndarray
    ky: np.ndarray
    kz: np.ndarray
    feat: np.ndarray


def c5_field(seed: int = 0, n: int = 256, corr: float = 8.0, sigma: float = 1.0,
             cbar: float = 50.0, purpose: int = P_DFIELD) -> DiffusionField:
    """log D = sigma * (unit Gaussian random field, correlation length corr
    voxels); scaled so that the mean of dt D / h^2 is cbar; faces by
    harmonic mean."""
    g = host.grf(seed, purpose, n, n, n, corr)
    d = np.exp(sigma * g.astype(np.float64))
    c = (cbar * d / d.mean()).astype(np.float32)
    kx, ky, kz = host.face_coefficients(c)
    lc = np.log(c.astype(np.float64))
    feat = ((lc - lc.mean()) / max(lc.std(), 1e-12)).astype(np.float32)
    return DiffusionField(n, c, kx, ky, kz, feat)


def point_sources(seed: int, n: int, count: int, purpose: int = P_PCSRC) -> np.ndarray:
    """`count` releases of random amplitude U(0.5, 1.5) at uniformly random
    voxels of an n^3 grid (float32, x fastest)."""
    u = host.uniform(seed, purpose, 2 * count, 0.0, 1.0)
    idx = np.minimum((u[:count] * float(n) ** 3).astype(np.int64), n ** 3 - 1)
    b = np.zeros(n ** 3, np.float64)
    np.add.at(b, idx, 0.5 + u[count:])
    return b.astype(np.float32)


def c4_sizes(seed: int = 0, count: int = 256):
    u = host.uniform(seed, P_NSIZE, count, 10.0, 16.0)
    return np.rint(2.0 ** u).astype(np.int64)


def c4_member(seed: int, index: int, n: int, lam_deg: float = 50.0) -> Network:
    s = (seed << 16) + index + 1
    deg = np.minimum(host.outdeg_poisson(s, n, lam_deg), n - 1)
    off = host.offsets_of(deg)
    tgt = host.uniform_targets(s, off, allow_self=False)
    w = host.uniform(s, P_W, len(tgt), 0.0, 0.1).astype(np.float32)
    mu = host.uniform(seed, P_MU, index + 1, 6.0, 14.0)[index]
    V0, m0, h0, n0, iext = initial_state(s, n, iext_mu=mu)
    pos = host.uniform(s, P_POS, 3 * n, 0.0, 64.0).astype(np.float32).reshape(n, 3)
    return Network(n, off, tgt.astype(np.int32), w, V0, m0, h0, n0, iext, pos,
                   grid=(64, 64, 64), h_grid=1.0, c=1.0, name=f"c4_{index}_n{n}")


@dataclass
class Ensemble:
    """Independent networks in one index space: network i owns neurons
    [begin[i], begin[i] + size[i]), begin[i] a multiple of 64; arrays are
    over the padded total n (a multiple of 64)."""
    n: int
    begin: np.ndarray
    size: np.ndarray
    off: np.ndarray
    tgt: np.ndarray
    w: np.ndarray
    V0: np.ndarray
    m0: np.ndarray
    h0: np.ndarray
    n0: np.ndarray
    iext: np.ndarray
    voxel_of: np.ndarray
    dt: float = 0.01
    members: list = field(default=None, repr=False)
    lattice: tuple = (1, 1, 1)   # networks' 64^3 grids per x, y, z of the release box
    c: float = 1.0               # dt D / h^2 of every network's grid

    @property
    def nnz(self) -> int:
        return int(self.off[-1])

    @property
    def grid(self) -> tuple:
        """The release box: the networks' 64^3 grids side by side."""
        return tuple(64 * k for k in self.lattice)

    def block_of(self, i: int) -> tuple:
        """Voxel offsets (x, y, z) of network i's grid in the box."""
        a, b, _ = self.lattice
        return 64 * (i % a), 64 * ((i // a) % b), 64 * (i // (a * b))


def ensemble_lattice(count: int) -> tuple:
    """Networks' grids per x, y, z: up to 4 along x and y, the rest along z
    (256 networks: 4 x 4 x 16, a 256 x 256 x 1024 box)."""
    a = max(1, min(4, count))
    b = max(1, min(4, -(-count // a)))
    return a, b, max(1, -(-count // (a * b)))


def c4_ensemble(seed: int = 0, count: int = 256, sizes=None, keep_members: bool = False) -> Ensemble:
    """The batched-ensemble configuration: `count` networks with
    log-uniform sizes 2^U(10,16), Poisson(50) out-degree, uniform targets,
    w ~ U[0, 0.1], I_ext ~ N(mu_i, 2) with mu_i ~ U(6, 14), each with its own
    64^3 release grid; the grids lie side by side in one box
    (ensemble_lattice), network i's at block_of(i), so that one solve of the
    box with Dirichlet faces every 64 voxels (nnx_grid.block) solves every
    network's grid."""
    sizes = c4_sizes(seed, count) if sizes is None else np.asarray(sizes, np.int64)
    begin = np.zeros(len(sizes), np.int64)
    pos = 0
    for i, s in enumerate(sizes):
        begin[i] = pos
        pos = (pos + int(s) + 63) // 64 * 64
    n = max(pos, 64)
    deg = np.zeros(n, np.int64)
    parts_t, parts_w, members = [], [], []
    V0 = np.full(n, -65.0)
    m0, h0, n0, iext = np.zeros(n), np.zeros(n), np.zeros(n), np.zeros(n)
    vox = np.zeros(n, np.int32)
    lat = ensemble_lattice(len(sizes))
    nx, ny = 64 * lat[0], 64 * lat[1]
    for i, s in enumerate(sizes):
        net = c4_member(seed, i, int(s))
        b = int(begin[i])
        deg[b:b + s] = np.diff(net.off)
        parts_t.append(net.tgt.astype(np.int32) + np.int32(b))
        parts_w.append(net.w)
        V0[b:b + s], m0[b:b + s], h0[b:b + s], n0[b:b + s], iext[b:b + s] = (
            net.V0, net.m0, net.h0, net.n0, net.iext)
        local = np.minimum(np.floor(net.pos).astype(np.int64), 63)
        ox, oy, oz = 64 * (i % lat[0]), 64 * ((i // lat[0]) % lat[1]), 64 * (i // (lat[0] * lat[1]))
        vox[b:b + s] = (ox + local[:, 0] + nx * (oy + local[:, 1] + ny * (oz + local[:, 2]))).astype(np.int32)
        if keep_members:
            members.append(net)
    off = np.zeros(n + 1, np.int32)
    np.cumsum(deg, out=off[1:])
    return Ensemble(n, begin, sizes.copy(), off, np.concatenate(parts_t) if parts_t else np.zeros(0, np.int32),
                    np.concatenate(parts_w) if parts_w else np.zeros(0, np.float32), V0, m0, h0, n0, iext, vox,
                    members=members if keep_members else None, lattice=lat)
