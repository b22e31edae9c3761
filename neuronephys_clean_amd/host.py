"""numpy wrappers of the host builders (include/nnx_host.h, libnnx_host.so)."""
from __future__ import annotations

import ctypes as C

import numpy as np

from ._lib import host


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def philox(ctr, key) -> np.ndarray:
    c = (C.c_uint64 * 4)(*[int(v) for v in ctr])
    k = (C.c_uint64 * 2)(*[int(v) for v in key])
    o = (C.c_uint64 * 4)()
    host().nnx_philox4x64_10(c, k, o)
    return np.array(list(o), dtype=np.uint64)


def uniform(seed, purpose, n, lo=0.0, hi=1.0) -> np.ndarray:
    out = np.empty(int(n), np.float64)
    host().nnx_gen_uniform(seed, purpose, int(n), lo, hi, _p(out))
    return out


def normal(seed, purpose, n, mu=0.0, sigma=1.0) -> np.ndarray:
    out = np.empty(int(n), np.float64)
    host().nnx_gen_normal(seed, purpose, int(n), mu, sigma, _p(out))
    return out


def lognormal(seed, purpose, n, mu=0.0, sigma=1.0) -> np.ndarray:
    out = np.empty(int(n), np.float64)
    host().nnx_gen_lognormal(seed, purpose, int(n), mu, sigma, _p(out))
    return out


def er_fixed(seed, n, k, allow_self=False):
    off = np.empty(n + 1, np.int32)
    tgt = np.empty(n * k, np.int32)
    r = host().nnx_gen_er_fixed(seed, n, k, int(allow_self), _p(off), _p(tgt))
    if r != 0:
        raise ValueError(f"out-degree {k} exceeds the candidates of {n} neurons")
    return off, tgt


def uniform_targets(seed, off, allow_self=False):
    n = len(off) - 1
    tgt = np.empty(int(off[-1]), np.int32)
    if host().nnx_gen_uniform_targets(seed, n, _p(off), int(allow_self), _p(tgt)) != 0:
        raise ValueError("an out-degree exceeds the number of candidates")
    return tgt


def outdeg_lognormal(seed, n, mu, sigma, cap):
    d = np.empty(n, np.int32)
    host().nnx_gen_outdeg_lognormal(seed, n, mu, sigma, cap, _p(d))
    return d


def outdeg_poisson(seed, n, lam):
    d = np.empty(n, np.int32)
    host().nnx_gen_outdeg_poisson(seed, n, lam, _p(d))
    return d


def offsets_of(deg) -> np.ndarray:
    off = np.zeros(len(deg) + 1, np.int64)
    np.cumsum(deg, out=off[1:])
    if off[-1] >= 2 ** 31:
        raise ValueError("more than 2^31 - 1 edges")
    return off.astype(np.int32)


def spatial_targets(seed, pos, box, lam, cutoff, off):
    pos = np.ascontiguousarray(pos, np.float32)
    n = pos.shape[0]
    tgt = np.empty(int(off[-1]), np.int32)
    r = host().nnx_gen_spatial_targets(seed, n, _p(pos), box, lam, cutoff, _p(off), _p(tgt))
    if r != 0:
        raise ValueError(f"{-r} sources had fewer candidates than their out-degree")
    return tgt


def csr_to_csc(n, off, tgt, w=None):
    nnz = len(tgt)
    in_off = np.empty(n + 1, np.int32)
    in_src = np.empty(nnz, np.int32)
    in_w = np.empty(nnz, np.float32)
    o2i = np.empty(nnz, np.int32)
    deg = np.empty(n, np.float32)
    w = None if w is None else np.ascontiguousarray(w, np.float32)
    host().nnx_csr_to_csc(n, nnz, _p(off), _p(tgt), _p(w), _p(in_off), _p(in_src), _p(in_w),
                          _p(o2i), _p(deg))
    return in_off, in_src, in_w, o2i, deg


def steady_state(V, v_shift):
    V = np.ascontiguousarray(V, np.float64)
    m, h, ng = (np.empty_like(V) for _ in range(3))
    host().nnx_hh_steady_state(len(V), _p(V), v_shift, _p(m), _p(h), _p(ng))
    return m, h, ng


def voxel_of(pos, h, nx, ny, nz):
    pos = np.ascontiguousarray(pos, np.float32)
    out = np.empty(pos.shape[0], np.int32)
    host().nnx_voxel_of(pos.shape[0], _p(pos), h, nx, ny, nz, _p(out))
    return out


def trilinear_matrix(x, y, z, nx, ny, nz, h):
    x, y, z = (np.ascontiguousarray(a, np.float32) for a in (x, y, z))
    n = len(x)
    off = np.empty(n + 1, np.int32)
    cols = np.empty(8 * n, np.int32)
    vals = np.empty(8 * n, np.float32)
    outside = host().nnx_trilinear_matrix(n, _p(x), _p(y), _p(z), nx, ny, nz, h, _p(off),
                                          _p(cols), _p(vals))
    return off, cols, vals, int(outside)


def csr_transpose(rows, cols, off, idx, val):
    nnz = int(off[-1])
    t_off = np.empty(cols + 1, np.int32)
    t_idx = np.empty(nnz, np.int32)
    t_val = np.empty(nnz, np.float32)
    host().nnx_csr_transpose(rows, cols, _p(off), _p(idx), _p(np.ascontiguousarray(val, np.float32)),
                             _p(t_off), _p(t_idx), _p(t_val))
    return t_off, t_idx, t_val


def grf(seed, purpose, nx, ny, nz, corr) -> np.ndarray:
    """Unit-variance Gaussian random field, covariance exp(-r^2/(2 corr^2)),
    shape (nz, ny, nx) (x fastest)."""
    out = np.empty((nz, ny, nx), np.float32)
    host().nnx_gen_grf(seed, purpose, nx, ny, nz, corr, _p(out))
    return out


def face_coefficients(c):
    """kx, ky, kz (nnx_grid layout) from per-voxel coefficients c (nz, ny, nx)."""
    c = np.ascontiguousarray(c, np.float32)
    nz, ny, nx = c.shape
    kx = np.empty((nz, ny, nx + 1), np.float32)
    ky = np.empty((nz, ny + 1, nx), np.float32)
    kz = np.empty((nz + 1, ny, nx), np.float32)
    host().nnx_face_coefficients(nx, ny, nz, _p(c), _p(kx), _p(ky), _p(kz))
    return kx, ky, kz


def spherical_regions(pos, k, center_idx=None, center_pos=None):
    """Electrode regions: the k nearest neurons of each center (neuron indices
    or points), nearest first, ties by lower index.  Returns (offsets, idx)."""
    pos = np.ascontiguousarray(pos, np.float32).reshape(-1, 3)
    ci = None if center_idx is None else np.ascontiguousarray(center_idx, np.int32)
    cp = None if center_pos is None else np.ascontiguousarray(center_pos, np.float32).reshape(-1, 3)
    nreg = len(ci) if ci is not None else len(cp)
    off = np.empty(nreg + 1, np.int32)
    idx = np.empty(nreg * int(k), np.int32)
    rc = host().nnx_spherical_regions(pos.shape[0], _p(pos), nreg, _p(ci), _p(cp), int(k), _p(off), _p(idx))
    if rc != 0:
        raise ValueError(f"nnx_spherical_regions: bad arguments (k={k}, n={pos.shape[0]})")
    return off, idx


H5_TYPES = {np.dtype(np.int32): 0, np.dtype(np.float32): 1, np.dtype(np.float64): 2, np.dtype(np.int64): 3}


def h5_read(path, name, dtype=np.float32) -> np.ndarray:
    """A whole HDF5 dataset, flattened row-major, converted to dtype (int32,
    int64, float32 or float64)."""
    err = C.create_string_buffer(512)
    n = host().nnx_h5_count(str(path).encode(), name.encode(), err, 512)
    if n < 0:
        raise IOError(err.value.decode())
    out = np.empty(max(n, 1), dtype)
    k = host().nnx_h5_read(str(path).encode(), name.encode(), H5_TYPES[np.dtype(dtype)], _p(out), n, err, 512)
    if k < 0:
        raise IOError(err.value.decode())
    return out[:k]


def csr_subsample(row_ptr, col_idx, N):
    """FileNetwork::GenerateFromHDF's sampling of N neurons: (offsets, targets, dropped)."""
    row_ptr = np.ascontiguousarray(row_ptr, np.int32)
    col_idx = np.ascontiguousarray(col_idx, np.int32)
    off = np.empty(N + 1, np.int32)
    tgt = np.empty(max(len(col_idx), 1), np.int32)
    dropped = C.c_int64()
    k = host().nnx_csr_subsample(len(row_ptr) - 1, _p(row_ptr), _p(col_idx), N, _p(off), _p(tgt),
                                 C.byref(dropped))
    if k < 0:
        raise ValueError(f"cannot sample {N} neurons from {len(row_ptr) - 1}")
    return off, tgt[:k].copy(), dropped.value


def morphology_from_axes(x, y, z, N):
    """FileMorphology::GenerateFromHDF's placement: (N, 3) positions in [-1, 1]."""
    ax = [np.ascontiguousarray(a, np.float32) for a in (x, y, z)]
    pos = np.empty((N, 3), np.float32)
    if host().nnx_morphology_from_axes(len(ax[0]), *(_p(a) for a in ax), N, _p(pos)) != 0:
        raise ValueError(f"cannot place {N} neurons from {len(ax[0])} points")
    return pos


def read_connection_file(path):
    """Csr_matrix::input_from_connection_file: (row_ptr, col, val)."""
    err = C.create_string_buffer(512)
    rows, nnz = C.c_int64(), C.c_int64()
    if host().nnx_read_connection_file(str(path).encode(), C.byref(rows), C.byref(nnz), None, None, None, err,
                                       512) < 0:
        raise IOError(err.value.decode())
    rp = np.empty(rows.value + 1, np.int32)
    col = np.empty(max(nnz.value, 1), np.int32)
    val = np.empty(max(nnz.value, 1), np.float32)
    if host().nnx_read_connection_file(str(path).encode(), C.byref(rows), C.byref(nnz), _p(rp), _p(col), _p(val),
                                       err, 512) < 0:
        raise IOError(err.value.decode())
    return rp, col[:nnz.value], val[:nnz.value]


def load_hdf5_network(path, N):
    """The reference's FileNetwork on an HDF5 file with /row_ptr and /col_idx."""
    return csr_subsample(h5_read(path, "/row_ptr", np.int32), h5_read(path, "/col_idx", np.int32), N)


def load_hdf5_morphology(path, N, regions=False):
    """The reference's FileMorphology: positions (N, 3) and, with regions,
    the region id of each neuron (float ids truncated to int, as there)."""
    x, y, z = (h5_read(path, a, np.float32) for a in ("/x", "/y", "/z"))
    pos = morphology_from_axes(x, y, z, N)
    if not regions:
        return pos
    rid = h5_read(path, "/region_id", np.float32)
    stride = len(x) // N
    return pos, rid[stride * np.arange(N)].astype(np.int32)

