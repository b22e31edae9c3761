"""Host builders (libnnx_host.so): Philox streams, generators, layouts."""
import numpy as np
import pytest

import oracle
from neuronephys_clean_amd import configs, host


def test_philox_known_answer_and_numpy():
    # Random123 known-answer vector: philox4x64_10(ctr = 0, key = 0)
    kat = host.philox([0, 0, 0, 0], [0, 0])
    assert [hex(int(v)) for v in kat] == ["0x16554d9eca36314c", "0xdb20fe9d672d0fdc",
                                          "0xd7e772cee186176b", "0x7e68b68aec7ba23b"]
    rng = np.random.default_rng(11)
    for _ in range(5):
        key = rng.integers(0, 2 ** 63, 2, dtype=np.uint64)
        ctr = rng.integers(0, 2 ** 63, 4, dtype=np.uint64)
        bg = np.random.Philox(key=key, counter=ctr)
        ref = bg.random_raw(8)
        # numpy advances the 256-bit counter before each block
        c1 = ctr.copy()
        c1[0] += np.uint64(1)
        c2 = c1.copy()
        c2[0] += np.uint64(1)
        assert np.array_equal(host.philox(c1, key), ref[:4])
        assert np.array_equal(host.philox(c2, key), ref[4:])


def test_uniform_normal_lognormal_statistics():
    n = 1_000_000
    u = host.uniform(3, 1, n, -2.0, 5.0)
    assert u.min() >= -2 and u.max() < 5
    assert abs(u.mean() - 1.5) < 4 * 7 / np.sqrt(12 * n)
    z = host.normal(3, 2, n, 10.0, 2.0)
    assert abs(z.mean() - 10) < 4 * 2 / np.sqrt(n) and abs(z.std() - 2) < 0.01
    ln = host.lognormal(3, 3, n, -3.0, 1.0)
    assert abs(np.log(ln).mean() + 3) < 0.01 and abs(np.log(ln).std() - 1) < 0.01
    # streams are reproducible and independent of the requested length
    assert np.array_equal(host.uniform(3, 1, 1000), host.uniform(3, 1, 5000)[:1000])


def test_er_fixed_out_degree():
    off, tgt = host.er_fixed(0, 1024, 100)
    assert np.array_equal(off, np.arange(1025) * 100)
    seg = tgt.reshape(1024, 100)
    assert np.all(np.diff(seg, axis=1) > 0)  # sorted and distinct
    assert not np.any(seg == np.arange(1024)[:, None])  # no self-loops
    assert tgt.min() >= 0 and tgt.max() < 1024
    indeg = np.bincount(tgt, minlength=1024)
    assert abs(indeg.mean() - 100) < 1e-9 and 60 < indeg.min() and indeg.max() < 145
    with pytest.raises(ValueError):
        host.er_fixed(0, 10, 10)


def test_uniform_targets_per_degree():
    deg = host.outdeg_poisson(5, 2000, 50.0)
    assert abs(deg.mean() - 50) < 1.0 and abs(deg.var() - 50) < 6
    off = host.offsets_of(deg)
    tgt = host.uniform_targets(5, off)
    for i in range(0, 2000, 97):
        s = tgt[off[i]:off[i + 1]]
        assert np.all(np.diff(s) > 0) and i not in s


def test_spatial_targets_distance_law():
    n, box, lam, cutoff = 20000, 34.0, 4.0, 12.0
    pos = host.uniform(9, 4, 3 * n, 0, box).astype(np.float32).reshape(n, 3)
    deg = np.full(n, 30, np.int32)
    off = host.offsets_of(deg)
    tgt = host.spatial_targets(9, pos, box, lam, cutoff, off)
    src = np.repeat(np.arange(n), 30)
    d = np.linalg.norm(pos[tgt].astype(np.float64) - pos[src], axis=1)
    assert d.max() <= cutoff + 1e-5 and not np.any(tgt == src)
    seg = tgt.reshape(n, 30)
    assert np.all(np.diff(seg, axis=1) > 0)
    # for sources away from the walls, the mean target distance equals the
    # exp(-d/lam)-weighted mean over all candidates within the cutoff
    inner = np.nonzero(np.all((pos > cutoff) & (pos < box - cutoff), axis=1))[0][:300]
    exp_mean, emp = [], []
    for i in inner:
        dd = np.linalg.norm(pos.astype(np.float64) - pos[i], axis=1)
        ok = (dd <= cutoff) & (np.arange(n) != i)
        wgt = np.exp(-dd[ok] / lam)
        exp_mean.append((dd[ok] * wgt).sum() / wgt.sum())
        emp.append(d[off[i]:off[i + 1]].mean())
    exp_mean, emp = np.mean(exp_mean), np.mean(emp)
    # 300 sources x 30 draws; the per-draw sd of d is ~2.5
    assert abs(emp - exp_mean) < 4 * 2.6 / np.sqrt(9000), (emp, exp_mean)


def test_csr_to_csc_matches_stable_sort():
    rng = np.random.default_rng(1)
    n = 500
    deg = rng.integers(0, 40, n).astype(np.int32)
    off = host.offsets_of(deg)
    tgt = rng.integers(0, n, off[-1]).astype(np.int32)  # duplicates allowed
    w = rng.random(off[-1]).astype(np.float32)
    in_off, in_src, in_w, o2i, indeg = host.csr_to_csc(n, off, tgt, w)
    src = np.repeat(np.arange(n), deg)
    order = np.lexsort((np.arange(len(tgt)), src, tgt))  # by target, source, CSR position
    assert np.array_equal(in_src, src[order]) and np.array_equal(in_w, w[order])
    assert np.array_equal(in_off, host.offsets_of(np.bincount(tgt, minlength=n)))
    assert np.array_equal(o2i[order], np.arange(len(tgt)))
    assert np.array_equal(indeg, np.bincount(tgt, minlength=n).astype(np.float32))


def test_steady_state_gates():
    V = np.linspace(-90, 30, 121)
    m, h, ng = host.steady_state(V, 5.0)
    am, bm, ah, bh, an, bn = oracle.rates(V + 5.0).T
    np.testing.assert_allclose(m, am / (am + bm), rtol=1e-9)  # a = 0 at V = -40
    np.testing.assert_allclose(h, ah / (ah + bh), rtol=1e-12)
    np.testing.assert_allclose(ng, an / (an + bn), rtol=1e-9)  # b = 0 at V = -55


def test_voxel_of():
    pos = np.array([[0, 0, 0], [31.99, 0.5, 2.0], [40, -1, 3.5]], np.float32)
    v = host.voxel_of(pos, 1.0, 32, 32, 32)
    assert v.tolist() == [0, 31 + 32 * (0 + 32 * 2), 31 + 32 * (0 + 32 * 3)]


def test_trilinear_matrix_partition_of_unity_and_linear_exact():
    nx = ny = nz = 9
    h = 0.25
    L = (nx - 1) * h
    rng = np.random.default_rng(2)
    xyz = (rng.random((3, 400)) - 0.5) * (L * 0.98)
    off, cols, vals, outside = host.trilinear_matrix(*xyz, nx, ny, nz, h)
    assert outside == 0
    W = vals.reshape(-1, 8).astype(np.float64)
    np.testing.assert_allclose(W.sum(1), 1.0, atol=1e-6)
    assert np.all(W >= -1e-7)
    # a linear field on the grid nodes is reproduced exactly
    i = cols % nx
    j = (cols // nx) % ny
    k = cols // (nx * ny)
    f = lambda a, b, c: 1.0 + 2.0 * a - 0.5 * b + 0.25 * c
    node = f(i * h - L / 2, j * h - L / 2, k * h - L / 2)
    interp = (W.ravel() * node).reshape(-1, 8).sum(1)
    np.testing.assert_allclose(interp, f(*xyz), atol=2e-5)


def test_config_c1_shape():
    net = configs.c1()
    assert net.n == 1024 and net.nnz == 102400
    assert net.V0.min() >= -70 and net.V0.max() < -60
    assert abs(net.iext.mean() - 10) < 0.3 and net.w.min() >= 0 and net.w.max() < 0.1
    v = net.ensure_voxel()
    assert v.min() >= 0 and v.max() < 32 ** 3
