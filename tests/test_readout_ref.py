"""CPU pins of the electrode readout: the oracle's region means, network
mean and synchrony statistics (oracle/oracle.hpp) against hand-derived
values, closed forms and exact sums, and the k-nearest region builder
(libnnx_host.so) against brute force."""
import math

import numpy as np
import pytest

import oracle
from neuronephys_clean_amd import host


def test_region_mean_known_answers():
    V = np.array([-60.0, -70.0, 0.5, 1.25, 1e8, 1.0, -1e8, 1.0], np.float32)
    off = np.array([0, 2, 4, 8, 8, 9], np.int32)
    idx = np.array([0, 1, 2, 3, 4, 5, 6, 7, 1], np.int32)
    fire = np.array([1, 0, 1, 1, 0, 0, 1, 0], np.int32)
    avg, cnt = oracle.electrode_measure(off, idx, V, fire)
    assert avg[0] == np.float32(-65.0) and avg[1] == np.float32(0.875)
    # the reference adds one at a time: ((1e8 + 1) - 1e8) + 1 = 1 in fp32
    # (1e8 + 1 rounds to 1e8), so the mean is 0.25; an exact sum would give
    # 0.5 and a pairwise one 0
    assert avg[2] == np.float32(0.25)
    assert np.isnan(avg[3])  # empty region: 0/0 as in the reference
    assert avg[4] == np.float32(-70.0)
    assert cnt.tolist() == [1, 2, 1, 0, 0]


def test_region_mean_order_is_sequential_fp64():
    # the same cancellation in FP64: 2^53 + 1 rounds to 2^53
    V = np.array([2.0 ** 53, 1.0, -(2.0 ** 53), 1.0])
    avg, _ = oracle.electrode_measure(np.array([0, 4]), np.arange(4), V)
    assert avg[0] == 0.25


@pytest.mark.parametrize("T", [np.float32, np.float64])
def test_region_mean_error_bound_and_counts(T):
    rng = np.random.default_rng(1)
    n = 5000
    V = rng.normal(-60, 15, n).astype(T)
    fire = (rng.random(n) < 0.1).astype(np.int32)
    sizes = rng.integers(0, 300, 40)
    off = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    idx = rng.integers(0, n, off[-1]).astype(np.int32)
    avg, cnt = oracle.electrode_measure(off, idx, V, fire)
    eps = np.finfo(T).eps
    for r in range(len(sizes)):
        sel = idx[off[r]:off[r + 1]]
        assert cnt[r] == fire[sel].sum()
        if len(sel) == 0:
            assert np.isnan(avg[r])
            continue
        exact = math.fsum(V[sel].astype(np.float64)) / len(sel)
        # recursive summation: |error| <= (k-1) eps sum|x| (+ the division)
        bound = len(sel) * eps * np.abs(V[sel].astype(np.float64)).sum() / len(sel) + eps * abs(exact)
        assert abs(float(avg[r]) - exact) <= bound


def test_region_out_of_range_index_gives_nan():
    V = np.ones(10, np.float32)
    avg, cnt = oracle.electrode_measure(np.array([0, 2, 3]), np.array([1, 10, -1]), V, np.ones(10))
    assert np.isnan(avg[0]) and np.isnan(avg[1]) and cnt.tolist() == [1, 0]


@pytest.mark.parametrize("T", [np.float32, np.float64])
@pytest.mark.parametrize("n", [1, 63, 262145, 600001])
def test_network_mean_against_exact_sum(T, n):
    rng = np.random.default_rng(n)
    V = rng.normal(-60, 20, n).astype(T)
    m = oracle.network_mean(V)
    exact = math.fsum(V.astype(np.float64)) / n
    if T == np.float32:
        # an FP64 sum of <= 2^20 fp32 values is within ~1e-10 relative: one
        # rounding to fp32 away from the exact mean
        assert abs(float(m) - exact) <= 0.5 * np.spacing(np.float32(abs(exact))) * 1.0001
    else:
        assert abs(m - exact) <= 1e-14 * abs(exact)
    assert oracle.network_mean(np.full(n, T(-61.25))) == T(-61.25)


def _sums(n_idx, n_sub):
    return [np.zeros(n_idx), np.zeros(n_idx), np.zeros(n_sub), np.zeros(n_sub)]


def test_synchrony_closed_forms():
    """Perfect synchrony gives chi^2 = 1 for every sub-population; independent
    neurons give chi^2 close to 1/N (Golomb's measure)."""
    rng = np.random.default_rng(2)
    n, steps = 64, 3000
    counts = np.array([1, 2, 8, 32, 64], np.int32)
    idx = rng.permutation(n).astype(np.int32)
    same, indep = _sums(n, 5), _sums(n, 5)
    t = np.arange(steps) * 0.05
    for k in range(steps):
        oracle.sync_accumulate(np.full(n, -60 + 10 * np.sin(t[k]), np.float64), idx, counts, same)
        oracle.sync_accumulate(rng.normal(-60, 5, n), idx, counts, indep)
    chi_same = oracle.sync_measures(n, counts, same, steps)
    chi_ind = oracle.sync_measures(n, counts, indep, steps)
    assert np.allclose(chi_same, 1.0, rtol=0, atol=1e-9)
    assert chi_ind[0] == pytest.approx(1.0, abs=1e-12)
    for c, x in zip(counts[1:], chi_ind[1:]):
        assert abs(x * c - 1.0) < 0.25, (c, x)


def test_synchrony_sub_mean_is_rounded_fp32_division():
    """Vavg = (fp32 sum) / count in fp32, widened (src_wholebrain/electrode_net.cu:88-90)."""
    V = np.array([0.1, 0.2, 0.7], np.float32)
    sums = _sums(3, 1)
    oracle.sync_accumulate(V, np.arange(3), [3], sums)
    vs = np.float32(math.fsum(V.astype(np.float64)))
    assert sums[2][0] == float(np.float32(vs / np.float32(3)))
    assert sums[0].tolist() == V.astype(np.float64).tolist()
    assert sums[1].tolist() == [float(np.float32(x * x)) for x in V]


def test_synchrony_degenerate_counts():
    sums = _sums(4, 2)
    oracle.sync_accumulate(np.ones(4), np.arange(4), [0, 5], sums)
    chi = oracle.sync_measures(4, [0, 5], sums, 1)
    assert np.isnan(sums[2]).all() and np.isnan(chi).all()
    chi0 = oracle.sync_measures(4, [2, 2], _sums(4, 2), 0)
    assert np.isnan(chi0).all()  # no steps: 0/0


def _brute_knn(pos, q, k):
    d2 = ((pos.astype(np.float64) - q.astype(np.float64)) ** 2).sum(1)
    return np.lexsort((np.arange(len(pos)), d2))[:k]


@pytest.mark.parametrize("k", [1, 7, 50])
def test_spherical_regions_match_brute_force(k):
    rng = np.random.default_rng(k)
    n = 3000
    pos = rng.uniform(0, 40, (n, 3)).astype(np.float32)
    pos[:500] = np.floor(pos[:500])  # lattice points: many equal distances
    centers = rng.integers(0, n, 60).astype(np.int32)
    off, idx = host.spherical_regions(pos, k, center_idx=centers)
    assert off.tolist() == list(range(0, 61 * k, k))
    for r, c in enumerate(centers):
        ref = _brute_knn(pos, pos[c], k)
        assert idx[r * k:(r + 1) * k].tolist() == ref.tolist()
        assert idx[r * k] == c or ((pos[idx[r * k]] == pos[c]).all() and idx[r * k] < c)
    pts = rng.uniform(-10, 50, (25, 3)).astype(np.float32)  # some outside the cloud
    off, idx = host.spherical_regions(pos, k, center_pos=pts)
    for r in range(len(pts)):
        assert idx[r * k:(r + 1) * k].tolist() == _brute_knn(pos, pts[r], k).tolist()


def test_spherical_regions_flat_cloud_and_errors():
    rng = np.random.default_rng(9)
    pos = np.zeros((400, 3), np.float32)
    pos[:, :2] = rng.uniform(0, 10, (400, 2))  # all in one plane
    off, idx = host.spherical_regions(pos, 400, center_idx=[3, 17])
    for r, c in enumerate([3, 17]):
        assert idx[r * 400:(r + 1) * 400].tolist() == _brute_knn(pos, pos[c], 400).tolist()
    with pytest.raises(ValueError):
        host.spherical_regions(pos, 401, center_idx=[0])
    with pytest.raises(ValueError):
        host.spherical_regions(pos, 3, center_idx=[400])
