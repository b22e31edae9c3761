"""The white-noise generator of the oracle (the restatement the device's
leapfrog noise is checked against): Philox4x64-10 against the Random123
known answer and numpy.random.Philox, ln and cos 2 pi u against extended
precision, the Box-Muller draw against its formula evaluated in extended
precision from the same words, and the normals' statistics."""
import numpy as np
import scipy.stats

import oracle

PURPOSE = 0x4e4f495345


def test_philox_known_answer_and_numpy():
    kat = oracle.philox4x64([0, 0, 0, 0], [0, 0])
    assert [hex(int(v)) for v in kat] == ["0x16554d9eca36314c", "0xdb20fe9d672d0fdc",
                                          "0xd7e772cee186176b", "0x7e68b68aec7ba23b"]
    rng = np.random.default_rng(3)
    for _ in range(5):
        key = rng.integers(0, 2 ** 63, 2, dtype=np.uint64)
        ctr = rng.integers(0, 2 ** 63, 4, dtype=np.uint64)
        ref = np.random.Philox(key=key, counter=ctr).random_raw(4)
        c1 = ctr.copy()
        c1[0] += np.uint64(1)  # numpy advances the counter before each block
        assert np.array_equal(oracle.philox4x64(c1, key), ref)


def _ulp(x):
    return np.spacing(np.abs(x).astype(np.float64))


def test_log_within_two_ulp():
    """ln u = e ln2 + ln m: the two terms cancel partly for u just below 1/2
    and near sqrt(1/2), so the result is within 2 ulp (1.96 measured)."""
    rng = np.random.default_rng(4)
    u = np.concatenate([rng.random(200_000) + 2.0 ** -53, 2.0 ** -np.arange(0, 54, dtype=np.float64),
                        np.nextafter(np.sqrt(0.5), [0.0, 1.0]), [1.0, np.nextafter(1.0, 0.0), 2.0 ** -53]])
    got = oracle.noise_log(u)
    ref = np.log(u.astype(np.longdouble))
    err = np.abs(got.astype(np.longdouble) - ref)
    assert np.all(got[u == 1.0] == 0.0)
    assert np.all(err <= 2.0 * _ulp(ref.astype(np.float64)) + 1e-300)


def test_cos2pi_accuracy_and_symmetry():
    rng = np.random.default_rng(5)
    u = np.concatenate([rng.random(200_000), np.arange(0, 1, 1 / 64), [0.125, 0.375, 0.625, 0.875]])
    got = oracle.noise_cos2pi(u)
    two_pi = 2 * np.arccos(np.longdouble(-1))
    ref = np.cos(two_pi * u.astype(np.longdouble))
    assert np.all(np.abs(got - ref.astype(np.float64)) <= 4e-16)
    # exact values of the octant ends and symmetry cos 2 pi u = cos 2 pi (1 - u)
    assert got[u == 0.0][0] == 1.0 and got[u == 0.5][0] == -1.0 and abs(got[u == 0.25][0]) < 1e-16
    v = rng.random(10_000)
    assert np.array_equal(oracle.noise_cos2pi(v), oracle.noise_cos2pi(1.0 - v)) or \
        np.max(np.abs(oracle.noise_cos2pi(v) - oracle.noise_cos2pi(1.0 - v))) <= 2e-16


def test_normal_is_box_muller_of_the_words():
    seed, step, n = 99, 1234, 2000
    z = oracle.noise_normal(seed, step, n)
    two_pi = 2 * np.arccos(np.longdouble(-1))
    for i in range(0, n, 97):
        w = oracle.philox4x64([step, i, PURPOSE, 0], [seed, 0])
        u1 = np.longdouble(int(w[0] >> np.uint64(11)) + 1) * np.longdouble(2.0) ** -53
        u2 = np.longdouble(int(w[1] >> np.uint64(11))) * np.longdouble(2.0) ** -53
        ref = np.sqrt(-2 * np.log(u1)) * np.cos(two_pi * u2)
        assert abs(z[i] - float(ref)) <= 4e-15 * max(1.0, abs(float(ref))), i
    # first index offsets the neuron number
    assert np.array_equal(oracle.noise_normal(seed, step, 10, first=500), z[500:510])


def test_normal_statistics():
    z = np.concatenate([oracle.noise_normal(0, s, 250_000) for s in range(4)])
    n = z.size
    assert abs(z.mean()) < 5 / np.sqrt(n)
    assert abs(z.var() - 1.0) < 5 * np.sqrt(2.0 / n)
    assert scipy.stats.kstest(z, "norm").pvalue > 1e-3
    # different steps and seeds are different streams
    assert not np.array_equal(oracle.noise_normal(0, 1, 100), oracle.noise_normal(0, 2, 100))
    assert not np.array_equal(oracle.noise_normal(0, 1, 100), oracle.noise_normal(1, 1, 100))
    # successive steps of one neuron are uncorrelated
    a, b = oracle.noise_normal(3, 10, 200_000), oracle.noise_normal(3, 11, 200_000)
    assert abs(np.corrcoef(a, b)[0, 1]) < 5 / np.sqrt(a.size)
