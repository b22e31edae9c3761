"""The numpy multigrid restatement (tests/mg_ref.py), pinned by what the
mathematics fixes: the coarse operator of constant coefficients is the
rediscretised one (k/4), restriction and prolongation preserve constants
away from the boundary, the operator is symmetric, and the V-cycle used as
a stationary iteration converges at a rate that does not grow with the
grid."""
import numpy as np

import mg_ref
from neuronephys_clean_amd import configs


def faces(n, k):
    return (np.full((n, n, n + 1), k), np.full((n, n + 1, n), k), np.full((n + 1, n, n), k))


def test_coarse_faces_of_constant_coefficients_are_rediscretised():
    cx, cy, cz = mg_ref.coarsen(*faces(16, 8.0))
    assert cx.shape == (8, 8, 9) and cy.shape == (8, 9, 8) and cz.shape == (9, 8, 8)
    for c in (cx, cy, cz):
        assert np.allclose(c, 2.0)  # dt D / (2h)^2 = k / 4


def test_transfer_operators_preserve_constants_inside():
    r = np.full((8, 8, 8), 3.0)
    assert np.allclose(mg_ref.restrict(np.full((16, 16, 16), 3.0)), 3.0)
    p = mg_ref.prolong(r)
    assert p.shape == (16, 16, 16) and np.allclose(p[1:-1, 1:-1, 1:-1], 3.0)
    # one cell from the boundary the missing neighbour weighs 1/4 per axis
    assert np.isclose(p[0, 5, 5], 0.75 * 3.0)


def test_operator_symmetric_positive():
    rng = np.random.default_rng(0)
    n = 10
    kx, ky, kz = (rng.random(s) * 5 for s in ((n, n, n + 1), (n, n + 1, n), (n + 1, n, n)))
    x, y = rng.normal(size=(n, n, n)), rng.normal(size=(n, n, n))
    assert np.isclose(np.sum(y * mg_ref.apply(x, kx, ky, kz)), np.sum(x * mg_ref.apply(y, kx, ky, kz)))
    assert np.sum(x * mg_ref.apply(x, kx, ky, kz)) > np.sum(x * x)


def _rate(kx, ky, kz, n, iters=6):
    rng = np.random.default_rng(1)
    b = rng.normal(size=(n, n, n))
    x = np.zeros_like(b)
    res = [np.linalg.norm(b)]
    for _ in range(iters):
        x = x + mg_ref.vcycle(b - mg_ref.apply(x, kx, ky, kz), kx, ky, kz)
        res.append(np.linalg.norm(b - mg_ref.apply(x, kx, ky, kz)))
    return (res[-1] / res[2]) ** (1.0 / (iters - 2))


def test_vcycle_rate_independent_of_grid():
    rates = {n: _rate(*faces(n, 50.0), n) for n in (16, 32, 64)}
    print(rates)
    assert max(rates.values()) < 0.35
    assert rates[64] < rates[16] + 0.1


def test_vcycle_converges_on_lognormal_field():
    f = configs.c5_field(seed=2, n=32)
    kx = f.kx.reshape(32, 32, 33).astype(np.float64)
    ky = f.ky.reshape(32, 33, 32).astype(np.float64)
    kz = f.kz.reshape(33, 32, 32).astype(np.float64)
    rate = _rate(kx, ky, kz, 32)
    print("log-normal rate", rate)
    assert rate < 0.5


def test_chebyshev_recurrence_is_the_chebyshev_residual_polynomial():
    """m sweeps from zero on a scalar equation lambda x = 1 leave the
    residual 1 - lambda x = T_m((theta - lambda)/delta) / T_m(theta/delta)
    (theta, delta: centre and half-width of [lmax/16, lmax])."""
    lmax, lmin = 2.0, 2.0 / 16
    th, de = (lmax + lmin) / 2, (lmax - lmin) / 2
    for m in (1, 2, 4, 7):
        co = mg_ref.chebyshev_coefficients(m)
        for lam in np.linspace(lmin, lmax, 9):
            x, d = 0.0, None
            for c1, c2 in co:
                res = 1.0 - lam * x
                d = c2 * res if d is None else c1 * d + c2 * res
                x = x + d
            T = np.polynomial.chebyshev.Chebyshev.basis(m)
            assert np.isclose(1.0 - lam * x, T((th - lam) / de) / T(th / de), rtol=1e-9, atol=1e-12)


def test_chebyshev_vcycle_rate():
    rates = {}
    for n in (16, 32):
        k = faces(n, 50.0)
        rng = np.random.default_rng(1)
        b = rng.normal(size=(n, n, n))
        x = np.zeros_like(b)
        res = [np.linalg.norm(b)]
        for _ in range(6):
            x = x + mg_ref.vcycle(b - mg_ref.apply(x, *k), *k, nu=4, smoother="chebyshev")
            res.append(np.linalg.norm(b - mg_ref.apply(x, *k)))
        rates[n] = (res[-1] / res[2]) ** 0.25
    print("chebyshev rates", rates)
    assert max(rates.values()) < 0.2


def test_wcycle_converges_faster_than_vcycle():
    """On the log-normal field the W-cycle (each coarser level visited twice)
    contracts more per cycle than the V-cycle (at 64^3: five levels)."""
    n = 64
    f = configs.c5_field(seed=2, n=n)
    k = (f.kx.astype(np.float64), f.ky.astype(np.float64), f.kz.astype(np.float64))
    rng = np.random.default_rng(1)
    b = rng.normal(size=(n, n, n))
    out = {}
    for gamma in (1, 2):
        x = np.zeros_like(b)
        res = [np.linalg.norm(b)]
        for _ in range(6):
            x = x + mg_ref.vcycle(b - mg_ref.apply(x, *k), *k, nu=3, smoother="chebyshev", gamma=gamma)
            res.append(np.linalg.norm(b - mg_ref.apply(x, *k)))
        out[gamma] = (res[-1] / res[2]) ** 0.25
    print("V/W rates", out)
    assert out[2] < 0.8 * out[1]
