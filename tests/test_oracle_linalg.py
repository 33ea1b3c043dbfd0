"""The likelihood oracle (oracle/linalg.py) against what the mathematics fixes: the
flat-prior marginal likelihood by brute-force quadrature, a hand-solved exact fit, and
finite differences of ll for the gradient."""
import numpy as np
import pytest

from oracle import linalg as OL


def _curve(g, n, Fs=900.0, Fb=150.0, sigma=12.0):
    t = np.sort(g.uniform(-30, 30, n))
    u = np.hypot(0.2, t / 12.0)
    A = (u**2 + 2) / (u * np.sqrt(u**2 + 4))
    f = Fs * A + Fb + g.normal(0, sigma, n)
    c = 1.0 / (sigma * g.uniform(0.8, 1.2, n)) ** 2
    return A, f, c


def test_exact_fit_known_answer():
    # A = (1, 2), f = A: beta = (1, 0), chi2 = 0, N = [[5, 3], [3, 2]], det N = 1,
    # ll = 1/2 log det(2 pi N^-1) = log(2 pi)
    beta, ll = OL.marginalized_log_likelihood([[1.0, 2.0]], [[1.0, 2.0]], [[1.0, 1.0]])
    assert np.allclose(beta[0], [1.0, 0.0], atol=1e-14)
    assert ll == pytest.approx(np.log(2 * np.pi), rel=1e-14)


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_equals_flat_prior_marginal_by_quadrature(seed):
    g = np.random.default_rng(seed)
    A, f, c = _curve(g, 40)
    _, ll = OL.marginalized_log_likelihood([A], [f], [c])
    q = OL.marginal_ll_quadrature(A, f, c)
    assert abs(ll - q) < 1e-9 * max(1.0, abs(q))


def test_sum_over_observatories():
    g = np.random.default_rng(3)
    curves = [_curve(g, n) for n in (30, 17, 55)]
    _, ll = OL.marginalized_log_likelihood(*[list(x) for x in zip(*curves)])
    parts = [OL.marginalized_log_likelihood([a], [f], [c])[1] for a, f, c in curves]
    assert ll == pytest.approx(sum(parts), rel=1e-14)


def test_grad_A_matches_central_differences():
    g = np.random.default_rng(4)
    A, f, c = _curve(g, 25)
    gA = OL.grad_A([A], [f], [c])[0]
    for i in range(0, 25, 3):
        h = 1e-6 * A[i]
        Ap, Am = A.copy(), A.copy()
        Ap[i] += h
        Am[i] -= h
        fd = (OL.marginalized_log_likelihood([Ap], [f], [c])[1]
              - OL.marginalized_log_likelihood([Am], [f], [c])[1]) / (2 * h)
        assert gA[i] == pytest.approx(fd, rel=1e-5, abs=1e-6 * np.abs(gA).max())
