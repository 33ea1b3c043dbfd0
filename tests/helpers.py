"""Shared checks for the parity tests."""
import numpy as np
from scipy.optimize import linear_sum_assignment

EPS = 2.0 ** -53


def match_rows(a, b):
    """Permute each row of b to best match the same row of a (assignment on |a_i-b_j|)."""
    a = np.asarray(a)
    b = np.asarray(b)
    out = b.copy()
    # rows already in the same order (relative test: roots may be far below 1)
    same = np.all(np.abs(a - b) <= 1e-8 * np.abs(a), axis=1)
    for i in np.nonzero(~same)[0]:
        d = np.abs(a[i][:, None] - b[i][None, :])
        r, c = linear_sum_assignment(d)
        out[i, r] = b[i, c]
    return out


def root_condition(coeffs_asc, z):
    """Relative condition number of each root: sum_k |p_k||z|^k / (|z| |p'(z)|)."""
    c = np.asarray(coeffs_asc)
    n = c.shape[-1] - 1
    k = np.arange(n + 1)
    az = np.abs(z)[..., None]
    num = (np.abs(c)[:, None, :] * az ** k).sum(-1)
    dp = (k[1:] * c[:, None, 1:] * z[..., None] ** (k[1:] - 1)).sum(-1)
    return num / np.maximum(np.abs(z) * np.abs(dp), 1e-300)


def residual_ratio(coeffs_asc, z):
    """|p(z)| / sum_k |p_k||z|^k (normwise backward error of each root)."""
    c = np.asarray(coeffs_asc)
    n = c.shape[-1] - 1
    k = np.arange(n + 1)
    pz = (c[:, None, :] * z[..., None] ** k).sum(-1)
    den = (np.abs(c)[:, None, :] * np.abs(z)[..., None] ** k).sum(-1)
    return np.abs(pz) / den


def roots_close(z_test, z_ref, coeffs_asc, rel=1e-12, kappa_factor=64.0, truth=True):
    """Matched root sets agree to `rel` relative, or within kappa_factor*kappa*eps
    for ill-conditioned roots.  With truth=True a row that fails is judged against
    its 40-digit roots instead: the tested roots must be within `rel`, the kappa
    bound, or 10x the oracle's own error of them (first-order kappa underestimates
    the error of nearly multiple roots).  Returns (ok mask, matched z_test, rel error)."""
    zt = match_rows(z_ref, z_test)
    err = np.abs(zt - z_ref) / np.maximum(np.abs(z_ref), 1e-300)
    kap = root_condition(coeffs_asc, z_ref)
    tol = np.maximum(rel, kappa_factor * kap * EPS)
    ok = err <= tol
    bad = np.nonzero(~ok.all(axis=1))[0]
    if truth and bad.size:
        from oracle import mp_lens

        zt_true = match_rows(z_ref[bad], mp_lens.roots_mp(np.asarray(coeffs_asc)[bad]))
        a = np.maximum(np.abs(zt_true), 1e-300)
        e_test = np.abs(match_rows(zt_true, zt[bad]) - zt_true) / a
        e_ref = np.abs(z_ref[bad] - zt_true) / a
        ok[bad] = e_test <= np.maximum(tol[bad], 10.0 * e_ref)
    return ok, zt, err


def agree_with_truth(test, oracle_runs, truth_fn, rel=1e-10, factor=10.0):
    """Per item: |test - oracle| <= rel |oracle|; where that fails the item is judged
    against the 40-digit value truth_fn(indices) -> (truth, undecided): it passes if
    |test - truth| <= max(rel |truth|, factor * max_k |oracle_k - truth|, undecided),
    the oracle runs being the plain FP64 oracle and copies with 1-ulp input
    perturbations (the size of FP64 error the route itself has at that point), and
    `undecided` the part of the value FP64 cannot decide (clustered images)."""
    test = np.asarray(test, float)
    runs = [np.asarray(r, float) for r in oracle_runs]
    ok = np.abs(test - runs[0]) <= rel * np.abs(runs[0])
    idx = np.nonzero(~ok.reshape(-1))[0]
    if idx.size:
        tr, amb = truth_fn(idx)
        tf, rf = test.reshape(-1)[idx], [r.reshape(-1)[idx] for r in runs]
        scale = np.max([np.abs(r - tr) for r in rf], axis=0)
        tol = np.maximum(np.maximum(rel * np.abs(tr), factor * scale), 1.000001 * amb)
        good = np.abs(tf - tr) <= tol
        ok.reshape(-1)[idx] = good
    return ok


def agree_condition_aware(test, plain, truth, rel=1e-10, factor=10.0):
    """Per item: the tested value is within `rel` of the plain FP64 oracle, or at least
    as close to a more accurate value (compensated-arithmetic roots) as the plain
    oracle is, up to `factor`.  Both FP64 solvers lose accuracy on ill-conditioned
    (clustered) roots by the same mechanism; this separates that from real errors."""
    test, plain, truth = map(np.asarray, (test, plain, truth))
    e_tp = np.abs(test - plain) / np.abs(plain)
    e_t = np.abs(test - truth) / np.abs(truth)
    e_p = np.abs(plain - truth) / np.abs(truth)
    return (e_tp <= rel) | (e_t <= np.maximum(rel, factor * e_p))


def agree_within_spread(test, plain, perturbed, rel=1e-10, factor=100.0):
    """Per item: |test - plain| <= max(rel |plain|, factor * spread), where `perturbed`
    holds oracle results for coefficients perturbed at the 1e-16 level (a list of
    arrays): the spread is how far the FP64 polynomial route itself is determined."""
    test, plain = np.asarray(test), np.asarray(plain)
    spread = np.max([np.abs(np.asarray(p) - plain) for p in perturbed], axis=0)
    return np.abs(test - plain) <= np.maximum(rel * np.abs(plain), factor * spread)
