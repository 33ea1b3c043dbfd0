"""Pins the C oracle solver (oracle/ea_oracle.c).

Checks: the reference's own test fixture (tests/test_ehrlich_aberth_primitive.py:19-35,
residual atol 1e-10), the reference op compiled from its sources (oracle/_ref, where
built: same root order, roots within 1e-14), closed forms (roots of prod (z - r_k)
built from known roots, Vieta), and numpy.roots on random polynomials."""
import numpy as np
import pytest

from oracle import ea, lens, ref_op
from caustics_amd import workloads as W

from helpers import match_rows, residual_ratio, root_condition


def reference_fixture():
    # tests/test_ehrlich_aberth_primitive.py:19-27: w = linspace(0.3, 0.35, 10), a=0.45, e1=0.8
    w = np.linspace(0.3, 0.35, 10).astype(complex)
    return lens.poly_coeffs_binary(w, 0.45, 0.8).reshape(5, 2, 6)


@pytest.mark.parametrize("comp", [False, True])
@pytest.mark.parametrize("init_mode", [0, 1])
def test_reference_fixture_residual(comp, init_mode):
    c = reference_fixture()
    r = ea.poly_roots(c, compensated=comp, init_mode=init_mode)
    for idx in [(3, 1), (4, 0), (0, 0)]:
        pe = np.polyval(c[idx], r[idx])
        np.testing.assert_allclose(pe, 0, atol=1e-10)


def test_known_roots_closed_form():
    g = np.random.default_rng(1)
    for deg in (5, 10):
        zs = g.normal(size=(200, deg)) + 1j * g.normal(size=(200, deg))
        c = np.array([np.poly(z) for z in zs])  # highest first, from the roots
        r = ea.poly_roots(c, init_mode=1)
        rm = match_rows(zs, r)
        assert np.max(np.abs(rm - zs) / np.abs(zs)) < 1e-10
        # Vieta: sum of roots = -p_{n-1}/p_n, product = (-1)^n p_0/p_n
        np.testing.assert_allclose(r.sum(-1), -c[:, 1] / c[:, 0], rtol=1e-12, atol=1e-12)
        np.testing.assert_allclose(np.prod(r, -1), (-1) ** deg * c[:, -1] / c[:, 0], rtol=1e-11)


@pytest.mark.parametrize("nlenses", [2, 3])
def test_lens_polys_vs_numpy_roots(nlenses):
    lf, w, _ = W.batched_polys(nlenses, 500, seed=7)
    c = W.coeffs_product(w, lf["a"], lf["e1"], lf["e2"], lf["r3"], nlenses)
    for init_mode in (0, 1):
        r, st, status = ea.roots_asc(c, init_mode=init_mode, return_stats=True)
        assert status.all()
        assert np.max(residual_ratio(c, r)) < 1e-14
        nr = np.array([np.roots(ci[::-1]) for ci in c])
        rm = match_rows(nr, r)
        # numpy's companion-matrix roots are less accurate: compare where well conditioned
        err = np.abs(rm - nr) / np.abs(nr)
        kap = root_condition(c, rm)
        assert np.median(err) < 1e-12
        assert np.max(err[kap < 1e3]) < 1e-10


@pytest.mark.skipif(not ref_op.available(), reason="reference op not built (oracle/_ref)")
@pytest.mark.parametrize("nlenses", [2, 3])
@pytest.mark.parametrize("comp", [False, True])
def test_oracle_matches_reference_op(nlenses, comp):
    lf, w, _ = W.batched_polys(nlenses, 2000, seed=11)
    c = W.coeffs_product(w, lf["a"], lf["e1"], lf["e2"], lf["r3"], nlenses)
    r_ref = ref_op.roots_asc(c, compensated=comp)
    r_or, st, status = ea.roots_asc(c, compensated=comp, init_mode=0, return_stats=True)
    ok = status.astype(bool)
    # same root ORDER (no matching) and the same values up to the rounding of pow(z,2),
    # amplified by the root's condition number
    d = np.abs(r_ref - r_or) / np.abs(r_ref)
    tol = np.maximum(1e-13, 64 * root_condition(c, r_ref) * 2.0**-53)
    assert np.all((d <= tol)[ok])
    # thrust and glibc complex log/exp differ in the last bit; many roots still match exactly
    assert np.mean(r_ref == r_or) > 0.3


@pytest.mark.skipif(not ref_op.available(), reason="reference op not built (oracle/_ref)")
def test_oracle_custom_init_matches_reference_op():
    lf, w, _ = W.batched_polys(2, 500, seed=12)
    c = W.coeffs_product(w, lf["a"], lf["e1"], lf["e2"], lf["r3"], 2)
    r0 = ea.roots_asc(c)
    g = np.random.default_rng(0)
    init = r0 + 1e-3 * (g.normal(size=r0.shape) + 1j * g.normal(size=r0.shape))
    r_ref = ref_op.roots_asc(c, roots_init=init)
    r_or = ea.roots_asc(c, roots_init=init, init_mode=0)
    d = np.abs(r_ref - r_or) / np.abs(r_ref)
    assert np.all(d <= np.maximum(1e-13, 64 * root_condition(c, r_ref) * 2.0**-53))


def test_non_convergence_flag():
    lf, w, _ = W.batched_polys(2, 50, seed=3)
    c = W.coeffs_product(w, lf["a"], lf["e1"], lf["e2"], lf["r3"], 2)
    _, st, status = ea.roots_asc(c, itmax=1, return_stats=True)
    assert not status.any()
    assert (st[:, 0] == 1).all()
