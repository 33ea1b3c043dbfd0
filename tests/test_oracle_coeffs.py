"""Pins the coefficient restatements (oracle/lens.py, caustics_amd/workloads.py).

Golden: tests/golden/coeffs_fixture.npz holds the reference notebook's printed
sympy polynomials (notebooks/ComplexPolynomialCoefficients.ipynb cells 10, 13, 19, 22)
evaluated at 50 digits.  Invariant: images z mapped forward by the lens equation give
w with p(z; w) = 0."""
import os

import numpy as np
import pytest

from oracle import lens
from caustics_amd import workloads as W

FIX = os.path.join(os.path.dirname(__file__), "golden", "coeffs_fixture.npz")


def rel(x, y):
    return np.max(np.abs(x - y) / np.max(np.abs(y), axis=-1, keepdims=True))


def test_binary_expanded_matches_notebook():
    f = np.load(FIX)
    c = np.array([lens.poly_coeffs_binary(f["w"][i], f["a"][i], f["e1"][i]) for i in range(64)])
    assert rel(c, f["lens2"]) < 1e-14


def test_product_form_matches_notebook_binary_and_triple():
    f = np.load(FIX)
    c2 = W.coeffs_product(f["w"], f["a"], f["e1"], 1 - f["e1"], 0 * f["r3"], 2)[..., ::-1]
    c3 = W.coeffs_product(f["w"], f["a"], f["e1"], f["e2"], f["r3"], 3)[..., ::-1]
    assert rel(c2, f["lens2"]) < 1e-14
    assert rel(c3, f["lens3"]) < 1e-14
    # the oracle's own triple builder agrees too
    c3o = np.array([lens.poly_coeffs_triple(f["w"][i], f["a"][i], f["r3"][i], f["e1"][i],
                                            f["e2"][i]) for i in range(64)])
    assert rel(c3o, f["lens3"]) < 1e-14


@pytest.mark.parametrize("nlenses", [2, 3])
def test_forward_map_invariant(nlenses):
    g = np.random.default_rng(5)
    n = 400
    z = g.uniform(-2, 2, n) + 1j * g.uniform(-2, 2, n)
    a = g.uniform(0.15, 1.5, n)
    e1 = g.uniform(0.05, 0.9, n)
    e2 = (1 - e1) * g.uniform(0.05, 0.95, n)
    r3 = g.uniform(0.3, 2, n) * np.exp(1j * g.uniform(0, 2 * np.pi, n))
    if nlenses == 2:
        w = lens.lens_eq(z, 2, a=a, e1=e1)
        c = W.coeffs_product(w, a, e1, 1 - e1, 0 * r3, 2)
    else:
        w = lens.lens_eq(z, 3, a=a, e1=e1, e2=e2, r3=r3)
        c = W.coeffs_product(w, a, e1, e2, r3, 3)
    k = np.arange(c.shape[-1])
    pz = (c * z[:, None] ** k).sum(-1)
    scale = (np.abs(c) * np.abs(z[:, None]) ** k).sum(-1)
    assert np.max(np.abs(pz) / scale) < 1e-13


def test_critical_curve_coefficients_match_notebook():
    """Critical-curve polynomials (point_source.py:1482-1534) against the notebook's
    printed sympy outputs (cells 19 and 22) evaluated at 50 digits: the oracle's
    restatement and the package's torch builder (evaluated here on CPU tensors)."""
    import torch

    from caustics_amd.point_source import _critical_coeffs

    f = np.load(FIX)
    phi, a, e1, e2, r3 = f["phi"], f["a"], f["e1"], f["e2"], f["r3"]
    c2 = np.array([lens.poly_coeffs_critical(phi[i], 2, a=a[i], e1=e1[i]) for i in range(64)])
    c3 = np.array([lens.poly_coeffs_critical(phi[i], 3, a=a[i], e1=e1[i], e2=e2[i], r3=r3[i])
                   for i in range(64)])
    assert rel(c2, f["crit2"]) < 1e-14
    assert rel(c3, f["crit3"]) < 1e-14
    t = lambda x: torch.as_tensor(x)[:, None]
    p2 = _critical_coeffs(t(phi), 2, t(a), t(e1))[:, 0].numpy()
    p3 = _critical_coeffs(t(phi), 3, t(a), t(e1), t(e2), t(r3))[:, 0].numpy()
    assert rel(p2, f["crit2"]) < 1e-14
    assert rel(p3, f["crit3"]) < 1e-14
