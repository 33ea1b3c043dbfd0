"""Hexadecapole magnification and contour gate (mag() decision) vs the oracle
(oracle/multipole.py restating multipole.py:8-244 and lightcurve.py:23-85)."""
import numpy as np
import pytest

from caustics_amd.lightcurve import hexadecapole_gate
from oracle import lens, multipole

from helpers import agree_within_spread

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("q", [0.2, 3e-3])
def test_gate_matches_oracle_near_caustics(gpu, q):
    s, rho = 0.9, 1e-2
    _, ca = lens.critical_and_caustic_curves(npts=120, nlenses=2, s=s, q=q)
    g = np.random.default_rng(4)
    ca = ca.reshape(-1)
    w = ca + g.uniform(0, 15 * rho, ca.shape) * np.exp(1j * g.uniform(-np.pi, np.pi, ca.shape))
    mu, dmu, use = hexadecapole_gate(w, rho, nlenses=2, s=s, q=q)
    mo, uo = multipole.hexadecapole_gate(w, rho, nlenses=2, s=s, q=q)
    mu, use = mu.cpu().numpy(), use.cpu().numpy()
    e = np.abs(mu - mo) / np.abs(mo)
    print(f"q={q}: accepted {use.mean():.2f}; mu max rel err {e.max():.2e}; "
          f"decision agreement {(use == uo).mean():.4f}")
    pert = [multipole.hexadecapole_gate(w, rho, nlenses=2, perturb=4e-16, seed=sd, s=s, q=q)[0]
            for sd in (1, 2)]
    ok = agree_within_spread(mu, mo, pert)
    print(f"agreement within the oracle's own coefficient spread {ok.mean():.4f}")
    assert ok.mean() >= 0.999
    assert (use == uo).mean() >= 0.995


def test_hexadecapole_u1_and_single_lens(gpu):
    w = np.array([0.5 + 0.1j, -0.3 + 0.4j, 1.2 - 0.2j])
    for u1 in (0.0, 0.6):
        mu, dmu, use = hexadecapole_gate(w, 0.02, nlenses=2, u1=u1, s=1.1, q=0.4)
        lp, x_cm = lens.lens_params(2, s=1.1, q=0.4)
        z, mask = lens.images_point_source(w + x_cm, 2, **lp)
        mo, dmo = multipole.mag_hexadecapole(z, mask, 0.02, u1, 2, **lp)
        np.testing.assert_allclose(mu.cpu().numpy(), mo, rtol=1e-11)
        np.testing.assert_allclose(dmu.cpu().numpy(), dmo, rtol=1e-9)
    mu, _, use = hexadecapole_gate(np.array([0.5, 0.01j]), 0.02, nlenses=1)
    assert use.cpu().numpy().tolist() == [True, False]
