"""Images of point sources and their sequential, warm-started form
(caus_images_point_source) against the oracle (oracle/lens.py, restating
point_source.py:1652-1759)."""
import numpy as np
import pytest
import torch

import caustics_amd as ca
from oracle import lens

from helpers import roots_close

pytestmark = pytest.mark.gpu

BIN = dict(a=0.45, e1=0.8)
TRI = dict(a=0.6, e1=0.5, e2=0.3, r3=0.2 + 0.7j)


def _track(n, u0=0.15, angle=0.4, half=1.6):
    t = np.linspace(-half, half, n)
    return (u0 * 1j + t) * np.exp(1j * angle)


@pytest.mark.parametrize("nlenses,lp", [(2, BIN), (3, TRI)])
def test_images_match_oracle(gpu, nlenses, lp):
    g = np.random.default_rng(nlenses)
    w = g.uniform(-1.5, 1.5, 3000) + 1j * g.uniform(-1.5, 1.5, 3000)
    z, m = ca.images_point_source(w, nlenses=nlenses, **lp)
    z, m = z.cpu().numpy(), m.cpu().numpy()
    zo, mo = lens.images_point_source(w, nlenses=nlenses, init_mode=1, **lp)
    c = lens.poly_coeffs(w, nlenses, **lp)[:, ::-1]
    ok, zm, _ = roots_close(z.T, zo.T, c)
    assert ok.all()
    # the 1e-6 filter on matched roots decides alike (except residuals at the threshold)
    agree = np.mean(np.sort(m.T, axis=1) == np.sort(mo.T, axis=1))
    assert agree > 0.999
    assert m.shape == (lens.poly_coeffs(w[:1], nlenses, **lp).shape[-1] - 1, 3000)


def test_images_custom_init_keeps_order(gpu):
    w = _track(200)
    zo, _ = lens.images_point_source(w, nlenses=3, init_mode=1, **TRI)
    g = np.random.default_rng(1)
    init = zo + 1e-5 * (g.normal(size=zo.shape) + 1j * g.normal(size=zo.shape))
    z, _ = ca.images_point_source(w, nlenses=3, custom_init=True, z_init=init, **TRI)
    z = z.cpu().numpy()
    assert np.max(np.abs(z - zo)) < 1e-10   # every root converges back to its own row


@pytest.mark.parametrize("nlenses,lp", [(2, BIN), (3, TRI)])
def test_sequential_matches_oracle_and_follows_images(gpu, nlenses, lp):
    w = _track(800)
    z, m = ca.images_point_source_sequential(w, nlenses=nlenses, **lp)
    z, m = z.cpu().numpy(), m.cpu().numpy()
    zo, mo = lens.images_point_source_sequential(w, nlenses=nlenses, init_mode=1, **lp)
    c = lens.poly_coeffs(w, nlenses, **lp)[:, ::-1]
    ok, _, _ = roots_close(z.T, zo.T, c)
    assert ok.all()
    # same row order as the oracle's warm-started sequence (root j follows one image)
    same = np.mean(np.all(np.abs(z - zo) <= 1e-9 * np.maximum(1, np.abs(zo)), axis=0))
    print(f"nlenses={nlenses}: positions with the oracle's row order {same:.4f}")
    assert same > 0.99
    assert np.mean(m == mo) > 0.999
    # continuity: each row's step is the smallest of the steps to the next roots
    step = np.abs(np.diff(z, axis=1))
    alt = np.abs(z[None, :, 1:] - z[:, None, :-1]).min(axis=1)
    assert np.mean(step <= alt + 1e-12) > 0.99


def test_sequential_batch_equals_single_sequences(gpu):
    ws = np.stack([_track(300, u0=u) for u in (0.1, 0.3, 0.5)])
    a = torch.tensor([0.45, 0.5, 0.4], dtype=torch.float64)
    e1 = torch.tensor([0.8, 0.7, 0.9], dtype=torch.float64)
    z, m = ca.images_point_source_sequential(ws, nlenses=2, a=a, e1=e1)
    for b in range(3):
        zb, mb = ca.images_point_source_sequential(ws[b], nlenses=2, a=float(a[b]),
                                                   e1=float(e1[b]))
        assert torch.equal(z[b], zb) and torch.equal(m[b], mb)


def test_single_lens_closed_form(gpu):
    w = _track(50)
    z, m = ca.images_point_source_sequential(w, nlenses=1)
    zo, _ = lens.images_point_source(w, nlenses=1)
    np.testing.assert_allclose(z.cpu().numpy(), zo, rtol=1e-14, atol=1e-15)
    assert m.all()


def test_python_numbers_and_lists_stay_64_bit(gpu):
    # python floats, complexes and lists must not pass through float32 on the way in
    w = list(_track(64))
    z1, _ = ca.images_point_source(w, nlenses=3, **TRI)
    z2, _ = ca.images_point_source(np.asarray(w), nlenses=3,
                                   **{k: np.full(64, v) for k, v in TRI.items()})
    assert torch.equal(z1, z2)
    z3, _ = ca.images_point_source_sequential(w, nlenses=2, a=[0.45], e1=[0.8])
    z4, _ = ca.images_point_source_sequential(np.asarray(w), nlenses=2, a=0.45, e1=0.8)
    assert torch.equal(z3, z4)
    zz = torch.tensor(w, dtype=torch.complex128, device="cuda")
    l1 = ca.lens_eq(zz, nlenses=3, **TRI)
    l2 = ca.lens_eq(zz, nlenses=3, a=0.6, e1=0.5, e2=0.3,
                    r3=torch.tensor(0.2 + 0.7j, dtype=torch.complex128, device="cuda"))
    assert torch.equal(l1, l2)
