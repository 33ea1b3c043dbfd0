"""Flux-marginalised log-likelihood kernel (caus_marginalized_log_likelihood) against the
oracle (oracle/linalg.py, a restatement of src/caustics/linalg.py:56-70), its exact
scale and shift identities at full size, and its gradient chained into the light-curve
VJP against finite differences."""

import numpy as np
import pytest
import torch

from oracle import linalg as OL

pytestmark = pytest.mark.gpu


def _curves(g, B, lens, sigma=10.0):
    """B magnification rows (binary-like Paczynski curves with per-row u0, tE) over
    ragged observatory segments, and one set of observations generated from row 0."""
    n = sum(lens)
    t = np.sort(g.uniform(-40, 40, n))
    u0 = g.uniform(0.05, 0.6, B)[:, None]
    tE = g.uniform(8, 40, B)[:, None]
    u = np.hypot(u0, t[None, :] / tE)
    A = (u**2 + 2) / (u * np.sqrt(u**2 + 4))
    Fs, Fb = 800.0, 200.0
    f = Fs * A[0] + Fb + g.normal(0, sigma, n)
    c = 1.0 / (sigma * g.uniform(0.7, 1.3, n)) ** 2
    off = np.concatenate([[0], np.cumsum(lens)]).astype(np.int64)
    return A, f, c, off


def _kernel(A, f, c, off, g_ll=None, want_grad=True):
    from caustics_amd import _lib
    dev = "cuda"
    At = torch.as_tensor(A, device=dev).contiguous()
    B, n = At.shape
    nseg = len(off) - 1
    ft, ct = torch.as_tensor(f, device=dev), torch.as_tensor(c, device=dev)
    ot = torch.as_tensor(off, device=dev)
    ll = torch.empty(B, dtype=torch.float64, device=dev)
    beta = torch.empty((B, nseg, 2), dtype=torch.float64, device=dev)
    gA = torch.empty((B, n), dtype=torch.float64, device=dev) if want_grad else None
    gt = None if g_ll is None else torch.as_tensor(g_ll, device=dev)
    rc = _lib.lib().caus_marginalized_log_likelihood(
        _lib.stream_ptr(), _lib.ptr(At), B, n, n, _lib.ptr(ft), _lib.ptr(ct), _lib.ptr(ot), nseg,
        _lib.ptr(ll), _lib.ptr(beta), _lib.ptr(gt), _lib.ptr(gA), n)
    assert rc == 0
    torch.cuda.synchronize()
    return ll.cpu().numpy(), beta.cpu().numpy(), None if gA is None else gA.cpu().numpy()


def test_matches_oracle_ragged_segments():
    g = np.random.default_rng(0)
    lens = [500, 300, 1200]
    A, f, c, off = _curves(g, 64, lens)
    g_ll = g.uniform(0.5, 2.0, 64)
    ll, beta, gA = _kernel(A, f, c, off, g_ll=g_ll)
    segs = [slice(off[k], off[k + 1]) for k in range(3)]
    worst = [0.0, 0.0, 0.0]
    for b in range(64):
        bl, llo = OL.marginalized_log_likelihood([A[b, s] for s in segs], [f[s] for s in segs],
                                                 [c[s] for s in segs])
        go = np.concatenate(OL.grad_A([A[b, s] for s in segs], [f[s] for s in segs],
                                      [c[s] for s in segs])) * g_ll[b]
        worst[0] = max(worst[0], abs(ll[b] - llo) / abs(llo))
        worst[1] = max(worst[1], max(np.max(np.abs(beta[b, k] - bl[k]) / np.abs(bl[k]))
                                     for k in range(3)))
        worst[2] = max(worst[2], np.max(np.abs(gA[b] - go)) / np.max(np.abs(go)))
    print(f"likelihood vs oracle: ll {worst[0]:.2e}, beta {worst[1]:.2e}, grad {worst[2]:.2e}")
    assert worst[0] < 1e-10
    assert worst[1] < 1e-10
    assert worst[2] < 1e-9


def test_exact_fit_and_singular_segment():
    # segment 0: exact fit A = f = (1, 2) -> ll = log(2 pi); segment 1: one epoch, the
    # normal matrix is singular and ll is not finite (as the reference's solve gives)
    A = np.array([[1.0, 2.0]])
    ll, beta, _ = _kernel(A, np.array([1.0, 2.0]), np.array([1.0, 1.0]), np.array([0, 2]))
    assert ll[0] == pytest.approx(np.log(2 * np.pi), rel=1e-14)
    assert np.allclose(beta[0, 0], [1.0, 0.0], atol=1e-14)
    A = np.array([[1.0, 2.0, 3.0]])
    ll, _, _ = _kernel(A, np.array([1.0, 2.0, 5.0]), np.ones(3), np.array([0, 2, 3]),
                       want_grad=False)
    assert not np.isfinite(ll[0])


def test_empty_batch_and_bad_arguments():
    from caustics_amd import _lib
    L = _lib.lib()
    s = _lib.stream_ptr()
    assert L.caus_marginalized_log_likelihood(s, None, 0, 10, 10, None, None, None, 1, None,
                                              None, None, None, 10) == 0
    x = torch.zeros(4, dtype=torch.float64, device="cuda")
    # lda < n, nseg < 1
    assert L.caus_marginalized_log_likelihood(s, _lib.ptr(x), 1, 4, 3, _lib.ptr(x), _lib.ptr(x),
                                              _lib.ptr(x), 1, _lib.ptr(x), None, None, None,
                                              4) == -1
    assert L.caus_marginalized_log_likelihood(s, _lib.ptr(x), 1, 4, 4, _lib.ptr(x), _lib.ptr(x),
                                              _lib.ptr(x), 0, _lib.ptr(x), None, None, None,
                                              4) == -1


def test_full_size_scale_and_shift_identities():
    # 1024 curves x 1500 epochs in two segments.  ll(lambda A) = ll(A) - nseg log lambda
    # and ll(A + mu) = ll(A) on every segment, so per curve
    #   sum_i A_i dll/dA_i = -nseg  and  sum_{i in seg} dll/dA_i = 0.
    g = np.random.default_rng(1)
    A, f, c, off = _curves(g, 1024, [900, 600])
    ll, _, gA = _kernel(A, f, c, off)
    scale = (gA * A).sum(1)
    scale_mag = np.abs(gA * A).sum(1)   # the sum cancels from this size down to 2
    shift = np.stack([gA[:, off[k]:off[k + 1]].sum(1) for k in range(2)], 1)
    shift_mag = np.stack([np.abs(gA[:, off[k]:off[k + 1]]).sum(1) for k in range(2)], 1)
    e_scale = (np.abs(scale + 2.0) / scale_mag).max()
    e_shift = (np.abs(shift) / shift_mag).max()
    print(f"scale identity rel err {e_scale:.2e}, shift identity rel err {e_shift:.2e}")
    assert e_scale < 1e-11
    assert e_shift < 1e-11
    ll2, _, _ = _kernel(A * 3.0, f, c, off, want_grad=False)
    assert np.allclose(ll2, ll - 2 * np.log(3.0), rtol=1e-12, atol=1e-10)
    # rows are independent and deterministic
    ll3, _, gA3 = _kernel(A[::-1].copy(), f, c, off)
    assert np.array_equal(ll3[::-1], ll) and np.array_equal(gA3[::-1], gA)


def test_autograd_api_and_lightcurve_chain():
    import caustics_amd as ca
    # reference signature, unbatched, two observatories
    g = np.random.default_rng(2)
    A, f, c, off = _curves(g, 1, [120, 80])
    segs = [slice(off[k], off[k + 1]) for k in range(2)]
    At = torch.tensor(A[0], device="cuda", requires_grad=True)
    beta, ll = ca.marginalized_log_likelihood([At[s] for s in segs], [f[s] for s in segs],
                                              [c[s] for s in segs])
    ll.backward()
    bl, llo = OL.marginalized_log_likelihood([A[0, s] for s in segs], [f[s] for s in segs],
                                             [c[s] for s in segs])
    assert float(ll) == pytest.approx(llo, rel=1e-10)
    assert np.allclose(beta[1].cpu().numpy(), bl[1], rtol=1e-10)
    go = np.concatenate(OL.grad_A([A[0, s] for s in segs], [f[s] for s in segs],
                                  [c[s] for s in segs]))
    assert np.allclose(At.grad.cpu().numpy(), go, rtol=1e-9, atol=1e-9 * np.abs(go).max())
    with pytest.raises(NotImplementedError):
        ca.marginalized_log_likelihood([At], [f], [c], dense_covariance=True)

    # ll of binary point-source light curves, differentiated through the fused
    # light-curve VJP, against central differences of the forward pipeline
    t = torch.linspace(-30.0, 30.0, 400, dtype=torch.float64, device="cuda")
    p0 = dict(t0=1.5, u0=0.35, tE=18.0, alpha=0.7, s=3.0, q=1e-3)  # no caustic near the track
    A_true = ca.mag_point_source_lightcurve(t, **p0)[0].detach().cpu().numpy()
    fo = 700.0 * A_true + 150.0 + g.normal(0, 8.0, 400)
    ci = np.full(400, 1.0 / 64.0)

    def ll_of(p, grad=False):
        ps = {k: torch.tensor([v], dtype=torch.float64, device="cuda", requires_grad=grad)
              for k, v in p.items()}
        Am = ca.mag_point_source_lightcurve(t, **ps)
        _, lv = ca.marginalized_log_likelihood([Am], [fo], [ci])
        if grad:
            lv.sum().backward()
            return float(lv.detach().sum()), {k: float(v.grad[0]) for k, v in ps.items()}
        return float(lv.detach().sum())

    p1 = dict(p0, t0=1.4, u0=0.37)
    _, gr = ll_of(p1, grad=True)
    for k in p1:
        h = 1e-6 * max(abs(p1[k]), 0.1)
        fd = (ll_of(dict(p1, **{k: p1[k] + h})) - ll_of(dict(p1, **{k: p1[k] - h}))) / (2 * h)
        print(f"{k}: autograd {gr[k]:.6e} fd {fd:.6e}")
        assert gr[k] == pytest.approx(fd, rel=1e-5, abs=1e-6 * max(abs(v) for v in gr.values()))
