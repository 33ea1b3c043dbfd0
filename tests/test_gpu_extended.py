"""Extended-source contour kernels vs the oracle (oracle/extended.py restating
extended_source.py / integrate.py) and vs closed forms.

Bars: every source within 1e-10 relative of the oracle (the largest deviation observed
is 5e-11), and within the reference tests' own rtol 1e-3 of the single-lens closed
forms."""
import numpy as np
import pytest
import torch
from scipy.integrate import quad

from caustics_amd.extended_source import mag_extended_source
from oracle import extended as E
from oracle import lens

pytestmark = pytest.mark.gpu


def single_lens_uniform(u, rho):
    """Uniform disk: (1/pi rho^2) int_0^{u+rho} (r^2+2)/sqrt(r^2+4) * arc(r) dr."""
    def arc(r):
        if r < abs(rho - u):
            return 2 * np.pi if u < rho else 0.0
        c = (r * r + u * u - rho * rho) / (2 * r * u)
        return 2 * np.arccos(np.clip(c, -1, 1))
    f = lambda r: (r * r + 2) / np.sqrt(r * r + 4) * arc(r)
    v, _ = quad(f, 0, u + rho, points=[abs(u - rho)], limit=200, epsabs=1e-13, epsrel=1e-12)
    return v / (np.pi * rho * rho)


def compare(m, mo, name):
    e = np.abs(m - mo) / np.abs(mo)
    print(f"{name}: n={len(m)} max rel err {e.max():.3e}, frac<=1e-10 {(e <= 1e-10).mean():.3f}")
    assert e.max() <= 1e-10


@pytest.mark.parametrize("rho", [1.0, 0.1, 0.01])
def test_single_lens_uniform(gpu, rho):
    w = np.linspace(0.0, 3.0 * rho, 11) + 1e-12
    m = mag_extended_source(w, rho, nlenses=1, npts_limb=150).cpu().numpy()
    mo = np.array([E.mag_extended_source(x, rho, nlenses=1, npts_limb=150) for x in w])
    compare(m, mo, f"single uniform rho={rho}")
    exact = np.array([single_lens_uniform(abs(x), rho) for x in w])
    np.testing.assert_allclose(m, exact, rtol=1e-3)


def test_single_lens_ld_reduces_to_uniform(gpu):
    rho = 0.1
    w = np.linspace(0.0, 3.0 * rho, 11) + 1e-12
    mu = mag_extended_source(w, rho, nlenses=1, npts_limb=150).cpu().numpy()
    ml = mag_extended_source(w, rho, nlenses=1, npts_limb=150, limb_darkening=True, u1=0.0,
                             npts_ld=50).cpu().numpy()
    np.testing.assert_allclose(ml, mu, rtol=1e-3)
    mo = np.array([E.mag_extended_source(x, rho, nlenses=1, npts_limb=150, limb_darkening=True,
                                         u1=0.7, npts_ld=100) for x in w])
    m7 = mag_extended_source(w, rho, nlenses=1, npts_limb=150, limb_darkening=True, u1=0.7,
                             npts_ld=100).cpu().numpy()
    compare(m7, mo, "single LD u1=0.7")


def caustic_points(rho, nlenses=2, npts=12, seed=42, **params):
    _, ca = lens.critical_and_caustic_curves(npts=npts, nlenses=nlenses, **params)
    ca = ca.reshape(-1)
    g = np.random.default_rng(seed)
    return ca + g.uniform(0, 2 * rho, ca.shape) * np.exp(1j * g.uniform(-np.pi, np.pi, ca.shape))


@pytest.mark.parametrize("rho", [1e-1, 1e-2, 1e-3])
def test_binary_uniform_near_caustics(gpu, rho):
    w = caustic_points(rho, s=0.9, q=0.2)
    m = mag_extended_source(w, rho, nlenses=2, npts_limb=300, s=0.9, q=0.2).cpu().numpy()
    mo = np.array([E.mag_extended_source(x, rho, nlenses=2, npts_limb=300, s=0.9, q=0.2)
                   for x in w])
    compare(m, mo, f"binary uniform rho={rho}")


def test_binary_ld_near_caustics(gpu):
    rho = 1e-2
    w = caustic_points(rho, npts=6, s=0.9, q=0.2)
    kw = dict(nlenses=2, npts_limb=300, limb_darkening=True, u1=0.7, npts_ld=100, s=0.9, q=0.2)
    m = mag_extended_source(w, rho, **kw).cpu().numpy()
    mo = np.array([E.mag_extended_source(x, rho, **kw) for x in w])
    compare(m, mo, "binary LD")


def test_triple_near_caustics(gpu):
    prm = dict(s=1.2, q=0.3, q3=0.01, r3=0.8, psi=1.0)
    rho = 1e-2
    w = caustic_points(rho, nlenses=3, npts=6, **prm)
    m = mag_extended_source(w, rho, nlenses=3, npts_limb=200, **prm).cpu().numpy()
    mo = np.array([E.mag_extended_source(x, rho, nlenses=3, npts_limb=200, **prm) for x in w])
    compare(m, mo, "triple uniform")


@pytest.mark.parametrize("nlenses", [2, 3])
def test_cf32_form_matches_fp64(gpu, nlenses):
    """The reduced-precision form (FP32 Aberth, compensated-FP32 sweep, FP64 polish) of
    the limb solves against the FP64 form, forward and VJP: within 1e-8 relative."""
    from caustics_amd.extended_source import extended_fwd_vjp

    if nlenses == 2:
        prm = dict(s=0.9, q=0.2)
    else:
        prm = dict(s=1.2, q=0.3, q3=0.01, r3=0.8, psi=1.0)
    rho = 1e-2
    w = caustic_points(rho, nlenses=nlenses, npts=6, **prm)
    kw = dict(nlenses=nlenses, npts_limb=256, limb_darkening=True, u1=0.5, npts_ld=100, **prm)
    g = np.ones(len(w))
    m64, g64, _ = extended_fwd_vjp(w, rho, g, mode="fp64", **kw)
    m32, g32, st = extended_fwd_vjp(w, rho, g, mode="cf32", **kw)
    m64, m32 = m64.cpu().numpy(), m32.cpu().numpy()
    g64, g32 = g64.cpu().numpy(), g32.cpu().numpy()
    assert st.cpu().numpy().all()
    e = np.abs(m32 - m64) / m64
    eg = np.abs(g32 - g64) / np.maximum(np.abs(g64), 1e-3 * m64[:, None])
    print(f"cf32 vs fp64 nl={nlenses}: mag max rel {e.max():.2e}, grad max rel {eg.max():.2e}")
    # both forms end each limb solve with the same canonical polish, which takes converged
    # roots to the doubles nearest the exact roots: the contours coincide (observed:
    # bit-identical)
    assert e.max() <= 1e-8
    assert eg.max() <= 1e-8
    # a larger batch across caustic crossings, clustered images included
    w2 = caustic_points(rho, nlenses=nlenses, npts=40, **prm)
    w2 = w2 + rho * 0.7 * np.exp(1j * np.linspace(0, 20, len(w2)))
    m64, _, _ = extended_fwd_vjp(w2, rho, np.ones(len(w2)), mode="fp64", **kw)
    m32, _, _ = extended_fwd_vjp(w2, rho, np.ones(len(w2)), mode="cf32", **kw)
    e2 = ((m32 - m64).abs() / m64).cpu().numpy()
    print(f"  {len(w2)} sources: frac <= 1e-8 {np.mean(e2 <= 1e-8):.4f}")
    assert e2.max() <= 1e-8


def test_long_contours_match_oracle(gpu):
    # npts_limb above 512: the gap sort of a refinement round runs on four waves
    rho = 2e-2
    w = caustic_points(rho, npts=4, s=0.8, q=0.5)[:12]
    m = mag_extended_source(w, rho, nlenses=2, npts_limb=1500, limb_darkening=True, u1=0.5,
                            s=0.8, q=0.5).cpu().numpy()
    mo = np.array([E.mag_extended_source(x, rho, nlenses=2, npts_limb=1500, limb_darkening=True,
                                         u1=0.5, s=0.8, q=0.5) for x in w])
    compare(m, mo, "binary LD npts_limb=1500")


def test_chunks_and_streams_do_not_change_results(gpu):
    # a batch cut into chunks that alternate over two streams, a ragged batch reordered
    # by size, and one launch sequence on one stream give the same bits
    from caustics_amd.extended_source import _run
    from caustics_amd import _lib
    g = np.random.default_rng(8)
    S = 3000
    w = torch.as_tensor(caustic_points(1e-2, npts=400, s=0.9, q=0.2)[:S], device="cuda")
    S = w.numel()
    rho = torch.as_tensor(10 ** g.uniform(-2.5, -1.5, S), device="cuda")
    prm = torch.as_tensor(np.tile([0.9, 0.2], (S, 1)), device="cuda").contiguous()
    u1 = torch.full((S,), 0.5, dtype=torch.float64, device="cuda")
    gm = torch.as_tensor(g.normal(size=S), device="cuda")
    npts = g.choice([64, 128, 256, 600], S).astype(np.int64)
    args = (2, w, rho, prm, npts, 600, True, u1, 40, 2500, _lib.MODE_FP64)
    one = _run(*args, g=gm, streams=1, max_chunk=S)
    many = _run(*args, g=gm, streams=2, max_chunk=257)
    torch.cuda.synchronize()
    for a, b in zip(one, many):
        assert torch.equal(a, b)
    # the tuner's knobs: workspace budget per chunk, three streams, and the flat kernels'
    # workgroups per CU (4 and 16), against the tuned table's schedule
    from caustics_amd import extended_source as X
    L = _lib.lib()
    try:
        for flat, cb, ns in ((4, 64 << 20, 3), (16, 8 << 30, 1)):
            _lib.check(L.caus_ext_set_flat_blocks_per_cu(flat), "schedule")
            X._flat_set[0] = flat
            alt = _run(*args, g=gm, streams=ns, chunk_bytes=cb)
            torch.cuda.synchronize()
            for a, b in zip(one, alt):
                assert torch.equal(a, b), (flat, cb, ns)
    finally:
        L.caus_ext_set_flat_blocks_per_cu(8)
        X._flat_set[0] = 8
    tuned = _run(*args, g=gm)
    for a, b in zip(one, tuned):
        assert torch.equal(a, b)
    assert L.caus_ext_set_flat_blocks_per_cu(0) == _lib.CAUS_ERR_ARG


def test_ragged_batch_matches_individual(gpu):
    rho = 2e-2
    w = caustic_points(rho, npts=5, s=1.1, q=0.3)
    npts = np.array([256, 300, 512, 150, 400] * (len(w) // 5 + 1))[:len(w)]
    m = mag_extended_source(w, rho, nlenses=2, npts_limb=npts, s=1.1, q=0.3).cpu().numpy()
    for i in range(len(w)):
        mi = mag_extended_source(w[i:i + 1], rho, nlenses=2, npts_limb=int(npts[i]), s=1.1,
                                 q=0.3).cpu().numpy()
        assert m[i] == mi[0]


def _fd(fun, x0, h):
    return (fun(x0 + h) - fun(x0 - h)) / (2 * h)


@pytest.mark.parametrize("ld", [False, True])
def test_extended_vjp_vs_finite_differences(gpu, ld):
    """d mag / d(Re w0, Im w0, rho, u1, s, q) near caustics against central differences
    of the forward kernel.  The discretised contour integral is only piecewise smooth
    (image counts and segment ends change as a limb point crosses the caustic), so the
    step is small (h = 1e-9, as the reference's test) and a point counts only where the
    differences at h and 2h agree to 1e-4; there the tolerance is the reference's rtol
    1e-3 (tests/test_extended_source.py:294-331)."""
    from caustics_amd.extended_source import extended_fwd_vjp

    s, q, rho, u1 = 0.9, 0.2, 1e-2, 0.7
    w = caustic_points(rho, npts=6, s=s, q=q)[::3]
    kw = dict(nlenses=2, npts_limb=300, limb_darkening=ld, npts_ld=100)
    mag, g, st = extended_fwd_vjp(w, rho, np.ones(len(w)), u1=u1, s=s, q=q, **kw)
    g = g.cpu().numpy()
    mag = mag.cpu().numpy()
    base = dict(w=w, rho=rho, s=s, q=q, u1=u1)

    def f(name, h):
        o = dict(base)
        if name == "re":
            o["w"] = w + h
        elif name == "im":
            o["w"] = w + 1j * h
        else:
            o[name] = base[name] + h
        return mag_extended_source(o["w"], o["rho"], u1=o["u1"], s=o["s"], q=o["q"],
                                   **kw).cpu().numpy()

    checks = {"re": 0, "im": 1, "rho": 2, "s": 4, "q": 5}
    if ld:
        checks["u1"] = 3
    for name, col in checks.items():
        h = 1e-9
        fd = (f(name, h) - f(name, -h)) / (2 * h)
        fd2 = (f(name, 2 * h) - f(name, -2 * h)) / (4 * h)
        stable = np.abs(fd - fd2) <= 1e-4 * np.maximum(np.abs(fd), 1e-3 * mag)
        err = np.abs(g[:, col] - fd) / np.maximum(np.abs(fd), 1e-3 * mag)
        print(f"{name}: stable {stable.mean():.2f}, max rel err vs FD {err[stable].max():.2e}")
        assert stable.mean() >= 0.75
        assert np.all(err[stable] < 2e-4)   # observed at most 2.1e-5


def test_extended_grad_reference_case(gpu):
    """The reference's own gradient test (tests/test_extended_source.py:294-331): the
    first caustic point of s=0.9, q=0.2 (50-point caustic), rho=1e-2, u1=0.7, 300 limb
    points, 100 LD nodes; jacobian in (s, q, rho, u1) against one-sided differences with
    epsilon 1e-9, rtol 1e-3."""
    from scipy.optimize import approx_fprime

    from caustics_amd import critical_and_caustic_curves
    from caustics_amd.extended_source import extended_fwd_vjp

    s, q, rho, u1 = 0.9, 0.2, 1e-2, 0.7
    _, ca = critical_and_caustic_curves(npts=50, nlenses=2, s=s, q=q)
    w0 = complex(ca.reshape(-1)[0].item())
    kw = dict(nlenses=2, npts_limb=300, limb_darkening=True, npts_ld=100)

    def fn(x):
        return mag_extended_source(np.array([w0]), x[2], u1=x[3], s=x[0], q=x[1],
                                   **kw).cpu().numpy()[0]

    x0 = np.array([s, q, rho, u1])
    jfd = approx_fprime(x0, fn, 1e-9)
    _, g, _ = extended_fwd_vjp(np.array([w0]), rho, np.ones(1), u1=u1, s=s, q=q, **kw)
    g = g.cpu().numpy()[0]
    jac = np.array([g[4], g[5], g[2], g[3]])
    err = np.abs((jac - jfd) / jfd)
    print("reference gradient case: rel err", err)
    assert np.all(err < 2e-4)   # observed at most 2.3e-5


def test_extended_autograd_matches_fused(gpu):
    import torch

    rho = 2e-2
    w = caustic_points(rho, npts=4, s=1.1, q=0.3)
    wt = torch.as_tensor(w, device="cuda").requires_grad_(True)
    st = torch.tensor(1.1, device="cuda", dtype=torch.float64, requires_grad=True)
    m = mag_extended_source(wt, rho, nlenses=2, npts_limb=200, s=st, q=0.3)
    m.sum().backward()
    from caustics_amd.extended_source import extended_fwd_vjp

    _, g, _ = extended_fwd_vjp(w, rho, np.ones(len(w)), nlenses=2, npts_limb=200, s=1.1, q=0.3)
    g = g.cpu().numpy()
    np.testing.assert_allclose(wt.grad.cpu().numpy(), g[:, 0] + 1j * g[:, 1], rtol=1e-12)
    np.testing.assert_allclose(st.grad.item(), g[:, 4].sum(), rtol=1e-12)


def test_mag_lightcurve_gate_and_contours(gpu):
    from caustics_amd import mag
    from oracle import multipole

    s, q, rho = 0.9, 0.2, 1e-2
    w = caustic_points(rho, npts=10, s=s, q=q)
    w = np.concatenate([w, w + 0.1])
    m = mag(w, rho, nlenses=2, npts_limb=200, s=s, q=q).cpu().numpy()
    mu, use = multipole.hexadecapole_gate(w, rho, nlenses=2, s=s, q=q)
    mo = np.where(use, mu, 0.0)
    for i in np.nonzero(~use)[0]:
        mo[i] = E.mag_extended_source(w[i], rho, nlenses=2, npts_limb=200, s=s, q=q)
    compare(m, mo, "mag() with gate")


def test_compensated_limb_solves(gpu):
    """roots_compensated=True (extended_source.py:750,839) runs the limb solves with the
    compensated stage; both forms end at the same canonical roots, so the contours and
    magnifications agree with the FP64 form."""
    prm = dict(s=0.9, q=0.2)
    rho = 1e-2
    w = caustic_points(rho, npts=6, **prm)
    kw = dict(nlenses=2, npts_limb=256, limb_darkening=True, u1=0.5, npts_ld=100, **prm)
    m64 = mag_extended_source(w, rho, **kw).cpu().numpy()
    mc = mag_extended_source(w, rho, roots_compensated=True, **kw).cpu().numpy()
    e = np.abs(mc - m64) / m64
    print(f"compensated vs plain limb solves: max rel {e.max():.2e}, identical {np.mean(mc == m64):.3f}")
    assert e.max() <= 1e-12
    from caustics_amd import mag
    mm = mag(w, rho, roots_compensated=True, **kw).cpu().numpy()
    assert np.all(np.isfinite(mm))


def test_skipped_sources_cost_nothing_and_leave_others_unchanged(gpu):
    """Sources marked to skip (npts_limb 0 for the kernels, set on the device) get
    magnification and gradient 0; every other source gives the same bits as without
    them."""
    from caustics_amd import _lib
    from caustics_amd.extended_source import _run

    rho = 1e-2
    w = torch.as_tensor(caustic_points(rho, npts=8, s=0.9, q=0.2), device="cuda")
    S = w.numel()
    rho_t = torch.full((S,), rho, dtype=torch.float64, device="cuda")
    prm = torch.as_tensor(np.tile([0.9, 0.2], (S, 1)), device="cuda").contiguous()
    u1 = torch.full((S,), 0.5, dtype=torch.float64, device="cuda")
    g = torch.ones(S, dtype=torch.float64, device="cuda")
    npts = np.random.default_rng(3).choice([128, 256, 300], S).astype(np.int64)
    skip = torch.as_tensor(np.arange(S) % 3 == 1, device="cuda")
    args = (2, w, rho_t, prm, npts, 300, True, u1, 50, 2500, _lib.MODE_FP64)
    m0, st0, g0 = _run(*args, g=g)
    m1, st1, g1 = _run(*args, g=g, skip=skip)
    keep = ~skip
    assert torch.equal(m1[keep], m0[keep]) and torch.equal(g1[keep], g0[keep])
    assert torch.all(m1[skip] == 0) and torch.all(g1[skip] == 0) and torch.all(st1 == 1)


def test_mag_autograd_matches_fused(gpu):
    """mag()'s backward reuses the Jacobian of its forward pass (per-point outputs): the
    gradients equal the fused forward+VJP of both branches."""
    from caustics_amd import mag
    from caustics_amd.lightcurve import mag_fwd_vjp

    s, q, rho = 0.9, 0.2, 1e-2
    w = caustic_points(rho, npts=6, s=s, q=q)
    w = np.concatenate([w, w + 0.15])
    wt = torch.as_tensor(w, device="cuda").requires_grad_(True)
    st = torch.tensor(s, dtype=torch.float64, device="cuda", requires_grad=True)
    rt = torch.tensor(rho, dtype=torch.float64, device="cuda", requires_grad=True)
    gm = torch.as_tensor(np.random.default_rng(2).normal(size=len(w)), device="cuda")
    m = mag(wt, rt, nlenses=2, npts_limb=256, limb_darkening=True, u1=0.5, s=st, q=q)
    (m * gm).sum().backward()
    mf, go, use = mag_fwd_vjp(w, rho, gm, nlenses=2, npts_limb=256, limb_darkening=True, u1=0.5,
                              s=s, q=q)
    assert torch.equal(m.detach(), mf)
    np.testing.assert_allclose(wt.grad.cpu().numpy(), (go[:, 0] + 1j * go[:, 1]).cpu().numpy(),
                               rtol=1e-13, atol=0)
    np.testing.assert_allclose(st.grad.item(), go[:, 4].sum().item(), rtol=1e-12)
    np.testing.assert_allclose(rt.grad.item(), go[:, 2].sum().item(), rtol=1e-12)
    assert use.any() and (~use).any()


@pytest.mark.parametrize("nlenses", [2, 3])
def test_workspace_contents_do_not_change_results(gpu, nlenses):
    """Every workspace read is of memory the same launch sequence wrote before: a
    workspace holding the 64-bit value 1 in every word gives the bits of a zero-filled
    one (a stale index would read 0 or 1 and a stale double a tiny finite number, so the
    check addresses nothing outside the workspace), forward and reverse mode, limb
    darkened, ragged, with a skipped source."""
    from caustics_amd import _lib
    from caustics_amd.extended_source import extended_workspace_bytes

    prm_d = dict(s=0.9, q=0.2) if nlenses == 2 else dict(s=1.2, q=0.3, q3=0.01, r3=0.8, psi=1.0)
    names = list(prm_d)
    rho = 1e-2
    w = caustic_points(rho, nlenses=nlenses, npts=6, **prm_d)
    S = len(w)
    g = np.random.default_rng(5)
    npts = g.choice([128, 256, 512], S).astype(np.int32)
    npts[1] = 0   # skipped
    nm = int(npts.max())
    wt = torch.as_tensor(w, dtype=torch.complex128, device="cuda")
    prm = torch.as_tensor(np.tile([prm_d[k] for k in names], (S, 1)), device="cuda").contiguous()
    rho_t = torch.full((S,), rho, dtype=torch.float64, device="cuda")
    u1 = torch.full((S,), 0.4, dtype=torch.float64, device="cuda")
    npts_t = torch.as_tensor(npts, device="cuda")
    gm = torch.as_tensor(g.normal(size=S), device="cuda")
    nb = extended_workspace_bytes(nlenses, S, nm)
    L = _lib.lib()
    out = []
    for fill in (0, 1):
        ws = torch.full(((nb + 7) // 8,), fill, dtype=torch.int64, device="cuda")
        mag = torch.empty(S, dtype=torch.float64, device="cuda")
        st = torch.empty(S, dtype=torch.uint8, device="cuda")
        go = torch.empty((S, 4 + len(names)), dtype=torch.float64, device="cuda")
        _lib.check(L.caus_mag_extended_source_work(
            _lib.stream_ptr(), nlenses, _lib.ptr(wt), S, _lib.ptr(rho_t), _lib.ptr(prm),
            len(names), _lib.ptr(npts_t), nm, 1, _lib.ptr(u1), 1, 50, _lib.ptr(gm),
            _lib.ptr(mag), _lib.ptr(go), _lib.ptr(st), _lib.ptr(ws), ws.numel() * 8, 2500,
            _lib.MODE_FP64, None), "caus_mag_extended_source_work")
        torch.cuda.synchronize()
        out.append((mag, go, st))
    for a, b in zip(*out):
        assert torch.equal(a, b)
    assert out[0][0][1].item() == 0 and bool(torch.isfinite(out[0][0]).all())
