"""GPU root solver vs the C oracle (oracle/ea_oracle.c) through the C ABI.

Bar: matched root sets within 1e-12 relative, or 64*kappa*eps for ill-conditioned
roots (kappa = sum|p_k||z|^k/(|z||p'(z)|)); with the reference's initial estimates
the root ORDER equals the oracle's for (almost) every polynomial."""
import ctypes

import numpy as np
import pytest
import torch

import caustics_amd as ca
from caustics_amd import _lib
from caustics_amd import workloads as W
from caustics_amd.ehrlich_aberth_primitive import solve_with_stats
from oracle import ea, lens

from helpers import residual_ratio, roots_close

pytestmark = pytest.mark.gpu


def lens_polys(nlenses, n, seed):
    lf, w, _ = W.batched_polys(nlenses, n, seed)
    return W.coeffs_product(w, lf["a"], lf["e1"], lf["e2"], lf["r3"], nlenses)


@pytest.mark.parametrize("nlenses,n", [(2, 20000), (3, 6000)])
@pytest.mark.parametrize("mode", ["fp64", "fp64_comp", "cf32"])
@pytest.mark.parametrize("init", ["reference", "complex"])
def test_lens_roots_match_oracle(gpu, nlenses, n, mode, init):
    if mode == "cf32" and init == "reference":
        pytest.skip("cf32 always uses complex initial estimates")
    c = lens_polys(nlenses, n, seed=100 + nlenses)
    z_or, _, st_or = ea.roots_asc(c, compensated=(mode == "fp64_comp"),
                                   init_mode=0 if init == "reference" else 1, return_stats=True)
    z, status, stats = solve_with_stats(c, itmax=2000, mode=mode, init=init)
    z = z.cpu().numpy()
    status = status.cpu().numpy()
    print(f"[{nlenses=} {mode=} {init=}] unconverged gpu={int((status == 0).sum())} "
          f"oracle={int((st_or == 0).sum())}")
    assert status.mean() > 0.999
    ok, zm, err = roots_close(z, z_or, c)
    good = status.astype(bool) & st_or.astype(bool)
    assert ok[good].all(), (err[good][~ok[good]][:10])
    assert np.max(residual_ratio(c, z)[good]) < 1e-13
    if init == "reference" and mode == "fp64":
        same_order = np.all(zm == z, axis=1)
        assert same_order.mean() > 0.99


def test_reference_fixture_real_coefficients(gpu):
    # tests/test_ehrlich_aberth_primitive.py:19-35: real source positions give real
    # coefficients; the reference init escapes the real axis through exp(2 log z)
    w = np.linspace(0.3, 0.35, 10).astype(complex)
    c = lens.poly_coeffs_binary(w, 0.45, 0.8).reshape(5, 2, 6)
    for comp in (False, True):
        r = ca.poly_roots(c, compensated=comp).cpu().numpy()
        assert r.shape == (5, 2, 5)
        for idx in [(3, 1), (4, 0)]:
            np.testing.assert_allclose(np.polyval(c[idx], r[idx]), 0, atol=1e-10)
        ro = ea.poly_roots(c, compensated=comp)
        ok, _, _ = roots_close(r.reshape(-1, 5), ro.reshape(-1, 5), c.reshape(-1, 6)[:, ::-1])
        assert ok.all()


def test_batching_shapes_equivalent(gpu):
    # vmap over either axis in the reference reduces to the flat batch
    w = np.linspace(0.3, 0.35, 10).astype(complex) + 0.05j
    c = lens.poly_coeffs_binary(w, 0.45, 0.8).reshape(5, 2, 6)
    r = ca.poly_roots(c).cpu().numpy()
    r2 = ca.poly_roots(np.swapaxes(c, 0, 1)).cpu().numpy()
    np.testing.assert_array_equal(np.swapaxes(r2, 0, 1), r)
    flat = ca.ehrlich_aberth(c.reshape(-1, 6)[:, ::-1].copy()).cpu().numpy()
    np.testing.assert_array_equal(flat, r.reshape(-1))


@pytest.mark.parametrize("deg", list(range(1, 13)))
def test_all_degrees_known_roots(gpu, deg):
    g = np.random.default_rng(deg)
    zs = g.normal(size=(2000, deg)) + 1j * g.normal(size=(2000, deg))
    c = np.array([np.poly(z) for z in zs])[:, ::-1].copy()
    for init in ("reference", "complex"):
        z, status, _ = solve_with_stats(c, mode="fp64", init=init)
        assert status.cpu().numpy().all()
        ok, _, err = roots_close(z.cpu().numpy(), zs, c, rel=1e-10)
        assert ok.all()


@pytest.mark.parametrize("deg", [2, 3, 5, 7, 10, 12])
def test_initial_estimates_match_oracle_hull(gpu, deg):
    """itmax = 0 returns the initial estimates.  Both forms place them on the upper convex
    hull of (i, log|p_i|) (init_est.h:57-102): the reference form by the monotone chain,
    the Fast form by its compile-time-indexed vertex test.  Against the oracle's
    restatement (ora_init_est, init_mode 0 real-axis, 1 complex): a wrong hull vertex or
    slot assignment moves an estimate by O(1), rounding by ~1e-15.  Coefficient
    magnitudes span 1e-13..1e13 with exact zeros, so hulls have many edges."""
    g = np.random.default_rng(100 + deg)
    n = 4000
    mag = np.exp(g.normal(0, 10, (n, deg + 1)))
    mag[g.random((n, deg + 1)) < 0.1] = 0.0
    mag[:, 0] = np.maximum(mag[:, 0], 1e-3)
    mag[:, deg] = np.maximum(mag[:, deg], 1e-3)
    c = mag * np.exp(2j * np.pi * g.random((n, deg + 1)))
    for init, mode in (("reference", 0), ("complex", 1)):
        z, _, _ = solve_with_stats(c, itmax=0, mode="fp64", init=init)
        zo = ea.roots_asc(c, itmax=0, init_mode=mode)
        rel = np.abs(z.cpu().numpy() - zo) / np.abs(zo)
        print(f"deg {deg} {init}: max rel diff {rel.max():.2e}")
        assert rel.max() <= 1e-12


def _known_root_polys(n, deg, seed, scale=1.0):
    g = np.random.default_rng(seed)
    zs = scale * (g.normal(size=(n, deg)) + 1j * g.normal(size=(n, deg)))
    c = np.zeros((n, deg + 1), complex)
    c[:, 0] = 1.0
    for k in range(deg):   # highest-first product (z - r_0)...(z - r_k)
        c[:, 1:k + 2] = c[:, 1:k + 2] - zs[:, k:k + 1] * c[:, 0:k + 1]
    return zs, c[:, ::-1].copy()   # ascending


@pytest.mark.parametrize("deg,mode,init", [(5, "fp64", "reference"), (5, "fp64_comp", "reference"),
                                           (5, "fp64_comp", "complex"), (4, "fp64", "complex"),
                                           (5, "fp64", "complex")])
def test_persistent_refill_large_batch(gpu, deg, mode, init):
    # more polynomials than the persistent kernel has lanes (8 workgroups of 256 per
    # compute unit), so lanes refill with second polynomials: the reference-compatible
    # and compensated forms, and the Fast form below degree 5, run k_roots_persistent;
    # the degree-5 Fast form runs the compacted kernel (one polynomial per lane)
    cus = torch.cuda.get_device_properties(0).multi_processor_count
    n = 8 * cus * 256 + 12345
    zs, c = _known_root_polys(n, deg, seed=7 + deg)
    z, status, _ = solve_with_stats(c, mode=mode, init=init)
    z, status = z.cpu().numpy(), status.cpu().numpy().astype(bool)
    # the reference's compensated stage leaves a few random polynomials unconverged
    # (its two termination tests can cycle): as many as the oracle, not more
    _, _, st_or = ea.roots_asc(c, compensated=(mode == "fp64_comp"),
                               init_mode=0 if init == "reference" else 1, return_stats=True,
                               nthreads=8)
    n_bad, n_bad_or = int((~status).sum()), int((st_or == 0).sum())
    print(f"[{deg=} {mode=} {init=}] unconverged gpu {n_bad}, oracle {n_bad_or} of {n}")
    assert n_bad <= 2 * n_bad_or + 10
    assert np.max(residual_ratio(c[status], z[status])) < 1e-13
    g = np.random.default_rng(1)
    pick = np.concatenate([np.arange(200), n - 200 + np.arange(200),
                           g.choice(n, 2000, replace=False)])
    pick = pick[status[pick]]
    ok, _, _ = roots_close(z[pick], zs[pick], c[pick], rel=1e-10)
    assert ok.all()


@pytest.mark.parametrize("case", ["tiny_coeffs", "huge_coeffs", "small_roots"])
@pytest.mark.parametrize("mode,init", [("fp64", "reference"), ("fp64", "complex"),
                                       ("fp64_comp", "reference"), ("cf32", "complex")])
def test_extreme_scales_match_oracle(gpu, case, mode, init):
    """Coefficients scaled by 1e-170 or 1e150, and degree-10 polynomials whose roots
    are ~1e-17: the squares of moduli leave the normal range, where the reference's
    hypot-based |.| and scaled division still work (the oracle restates them)."""
    if case == "small_roots":
        zs, c = _known_root_polys(3000, 10, seed=21, scale=1e-17)
    else:
        c = lens_polys(2, 3000, seed=22) * (1e-170 if case == "tiny_coeffs" else 1e150)
        zs = None
    z_or, _, st_or = ea.roots_asc(c, compensated=(mode == "fp64_comp"),
                                   init_mode=0 if init == "reference" else 1, return_stats=True)
    assert st_or.all()
    z, status, _ = solve_with_stats(c, itmax=2000, mode=mode, init=init)
    z = z.cpu().numpy()
    assert status.cpu().numpy().all()
    assert np.all(np.isfinite(z))
    ok, _, err = roots_close(z, z_or, c)
    print(f"[{case} {mode} {init}] max rel err vs oracle {err.max():.2e}")
    assert ok.all()
    if zs is not None:
        ok, _, _ = roots_close(z, zs, c, rel=1e-10)
        assert ok.all()


def test_compacted_sweep_keeps_root_slots(gpu):
    # degree 10 runs the compacted sweep (converged roots partitioned behind the others
    # between sweeps); the output must still put every root in the slot of the initial
    # estimate it came from.  Alternate slots start at the root (they pass the stopping
    # test at their first step and are partitioned away after sweep 1) or 1e-2 from it
    # (several sweeps), so every solve runs partitions; each root must come back in its
    # own slot.
    g = np.random.default_rng(31)
    n, deg = 4000, 10
    zs = g.normal(size=(n, deg)) + 1j * g.normal(size=(n, deg))
    c = np.array([np.poly(z) for z in zs])[:, ::-1].copy()
    scale = np.where(np.arange(deg) % 2 == 0, 1e-2, 1e-15)
    init = zs + scale * (g.normal(size=zs.shape) + 1j * g.normal(size=zs.shape))
    z, status, stats = solve_with_stats(c, mode="fp64", init="complex", roots_init=init)
    z, st = z.cpu().numpy(), stats.cpu().numpy()
    assert status.cpu().numpy().all()
    assert np.mean(st[:, 0] >= 2) > 0.9          # solves of several sweeps
    assert np.mean(st[:, 1] < deg * st[:, 0]) > 0.9   # ... with roots converged early
    err = (np.abs(z - zs) / np.abs(zs)).max(axis=1)
    assert np.mean(err < 1e-8) > 0.98
    ok, _, _ = roots_close(z, zs, c, rel=1e-10)
    assert ok.all()


def test_custom_init(gpu):
    c = lens_polys(2, 4000, seed=5)
    z0 = ea.roots_asc(c)
    g = np.random.default_rng(0)
    init = z0 + 1e-4 * (g.normal(size=z0.shape) + 1j * g.normal(size=z0.shape))
    zo = ea.roots_asc(c, roots_init=init)
    zg = ca.ehrlich_aberth(c, roots_init=init, custom_init=True).cpu().numpy().reshape(zo.shape)
    ok, zm, _ = roots_close(zg, zo, c)
    assert ok.all()
    assert np.mean(np.all(zm == zg, axis=1)) > 0.99  # warm start keeps the order


def test_non_convergence_reported(gpu):
    c = lens_polys(2, 1000, seed=9)
    _, status, stats = solve_with_stats(c, itmax=1, mode="fp64", init="complex")
    assert not status.cpu().numpy().any()
    assert (stats.cpu().numpy()[:, 0] == 1).all()


def test_descriptor_entry_point(gpu):
    c = lens_polys(2, 1000, seed=13)
    ct = torch.as_tensor(c, device="cuda")
    out = torch.zeros((1000, 5), dtype=torch.complex128, device="cuda")
    desc = struct_bytes = (ctypes.c_int * 3)(1000, 5, 2000)
    raw = bytes(desc) + bytes([0]) + bytes(3)  # int size, deg, itmax; bool compensated; pad
    bufs = (ctypes.c_void_p * 2)(ct.data_ptr(), out.data_ptr())
    rc = _lib.lib().caus_gpu_ehrlich_aberth(_lib.stream_ptr(), bufs, raw, len(raw))
    assert rc == 0
    z = ca.ehrlich_aberth(c).cpu().numpy().reshape(1000, 5)
    np.testing.assert_array_equal(out.cpu().numpy(), z)
    # the void-returning form of the reference target's signature
    L = _lib.lib()
    out.zero_()
    L.caus_gpu_ehrlich_aberth_void(_lib.stream_ptr(), bufs, raw, len(raw))
    assert L.caus_gpu_ehrlich_aberth_status() == 0
    np.testing.assert_array_equal(out.cpu().numpy(), z)
    L.caus_gpu_ehrlich_aberth_void(_lib.stream_ptr(), bufs, raw, len(raw) - 1)   # bad size
    assert L.caus_gpu_ehrlich_aberth_status() == _lib.CAUS_ERR_ARG
    assert L.caus_gpu_ehrlich_aberth_status() == 0   # read resets


def test_roots_vjp_matches_implicit_derivative_and_fd(gpu):
    # well-separated roots, so that a 1e-6 coefficient step stays in the linear regime
    g0 = np.random.default_rng(21)
    zs = g0.normal(size=(64, 5)) + 1j * g0.normal(size=(64, 5))
    c = np.array([np.poly(z) for z in zs])[:, ::-1].copy()
    ct = torch.as_tensor(c, device="cuda").requires_grad_(True)
    z = ca.ehrlich_aberth(ct).reshape(64, 5)
    g = np.random.default_rng(1)
    v = g.normal(size=(64, 5)) + 1j * g.normal(size=(64, 5))
    loss = (torch.as_tensor(v, device="cuda") * z).real.sum()
    loss.backward()
    grad = ct.grad.cpu().numpy()
    # closed form: dL/dRe(p_k) + i dL/dIm(p_k) with dz/dp_k = -z^k/p'(z)
    zn = z.detach().cpu().numpy()
    k = np.arange(6)
    dp = (k[1:] * c[:, None, 1:] * zn[..., None] ** (k[1:] - 1)).sum(-1)
    J = -(zn[..., None] ** k) / dp[..., None]           # (N, 5, 6)
    gre = np.real(v[..., None] * J).sum(1)
    gim = np.real(v[..., None] * 1j * J).sum(1)
    np.testing.assert_allclose(grad, gre + 1j * gim, rtol=1e-9, atol=1e-12)
    # the custom-call-shaped target: buffers {coeffs, roots, g_roots, g_coeffs}; the
    # cotangent of z for L = Re(v z) is conj(v) in torch's convention
    zt = z.detach().contiguous()
    gz = torch.as_tensor(np.conj(v), device="cuda")
    gc = torch.zeros_like(ct)
    bufs = (ctypes.c_void_p * 4)(ct.data_ptr(), zt.data_ptr(), gz.data_ptr(), gc.data_ptr())
    raw = bytes((ctypes.c_int * 3)(64, 5, 2000)) + bytes(4)
    L = _lib.lib()
    L.caus_gpu_ehrlich_aberth_vjp_void(_lib.stream_ptr(), bufs, raw, len(raw))
    assert L.caus_gpu_ehrlich_aberth_status() == 0
    np.testing.assert_array_equal(gc.cpu().numpy(), grad)
    # central finite differences through the oracle solver (h = 1e-6)
    h = 1e-6
    for kk in (0, 3, 5):
        cp, cm = c.copy(), c.copy()
        cp[:, kk] += h
        cm[:, kk] -= h
        zp, zm_ = ea.roots_asc(cp), ea.roots_asc(cm)
        from helpers import match_rows
        zp, zm_ = match_rows(zn, zp), match_rows(zn, zm_)
        fd = np.real(v * (zp - zm_)).sum(1) / (2 * h)
        np.testing.assert_allclose(grad[:, kk].real, fd, rtol=1e-4, atol=1e-6)


def test_cf32_nearly_double_roots_regression(gpu):
    """Limb points of a triple-lens source crossing a caustic (s=2.59, q=0.20, q3=0.022):
    their polynomials have a nearly double root that FP32 cannot resolve, which used to
    turn the CF32 form's Aberth sums infinite and return NaN roots flagged converged.
    The CF32 result must be finite and match the FP64 form's root sets."""
    from caustics_amd.ehrlich_aberth_primitive import solve_with_stats

    prm = dict(s=2.5915722237180048, q=0.20148953757052693, q3=0.022476073987355016,
               r3=1.0905260567197013, psi=3.2357322003521114)
    w0, rho = 0.16212946846468673 - 0.9840604288897297j, 0.003657724979001628
    lp, x_cm = lens.lens_params(3, **prm)
    r, e = lens.lens_positions_masses(3, **lp)
    th = np.linspace(-np.pi, np.pi, 2048)
    c = lens.poly_coeffs_product(w0 + x_cm + rho * np.exp(1j * th), r, e)[:, ::-1].copy()
    z64, st64, _ = solve_with_stats(c, mode="fp64", init="complex")
    z32, st32, _ = solve_with_stats(c, mode="cf32", init="complex")
    z64, z32 = z64.cpu().numpy(), z32.cpu().numpy()
    assert np.isfinite(z32).all() and st32.cpu().numpy().all()
    ok, _, _ = roots_close(z32, z64, c)
    assert ok.all()


def test_cf32_nan_iterates_stop_at_once(gpu):
    """Three degree-10 polynomials of the seeded triple-lens batch (W.batched_polys(3,
    2^18, seed=3), rows 81278, 102236, 248030) whose CF32 FP64 stage meets a NaN Aberth
    sum.  A NaN root must stop iterating at once and hand the polynomial to the FP64
    fallback, not run the stage's 2000 sweeps (one such lane held its wave ~7 ms)."""
    lf, w, _ = W.batched_polys(3, 1 << 18, seed=3)
    rows = np.array([81278, 102236, 248030])
    c = W.coeffs_product(w[rows], lf["a"][rows], lf["e1"][rows], lf["e2"][rows], lf["r3"][rows], 3)
    z64, st64, _ = solve_with_stats(c, mode="fp64", init="complex")
    z32, st32, stats = solve_with_stats(c, mode="cf32", init="complex")
    s = stats.cpu().numpy()
    print("CF32 stats (sweeps, FP64 iterations, compensated, FP32):", s.tolist())
    assert st32.cpu().numpy().all() and np.isfinite(z32.cpu().numpy()).all()
    assert (s[:, 1] < 1000).all()   # observed with the stuck stage: ~20,100
    ok, _, _ = roots_close(z32.cpu().numpy(), z64.cpu().numpy(), c)
    assert ok.all()
