"""Seeded synthetic inputs of the benchmark workloads (lens configurations,
source positions, light-curve parameters).

Every generator uses numpy ``Generator(PCG64(SeedSequence(seed)))`` so tests and
the benchmark see identical inputs.  Parameter maps follow the reference API
(src/caustics/point_source.py:1798-1817): a = s/2, e1 = 1/(1+q) for two lenses;
e1 = q/(1+q+q3), e2 = q*e1, r3 -> r3*exp(i*psi) for three; source positions are
shifted by x_cm = a(1-q)/(1+q) into the lens frame.
"""
import numpy as np


def rng(seed):
    return np.random.Generator(np.random.PCG64(np.random.SeedSequence(seed)))


def loguniform(g, lo, hi, n):
    return 10.0 ** g.uniform(np.log10(lo), np.log10(hi), n)


def lens_configs(g, n, nlenses):
    """Per-item user parameters: s, q (and q3, r3, psi)."""
    p = {"s": loguniform(g, 0.3, 3.0, n), "q": loguniform(g, 1e-4, 1.0, n)}
    if nlenses == 3:
        p["r3"] = g.uniform(0.3, 2.0, n)
        p["psi"] = g.uniform(0.0, 2 * np.pi, n)
        p["q3"] = loguniform(g, 1e-4, 1e-1, n)
    return p


def lens_frame(p, nlenses):
    """User parameters -> lens-frame (a, e1, e2, r3 complex) and x_cm, vectorised."""
    s, q = np.asarray(p["s"], float), np.asarray(p["q"], float)
    a = 0.5 * s
    x_cm = a * (1 - q) / (1 + q)
    if nlenses == 2:
        e1 = 1 / (1 + q)
        return dict(a=a, e1=e1, e2=1 - e1, r3=np.zeros_like(a, complex)), x_cm
    e1 = q / (1 + q + np.asarray(p["q3"], float))
    e2 = q * e1
    r3 = np.asarray(p["r3"], float) * np.exp(1j * np.asarray(p["psi"], float))
    return dict(a=a, e1=e1, e2=e2, r3=r3), x_cm


def disk(g, n, radius=2.0):
    return radius * np.sqrt(g.uniform(0, 1, n)) * np.exp(2j * np.pi * g.uniform(0, 1, n))


def batched_polys(nlenses, n, seed):
    """Batched root-solve workload: per polynomial a lens configuration and a source
    position uniform in |w| <= 2 (Einstein units), in the lens frame.
    Returns (lens-frame dict, w_lens_frame, user-parameter dict)."""
    g = rng(seed)
    p = lens_configs(g, n, nlenses)
    w = disk(g, n)
    lf, x_cm = lens_frame(p, nlenses)
    return lf, w + x_cm, p


def coeffs_product(w, a, e1, e2, r3, nlenses):
    """Lens polynomial coefficients, ascending, vectorised over items (numpy FP64).

    P(z) = (z - w) prod_j M_j - D sum_j e_j prod_{k!=j} M_k with D = prod_j (z - r_j),
    N = sum_j e_j prod_{k!=j} (z - r_j), M_j = (conj(w) - conj(r_j)) D + N — the
    polynomial of point_source.py:17-86 (two lenses) and :89-1479 (three lenses)."""
    w = np.asarray(w, complex)
    a = np.broadcast_to(np.asarray(a, float), w.shape)
    e1 = np.broadcast_to(np.asarray(e1, float), w.shape)
    if nlenses == 2:
        r = [a + 0j, -a + 0j]
        e = [e1, 1 - e1]
    else:
        e2 = np.broadcast_to(np.asarray(e2, float), w.shape)
        r = [a + 0j, -a + 0j, np.broadcast_to(np.asarray(r3, complex), w.shape)]
        e = [e1, e2, 1 - e1 - e2]

    def mul(p, q):
        out = np.zeros(w.shape + (p.shape[-1] + q.shape[-1] - 1,), complex)
        for i in range(p.shape[-1]):
            out[..., i:i + q.shape[-1]] += p[..., i:i + 1] * q
        return out

    def add(p, q):
        n = max(p.shape[-1], q.shape[-1])
        out = np.zeros(w.shape + (n,), complex)
        out[..., :p.shape[-1]] += p
        out[..., :q.shape[-1]] += q
        return out

    one = np.ones(w.shape + (1,), complex)
    m = len(r)
    lin = [np.concatenate([-r[j][..., None], one], -1) for j in range(m)]
    D = one
    for j in range(m):
        D = mul(D, lin[j])
    Nn = np.zeros(w.shape + (1,), complex)
    for j in range(m):
        t = e[j][..., None] * one
        for k in range(m):
            if k != j:
                t = mul(t, lin[k])
        Nn = add(Nn, t)
    wb = np.conj(w)[..., None]
    M = [add((wb - np.conj(r[j])[..., None]) * D, Nn) for j in range(m)]
    prodM = one
    for j in range(m):
        prodM = mul(prodM, M[j])
    P = mul(np.concatenate([-w[..., None], one], -1), prodM)
    S = np.zeros(w.shape + (1,), complex)
    for j in range(m):
        t = e[j][..., None] * one
        for k in range(m):
            if k != j:
                t = mul(t, M[k])
        S = add(S, t)
    return add(P, -mul(D, S))


def lightcurve_params(n_lc, seed):
    """Point-source light-curve batch: t0 ~ U[-10,10], u0 ~ U[-0.5,0.5],
    tE log-uniform [5,100] d, alpha ~ U[0,2pi), s, q as in lens_configs."""
    g = rng(seed)
    p = lens_configs(g, n_lc, 2)
    return dict(t0=g.uniform(-10, 10, n_lc), u0=g.uniform(-0.5, 0.5, n_lc),
                tE=loguniform(g, 5.0, 100.0, n_lc), alpha=g.uniform(0, 2 * np.pi, n_lc),
                s=p["s"], q=p["q"])


def trajectory(t, t0, u0, tE, alpha):
    """Rectilinear source trajectory: src/caustics/trajectory.py:152-155 with piE = 0
    and the angle psi -> alpha."""
    tau = (np.asarray(t)[None, :] - np.asarray(t0)[:, None]) / np.asarray(tE)[:, None]
    ca, sa = np.cos(alpha)[:, None], np.sin(alpha)[:, None]
    u0 = np.asarray(u0)[:, None]
    return (u0 * ca + tau * sa) + 1j * (-u0 * sa + tau * ca)


# ---------------------------------------------------------------------------
# extended-source workloads (GPU-side generation: caustic points come from the
# GPU critical-curve solver, far sources are rejection-sampled with the gate)
# ---------------------------------------------------------------------------
def extended_binary_batch(n_near, n_far, seed=4, npts_choices=(256, 512, 1024, 2048, 4096),
                          u1=0.5):
    """Binary limb-darkened sources: n_near placed within 2 rho of a caustic point of
    their own lens (as in the reference's tests/test_extended_source.py:82-97), n_far
    uniform in |w| <= 2 and accepted by the hexadecapole gate; rho log-uniform
    [1e-3, 1e-1]; s, q as lens_configs.  Returns a dict of numpy arrays."""
    from .lightcurve import hexadecapole_gate
    from .point_source import critical_and_caustic_curves

    g = rng(seed)
    p = lens_configs(g, n_near + n_far, 2)
    rho = loguniform(g, 1e-3, 1e-1, n_near + n_far)
    s, q = p["s"], p["q"]
    _, ca = critical_and_caustic_curves(npts=64, nlenses=2, s=s[:n_near], q=q[:n_near])
    ca = ca.reshape(n_near, -1).cpu().numpy()
    pick = ca[np.arange(n_near), g.integers(0, ca.shape[1], n_near)]
    r = g.uniform(0, 2 * rho[:n_near])
    w_near = pick + r * np.exp(1j * g.uniform(-np.pi, np.pi, n_near))
    w_far = np.zeros(n_far, complex)
    todo = np.arange(n_far)
    for attempt in range(400):
        if not len(todo):
            break
        if attempt and attempt % 8 == 0:
            # some (s, q, rho) admit no gate-accepted point in |w| <= 2 (the quadrupole
            # term alone exceeds the gate's 1e-2 for large rho): redraw those systems
            k = n_near + todo
            pr = lens_configs(g, len(todo), 2)
            s[k], q[k] = pr["s"], pr["q"]
            rho[k] = loguniform(g, 1e-3, 1e-1, len(todo))
        cand = disk(g, len(todo))
        _, _, use = hexadecapole_gate(cand, rho[n_near + todo], nlenses=2, s=s[n_near + todo],
                                      q=q[n_near + todo])
        ok = use.cpu().numpy()
        w_far[todo[ok]] = cand[ok]
        todo = todo[~ok]
    if len(todo):
        raise RuntimeError(f"{len(todo)} far sources found no gate-accepted position")
    npts = g.choice(np.asarray(npts_choices), n_near)
    return dict(w=np.concatenate([w_near, w_far]), rho=rho, s=s, q=q, u1=np.full(len(rho), u1),
                npts=npts, n_near=n_near)


def triple_caustic_lightcurves(n_sets, n_epochs, seed=5, u1=0.4):
    """Triple-lens light curves crossing a caustic: per parameter set s, q, q3, r3, psi as
    lens_configs, rho log-uniform [1e-3, 3e-2], tE = 20 d, and a rectilinear trajectory
    (angle alpha random) through a caustic point at t = 0; epochs t in [-2 tE, 2 tE]."""
    from .point_source import critical_and_caustic_curves

    g = rng(seed)
    p = lens_configs(g, n_sets, 3)
    rho = loguniform(g, 1e-3, 3e-2, n_sets)
    _, ca = critical_and_caustic_curves(npts=64, nlenses=3, **p)
    ca = ca.reshape(n_sets, -1).cpu().numpy()
    c = ca[np.arange(n_sets), g.integers(0, ca.shape[1], n_sets)]
    alpha = g.uniform(0, 2 * np.pi, n_sets)
    tE = np.full(n_sets, 20.0)
    rot = c * np.exp(1j * alpha)            # = u0 + i tau_c
    u0, tau_c = rot.real, rot.imag
    t0 = -tau_c * tE
    t = np.linspace(-2 * 20.0, 2 * 20.0, n_epochs)
    w = trajectory(t, t0, u0, tE, alpha)    # (n_sets, n_epochs)
    return dict(w=w, rho=rho, u1=np.full(n_sets, u1), t=t, **p)


def crossing_npts(b, lo=256, hi=4096):
    """Uneven limb resolutions for triple_caustic_lightcurves: hi within one source radius
    of the caustic point the trajectories cross at t = 0, halved per doubling of the
    distance (in source radii) down to lo beyond 8 radii (the adaptive resolution an
    inference run would give epochs near a crossing).  Returns (n_sets, n_epochs) int32."""
    tE = 20.0
    x = np.abs(np.asarray(b["t"]))[None, :] / tE / np.asarray(b["rho"])[:, None]
    k = np.clip(np.ceil(np.log2(np.maximum(x, 1e-300))), 0, None)
    return np.clip(hi / 2.0 ** k, lo, hi).astype(np.int32)
