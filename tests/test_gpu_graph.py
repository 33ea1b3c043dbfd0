"""The fused entry points enqueue only kernels (no allocation, host synchronisation or
memset), so a HIP graph can capture them: each test captures one call, replays it
several times with other eager work of the same entry point in between, and requires
every replay to reproduce the eager result bit for bit."""
import numpy as np
import pytest
import torch

from caustics_amd import _lib
from caustics_amd import workloads as W

pytestmark = pytest.mark.gpu


def _replay_check(fn, outputs, eager, other):
    """fn() enqueues into preallocated outputs; eager is the list of reference results."""
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        fn()   # warm-up on a side stream, as torch requires before capture
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        fn()
    torch.cuda.synchronize()
    for rep in range(3):
        for o in outputs:
            o.fill_(float("nan")) if o.is_floating_point() else o.fill_(0)
        other()             # the entry point's own eager work between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for o, e in zip(outputs, eager):
            assert torch.equal(o, e), f"replay {rep} differs"


def test_graph_roots():
    c = torch.as_tensor(W.coeffs_product(*_lens_args(3, 4096, 5)), device="cuda")
    r = torch.empty((4096, 10), dtype=torch.complex128, device="cuda")
    ok = torch.empty(4096, dtype=torch.uint8, device="cuda")
    L = _lib.lib()

    def call(cc, rr, oo):
        assert L.caus_roots(_lib.stream_ptr(), _lib.ptr(cc), None, _lib.ptr(rr), None,
                            _lib.ptr(oo), cc.shape[0], 10, 2000, 0, 1) == 0

    call(c, r, ok)
    torch.cuda.synchronize()
    eager = [r.clone(), ok.clone()]
    c2 = torch.as_tensor(W.coeffs_product(*_lens_args(3, 4096, 6)), device="cuda")
    r2, ok2 = torch.empty_like(r), torch.empty_like(ok)
    _replay_check(lambda: call(c, r, ok), [r, ok], eager, lambda: call(c2, r2, ok2))


def _lens_args(nlenses, n, seed):
    lf, w, _ = W.batched_polys(nlenses, n, seed)
    return w, lf["a"], lf["e1"], lf["e2"], lf["r3"], nlenses


def test_graph_lightcurve_fwd_vjp():
    from caustics_amd.lightcurve import _call
    t = torch.linspace(-40, 40, 1500, dtype=torch.float64, device="cuda")
    g = np.random.default_rng(3)
    L = 64
    lcp = torch.as_tensor(np.stack([g.uniform(-5, 5, L), g.uniform(-0.4, 0.4, L),
                                    g.uniform(10, 60, L), g.uniform(0, 6.28, L),
                                    g.uniform(0.5, 2.0, L), 10 ** g.uniform(-3, -1, L)], 1),
                          device="cuda").contiguous()
    gm = torch.as_tensor(g.normal(size=(L, 1500)), device="cuda")
    Lb = _lib.lib()
    wsb = Lb.caus_lightcurve_workspace_bytes(2, 1500, L)
    ws = torch.empty(max(wsb // 8, 1), dtype=torch.float64, device="cuda")
    mag = torch.empty((L, 1500), dtype=torch.float64, device="cuda")
    gp = torch.empty_like(lcp)

    def call():
        rc = Lb.caus_lightcurve_point_source(_lib.stream_ptr(), 2, _lib.ptr(t), 1500,
                                             _lib.ptr(lcp), L, _lib.ptr(gm), _lib.ptr(mag),
                                             _lib.ptr(gp), _lib.ptr(ws), wsb, None, 2500, 0)
        assert rc == 0

    call()
    torch.cuda.synchronize()
    eager = [mag.clone(), gp.clone()]
    lcp2 = lcp.clone()
    lcp2[:, 1] += 0.05

    def other():
        _call(t, lcp2, gm, 2, 2500, 0)

    _replay_check(call, [mag, gp], eager, other)


@pytest.mark.parametrize("nlenses", [2, 3])
def test_graph_extended_ld_fwd_vjp(nlenses):
    from caustics_amd.extended_source import extended_workspace_bytes
    g = np.random.default_rng(10 + nlenses)
    S, npts = 48, 128
    if nlenses == 2:
        prm = np.tile([0.9, 0.3], (S, 1))
    else:
        prm = np.tile([1.0, 0.4, 0.05, 0.9, 0.7], (S, 1))
    w = torch.as_tensor(g.uniform(-0.4, 0.4, S) + 1j * g.uniform(-0.4, 0.4, S), device="cuda")
    rho = torch.as_tensor(10 ** g.uniform(-2, -1, S), device="cuda")
    u1 = torch.full((S,), 0.5, dtype=torch.float64, device="cuda")
    prm_t = torch.as_tensor(prm, device="cuda").contiguous()
    gm = torch.ones(S, dtype=torch.float64, device="cuda")
    P = prm.shape[1]
    mag = torch.empty(S, dtype=torch.float64, device="cuda")
    gout = torch.empty((S, 4 + P), dtype=torch.float64, device="cuda")
    st = torch.empty(S, dtype=torch.uint8, device="cuda")
    wsb = extended_workspace_bytes(nlenses, S, npts)
    ws = torch.empty((wsb + 7) // 8, dtype=torch.float64, device="cuda")
    Lb = _lib.lib()

    def call(ww, mm, go, ss):
        rc = Lb.caus_mag_extended_source_vjp(_lib.stream_ptr(), nlenses, _lib.ptr(ww), S,
                                             _lib.ptr(rho), _lib.ptr(prm_t), P, None, npts, 1,
                                             _lib.ptr(u1), 1, 64, _lib.ptr(gm), _lib.ptr(mm),
                                             _lib.ptr(go), _lib.ptr(ss), _lib.ptr(ws), wsb, 2500, 0)
        assert rc == 0

    call(w, mag, gout, st)
    torch.cuda.synchronize()
    eager = [mag.clone(), gout.clone(), st.clone()]
    w2 = w + 0.01
    m2, g2, s2 = torch.empty_like(mag), torch.empty_like(gout), torch.empty_like(st)
    # the eager call in between uses the same workspace, as a caller reusing it would
    _replay_check(lambda: call(w, mag, gout, st), [mag, gout, st], eager,
                  lambda: call(w2, m2, g2, s2))


def test_graph_trajectory_fwd_vjp():
    g = np.random.default_rng(14)
    T, M, L = 1500, 401, 64
    t = torch.as_tensor(np.sort(g.uniform(8000.0, 8400.0, T)), device="cuda")
    xp_h = np.arange(8000.0, 8401.0, 1.0)
    om = 2 * np.pi / 365.25
    tab = torch.as_tensor(np.stack([np.sin(om * xp_h), 0.4 * np.cos(om * xp_h),
                                    om * np.cos(om * xp_h), -0.4 * om * np.sin(om * xp_h)]),
                          device="cuda").contiguous()
    xp = torch.as_tensor(xp_h, device="cuda")
    P = torch.as_tensor(np.stack([g.uniform(8050, 8350, L), g.uniform(-1, 1, L),
                                  g.uniform(20, 200, L), g.uniform(-0.5, 0.5, L),
                                  g.uniform(-0.5, 0.5, L)], 1), device="cuda").contiguous()
    gw = torch.as_tensor(g.normal(size=(L, T, 2)), device="cuda")
    w = torch.empty((L, T, 2), dtype=torch.float64, device="cuda")
    gP = torch.empty_like(P)
    Lb = _lib.lib()

    def call(PP, ww, gg):
        assert Lb.caus_trajectory_annual_parallax(_lib.stream_ptr(), _lib.ptr(t), T, _lib.ptr(xp),
                                                  _lib.ptr(tab), M, _lib.ptr(PP), L, 0,
                                                  _lib.ptr(ww), _lib.ptr(gw), _lib.ptr(gg)) == 0

    call(P, w, gP)
    torch.cuda.synchronize()
    eager = [w.clone(), gP.clone()]
    P2 = P.clone()
    P2[:, 1] += 0.1
    w2, g2 = torch.empty_like(w), torch.empty_like(gP)
    _replay_check(lambda: call(P, w, gP), [w, gP], eager, lambda: call(P2, w2, g2))


def test_graph_mag_ragged_fwd_vjp():
    """mag() forward + VJP on a ragged batch (gate, hexadecapole VJP, contour pipeline
    with the gated points skipped on the device) waits for nothing on the host, so one
    call can be captured and replayed."""
    from caustics_amd.lightcurve import _mag_args, _mag_fwd

    b = W.extended_binary_batch(48, 48, seed=9)
    S = len(b["w"])
    npts = np.concatenate([b["npts"], np.full(48, 256)]) % 1024 + 64
    cfg, _, w, rho, u1, ps = _mag_args(b["w"], b["rho"], 2, npts, True, b["u1"], 60, 2500,
                                       False, None, dict(s=b["s"], q=b["q"]))
    prm = torch.stack(ps, -1).contiguous()
    g = torch.as_tensor(np.random.default_rng(1).normal(size=S), device="cuda")
    out = {}

    def call():
        m, use, go = _mag_fwd(cfg, w, rho, u1, prm, g)
        out["m"], out["go"] = m, go

    call()
    torch.cuda.synchronize()
    eager = [out["m"].clone(), out["go"].clone()]
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        call()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    graph = torch.cuda.CUDAGraph()
    with torch.cuda.graph(graph):
        call()
    captured = [out["m"], out["go"]]
    torch.cuda.synchronize()
    for rep in range(3):
        for o in captured:
            o.fill_(float("nan"))
        call()   # eager work of the same entry points between replays
        torch.cuda.synchronize()
        graph.replay()
        torch.cuda.synchronize()
        for o, e in zip(captured, eager):
            assert torch.equal(o, e), f"replay {rep} differs"
