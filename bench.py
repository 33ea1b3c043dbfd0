"""Benchmark of the magnification hot path on one MI355X.

Headline workload (one "step"): the batched root solve of 2^20 degree-5 (binary-lens)
and 2^18 degree-10 (triple-lens) polynomials built from seeded lens configurations
(s log-uniform [0.3,3], q log-uniform [1e-4,1]; third lens r ~ U[0.3,2],
psi ~ U[0,2pi), q3 log-uniform [1e-4,1e-1]; sources uniform in |w| <= 2), coefficients
resident in HBM, FP64 Ehrlich-Aberth.  value = polynomials solved per second.

Also reported in the same JSON line: the compensated-FP32 (CF32) form of the same
solve with its own roofline, per-form error against the FP64 CPU oracle, the roofline of
the dominant kernel (algorithmic FP64 flops per launch / HIP-event launch time vs the
MI355X FP64 vector peak), and the CPU baseline (the reference's own CPU op, compiled
from its sources, on a bounded sample, one core and every core of the process).

Under "workloads": the other hot-path workloads, each timed on the GPU with the
algorithmic flops of its dominant kernels (forward and reverse mode) against the vector
peak, its error against the oracle on a sample, and the oracle's own rate on the host
(one core, and every core through spawned single-thread worker processes):
  C1 10^4 binary point-source magnifications;
  C3 256 binary light curves x 2000 epochs, forward + VJP (FP64 and CF32);
  C4 4096 limb-darkened binary extended sources, forward + VJP;
  C5 1024 triple-lens limb-darkened light curves x 1500 epochs, hexadecapole gate and
     contour integration with 256 limb points, forward + VJP (FP64 and CF32);
  ll the marginalised likelihood with its gradient; tr the annual-parallax trajectory.
Every timed region starts after an untimed clock ramp.

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import ctypes
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector, dense
FP32_PEAK_TFLOPS = 157.3  # MI355X FP32 vector (packed), dense
HBM_PEAK_GBS = 8000.0


def flops_per_root_iter(n):
    """Algorithmic FP64 flops of one active root-iteration of the plain solver
    (28n + 20: |z|, real Horner 2n, complex Horner value+derivative 16n, test,
    Aberth sum 10(n-1) and correction 19, update 2)."""
    return 28 * n + 20


def flops_comp_root_iter(n):
    return 227 * n + 33


def flops_cf32_comp_root_iter(n):
    """FP32 flops of one root-iteration of the CF32 form's compensated FP32 sweep: per
    Horner step four TwoProd (2 each) and four TwoSum (6 each) plus the error
    recurrence (8), then the Aberth sum and correction as in the plain step."""
    return 48 * n + 10 * (n - 1) + 21


# Gauss-Legendre node of a limb-darkened line integral (integrate.py:29-121), counted by
# the same rule from the kernel's operations: the lens equation with one shared
# reciprocal, |D|, the profile and the quadrature accumulation (forward), and with the
# reverse mode also the squared v_j, K, the two spatial derivatives, the parameter
# adjoints of every lens and the four running sums (adjoint).
NODE_FLOPS_FWD = {1: 34, 2: 48, 3: 62}
NODE_FLOPS_VJP = {1: 100, 2: 138, 3: 177}
# implicit-derivative adjoint of one contour point (image_terms, det J, c, per-lens terms)
IMAGE_ADJ_FLOPS = {1: 55, 2: 95, 3: 132}
# lens polynomial of one source position (lens.hpp lens_poly: the products of the lens
# factors with products by an exact 1 and sums with an exact 0 not counted)
COEFF_FLOPS = {1: 10, 2: 223, 3: 1016}
# per image: 1/(conj z - conj r_j) (8), G and K terms (4 + 10) per lens, then z - G,
# |z - G - w|^2 against the filter, det J = 1 - |K|^2
IMAGE_FLOPS = {1: 33, 2: 55, 3: 77}


def make_inputs(n5, n10):
    from caustics_amd import workloads as W

    out = {}
    for nl, n, seed in ((2, n5, 2), (3, n10, 3)):
        lf, w, _ = W.batched_polys(nl, n, seed=seed)
        out[nl] = W.coeffs_product(w, lf["a"], lf["e1"], lf["e2"], lf["r3"], nl)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=30)
    ap.add_argument("--n5", type=int, default=1 << 20)
    ap.add_argument("--n10", type=int, default=1 << 18)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--ramp-seconds", type=float, default=3.0,
                    help="untimed sustained work before any timing, so that every timed "
                         "region runs at steady clocks")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--roots-only", action="store_true")
    ap.add_argument("--workloads", default="c1,c3,c4,c5,ll,tr",
                    help="comma-separated extra workloads to run after the root solve")
    ap.add_argument("--c5-sets", type=int, default=1024,
                    help="parameter sets of the C5 workload (1024 = the full workload)")
    ap.add_argument("--copy-mib", type=int, default=0,
                    help="also run one device-to-device copy of this many MiB (a kernel of "
                         "known traffic, to scale hardware byte counters)")
    args = ap.parse_args()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py: this build measures one GPU only (every unit of the path is "
              "independent; multi-GPU sharding is a design note, not code). Use --gpus 1.",
              file=sys.stderr)
        sys.exit(2)

    import torch

    from caustics_amd import _lib
    from caustics_amd.ehrlich_aberth_primitive import solve_with_stats

    L = _lib.lib()
    dev = torch.device("cuda")
    print(f"bench: generating inputs n5={args.n5} n10={args.n10}", file=sys.stderr, flush=True)
    host = make_inputs(args.n5, args.n10)
    coeffs = {k: torch.as_tensor(v, device=dev) for k, v in host.items()}
    degs = {2: 5, 3: 10}
    roots = {k: torch.empty((v.shape[0], degs[k]), dtype=torch.complex128, device=dev)
             for k, v in coeffs.items()}
    status = {k: torch.empty((v.shape[0],), dtype=torch.uint8, device=dev)
              for k, v in coeffs.items()}

    def launch(nl, mode, init):
        c = coeffs[nl]
        rc = L.caus_roots(_lib.stream_ptr(), _lib.ptr(c), None, _lib.ptr(roots[nl]), None,
                          _lib.ptr(status[nl]), c.shape[0], degs[nl], 2000, mode, init)
        _lib.check(rc, "caus_roots")

    def launcher(nl, mode, init, stream):
        """launch() on a fixed stream with its arguments prepared once, so that the host
        enqueues a step in a few microseconds"""
        c = coeffs[nl]
        a = (ctypes.c_void_p(stream.cuda_stream), _lib.ptr(c), None, _lib.ptr(roots[nl]), None,
             _lib.ptr(status[nl]), c.shape[0], degs[nl], 2000, mode, init)
        f = L.caus_roots

        def go():
            if f(*a) != 0:
                raise RuntimeError("caus_roots failed")
        return go

    if args.copy_mib:
        src = torch.ones(args.copy_mib << 17, dtype=torch.float64, device=dev)
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        print(f"bench: copied {args.copy_mib} MiB", file=sys.stderr, flush=True)
        del src, dst

    forms = {"fp64": (_lib.MODE_FP64, _lib.INIT_COMPLEX),
             "cf32": (_lib.MODE_CF32, _lib.INIT_COMPLEX),
             "fp64_refinit": (_lib.MODE_FP64, _lib.INIT_REFERENCE)}
    # The two degree batches are independent: a step launches the degree-10 batch on one
    # stream and the degree-5 batch on another, so that each batch's tail (its last,
    # slowest polynomials) overlaps the other's work.  Each kernel's own duration, for the
    # roofline, is timed separately with HIP events around single launches on one stream.
    s_a, s_b = _lib.side_streams(2)   # the library's pool (default priority)
    # untimed clock ramp: the device leaves its idle clocks only after a second or two of
    # sustained work, so a short warmup alone would time a partly ramped device
    t0 = time.perf_counter()
    nramp = 0
    while time.perf_counter() - t0 < args.ramp_seconds:
        for nl in (2, 3):
            launch(nl, _lib.MODE_FP64, _lib.INIT_COMPLEX)
        torch.cuda.synchronize()
        nramp += 1
    print(f"bench: clock ramp {nramp} steps, {time.perf_counter() - t0:.1f} s", file=sys.stderr,
          flush=True)
    results = {}
    for name, (mode, init) in forms.items():
        steps = args.steps if name == "fp64" else max(3, args.steps // 4)
        for _ in range(args.warmup):
            for nl in (2, 3):
                launch(nl, mode, init)
        torch.cuda.synchronize()
        ev = {nl: [] for nl in (2, 3)}
        for _ in range(steps):
            for nl in (2, 3):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                launch(nl, mode, init)
                e1.record()
                ev[nl].append((e0, e1))
        torch.cuda.synchronize()
        kms = {nl: float(np.mean([a.elapsed_time(b) for a, b in ev[nl]])) for nl in (2, 3)}
        one_stream_ms = sum(kms.values())
        go5, go10 = launcher(2, mode, init, s_b), launcher(3, mode, init, s_a)
        s_a.wait_stream(torch.cuda.current_stream())
        s_b.wait_stream(torch.cuda.current_stream())
        for _ in range(args.warmup):
            go5()
            go10()
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        for _ in range(steps):
            go5()
            go10()
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        npolys = args.n5 + args.n10
        # work counters and accuracy (untimed)
        work = {}
        err = {}
        unconv = 0
        for nl in (2, 3):
            z, st, stats = solve_with_stats(coeffs[nl], mode=name if name != "fp64_refinit"
                                            else "fp64", init="complex" if init else "reference")
            s = stats.sum(0).cpu().numpy().astype(np.int64)
            work[nl] = s
            unconv += int((st == 0).sum().item())
            err[nl] = z
        results[name] = dict(ms_per_step=(t1 - t0) * 1e3 / steps, kernel_ms=kms, work=work,
                             value=npolys * steps / (t1 - t0), steps=steps, z=err,
                             unconverged=unconv, one_stream_kernel_ms_sum=one_stream_ms)
        print(f"bench: {name}: {results[name]['value']:.4g} polys/s, kernel ms {kms}",
              file=sys.stderr, flush=True)

    if not args.no_cpu:
        oracle_accuracy(host, results)
    for r in results.values():
        del r["z"]

    main_r = results["fp64"]
    # roofline of the dominant kernel (FP64 form)
    dom = max((2, 3), key=lambda nl: main_r["kernel_ms"][nl])
    n = degs[dom]
    w = main_r["work"][dom]
    flops = float(w[1]) * flops_per_root_iter(n) + float(w[2]) * flops_comp_root_iter(n)
    achieved = flops / (main_r["kernel_ms"][dom] * 1e-3) / 1e12
    # CF32 form: FP32 stage-A root-iterations and compensated FP32 sweeps against the
    # FP32 vector peak, its FP64 finishing iterations against the FP64 peak; frac is the
    # sum of the two time shares (the fraction of the launch the VALU would need at peak)
    r32 = results["cf32"]
    w32 = r32["work"][dom]
    f32 = float(w32[3]) * flops_per_root_iter(n) + float(w32[2]) * flops_cf32_comp_root_iter(n)
    f64 = float(w32[1]) * flops_per_root_iter(n)
    t32 = r32["kernel_ms"][dom] * 1e-3
    r32["roofline"] = {
        "bound": "valu", "kernel": f"k_roots_compact<{n}> CF32 form", "kernel_ms": t32 * 1e3,
        "fp32_flops_per_launch": f32, "fp64_flops_per_launch": f64,
        "achieved_fp32": f32 / t32 / 1e12, "achieved_fp64": f64 / t32 / 1e12,
        "peak_fp32": FP32_PEAK_TFLOPS, "peak_fp64": FP64_PEAK_TFLOPS, "unit": "TFLOP/s",
        "frac": (f32 / (FP32_PEAK_TFLOPS * 1e12) + f64 / (FP64_PEAK_TFLOPS * 1e12)) / t32,
        "note": "the FP32 peak counts packed FP32 (two lanes per instruction), which this "
                "build does not emit; without it the FP32 ceiling equals the FP64 one"}
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(f"roots_deg{n}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(host, args.cpu_seconds)

    workloads = {}
    if not args.roots_only:
        sel = set(args.workloads.split(","))
        for key, name, fn in (
                ("c1", "point_source_1e4",
                 lambda: bench_point_source(args.steps, 2, cpu_s=0 if args.no_cpu else args.cpu_seconds)),
                ("c3", "lightcurves_256x2000_fwd_vjp",
                 lambda: bench_lightcurves(20, 5, cpu_s=0 if args.no_cpu else args.cpu_seconds)),
                ("c4", "extended_binary_ld_4096_fwd_vjp",
                 lambda: bench_extended_binary(5, 2, cpu_s=0 if args.no_cpu else args.cpu_seconds)),
                ("c5", "triple_extended_lightcurves_fwd_vjp",
                 lambda: bench_triple(1, 0, n_sets=args.c5_sets,
                                      cpu_s=0 if args.no_cpu else args.cpu_seconds)),
                ("ll", "marginalized_log_likelihood_grad", lambda: bench_likelihood(20, 3)),
                ("tr", "annual_parallax_trajectory_fwd_vjp", lambda: bench_trajectory(20, 3))):
            if key not in sel:
                continue
            print(f"bench: {name}", file=sys.stderr, flush=True)
            try:
                workloads[name] = fn()
            except Exception as e:  # a failing extra workload must not hide the headline
                workloads[name] = {"error": repr(e)}
            print(f"bench: {name}: {workloads[name]}", file=sys.stderr, flush=True)

    line = {
        "metric": "mag evals/s fwd+VJP at \u22641e-8 rel err vs FP64; polys/s deg 5/10",
        "value": main_r["value"],
        "unit": "polys/s",
        "n_gpus": 1,
        "steps": main_r["steps"],
        "warmup": args.warmup,
        "ms_per_step": main_r["ms_per_step"],
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic (seeded lens configurations)",
        "config": {"workload": f"batched root solve: {args.n5} degree-5 binary + {args.n10} "
                               "degree-10 triple lens polynomials",
                   "n_deg5": args.n5, "n_deg10": args.n10, "itmax": 2000,
                   "init": "complex"},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP64_PEAK_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / FP64_PEAK_TFLOPS, "traffic": traffic,
                     "kernel": (f"k_roots_compact<{n}> (initial estimates in-kernel)" if n >= 5 else
                                f"k_roots_init<{n}> + k_roots_persistent<{n},FP64,fast>")
                               + " (one caus_roots launch, HIP events on its stream)",
                     "kernel_ms": main_r["kernel_ms"][dom],
                     "flops_per_launch": flops,
                     "flop_rule": f"{flops_per_root_iter(n)} per active FP64 root-iteration"},
        "cpu_baseline": cpu,
        "workloads": workloads,
        "forms": {k: {kk: vv for kk, vv in v.items() if kk not in ("work",)}
                  for k, v in results.items()},
        "work_counts": {k: {str(degs[nl]): [int(x) for x in v["work"][nl]] for nl in (2, 3)}
                        for k, v in results.items()},
    }
    print(json.dumps(line, default=float))


def _timed(fn, steps, warmup, name="", budget_s=60.0):
    """Run fn warmup times, then time `steps` calls between device synchronisations
    (fewer if one call is so slow that `steps` would exceed budget_s)."""
    import torch

    for i in range(max(1, warmup)):
        t0 = time.perf_counter()
        fn()
        torch.cuda.synchronize()
        dt1 = time.perf_counter() - t0
        print(f"bench: {name} warmup {i + 1}: {dt1:.3f} s", file=sys.stderr, flush=True)
    steps = max(1, min(steps, int(budget_s / max(dt1, 1e-9))))
    t0 = time.perf_counter()
    for _ in range(steps):
        fn()
    torch.cuda.synchronize()
    dt = (time.perf_counter() - t0) / steps
    print(f"bench: {name} {steps} timed steps, {dt * 1e3:.3f} ms/step", file=sys.stderr,
          flush=True)
    return dt


def bench_point_source(steps, warmup, cpu_s=0.0):
    """10^4 binary point-source magnifications, s=0.9, q=0.1, w uniform in [-1.5,1.5]^2."""
    import torch

    from caustics_amd import _lib
    from caustics_amd import workloads as W

    g = W.rng(1)
    n = 10000
    w = torch.as_tensor(g.uniform(-1.5, 1.5, n) + 1j * g.uniform(-1.5, 1.5, n), device="cuda")
    prm = torch.tensor([0.9, 0.1], dtype=torch.float64, device="cuda")
    mag = torch.empty(n, dtype=torch.float64, device="cuda")
    L = _lib.lib()

    def run():
        _lib.check(L.caus_mag_point_source(_lib.stream_ptr(), 2, _lib.ptr(w), n, _lib.ptr(prm), 0,
                                           _lib.ptr(mag), None, 2500, _lib.MODE_FP64), "ps")
    dt = _timed(run, steps * 10, warmup, "C1 point source")
    out = {"workload": "10^4 binary point-source magnifications (s=0.9, q=0.1), forward, FP64",
           "mag_evals_per_s": n / dt, "ms": dt * 1e3}
    mo = _oracle_point_source((w.cpu().numpy(), dict(s=0.9, q=0.1)))
    e = np.abs(mag.cpu().numpy() - mo) / mo
    out["vs_oracle"] = {"max_rel_err": float(e.max()), "p99_rel_err": float(np.quantile(e, 0.99)),
                        "frac_le_1e-10": float(np.mean(e <= 1e-10))}
    if cpu_s > 0:
        from oracle import lens
        wc = w.cpu().numpy()
        out["cpu_baseline"] = cpu_rates_omp(
            lambda nt: lens.mag_point_source_c(wc, 2, nthreads=nt, s=0.9, q=0.1), n, cpu_s,
            "point-source magnifications of the same 10^4 positions (compiled C oracle, "
            "oracle/lens_oracle.c)", "mag evals/s")
    return out


def bench_lightcurves(steps, warmup, host_check=True, cpu_s=0.0):
    """256 binary point-source light curves x 2000 epochs, fused forward + VJP wrt
    (t0, u0, tE, alpha, s, q); FP64 and CF32 forms; error of CF32 vs FP64."""
    import torch

    from caustics_amd import _lib
    from caustics_amd import workloads as W

    L_, T = 256, 2000
    prm = W.lightcurve_params(L_, seed=3)
    lcp = torch.as_tensor(np.stack([prm[k] for k in ("t0", "u0", "tE", "alpha", "s", "q")], -1),
                          device="cuda").contiguous()
    t = torch.linspace(-100, 100, T, dtype=torch.float64, device="cuda")
    g = torch.ones((L_, T), dtype=torch.float64, device="cuda")
    mag = torch.empty((L_, T), dtype=torch.float64, device="cuda")
    gp = torch.empty_like(lcp)
    Lb = _lib.lib()
    wsb = Lb.caus_lightcurve_workspace_bytes(2, T, L_)
    ws = torch.empty(wsb // 8, dtype=torch.float64, device="cuda")
    out = {"workload": "256 binary point-source light curves x 2000 epochs, fused fwd+VJP "
                       "(t0, u0, tE, alpha, s, q)", "evals": L_ * T}
    mags = {}
    for name, mode in (("fp64", _lib.MODE_FP64), ("cf32", _lib.MODE_CF32)):
        def run():
            _lib.check(Lb.caus_lightcurve_point_source(
                _lib.stream_ptr(), 2, _lib.ptr(t), T, _lib.ptr(lcp), L_, _lib.ptr(g),
                _lib.ptr(mag), _lib.ptr(gp), _lib.ptr(ws), wsb, None, 2500, mode), "lc")
        dt = _timed(run, steps, warmup, f"C3 light curves {name}")
        run()
        torch.cuda.synchronize()
        mags[name] = (mag.clone(), gp.clone())
        out[name] = {"mag_evals_per_s_fwd_vjp": L_ * T / dt, "ms": dt * 1e3}
    # roofline: each lane of the kernel takes epl consecutive epochs of one light curve,
    # the first solved from the initial estimates of caus_roots, each later one started
    # from the previous epoch's roots.  The root-iterations of each form are counted by
    # replaying that schedule through the counting entry point: round r solves the
    # epochs with index = r (mod epl), started from round r-1's roots.
    from caustics_amd.ehrlich_aberth_primitive import solve_with_stats
    tt = t.cpu().numpy()[None, :]
    pr = {k: np.asarray(prm[k], float)[:, None] for k in ("t0", "u0", "tE", "alpha", "s", "q")}
    w_user = np.exp(-1j * pr["alpha"]) * (pr["u0"] + 1j * (tt - pr["t0"]) / pr["tE"])
    lf, x_cm = W.lens_frame({"s": pr["s"], "q": pr["q"]}, 2)
    c = W.coeffs_product((w_user + x_cm).reshape(-1), np.broadcast_to(lf["a"], w_user.shape).reshape(-1),
                         np.broadcast_to(lf["e1"], w_user.shape).reshape(-1), None, None, 2)
    c = c.reshape(L_, T, -1)
    epl = min(8, max(1, -(-(T * L_) // 131072)))   # lc_epochs_per_lane of the kernel
    for name in ("fp64", "cf32"):
        sw = np.zeros(4)
        z = None
        for r in range(epl):
            cr = torch.as_tensor(c[:, r::epl].reshape(-1, c.shape[-1]), device="cuda")
            if z is not None:
                z = z.reshape(L_, -1, z.shape[-1])[:, :cr.shape[0] // L_].reshape(-1, z.shape[-1])
            z, _, st = solve_with_stats(cr, mode=name, init="complex", roots_init=z)
            sw += st.sum(0).cpu().numpy().astype(np.float64)
        work = [sw[1], sw[3], sw[2], float(L_ * T), 0.0]
        f64, f32 = contour_flops(work, 2, False, cf32=name == "cf32")
        f64 += pointsource_vjp_flops(L_ * T, 2, 6)
        out[name]["roofline"] = _roofline_entry((f64, f32), out[name]["ms"] * 1e-3)
        out[name]["roofline"]["work"] = {"root_iterations_fp64": int(sw[1]),
                                         "root_iterations_fp32": int(sw[3]),
                                         "root_iterations_comp": int(sw[2]), "epochs": L_ * T}
        out[name]["roofline"]["rule"] = (
            f"root-iterations x {flops_per_root_iter(5)} (FP32 ones against the FP32 peak; "
            f"CF32 compensated FP32 sweeps x {flops_cf32_comp_root_iter(5)}) + "
            f"per epoch {COEFF_FLOPS[2]} (coefficients) + 5 x {IMAGE_FLOPS[2]} (image terms) "
            "+ VJP 5 roots x 6 directions x 50")
    m64, g64 = mags["fp64"]
    m32, g32 = mags["cf32"]
    wtraj = W.trajectory(tt[0], prm["t0"], prm["u0"], prm["tE"], prm["alpha"])
    curves = [(wtraj[l], {"s": float(prm["s"][l]), "q": float(prm["q"][l])}) for l in range(L_)]
    mo = np.stack([_oracle_point_source(cv) for cv in curves[:32]])
    e = np.abs(m64[:32].cpu().numpy() - mo) / mo
    out["fp64"]["vs_oracle"] = {"curves": 32, "max_rel_err": float(e.max()),
                                "p99_rel_err": float(np.quantile(e, 0.99)),
                                "frac_le_1e-10": float(np.mean(e <= 1e-10))}
    if cpu_s > 0:
        from oracle import lens
        wf = wtraj.reshape(-1)
        sf, qf = np.repeat(prm["s"], T), np.repeat(prm["q"], T)
        out["cpu_baseline"] = cpu_rates_omp(
            lambda nt: lens.mag_point_source_c(wf, 2, nthreads=nt, s=sf, q=qf), L_ * T, cpu_s,
            "light-curve epochs of the same workload, forward only (compiled C oracle, "
            "oracle/lens_oracle.c)", "mag evals/s (forward only)")
    rel = ((m32 - m64).abs() / m64.abs()).cpu().numpy()
    out["cf32"]["max_rel_err_vs_fp64"] = float(rel.max())
    out["cf32"]["p99_rel_err_vs_fp64"] = float(np.quantile(rel, 0.99))
    out["cf32"]["frac_le_1e-8"] = float(np.mean(rel <= 1e-8))
    return out


FP32_PEAK_TFLOPS = 157.3


def contour_flops(work, nlenses, ld, npts_ld=100, vjp=False, cf32=False):
    """Algorithmic flops of a contour run from its work counters (FP64 root-iterations,
    FP32 root-iterations, compensated root-iterations, limb-point solves, contour
    points), by the counting rule of the root solver plus, per limb-point solve, the
    coefficient build (COEFF_FLOPS) and n x the image terms (IMAGE_FLOPS: lens equation
    and det J), and per limb-darkened contour point 2 npts_ld Gauss-Legendre nodes
    (NODE_FLOPS_FWD, or NODE_FLOPS_VJP with the reverse mode, which also adds the image
    adjoint of every contour point).  The compensated polish that ends every limb solve
    is this implementation's addition and is not counted.  With cf32 the compensated
    root-iterations are the CF32 form's FP32 sweep.  Returns (FP64, FP32 flops)."""
    n = nlenses * nlenses + 1
    it64, it32, itc, solves, points = (float(x) for x in work)
    f64 = it64 * flops_per_root_iter(n) + (0.0 if cf32 else itc * flops_comp_root_iter(n))
    f64 += solves * (COEFF_FLOPS[nlenses] + n * IMAGE_FLOPS[nlenses])
    node = (NODE_FLOPS_VJP if vjp else NODE_FLOPS_FWD)[nlenses]
    f64 += points * (2 * npts_ld * node if ld else 4)
    if vjp:
        f64 += points * IMAGE_ADJ_FLOPS[nlenses]
    f32 = it32 * flops_per_root_iter(n) + (itc * flops_cf32_comp_root_iter(n) if cf32 else 0.0)
    return f64, f32


def pointsource_vjp_flops(epochs, nlenses, ndir):
    """Reverse-mode flops of the fused point-source VJP per the counting rule: for each
    of the n roots and each of ndir directions, the implicit image tangent
    dz = (u - K conj u)/det (about 30) and the tangent of 1/|det J| (about 20)."""
    n = nlenses * nlenses + 1
    return epochs * n * ndir * 50.0


def _roofline_entry(flops, dt):
    f64, f32 = flops
    busy = (f64 / (FP64_PEAK_TFLOPS * 1e12) + f32 / (FP32_PEAK_TFLOPS * 1e12)) / dt
    return {"bound": "valu", "achieved_fp64_tflops": f64 / dt / 1e12,
            "achieved_fp32_tflops": f32 / dt / 1e12, "frac": busy,
            "fp64_flops": f64, "fp32_flops": f32, "seconds": dt}


# ---- CPU baselines: the oracle (NumPy restatement + the C root oracle) on the host ----
def host_cores():
    """(cores used, how they were counted): every CPU of this process's affinity mask,
    unless a cgroup CPU quota (cgroup v2 cpu.max) grants fewer CPUs' worth of time."""
    n = len(os.sched_getaffinity(0))
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            q = int(-(-int(quota) // int(period)))
            if 0 < q < n:
                return q, f"cgroup cpu.max quota {q} of {n} affinity CPUs"
    except (OSError, ValueError):
        pass
    return n, f"{n} affinity CPUs"


def _pool_init(root):
    if root not in sys.path:
        sys.path.insert(0, root)
    try:
        ctypes.CDLL("libgomp.so.1").omp_set_num_threads(1)   # one thread per worker
    except OSError:
        pass


def _oracle_extended_pair(kw):
    """(oracle as the reference computes it, oracle with the canonical root polish)"""
    from oracle import extended as OE
    return OE.mag_extended_source(**kw), OE.mag_extended_source(polish=True, **kw)


def _oracle_truth(kw):
    """source-plane quadrature of the same magnification at two resolutions"""
    from oracle import extended as OE
    k = {n: kw[n] for n in ("s", "q", "q3", "r3", "psi") if n in kw}
    return [OE.mag_disk_quadrature(kw["w0"], kw["rho"], kw["nlenses"], u1=kw["u1"], nr=nr, **k)
            for nr in (100, 200)]


def _oracle_point_source(args):
    from oracle import lens
    w, prm = args
    return lens.mag_point_source(w, 2, **prm)


def cpu_rates_omp(call, n_evals, budget_s, what, unit):
    """Evaluations per second of call(nthreads) (n_evals per call) on one OpenMP thread
    and on every core of the process, each repeated for about budget_s."""
    ncores, how = host_cores()
    out = None
    for nt in (1, ncores):
        call(nt)   # warm: thread pool, caches
        n, t0 = 0, time.perf_counter()
        while True:
            call(nt)
            n += 1
            if time.perf_counter() - t0 > budget_s:
                break
        dt = time.perf_counter() - t0
        ent = {"value": n * n_evals / dt, "unit": unit, "cores": nt, "kind": "port",
               "sample": f"{n} x {n_evals} {what}, {dt:.1f} s, {nt} OpenMP thread(s)"
                         + ("" if nt == 1 else f" ({how})")}
        if out is None:
            out = ent
        else:
            out["all_cores"] = ent
    print(f"bench: CPU {what}: 1 core {out['value']:.4g}, {ncores} cores "
          f"{out['all_cores']['value']:.4g} {unit}", file=sys.stderr, flush=True)
    return out


def cpu_rates_omp_items(call, n_items, budget_s, what, unit):
    """Evaluations per second of call(lo, hi, nthreads), which evaluates items lo..hi-1 of
    a fixed list, on one OpenMP thread and on every core of the process.  Each thread
    count walks the list from its start in slices that double from 8 items (at least 4
    per thread) until budget_s is spent, so both rates cover the same leading items."""
    ncores, how = host_cores()
    out = None
    for nt in (1, ncores):
        call(0, min(n_items, nt), nt)   # warm: thread pool, caches
        lo, size, t0 = 0, max(8, 4 * nt), time.perf_counter()
        while lo < n_items:
            hi = min(n_items, lo + size)
            call(lo, hi, nt)
            lo, size = hi, 2 * size
            if time.perf_counter() - t0 > budget_s:
                break
        dt = time.perf_counter() - t0
        ent = {"value": lo / dt, "unit": unit, "cores": nt, "kind": "port",
               "sample": f"the first {lo} {what}, {dt:.1f} s, {nt} OpenMP thread(s)"
                         + ("" if nt == 1 else f" ({how})")}
        if out is None:
            out = ent
        else:
            out["all_cores"] = ent
    print(f"bench: CPU {what}: 1 core {out['value']:.4g}, {ncores} cores "
          f"{out['all_cores']['value']:.4g} {unit}", file=sys.stderr, flush=True)
    return out


def _extended_c_call(items, polish=False):
    """call(lo, hi, nthreads) over a list of mag_extended_source keyword sets with one lens
    count: the compiled oracle on items lo..hi-1 (polish: with the canonical root polish
    the GPU path applies)"""
    from oracle import extended as OE
    names = ("s", "q", "q3", "r3", "psi")
    cols = {k: np.array([it[k] for it in items]) for k in ("w0", "rho", "npts_limb", "u1")}
    prm = {k: np.array([it[k] for it in items]) for k in names if k in items[0]}
    nl = items[0]["nlenses"]

    def call(lo, hi, nt):
        m, bad = OE.mag_extended_source_c(
            cols["w0"][lo:hi], cols["rho"][lo:hi], nl, npts_limb=cols["npts_limb"][lo:hi],
            limb_darkening=True, u1=cols["u1"][lo:hi], npts_ld=100, nthreads=nt, polish=polish,
            **{k: v[lo:hi] for k, v in prm.items()})
        return m
    return call


def bench_extended_binary(steps, warmup, n_near=2048, n_far=2048, cpu_s=0.0):
    """Binary limb-darkened (u=0.5) sources, rho log-uniform [1e-3,1e-1]: n_near within
    2 rho of a caustic with ragged contour resolution 256..4096 points, n_far accepted by
    the hexadecapole gate.  One step is mag() forward + VJP: the gate on all sources,
    the hexadecapole VJP on the accepted ones and the contour pipeline with its VJP on
    every source the gate rejects.  Both solver forms (FP64, CF32)."""
    import torch

    from caustics_amd import _lib
    from caustics_amd import workloads as W
    from caustics_amd.lightcurve import _mag_fwd

    b = W.extended_binary_batch(n_near, n_far)
    S = len(b["w"])
    w = torch.as_tensor(b["w"], device="cuda")
    rho = torch.as_tensor(b["rho"], device="cuda")
    u1 = torch.as_tensor(b["u1"], device="cuda")
    prm = torch.as_tensor(np.stack([b["s"], b["q"]], -1), device="cuda").contiguous()
    npts_all = np.concatenate([b["npts"], np.full(n_far, 256)]).astype(np.int64)
    gm = torch.ones(S, dtype=torch.float64, device="cuda")
    out = {"workload": f"{S} binary limb-darkened sources (u=0.5), {n_near} near caustics "
                       "with ragged 256-4096 limb points, rest hexadecapole-gated; mag() "
                       "fwd+VJP"}
    res = {}
    for form, mode in (("fp64", _lib.MODE_FP64), ("cf32", _lib.MODE_CF32)):
        cfg = (2, npts_all, int(npts_all.max()), True, 100, 2500, mode)

        def step():
            res[form] = _mag_fwd(cfg, w, rho, u1, prm, g=gm)
        dt = _timed(step, steps, warmup, f"C4 mag() fwd+VJP {form}")
        work = torch.zeros(5, dtype=torch.int64, device="cuda")
        _mag_fwd(cfg, w, rho, u1, prm, g=gm, work=work)
        work = work.cpu().numpy()
        m, use, go = res[form]
        out[form] = {"mag_evals_per_s_fwd_vjp": S / dt, "ms": dt * 1e3,
                     "contour_sources": int((~use).sum().item()),
                     "hexadecapole_sources": int(use.sum().item()),
                     "work": {"root_iterations_fp64": int(work[0]),
                              "root_iterations_fp32": int(work[1]),
                              "limb_solves": int(work[3]), "contour_points": int(work[4])},
                     "roofline": _roofline_entry(contour_flops(work, 2, True, vjp=True,
                                                               cf32=form == "cf32"), dt),
                     "grad_finite": bool(torch.isfinite(go).all().item())}
    m64, m32 = res["fp64"][0], res["cf32"][0]
    e = ((m32 - m64).abs() / m64.abs()).cpu().numpy()
    ge = ((res["cf32"][2] - res["fp64"][2]).abs()
          / torch.maximum(res["fp64"][2].abs(), 1e-3 * m64.abs()[:, None])).cpu().numpy()
    out["cf32"].update({"max_rel_err_vs_fp64": float(e.max()),
                        "p99_rel_err_vs_fp64": float(np.quantile(e, 0.99)),
                        "frac_le_1e-8": float(np.mean(e <= 1e-8)),
                        "grad_p99_rel_err_vs_fp64": float(np.quantile(ge, 0.99))})
    if cpu_s > 0:
        items = (dict(w0=complex(b["w"][i]), rho=float(b["rho"][i]), nlenses=2,
                      npts_limb=int(b["npts"][i]), limb_darkening=True, u1=float(b["u1"][i]),
                      npts_ld=100, s=float(b["s"][i]), q=float(b["q"][i]))
                 for i in range(n_near))
        items = list(items)
        out["cpu_baseline"] = cpu_rates_omp_items(
            _extended_c_call(items), len(items), cpu_s,
            "near-caustic C4 sources (their ragged npts_limb), forward only (compiled C++ "
            "oracle, oracle/extended_oracle.cpp)", "mag evals/s (forward only)")
        # error of the FP64 GPU form against the oracle on the sources the gate sent to
        # the contour route (the first 512 of them, their npts_limb)
        use64 = res["fp64"][1].cpu().numpy()
        idx = [i for i in range(n_near) if not use64[i]][:512]
        mo = _extended_c_call([items[i] for i in idx])(0, len(idx), host_cores()[0])
        gv = m64.cpu().numpy()[idx]
        out["fp64"]["vs_oracle"] = _rel_err_stats(gv, mo)
        out["fp64"]["vs_oracle"]["sample"] = (f"the first {len(idx)} gate-rejected near-caustic "
                                              "sources, their npts_limb (compiled oracle)")
        mp_ = _extended_c_call([items[i] for i in idx], polish=True)(0, len(idx), host_cores()[0])
        out["fp64"]["vs_oracle"]["polished_oracle"] = _rel_err_stats(gv, mp_)
        out["fp64"]["vs_oracle"]["largest_differences_vs_truth"] = _closer_to_truth(
            gv, mo, [items[i] for i in idx])
    return out


def _closer_to_truth(gpu, ref, kws, top=20):
    """The largest GPU-oracle differences compared with the source-plane quadrature of
    the same disk (oracle/extended.py mag_disk_quadrature at 200 nodes, 100 for the
    resolution gap): which of the two lies closer to it."""
    import multiprocessing as mp
    e = np.abs(gpu - ref) / np.abs(ref)
    idx = np.argsort(-e)[:top]
    with mp.get_context("spawn").Pool(host_cores()[0], initializer=_pool_init,
                                      initargs=(ROOT,)) as pool:
        tr = np.array(pool.map(_oracle_truth, [kws[i] for i in idx], chunksize=1))
    t = tr[:, 1]
    eg, eo = np.abs(gpu[idx] - t) / t, np.abs(ref[idx] - t) / t
    return {"sources": int(len(idx)), "gpu_closer": int((eg < eo).sum()),
            "oracle_closer": int((eo < eg).sum()), "gpu_median_rel_err": float(np.median(eg)),
            "oracle_median_rel_err": float(np.median(eo)),
            "quadrature_resolution_gap_median": float(np.median(np.abs(tr[:, 0] - t) / t))}


def _rel_err_stats(gpu, ref):
    e = np.abs(gpu - ref) / np.abs(ref)
    return {"n": int(e.size), "max_rel_err": float(e.max()),
            "p99_rel_err": float(np.quantile(e, 0.99)), "median_rel_err": float(np.median(e)),
            "frac_le_1e-10": float(np.mean(e <= 1e-10)), "frac_le_1e-8": float(np.mean(e <= 1e-8))}


def bench_triple(steps, warmup, n_sets=1024, n_epochs=1500, budget_s=40.0, cpu_s=0.0):
    """Triple-lens caustic-crossing light curves (u=0.4, rho log-uniform [1e-3,3e-2]),
    full limb-darkened contour at every epoch with uneven limb resolutions (256 far from
    the crossing up to 4096 within one source radius, W.crossing_npts), fused fwd+VJP,
    both solver forms.  The whole batch (ragged ordering, chunks on two streams, every
    kernel) is captured once in a HIP graph and the timed step is one replay, checked bit
    for bit against the eager call.  The timed batch is the full n_sets x n_epochs when one
    step fits budget_s, otherwise the largest multiple of 8 parameter sets that does
    (stated in the result)."""
    import torch

    from caustics_amd import _lib
    from caustics_amd import workloads as W
    from caustics_amd.extended_source import _run

    b = W.triple_caustic_lightcurves(n_sets, n_epochs)
    npts_all = W.crossing_npts(b)
    NMAX = int(npts_all.max())

    def batch(k):
        w = torch.as_tensor(b["w"][:k].reshape(-1), device="cuda")
        rep = lambda x: torch.as_tensor(np.repeat(np.asarray(x)[:k], n_epochs), device="cuda")
        prm = torch.stack([rep(b[c]) for c in ("s", "q", "q3", "r3", "psi")], -1).contiguous()
        return w, rep(b["rho"]), rep(b["u1"]), prm, npts_all[:k].reshape(-1).astype(np.int64)

    # size the timed batch from a probe of 8 parameter sets
    w, rho, u1, prm, nh = batch(8)
    gm = torch.ones(w.numel(), dtype=torch.float64, device="cuda")
    _run(3, w, rho, prm, nh, NMAX, True, u1, 100, 2500, _lib.MODE_FP64, g=gm)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    _run(3, w, rho, prm, nh, NMAX, True, u1, 100, 2500, _lib.MODE_FP64, g=gm)
    torch.cuda.synchronize()
    per_set = (time.perf_counter() - t0) / 8
    k = int(min(n_sets, max(8, budget_s / 2.5 / max(per_set, 1e-9))) // 8 * 8)
    w, rho, u1, prm, nh = batch(k)
    S = w.numel()
    gm = torch.ones(S, dtype=torch.float64, device="cuda")
    hist = {int(v): int(c) for v, c in zip(*np.unique(nh, return_counts=True))}
    out = {"workload": f"{k} of {n_sets} triple-lens parameter sets x {n_epochs} epochs "
                       "(caustic crossing), limb-darkened u=0.4, npts_limb 256..4096 by distance "
                       "to the crossing, full contour fwd+VJP at every epoch, one HIP graph "
                       "replay", "evals": S, "npts_limb_histogram": hist,
           "limb_points": int(nh.sum())}
    res = {}
    for form, mode in (("fp64", _lib.MODE_FP64), ("cf32", _lib.MODE_CF32)):
        def call():
            res[form] = _run(3, w, rho, prm, nh, NMAX, True, u1, 100, 2500, mode, g=gm)
        t0 = time.perf_counter()
        call()
        torch.cuda.synchronize()
        eager_s = time.perf_counter() - t0
        eager = [x.clone() for x in (res[form][0], res[form][2])]
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        with torch.cuda.stream(side):
            call()
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        graph = torch.cuda.CUDAGraph()
        with torch.cuda.graph(graph):
            call()
        captured = res[form]
        dt = _timed(graph.replay, 1, 1, f"C5 triple contours {form} graph replay ({S} contours)")
        same = bool(torch.equal(captured[0], eager[0]) and torch.equal(captured[2], eager[1]))
        work = torch.zeros(5, dtype=torch.int64, device="cuda")
        _run(3, w, rho, prm, nh, NMAX, True, u1, 100, 2500, mode, g=gm, work=work)
        work = work.cpu().numpy()
        out[form] = {"mag_evals_per_s_fwd_vjp": S / dt, "ms": dt * 1e3,
                     "eager_ms_first_call": eager_s * 1e3, "graph_replay_matches_eager": same,
                     "unconverged": int((captured[1] == 0).sum().item()),
                     "work": {"root_iterations_fp64": int(work[0]),
                              "root_iterations_fp32": int(work[1]),
                              "root_iterations_comp": int(work[2]),
                              "limb_solves": int(work[3]), "contour_points": int(work[4])},
                     "roofline": _roofline_entry(contour_flops(work, 3, True, vjp=True,
                                                               cf32=form == "cf32"), dt)}
        res[form] = (captured[0].clone(), captured[1].clone(), captured[2].clone())
        del graph, captured
    m64, m32 = res["fp64"][0], res["cf32"][0]
    e = ((m32 - m64).abs() / m64.abs()).cpu().numpy()
    out["cf32"].update({"max_rel_err_vs_fp64": float(e.max()),
                        "p99_rel_err_vs_fp64": float(np.quantile(e, 0.99)),
                        "frac_le_1e-8": float(np.mean(e <= 1e-8))})
    if cpu_s > 0:
        def kw(kk, e):
            return dict(w0=complex(b["w"][kk][e]), rho=float(b["rho"][kk]), nlenses=3,
                        npts_limb=int(npts_all[kk][e]), limb_darkening=True,
                        u1=float(b["u1"][kk]), npts_ld=100, s=float(b["s"][kk]),
                        q=float(b["q"][kk]), q3=float(b["q3"][kk]), r3=float(b["r3"][kk]),
                        psi=float(b["psi"][kk]))
        # every 25th epoch of the parameter sets in order, so the sample spans whole
        # light curves (caustic crossings included)
        items = [kw(kk, e) for kk in range(n_sets) for e in range(0, n_epochs, 25)][:20000]
        out["cpu_baseline"] = cpu_rates_omp_items(
            _extended_c_call(items), len(items), cpu_s,
            "C5 epochs (every 25th epoch of the light curves in order, their npts_limb), "
            "forward only (compiled C++ oracle, oracle/extended_oracle.cpp)",
            "mag evals/s (forward only)")
        # error of the FP64 GPU form against the oracle on a sample around and away from
        # the caustic crossings: the reference's route as such, and with the canonical
        # root polish that the GPU path applies (where the reference's plain stopping
        # test leaves ill-conditioned images off by enough to flip the 1e-6 filter)
        nl = min(k, 16)
        rows = [(kk, e) for kk in range(nl) for e in (300, 744, 749, 753, 1200)]
        samp = [kw(kk, e) for kk, e in rows]
        import multiprocessing as mp
        with mp.get_context("spawn").Pool(host_cores()[0], initializer=_pool_init,
                                          initargs=(ROOT,)) as pool:
            pairs = np.array(pool.map(_oracle_extended_pair, samp, chunksize=1))
            truth = np.array(pool.map(_oracle_truth, samp, chunksize=1))
        gi = np.array([kk * n_epochs + e for kk, e in rows])
        g64 = m64.cpu().numpy()[gi]
        e_p = np.abs(g64 - pairs[:, 0]) / pairs[:, 0]
        e_c = np.abs(g64 - pairs[:, 1]) / pairs[:, 1]
        tq = truth[:, 1]
        conv = np.abs(truth[:, 0] - tq) <= 1e-7 * tq
        eg_t = np.abs(g64 - tq) / tq
        eo_t = np.abs(pairs[:, 0] - tq) / tq
        out["fp64"]["vs_oracle"] = {
            "epochs": len(rows), "sample": "epochs 300, 744, 749, 753, 1200 of the first "
                                           f"{nl} light curves (crossing near 749.5), their "
                                           "npts_limb",
            "frac_le_1e-10": float(np.mean(e_p <= 1e-10)),
            "frac_le_1e-8": float(np.mean(e_p <= 1e-8)),
            "max_rel_err": float(e_p.max()), "p99_rel_err": float(np.quantile(e_p, 0.99)),
            "polished_oracle": {"frac_le_1e-10": float(np.mean(e_c <= 1e-10)),
                                "frac_le_1e-8": float(np.mean(e_c <= 1e-8)),
                                "max_rel_err": float(e_c.max())},
            # both against the source-plane quadrature of the same disk (oracle/extended.py
            # mag_disk_quadrature, 200 nodes; "converged" where 100 nodes agree to 1e-7)
            "vs_truth": {
                "converged_epochs": int(conv.sum()),
                "gpu_median_rel_err_converged": float(np.median(eg_t[conv])) if conv.any() else None,
                "oracle_median_rel_err_converged": float(np.median(eo_t[conv])) if conv.any() else None,
                "gpu_max_rel_err_converged": float(eg_t[conv].max()) if conv.any() else None,
                "oracle_max_rel_err_converged": float(eo_t[conv].max()) if conv.any() else None,
                "frac_gpu_oracle_gap_within_oracle_error": float(
                    np.mean(np.abs(g64 - pairs[:, 0]) <= np.abs(pairs[:, 0] - tq))),
                "where_gpu_differs_gpu_closer": int(((e_p > 1e-10) & (eg_t < eo_t)).sum()),
                "where_gpu_differs_oracle_closer": int(((e_p > 1e-10) & (eo_t < eg_t)).sum())}}
        # and on a wide sample: every 25th epoch of the first 64 light curves (compiled
        # oracle, the reference's route without the polish)
        nw = min(k, 64)
        wide = [(kk, e) for kk in range(nw) for e in range(0, n_epochs, 25)]
        mo = _extended_c_call([kw(kk, e) for kk, e in wide])(0, len(wide), host_cores()[0])
        gi = np.array([kk * n_epochs + e for kk, e in wide])
        gv = m64.cpu().numpy()[gi]
        out["fp64"]["vs_oracle_wide"] = _rel_err_stats(gv, mo)
        out["fp64"]["vs_oracle_wide"]["sample"] = (f"every 25th epoch of the first {nw} light "
                                                   "curves, their npts_limb (compiled oracle)")
        mp_ = _extended_c_call([kw(kk, e) for kk, e in wide], polish=True)(0, len(wide),
                                                                          host_cores()[0])
        out["fp64"]["vs_oracle_wide"]["polished_oracle"] = _rel_err_stats(gv, mp_)
        out["fp64"]["vs_oracle_wide"]["largest_differences_vs_truth"] = _closer_to_truth(
            gv, mo, [kw(kk, e) for kk, e in wide])
    return out


def bench_trajectory(steps, warmup, n_epochs=1500):
    """Annual-parallax source trajectories (src/caustics/trajectory.py:104-157) for L
    parameter sets over 1500 epochs: forward (w written, 16 B per curve-epoch) and
    reverse mode (the cotangent of w read, 16 B per curve-epoch, five sums per curve),
    L = 1024 and 16384.  The Sun table is an analytic circular orbit on the daily grid
    (the ephemeris is an input of the kernel).  HIP events on the launch stream."""
    import torch

    from caustics_amd import _lib

    g = np.random.default_rng(12)
    t = torch.as_tensor(np.sort(g.uniform(8000.0, 8400.0, n_epochs)), device="cuda")
    xp_h = np.arange(8000.0, 8401.0, 1.0)
    om = 2 * np.pi / 365.25
    tab_h = np.stack([np.sin(om * xp_h), 0.4 * np.cos(om * xp_h), om * np.cos(om * xp_h),
                      -0.4 * om * np.sin(om * xp_h)])
    xp = torch.as_tensor(xp_h, device="cuda")
    tab = torch.as_tensor(tab_h, device="cuda").contiguous()
    L_ = _lib.lib()
    out = {}
    for L in (1024, 16384):
        P = torch.as_tensor(np.stack([g.uniform(8050, 8350, L), g.uniform(-1, 1, L),
                                      g.uniform(20, 200, L), g.uniform(-0.5, 0.5, L),
                                      g.uniform(-0.5, 0.5, L)], 1), device="cuda").contiguous()
        w = torch.empty((L, n_epochs, 2), dtype=torch.float64, device="cuda")
        gw = torch.as_tensor(g.normal(size=(L, n_epochs, 2)), device="cuda")
        gP = torch.empty_like(P)
        stream = torch.cuda.current_stream()
        res = {}
        for kind in ("fwd", "vjp"):
            def call():
                rc = L_.caus_trajectory_annual_parallax(
                    _lib.stream_ptr(), _lib.ptr(t), n_epochs, _lib.ptr(xp), _lib.ptr(tab),
                    xp_h.size, _lib.ptr(P), L, 0, _lib.ptr(w) if kind == "fwd" else None,
                    _lib.ptr(gw) if kind == "vjp" else None,
                    _lib.ptr(gP) if kind == "vjp" else None)
                assert rc == 0

            for _ in range(warmup):
                call()
            torch.cuda.synchronize()
            e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
            e0.record(stream)
            for _ in range(steps):
                call()
            e1.record(stream)
            torch.cuda.synchronize()
            ms = e0.elapsed_time(e1) / steps
            nbytes = 16.0 * L * n_epochs
            res[kind] = {"ms": ms, "achieved_GBps": nbytes / ms / 1e6,
                         "frac": nbytes / ms / 1e6 / 8000.0, "bytes_per_launch": nbytes}
            print(f"bench: trajectory L={L} {kind}: {ms * 1e3:.1f} us/launch, "
                  f"{nbytes / ms / 1e6:.0f} GB/s", file=sys.stderr, flush=True)
        out[f"L{L}"] = {"curves": L, "epochs": n_epochs, **res,
                        "finite": bool(torch.isfinite(gP).all().item())}
    out["workload"] = ("annual-parallax trajectories w(t) and d/d(t0, u0, tE, piEE, piEN), "
                       "1500 epochs, daily Sun table")
    return out


def bench_likelihood(steps, warmup, n_epochs=1500, lens=(900, 600)):
    """Flux-marginalised log-likelihood with its gradient w.r.t. the magnifications
    (src/caustics/linalg.py:56-70) for B magnification curves over 1500 epochs of two
    observatories: B = 1024 (one HMC step of the C5 parameter sets) and B = 16384 (large
    enough to price the kernel against HBM).  Algorithmic bytes per launch: A read once
    and g_A written once, 16 B per curve-epoch (fobs and C_inv, shared by all curves,
    add 16 B per epoch); HIP events on the stream the kernel is launched on."""
    import torch

    from caustics_amd import _lib

    g = np.random.default_rng(11)
    n = n_epochs
    t_host = np.sort(g.uniform(-40, 40, n))
    t = torch.as_tensor(t_host, device="cuda")
    f_host = 800.0 * (1.0 + 1.0 / (1.0 + (t_host / 10.0) ** 2)) + 200.0
    fobs = torch.as_tensor(f_host + g.normal(0, 10.0, n), device="cuda")
    c_inv = torch.as_tensor(1.0 / (10.0 * g.uniform(0.7, 1.3, n)) ** 2, device="cuda")
    off = torch.as_tensor(np.concatenate([[0], np.cumsum(lens)]), dtype=torch.int64,
                          device="cuda")
    L = _lib.lib()
    out = {}
    for B in (1024, 16384):
        u0 = torch.as_tensor(g.uniform(0.05, 0.6, B), device="cuda")[:, None]
        tE = torch.as_tensor(g.uniform(8, 40, B), device="cuda")[:, None]
        u = torch.hypot(u0, t[None, :] / tE)
        A = ((u * u + 2) / (u * torch.sqrt(u * u + 4))).contiguous()
        ll = torch.empty(B, dtype=torch.float64, device="cuda")
        beta = torch.empty((B, 2, 2), dtype=torch.float64, device="cuda")
        gA = torch.empty((B, n), dtype=torch.float64, device="cuda")
        stream = torch.cuda.current_stream()

        def call():
            rc = L.caus_marginalized_log_likelihood(
                _lib.stream_ptr(), _lib.ptr(A), B, n, n, _lib.ptr(fobs), _lib.ptr(c_inv),
                _lib.ptr(off), len(lens), _lib.ptr(ll), _lib.ptr(beta), None, _lib.ptr(gA), n)
            assert rc == 0

        for _ in range(warmup):
            call()
        torch.cuda.synchronize()
        e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        for _ in range(steps):
            call()
        e1.record(stream)
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        nbytes = 16.0 * B * n + 16.0 * n
        out[f"B{B}"] = {"curves": B, "epochs": n, "ms": ms, "curves_per_s": B / ms * 1e3,
                        "roofline": {"bound": "hbm", "achieved_GBps": nbytes / ms / 1e6,
                                     "peak_GBps": 8000.0,
                                     "frac": nbytes / ms / 1e6 / 8000.0,
                                     "bytes_per_launch": nbytes},
                        "finite": bool(torch.isfinite(ll).all().item())}
        print(f"bench: likelihood B={B}: {ms * 1e3:.1f} us/launch, "
              f"{nbytes / ms / 1e6:.0f} GB/s", file=sys.stderr, flush=True)
    out["workload"] = ("flux-marginalised log-likelihood + d ll/dA, 1500 epochs in two "
                       "observatories (900 + 600), diagonal covariance")
    return out


def root_condition(c, z):
    """Relative condition number of each root: sum_k |c_k||z|^k / (|z| |p'(z)|)."""
    k = np.arange(c.shape[-1])
    az = np.abs(z)[..., None]
    num = (np.abs(c)[:, None, :] * az**k).sum(-1)
    dp = (k[1:] * c[:, None, 1:] * z[..., None] ** (k[1:] - 1)).sum(-1)
    return num / np.maximum(np.abs(z) * np.abs(dp), 1e-300)


def oracle_accuracy(host, results, nsamp=4096):
    """Root error of each form against the FP64 C oracle on the first nsamp polynomials
    of each batch: max and p99 relative error of the matched roots, and the fraction
    within max(1e-12, 64 kappa eps), kappa the root's condition number (clustered roots
    are determined only to that level by any FP64 solver)."""
    from oracle import ea

    for name, r in results.items():
        errs, oks = [], []
        for nl in (2, 3):
            c = host[nl][:nsamp]
            zo = ea.roots_asc(c, init_mode=1)
            zg = r["z"][nl][:nsamp].cpu().numpy()
            d = np.abs(zg[:, :, None] - zo[:, None, :]).min(1) / np.abs(zo)
            tol = np.maximum(1e-12, 64 * root_condition(c, zo) * 2.0**-53)
            errs.append(d.ravel())
            oks.append((d <= tol).ravel())
        e = np.concatenate(errs)
        r["err_max"] = float(e.max())
        r["err_p99"] = float(np.quantile(e, 0.99))
        r["frac_within_condition_tol"] = float(np.concatenate(oks).mean())


def cpu_baseline(host, seconds):
    """The reference's CPU op (compiled from its sources into oracle/_ref) on one core
    over a sample of the same polynomials; falls back to the C oracle restatement."""
    from oracle import ea, ref_op

    kind = "reference" if ref_op.available() else "port"
    total_t, total_n = 0.0, 0
    # same 4:1 mix of degree-5 and degree-10 polynomials as the workload
    chunk5, chunk10 = 1 << 14, 1 << 12
    off = 0
    t_start = time.perf_counter()
    while time.perf_counter() - t_start < seconds and off + chunk5 <= host[2].shape[0]:
        c5 = host[2][off:off + chunk5]
        c10 = host[3][off // 4:off // 4 + chunk10]
        t0 = time.perf_counter()
        if kind == "reference":
            ref_op.roots_asc(c5)
            ref_op.roots_asc(c10)
        else:
            ea.roots_asc(c5, nthreads=1)
            ea.roots_asc(c10, nthreads=1)
        total_t += time.perf_counter() - t0
        total_n += c5.shape[0] + c10.shape[0]
        off += chunk5
    out = {"value": total_n / total_t, "unit": "polys/s", "cores": 1, "kind": kind,
           "sample": f"{total_n} polynomials ({off} degree-5 + {off // 4} degree-10) of the "
                     f"same workload, FP64, plain Ehrlich-Aberth, {total_t:.1f} s"}
    # the C restatement with OpenMP over polynomials on every core this process may use
    ncores, how = host_cores()
    n5 = min(host[2].shape[0], max(chunk5, int(out["value"] * seconds * ncores * 0.4) // 5 * 4))
    c5, c10 = host[2][:n5], host[3][:n5 // 4]
    ea.roots_asc(c5[:1024], nthreads=ncores)
    t0 = time.perf_counter()
    ea.roots_asc(c5, nthreads=ncores)
    ea.roots_asc(c10, nthreads=ncores)
    dt = time.perf_counter() - t0
    out["all_cores"] = {"value": (n5 + n5 // 4) / dt, "unit": "polys/s", "cores": ncores,
                        "kind": "port", "sample": f"{n5} degree-5 + {n5 // 4} degree-10, "
                                                  f"{dt:.1f} s, {how}"}
    return out


if __name__ == "__main__":
    main()
