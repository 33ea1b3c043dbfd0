"""Benchmark of the magnification hot path on one MI355X.

Headline workload (one "step"): the batched root solve of 2^20 degree-5 (binary-lens)
and 2^18 degree-10 (triple-lens) polynomials built from seeded lens configurations
(s log-uniform [0.3,3], q log-uniform [1e-4,1]; third lens r ~ U[0.3,2],
psi ~ U[0,2pi), q3 log-uniform [1e-4,1e-1]; sources uniform in |w| <= 2), coefficients
resident in HBM, FP64 Ehrlich-Aberth.  value = polynomials solved per second.

Also reported in the same JSON line: the compensated-FP32 (CF32) form of the same
solve, per-form error against the FP64 CPU oracle, the roofline of the dominant
kernel (algorithmic FP64 flops per launch / HIP-event launch time vs the MI355X FP64
vector peak), and the CPU baseline (the reference's own CPU op, compiled from its
sources, one core, on a bounded sample).

usage: python bench.py [--gpus 1] [--steps K] [--warmup W]
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

FP64_PEAK_TFLOPS = 78.6   # MI355X FP64 vector, dense
HBM_PEAK_GBS = 8000.0


def flops_per_root_iter(n):
    """Algorithmic FP64 flops of one active root-iteration of the plain solver
    (28n + 20: |z|, real Horner 2n, complex Horner value+derivative 16n, test,
    Aberth sum 10(n-1) and correction 19, update 2)."""
    return 28 * n + 20


def flops_comp_root_iter(n):
    return 227 * n + 33


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
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--n5", type=int, default=1 << 20)
    ap.add_argument("--n10", type=int, default=1 << 18)
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--no-cpu", action="store_true")
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

    forms = {"fp64": (_lib.MODE_FP64, _lib.INIT_COMPLEX),
             "cf32": (_lib.MODE_CF32, _lib.INIT_COMPLEX),
             "fp64_refinit": (_lib.MODE_FP64, _lib.INIT_REFERENCE)}
    results = {}
    for name, (mode, init) in forms.items():
        steps = args.steps if name == "fp64" else max(3, args.steps // 4)
        for _ in range(args.warmup):
            for nl in (2, 3):
                launch(nl, mode, init)
        torch.cuda.synchronize()
        ev = {nl: [] for nl in (2, 3)}
        t0 = time.perf_counter()
        for _ in range(steps):
            for nl in (2, 3):
                e0 = torch.cuda.Event(enable_timing=True)
                e1 = torch.cuda.Event(enable_timing=True)
                e0.record()
                launch(nl, mode, init)
                e1.record()
                ev[nl].append((e0, e1))
        torch.cuda.synchronize()
        t1 = time.perf_counter()
        kms = {nl: float(np.mean([a.elapsed_time(b) for a, b in ev[nl]])) for nl in (2, 3)}
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
                             unconverged=unconv)
        print(f"bench: {name}: {results[name]['value']:.4g} polys/s, kernel ms {kms}",
              file=sys.stderr, flush=True)

    # accuracy vs the FP64 CPU oracle on a sample of each batch
    from oracle import ea

    sys.path.insert(0, os.path.join(ROOT, "tests"))
    nsamp = 4096
    for name, r in results.items():
        errs = []
        for nl in (2, 3):
            c = host[nl][:nsamp]
            zo = ea.roots_asc(c, init_mode=1)
            zg = r["z"][nl][:nsamp].cpu().numpy()
            d = np.abs(zg[:, :, None] - zo[:, None, :]).min(-1) / np.abs(zo)
            errs.append(d.ravel())
        e = np.concatenate(errs)
        r["err_max"] = float(e.max())
        r["err_p99"] = float(np.quantile(e, 0.99))
        del r["z"]

    main_r = results["fp64"]
    # roofline of the dominant kernel (FP64 form)
    dom = max((2, 3), key=lambda nl: main_r["kernel_ms"][nl])
    n = degs[dom]
    w = main_r["work"][dom]
    flops = float(w[1]) * flops_per_root_iter(n) + float(w[2]) * flops_comp_root_iter(n)
    achieved = flops / (main_r["kernel_ms"][dom] * 1e-3) / 1e12
    traffic = None
    tpath = os.path.join(ROOT, "profiles", "traffic.json")
    if os.path.exists(tpath):
        traffic = json.load(open(tpath)).get(f"roots_deg{n}")

    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(host, args.cpu_seconds)

    line = {
        "metric": "polys/s deg 5/10 batched root solve (FP64, roots matched to FP64 oracle); "
                  "mag evals/s fwd+VJP at <=1e-8 rel err vs FP64",
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
                     "kernel": f"k_roots<{n},FP64,fast>", "kernel_ms": main_r["kernel_ms"][dom],
                     "flops_per_launch": flops,
                     "flop_rule": f"{flops_per_root_iter(n)} per active FP64 root-iteration"},
        "cpu_baseline": cpu,
        "forms": {k: {kk: vv for kk, vv in v.items() if kk not in ("work",)}
                  for k, v in results.items()},
        "work_counts": {k: {str(degs[nl]): [int(x) for x in v["work"][nl]] for nl in (2, 3)}
                        for k, v in results.items()},
    }
    print(json.dumps(line, default=float))


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
    return {"value": total_n / total_t, "unit": "polys/s", "cores": 1, "kind": kind,
            "sample": f"{total_n} polynomials ({off} degree-5 + {off // 4} degree-10) of the "
                      f"same workload, FP64, plain Ehrlich-Aberth, {total_t:.1f} s"}


if __name__ == "__main__":
    main()
