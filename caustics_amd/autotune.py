"""Launch-schedule tuner of the extended-source contour pipeline.

The pipeline's results do not depend on how a batch is scheduled, only its speed does.
Three knobs set the schedule:
  chunk_bytes  workspace budget of one chunk (a launch sequence over as many sources as
               fit): larger chunks fill the per-source kernels better and launch less
               often, smaller ones let two streams overlap sooner;
  streams      streams the chunks alternate over (each with its own workspace);
  flat_per_cu  workgroups per CU of the grid-stride limb-solve kernels
               (caus_ext_set_flat_blocks_per_cu).
The tuner times candidate schedules on a representative batch per key
(nlenses, limb darkening, reverse mode, limb-resolution bucket) by coordinate descent
(flat_per_cu, then streams, then chunk_bytes, each over its candidates with the others
held) and stores the fastest in tuning.json beside this module; _run looks the key up
for every batch that does not fix its schedule.

    python -m caustics_amd.autotune [--nlenses 2 3] [--buckets 256 4096] [--out PATH]

runs on the GPU and rewrites the table (the committed table was measured on one
MI355X).  The reference has no counterpart: it maps mag_extended_source over sources
one at a time (lightcurve.py:245-254).
"""
import argparse
import json
import os
import time

TABLE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "tuning.json")
BUCKETS = (256, 512, 1024, 2048, 4096)
DEFAULT = {"chunk_bytes": 2 << 30, "streams": 2, "flat_per_cu": 8}
CANDIDATES = {"flat_per_cu": (4, 8, 16),
              "streams": (1, 2, 3),
              "chunk_bytes": (512 << 20, 1 << 30, 2 << 30, 4 << 30, 8 << 30)}

_table = None


def bucket(npts_max):
    """The smallest limb-resolution bucket that holds npts_max."""
    for b in BUCKETS:
        if npts_max <= b:
            return b
    return BUCKETS[-1]


def key(nlenses, ld, vjp, npts_max):
    return f"nl{int(nlenses)}_ld{int(bool(ld))}_vjp{int(bool(vjp))}_n{bucket(npts_max)}"


def load(path=TABLE):
    global _table
    if _table is None or path != TABLE:
        try:
            with open(path) as f:
                t = json.load(f)
        except (OSError, ValueError):
            t = {}
        if path != TABLE:
            return t
        _table = t
    return _table


def lookup(nlenses, ld, vjp, npts_max):
    """Schedule for a batch: the tuned entry of its key, else the nearest tuned bucket
    of the same (nlenses, ld, vjp), else DEFAULT."""
    ent = load().get("entries", {})
    k = key(nlenses, ld, vjp, npts_max)
    if k in ent:
        return {**DEFAULT, **ent[k]["schedule"]}
    b0 = bucket(npts_max)
    near = [(abs(BUCKETS.index(b) - BUCKETS.index(b0)), b) for b in BUCKETS
            if key(nlenses, ld, vjp, b) in ent]
    if near:
        return {**DEFAULT, **ent[key(nlenses, ld, vjp, min(near)[1])]["schedule"]}
    return dict(DEFAULT)


def _batch(nlenses, npts, n_sources, seed):
    """Representative sources near caustics: binary sources within 2 rho of a caustic
    point with ragged limb resolutions up to npts, or epochs of triple-lens crossing
    light curves with W.crossing_npts capped at npts."""
    import numpy as np
    import torch

    from . import workloads as W
    if nlenses == 2:
        choices = tuple(b for b in BUCKETS if b <= npts) or (npts,)
        b = W.extended_binary_batch(n_sources, 0, seed=seed, npts_choices=choices)
        prm = np.stack([b["s"], b["q"]], -1)
        nh = np.asarray(b["npts"], np.int64)
        w, rho, u1 = b["w"], b["rho"], b["u1"]
    else:
        n_sets = max(1, n_sources // 1500)
        b = W.triple_caustic_lightcurves(n_sets, 1500, seed=seed)
        nh = np.minimum(W.crossing_npts(b), npts).reshape(-1).astype(np.int64)
        w = b["w"].reshape(-1)
        rep = lambda x: np.repeat(np.asarray(x), 1500)
        prm = np.stack([rep(b[c]) for c in ("s", "q", "q3", "r3", "psi")], -1)
        rho, u1 = rep(b["rho"]), rep(b["u1"])
    cu = lambda x: torch.as_tensor(np.ascontiguousarray(x), device="cuda")
    return cu(w), cu(rho), cu(u1), cu(prm), nh


def _time(run, reps):
    import torch
    run()
    torch.cuda.synchronize()
    best = float("inf")
    for _ in range(reps):
        t0 = time.perf_counter()
        run()
        torch.cuda.synchronize()
        best = min(best, time.perf_counter() - t0)
    return best


def tune(nlenses, ld=True, vjp=True, npts=4096, n_sources=None, reps=2, seed=11, log=print):
    """Coordinate-descent search of the schedule for one key; returns (schedule,
    {candidate: seconds}) measured on the current GPU."""
    import torch

    from . import _lib
    from .extended_source import _run
    n_sources = n_sources or (4096 if nlenses == 2 else 48000)
    w, rho, u1, prm, nh = _batch(nlenses, npts, n_sources, seed)
    g = torch.ones(w.numel(), dtype=torch.float64, device="cuda") if vjp else None
    nmax = int(nh.max())
    L = _lib.lib()
    cur = dict(DEFAULT)
    times = {}

    from . import extended_source as X

    def run_with(sch):
        _lib.check(L.caus_ext_set_flat_blocks_per_cu(int(sch["flat_per_cu"])), "schedule")
        X._flat_set[0] = sch["flat_per_cu"]
        return _time(lambda: _run(nlenses, w, rho, prm, nh, nmax, ld, u1 if ld else None, 100,
                                  2500, _lib.MODE_FP64, g=g, chunk_bytes=sch["chunk_bytes"],
                                  streams=sch["streams"]), reps)
    try:
        for knob in ("flat_per_cu", "streams", "chunk_bytes"):
            best = None
            for v in CANDIDATES[knob]:
                sch = {**cur, knob: v}
                tk = json.dumps(sch, sort_keys=True)
                if tk not in times:
                    times[tk] = run_with(sch)
                    log(f"  {key(nlenses, ld, vjp, nmax)} {sch}: {times[tk] * 1e3:.2f} ms")
                if best is None or times[tk] < best[0]:
                    best = (times[tk], v)
            cur[knob] = best[1]
    finally:
        L.caus_ext_set_flat_blocks_per_cu(DEFAULT["flat_per_cu"])
        X._flat_set[0] = DEFAULT["flat_per_cu"]
    return cur, times


def main(argv=None):
    ap = argparse.ArgumentParser(description=__doc__.split("\n")[0])
    ap.add_argument("--nlenses", type=int, nargs="+", default=[2, 3])
    ap.add_argument("--buckets", type=int, nargs="+", default=[256, 4096])
    ap.add_argument("--reps", type=int, default=2)
    ap.add_argument("--dry-run", action="store_true", help="do not rewrite the table")
    ap.add_argument("--out", default=TABLE, help="where to write the table")
    args = ap.parse_args(argv)
    import torch
    table = load(TABLE) if os.path.exists(TABLE) else {}
    ent = dict(table.get("entries", {}))
    for nl in args.nlenses:
        for b in args.buckets:
            t0 = time.perf_counter()
            sch, times = tune(nl, True, True, b, reps=args.reps)
            k = key(nl, True, True, b)
            base = times[json.dumps(DEFAULT, sort_keys=True)]
            best = times[json.dumps(sch, sort_keys=True)]
            ent[k] = {"schedule": sch, "ms": best * 1e3, "default_ms": base * 1e3,
                      "candidates_timed": len(times)}
            print(f"{k}: {sch} {best * 1e3:.2f} ms (default {base * 1e3:.2f} ms), "
                  f"{time.perf_counter() - t0:.0f} s", flush=True)
    pr = torch.cuda.get_device_properties(0)
    out = {"device": f"{pr.name} ({getattr(pr, 'gcnArchName', '?')}, "
                     f"{pr.multi_processor_count} CUs)", "entries": ent}
    if not args.dry_run:
        with open(args.out, "w") as f:
            json.dump(out, f, indent=1, sort_keys=True)
            f.write("\n")
    print(json.dumps(out, sort_keys=True))


if __name__ == "__main__":
    main()
