"""Autotuner of the factorization's launch shapes (rskx_set_tuning).

Which kernel shape a box runs on is a choice by size: the 8-wave or the
16-wave register elimination, the Gram-ID class (256 or 512 threads,
2 or 5 register tiles per thread), the shared-memory budgets that send
boxes between kernels, and whether a phase's size classes run on
concurrent streams.  These only change how fast the same arithmetic runs.
autotune() measures a workload (a callable running factor + solve) over
each parameter's candidates in turn (coordinate descent, median of `reps`
timings, keeping a change only when it is at least 1% faster) and returns
the table; save() / load() keep it as JSON under flam_amd/tuning/, and
bench.py applies the committed table of its workload.

    python -m flam_amd.tune --N 1048576 --out flam_amd/tuning/circle_N1048576.json
"""
import argparse
import json
import os
import statistics
import sys
import time

from . import _lib as L

HERE = os.path.dirname(os.path.abspath(__file__))

# candidates per parameter (FP32 register shapes: the 8-wave tile holds
# n <= 192, the Gram classes 0 / 1 hold n <= 124 / 176)
SPACE = {
    "concurrent": [0, 1],
    "elim8_nmax": [128, 160, 192],
    "elim8_lds": [64 * 1024, 72 * 1024, 80 * 1024],
    "gram_n0": [0, 96, 124],
    "gram_n1": [0, 148, 176],
    "sweep_stage": [48 * 1024, 64 * 1024, 76 * 1024],
    "elim_mfma": [0, 1],
    "gram_mfma": [0, 1],
}


def set_tuning(name, value):
    rc = L.lib().rskx_set_tuning(name.encode(), float(value))
    if rc != L.RSKX_OK:
        raise ValueError(L.last_error())


def get_tuning(name):
    import ctypes
    v = ctypes.c_double()
    rc = L.lib().rskx_get_tuning(name.encode(), ctypes.byref(v))
    if rc != L.RSKX_OK:
        raise ValueError(L.last_error())
    return int(v.value)


def apply(table):
    for k, v in table.items():
        set_tuning(k, v)


def load(path):
    with open(path) as fh:
        rec = json.load(fh)
    apply(rec["table"])
    return rec


def save(path, table, meta):
    with open(path, "w") as fh:
        json.dump(dict(table=table, **meta), fh, indent=1, sort_keys=True)
        fh.write("\n")


def autotune(run, space=SPACE, reps=5, log=print):
    """Coordinate descent over `space`; run() does one timed unit of work
    (it synchronizes itself).  Returns (table, trace)."""
    def measure():
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            run()
            ts.append(time.perf_counter() - t0)
        return statistics.median(ts)
    best = {k: get_tuning(k) for k in space}
    apply(best)
    run()                                   # warm-up at the defaults
    best_t = measure()
    trace = [dict(params=dict(best), ms=1e3 * best_t)]
    log(f"defaults {best}: {1e3 * best_t:.2f} ms")
    for name, cands in space.items():
        for v in cands:
            if v == best[name]:
                continue
            set_tuning(name, v)
            t = measure()
            trace.append(dict(params={**best, name: v}, ms=1e3 * t))
            log(f"  {name}={v}: {1e3 * t:.2f} ms")
            if t < 0.99 * best_t:
                best[name], best_t = v, t
        set_tuning(name, best[name])
    log(f"chosen {best}: {1e3 * best_t:.2f} ms")
    return best, trace


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--N", type=int, default=1 << 20)
    ap.add_argument("--reps", type=int, default=5)
    ap.add_argument("--out", default=os.path.join(HERE, "tuning", "circle_N1048576.json"))
    args = ap.parse_args()
    import numpy as np
    import torch

    import flam_amd
    from flam_amd import geometry
    from flam_amd.refine import fgmres
    g = geometry.circle(args.N)
    B = torch.from_numpy(np.random.default_rng(0).standard_normal((args.N, 8))).cuda()

    def run():
        F = flam_amd.rskelf(g.kernel, g.x, 256, 1e-6, g.proxy, dict(symm="p", precision="fp32"))
        fgmres(g.circulant_matvec, B, F, tol=1e-10, inner_matvec=g.closed_form_matvec)
        torch.cuda.synchronize()
    log = lambda m: print(m, file=sys.stderr, flush=True)
    table, trace = autotune(run, reps=args.reps, log=log)
    save(args.out, table, dict(workload=f"circle N={args.N}, occ 256, tol 1e-6, symm p, FP32 "
                                        "factor + FGMRES (8 RHS) to 1e-10",
                               device=torch.cuda.get_device_name(0), trace=trace))
    print(json.dumps(table))


if __name__ == "__main__":
    main()
