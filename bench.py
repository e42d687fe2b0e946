"""Benchmark: rskelf factor + solve throughput at relative residual 1e-10.

Workload (default): second-kind Laplace double-layer integral equation on
the unit circle, N = 2^20 points, occ = 256, proxy ring of 64 points at 1.5
box sizes, 8 right-hand sides ~ N(0,1) from seed 0.

  mixed  FP32 factorization at tol 1e-6 (FP64 ID decisions), used as the
         preconditioner of flexible GMRES with an exact FP64 matvec, run to
         ||A x - b|| / ||b|| <= 1e-10 for every column   -> `value`
  plain  the same pipeline with an FP64 factorization (reported beside it)

One step = tree + factorization + solve + refinement for the 8 columns,
inputs (points, normals, weights, RHS) already resident in HBM.  The metric
is N / step time (points per second).  The dominant kernel's roofline is
measured live with HIP events inside the library, and the CPU oracle (the
FP64 restatement of the reference algorithm) is timed on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

PEAK = {"fp64": 78.6e12, "fp32": 157.3e12, "hbm": 8.0e12}


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--N", type=int, default=1 << 20)
    ap.add_argument("--occ", type=int, default=256)
    ap.add_argument("--tol", type=float, default=1e-6)
    ap.add_argument("--nrhs", type=int, default=8)
    ap.add_argument("--target", type=float, default=1e-10)
    ap.add_argument("--no-plain", action="store_true")
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-N", type=int, default=1 << 20)
    ap.add_argument("--workload", choices=["circle", "ellipses", "sphere", "helmholtz"],
                    default="circle",
                    help="circle: the metric's configuration (default); ellipses: 96 uneven "
                         "ellipses in one batched factorization; sphere: 3D single layer; "
                         "helmholtz: complex combined field on the star curve")
    ap.add_argument("--hN", type=int, default=262144, help="helmholtz: points on the star")
    ap.add_argument("--k", type=float, default=200.0, help="helmholtz: wavenumber")
    ap.add_argument("--nprob", type=int, default=96, help="ellipses: number of problems")
    ap.add_argument("--nface", type=int, default=327680, help="sphere: triangles")
    ap.add_argument("--socc", type=int, default=512, help="sphere: octree occupancy")
    ap.add_argument("--exact-op", action="store_true",
                    help="ellipses: run FGMRES on the exact on-the-fly FP64 product instead of "
                         "a tol-1e-12 FP64 factor's forward sweep")
    ap.add_argument("--calibrate-copy", type=int, default=0, metavar="MIB",
                    help="first run one device-to-device copy of MIB MiB (a known-size "
                         "reference dispatch for FETCH_SIZE / WRITE_SIZE counters)")
    return ap.parse_args()


def run_path(flam_amd, g, B, args, precision, steps, warmup, log):
    import torch
    from flam_amd.refine import fgmres
    # the timed steps time only the elimination class with HIP events (so
    # no event markers sit between the other launches); one more, untimed
    # step times every class for the breakdown
    times, infos, F = [], [], None
    for it in range(warmup + steps + 1):
        full = it == warmup + steps
        # both legs take the same ID route (the Gram-matrix ID at tol >= 1e-7),
        # so the ratio measures precision, not a change of algorithm
        opts = dict(symm="p", precision=precision, profile=1 if full else 2, id_route="gram")
        if full:
            F_timed, F = F, None
        if F is not None:
            del F
        torch.cuda.synchronize()
        t0 = time.perf_counter()
        F = flam_amd.rskelf(g.kernel, g.x, args.occ, args.tol, g.proxy, opts)
        t1 = time.perf_counter()
        # Arnoldi steps with the circle's closed form (one pass over X), true
        # residuals with the FFT circulant product of the evaluated matrix
        X, info = fgmres(g.circulant_matvec, B, F, tol=args.target,
                         inner_matvec=g.closed_form_matvec)
        torch.cuda.synchronize()
        t2 = time.perf_counter()
        R = B - g.circulant_matvec(X)
        rel = (torch.linalg.vector_norm(R, dim=0) / torch.linalg.vector_norm(B, dim=0)).max().item()
        log(f"{precision} step {it}{' (profiled, untimed)' if full else ''}: "
            f"factor {1e3*(t1-t0):.1f} ms, refine {1e3*(t2-t1):.1f} ms, "
            f"gmres iters {info['iters']}, relres {rel:.2e}")
        if warmup <= it < warmup + steps:
            times.append(t2 - t0)
            infos.append(dict(factor_ms=1e3 * (t1 - t0), refine_ms=1e3 * (t2 - t1),
                              iters=info["iters"], relres=rel))
    return times, infos, F_timed, F


def measured_traffic(kernel, workload="circle_N1048576_fp32"):
    """HBM bytes per launch of a kernel class on one workload from the
    calibrated rocprofv3 FETCH_SIZE / WRITE_SIZE summary kept under
    profiles/ (None if that workload has none)."""
    import glob
    for path in sorted(glob.glob(os.path.join(ROOT, "profiles", "*_traffic.md"))):
        with open(path) as fh:
            for line in fh:
                if line.startswith("traffic_json:"):
                    rec = json.loads(line.split(":", 1)[1])
                    if rec.get("kernel") == kernel and rec.get("workload") == workload:
                        return rec["bytes_per_launch"], os.path.basename(path)
    return None, None


def roofline_of(F, F_full, precision, workload=None, name=None):
    """Roofline of one kernel class (default: the dominant class by time in
    the fully profiled step F_full), with algorithmic flops / bytes; its
    launch times come from the last timed step F when that step timed it
    (the elimination class), else from F_full.  `pipe` names the unit the
    peak belongs to: the elimination runs on the matrix cores (MFMA, FP32
    or FP64), the Gram ID on the vector FP64 units (whose peak on MI355X
    equals the FP64 matrix-core peak), the sweeps are bound by HBM."""
    kt = F_full.kernel_times()
    if name is None:
        name, st = max(((k, v) for k, v in kt.items() if k != "tree" and not k.startswith("h:")),
                       key=lambda kv: kv[1]["ms"])
    elif name not in kt:
        return None, kt
    else:
        st = kt[name]
    st = F.kernel_times().get(name, st)
    avg_ms = st["ms"] / max(st["launches"], 1)
    if st["flops"] > 0:
        ach = st["flops"] / st["launches"] / (avg_ms * 1e-3)
        # the wide boxes' G = K'K (gram) runs on the FP64 matrix cores in
        # both precisions
        peak = PEAK["fp64"] if name in ("id", "gram") or precision == "fp64" else PEAK["fp32"]
        pipe = ("valu-fp64" if name == "id" else
                "mfma-fp64" if name == "gram" or precision == "fp64" else "mfma-fp32")
        traffic, src = measured_traffic(name, workload) if workload else (None, None)
        return dict(kernel=name, bound="mfma", pipe=pipe, achieved=ach / 1e12, peak=peak / 1e12,
                    unit="TFLOP/s", frac=ach / peak, traffic=traffic, traffic_source=src,
                    avg_launch_ms=avg_ms, launches=st["launches"]), kt
    ach = st["bytes"] / st["launches"] / (avg_ms * 1e-3)
    traffic, src = measured_traffic(name, workload) if workload else (None, None)
    return dict(kernel=name, bound="hbm", pipe="hbm", achieved=ach / 1e9, peak=PEAK["hbm"] / 1e9,
                unit="GB/s", frac=ach / PEAK["hbm"], traffic=traffic, traffic_source=src,
                avg_launch_ms=avg_ms, launches=st["launches"]), kt


def class_rooflines(F, F_full, precision, workload):
    """Rooflines of the main kernel classes: elimination (factor), Gram ID
    (id), the wide boxes' G = K'K (gram, 3D top levels only) and sweeps
    (sweep)."""
    out = {}
    for name in ("factor", "id", "gram", "sweep"):
        roof, _ = roofline_of(F, F_full, precision, workload, name)
        if roof is not None:
            out[name] = roof
    return out


def cpu_baseline(N, occ, tol, nrhs, target, log, runs=3):
    """The FP64 CPU oracle on the host's cores: the restated factorization,
    then the same flexible GMRES (restated in NumPy) to the same target on
    the exact FFT product; median of `runs` runs."""
    import numpy as np
    from oracle import flam as ofl, problems as opr
    try:
        from threadpoolctl import threadpool_info
        threads = max([i.get("num_threads", 1) for i in threadpool_info()] or [1])
    except ImportError:
        threads = os.cpu_count() or 1
    def one(n):
        pr = opr.CurveDL(n, "circle")
        B = np.random.default_rng(0).standard_normal((n, nrhs))
        t0 = time.perf_counter()
        Fo = ofl.rskelf(pr.afun, pr.x, occ, tol, pr.pxyfun, dict(symm="p", order="level"))
        X, iters, rel = ofl.fgmres(pr.matvec, B, lambda V: ofl.rskelf_sv(Fo, V), tol=target)
        dt = time.perf_counter() - t0
        log(f"cpu oracle N={n}: {dt:.2f} s, fgmres iters {iters}, relres {rel.max():.2e}")
        return dt

    # the oracle's per-box BLAS calls are small, so all BLAS threads can be
    # slower than one: a probe at N/4 picks the faster setting
    try:
        from threadpoolctl import threadpool_limits
    except ImportError:
        threadpool_limits = None
    use = threads
    if threadpool_limits is not None and threads > 1:
        with threadpool_limits(limits=1):
            t1 = one(max(N // 4, 1024))
        tall = one(max(N // 4, 1024))
        use = 1 if t1 < tall else threads
    times = []
    for _ in range(runs):
        if threadpool_limits is not None:
            with threadpool_limits(limits=use):
                times.append(one(N))
        else:
            times.append(one(N))
    dt = float(np.median(times))
    return dict(value=N / dt, unit="points/s", cores=int(use), kind="port",
                sample=f"circle N={N}, occ={occ}, tol={tol:g}, symm p, {nrhs} RHS, FP64 "
                       f"NumPy/SciPy restatement of rskelf (level order) + the same FGMRES to "
                       f"{target:g}, BLAS threads {use} of {os.cpu_count()} host cores (the "
                       f"faster of 1 and {threads} at N/4), median of {runs} runs {dt:.1f} s")


def other_workload(args, log):
    """Extra measurements (not the metric's configuration): the batched
    ellipses and the 3D sphere, mixed FP32 factor + FP64 refinement beside
    the plain FP64 factor.  The sphere refines on the exact on-the-fly FP64
    product.  The ellipses refine on the forward sweep of one FP64 factor
    at tol 1e-12, built inside every timed step (refine_ms includes it,
    operator_ms reports it), since their exact product is O(N^2) per
    problem; the reported residuals are then taken with the exact product."""
    import numpy as np
    import torch

    import flam_amd
    from flam_amd import geometry
    from flam_amd.refine import fgmres
    if args.workload == "ellipses":
        # 96 ellipses, a = 1, b ~ U[0.3, 1], N log-uniform in [2048, 131072], seed 0
        rng = np.random.default_rng(0)
        bs = rng.uniform(0.3, 1.0, args.nprob)
        Ns = np.round(np.exp(rng.uniform(np.log(2048), np.log(131072), args.nprob))).astype(int)
        batch = flam_amd.stack_problems([geometry.ellipse(int(n), 1.0, float(b))
                                         for n, b in zip(Ns, bs)])
        N, occ, tol, target, symm = batch.N, 256, 1e-6, 1e-10, "n"
        segs = list(zip(batch.offsets[:-1], batch.offsets[1:]))
        factor = lambda o: flam_amd.rskelf_batch(batch, occ, tol, o)
        mv = lambda V: flam_amd.kernel_matvec_batch(batch, V)
        exact = mv
        if not args.exact_op:
            from flam_amd.refine import factor_operator

            def make_op():
                # the refinement operator: a tol-1e-12 FP64 factor's forward
                # sweep, built inside every timed step
                return factor_operator(flam_amd.rskelf_batch(batch, occ, 1e-12, dict(symm="n")))
        desc = (f"{args.nprob} ellipses (a=1, b~U[0.3,1], N log-uniform [2048,131072], seed 0), "
                f"N total {N}, one batched factorization, occ={occ}, tol={tol:g}, symm n, "
                f"{args.nrhs} RHS, FP32 factor + FP64 FGMRES to {target:g} per problem")
    elif args.workload == "helmholtz":
        g = geometry.helmholtz_curve(args.hN, "star", k=args.k)
        N, occ, tol, target, symm = g.N, 256, 1e-6, 1e-8, "n"
        segs = None
        factor = lambda o: flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, o)
        mv = g.matvec
        exact = mv
        desc = (f"2D Helmholtz combined field (eta = k) on the star r = 1 + 0.3 cos 5t, N={N}, "
                f"k={args.k:g}, occ={occ}, tol={tol:g}, symm n, {args.nrhs} complex RHS, complex64 "
                f"factor + complex FP64 FGMRES (exact on-the-fly product) to {target:g}")
    else:
        g = geometry.sphere(args.nface)
        N, occ, tol, target, symm = g.N, args.socc, 1e-5, 1e-9, "n"
        segs = None
        factor = lambda o: flam_amd.rskelf(g.kernel, g.x, occ, tol, g.proxy, o)
        mv = lambda V: flam_amd.kernel_matvec(g.kernel, g.x, V)
        exact = mv
        desc = (f"3D Laplace single layer on the unit sphere, {args.nface} triangles, occ={occ}, "
                f"tol={tol:g}, symm n, Gram ID, {args.nrhs} RHS, FP32 factor + FP64 FGMRES "
                f"(exact on-the-fly product) to {target:g}")
    if args.workload == "ellipses" and not args.exact_op:
        desc += ("; FGMRES on a tol-1e-12 FP64 factor's rskelf_mv (built inside every step), "
                 "residuals with the exact product")
    else:
        make_op = None
    log(f"workload: {desc}")
    rng = np.random.default_rng(0)
    B = torch.from_numpy(rng.standard_normal((N, args.nrhs))).cuda()
    if args.workload == "helmholtz":
        B = torch.complex(B, torch.from_numpy(rng.standard_normal((N, args.nrhs))).cuda())

    def exact_relres(X):
        """max over problems and columns of ||B - A X|| / ||B|| with the
        exact on-the-fly product (outside the timed region)"""
        R = B - exact(X)
        parts = segs if segs is not None else [(0, N)]
        return max((torch.linalg.vector_norm(R[a:b], dim=0) /
                    torch.linalg.vector_norm(B[a:b], dim=0)).max().item() for a, b in parts)
    # the sphere takes the Gram ID in both precisions (tol 1e-5 >= 1e-7), so
    # the two factorizations differ only in the precision of the rest
    route = dict(id_route="gram") if args.workload == "sphere" else {}
    res = {}
    for prec in (("fp32",) if args.no_plain else ("fp32", "fp64")):
        times, F = [], None
        for it in range(args.warmup + args.steps + 1):    # the last: profiled, untimed
            full = it == args.warmup + args.steps
            if full:
                F_timed, F = F, None
            if F is not None:
                del F
            torch.cuda.synchronize()
            t0 = time.perf_counter()
            F = factor(dict(symm=symm, precision=prec, profile=1 if full else 2, **route))
            t1 = time.perf_counter()
            op = make_op() if make_op is not None else mv
            t_op = time.perf_counter()
            X, info = fgmres(op, B, F, tol=target, segments=segs)
            del op
            torch.cuda.synchronize()
            t2 = time.perf_counter()
            rel = exact_relres(X)
            log(f"{prec} step {it}: factor {1e3*(t1-t0):.1f} ms, operator {1e3*(t_op-t1):.1f} ms, "
                f"refine {1e3*(t2-t_op):.1f} ms, gmres iters {info['iters']}, "
                f"relres {max(info['relres']):.2e} (exact product {rel:.2e})")
            if args.warmup <= it < args.warmup + args.steps:
                times.append((t2 - t0, 1e3 * (t1 - t0), 1e3 * (t2 - t1), info["iters"], rel,
                              1e3 * (t_op - t1)))
        roof, kt = roofline_of(F_timed, F, prec, None)
        rooflines = class_rooflines(F_timed, F, prec, None)
        res[prec] = dict(value=N / float(np.mean([t[0] for t in times])),
                         ms_per_step=1e3 * float(np.mean([t[0] for t in times])),
                         factor_ms=float(np.mean([t[1] for t in times])),
                         refine_ms=float(np.mean([t[2] for t in times])),
                         operator_ms=float(np.mean([t[5] for t in times])),
                         iters=times[-1][3], relres_exact=max(t[4] for t in times),
                         levels=F.level_table(), roofline=roof, rooflines=rooflines,
                         kernels={k: dict(ms=v["ms"], launches=v["launches"]) for k, v in kt.items()})
        del F, F_timed
    dtype = "c64+c128" if args.workload == "helmholtz" else "f32+f64"
    out = {"metric": f"{args.workload} factor+solve points/sec at the target residual",
           "value": res["fp32"]["value"], "unit": "points/s", "n_gpus": 1, "steps": args.steps,
           "warmup": args.warmup, "ms_per_step": res["fp32"]["ms_per_step"],
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": dtype,
           "data": "synthetic", "config": {"workload": desc, "N": N},
           "mixed": res["fp32"], "plain_fp64": res.get("fp64"),
           "mixed_over_plain": res["fp32"]["value"] / res["fp64"]["value"] if "fp64" in res else None}
    print(json.dumps(out))


def main():
    args = parse()
    if args.gpus != 1:
        print(json.dumps({"error": "this build runs on one GPU; --gpus must be 1 "
                                   "(the multi-GPU design is described, not built)"}))
        sys.exit(2)
    import numpy as np
    import torch

    import flam_amd
    from flam_amd import geometry

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    if args.workload != "circle":
        return other_workload(args, log)

    if args.calibrate_copy:
        # a device-to-device copy of known size; 2 GiB written in between
        # evicts the source from the last-level caches first
        src = torch.ones(args.calibrate_copy << 18, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        junk = torch.empty(1 << 29, dtype=torch.float32, device="cuda")
        junk.fill_(1.0)
        torch.cuda.synchronize()
        dst.copy_(src)
        torch.cuda.synchronize()
        log(f"calibration copy: {args.calibrate_copy} MiB read + {args.calibrate_copy} MiB written")
        del src, dst, junk
    N = args.N
    log(f"workload: circle N={N} occ={args.occ} tol={args.tol:g} nrhs={args.nrhs} "
        f"target {args.target:g}")
    # launch shapes chosen by the autotuner for this workload (flam_amd/tune.py)
    tuned = os.path.join(ROOT, "flam_amd", "tuning", f"circle_N{N}.json")
    if os.path.exists(tuned):
        from flam_amd import tune
        log(f"tuning table {os.path.basename(tuned)}: {tune.load(tuned)['table']}")
    g = geometry.circle(N)
    B = torch.from_numpy(np.random.default_rng(0).standard_normal((N, args.nrhs))).cuda()
    torch.cuda.synchronize()

    t_mix, i_mix, F_mix, Ff_mix = run_path(flam_amd, g, B, args, "fp32", args.steps, args.warmup,
                                           log)
    roof, kt = roofline_of(F_mix, Ff_mix, "fp32", f"circle_N{N}_fp32")
    roofs = class_rooflines(F_mix, Ff_mix, "fp32", f"circle_N{N}_fp32")
    del Ff_mix
    info_mix = F_mix.info()
    table = F_mix.level_table()
    plain = None
    if not args.no_plain:
        t_pl, i_pl, F_pl, Ff_pl = run_path(flam_amd, g, B, args, "fp64", args.steps, args.warmup,
                                           log)
        roof64, kt64 = roofline_of(F_pl, Ff_pl, "fp64", f"circle_N{N}_fp64")
        roofs64 = class_rooflines(F_pl, Ff_pl, "fp64", f"circle_N{N}_fp64")
        del Ff_pl
        plain = dict(value=N / float(np.mean(t_pl)), ms_per_step=1e3 * float(np.mean(t_pl)),
                     steps=i_pl, roofline=roof64, rooflines=roofs64,
                     kernels={k: dict(ms=v["ms"], launches=v["launches"]) for k, v in kt64.items()})
        del F_pl
    cpu = None
    if not args.no_cpu:
        cpu = cpu_baseline(args.cpu_N, args.occ, args.tol, args.nrhs, args.target, log)
    ms = 1e3 * float(np.mean(t_mix))
    value = N / float(np.mean(t_mix))
    out = {
        "metric": "rskelf factor+solve points/sec at rel. residual 1e-10, mixed vs plain FP64 GPU",
        "value": value, "unit": "points/s", "n_gpus": 1, "steps": args.steps,
        "warmup": args.warmup, "ms_per_step": ms, "higher_is_better": True,
        "scaling": "weak", "vs_baseline": None, "dtype": "f32+f64",
        "data": "synthetic (unit-circle Nystrom discretization, RHS N(0,1) seed 0)",
        "config": {"workload": f"circle Laplace double-layer IE, N={N}, occ={args.occ}, "
                               f"tol={args.tol:g}, {args.nrhs} RHS, FP32 factor + FP64 "
                               f"flexible GMRES to {args.target:g}",
                   "N": N, "occ": args.occ, "tol": args.tol, "nrhs": args.nrhs,
                   "proxy": 64, "symm": "p"},
        "steps_detail": i_mix,
        "factor": dict(info_mix, levels=table),
        "kernels": {k: dict(ms=v["ms"], launches=v["launches"]) for k, v in kt.items()},
        "roofline": roof,
        "rooflines": roofs,
        "plain_fp64": plain,
        "mixed_over_plain": (value / plain["value"]) if plain else None,
        "cpu_baseline": cpu,
    }
    print(json.dumps(out))


if __name__ == "__main__":
    main()
