#!/usr/bin/env python3
"""Solves/s of the H^2 solve path on one MI355X.

Workload (the metric's configuration): N = 2^20 points uniform on the unit
sphere (seed 2), Yukawa kernel exp(-|x-y|)/|x-y| with K(0) = 1e3, theta = 2,
leaf 256. The operator is an interpolative H^2 matrix at tolerance OP_EPS
(far-body sampling 2000 per box); the preconditioner is an orthogonal H^2
matrix truncated at 1e-6 with rank cap 64, ULV-factorized. A step is one
step: GMRES for a batch of BATCH right-hand sides (U(-1, 1), fresh seeds), each
to ||b - A x|| / ||b|| <= 1e-10 on the FP64 H^2 operator, restart 30, with
the factorization resident in HBM; the batch is solved in lockstep, so every
GMRES step streams the H^2 and ULV data once for all its right-hand sides.
A solve's latency is the latency of its batch. The mixed path factorizes the ULV in FP32,
stores the factors in bfloat16 for the substitution (FP32 accumulation) and
runs the inner operator applications in FP32 (FP64 Arnoldi, Givens, outer
residual);
the FP64 path runs everything in FP64. The mixed path's inner iterations
apply the operator built at the looser tolerance INNER_EPS (GMRES iterative
refinement: every cycle's residual uses the FP64 operator at OP_EPS).
`value` is the mixed path's solves/s;
the FP64 path is reported beside it. After the timed region every solution is
checked with the matrix-free FP64 dense matvec.

Prints one JSON line (see the keys below).
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0  # MI355X HBM3E, vendor figure


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--batch", type=int, default=8,
                    help="right-hand sides per step, solved together (one pass over the matrices per GMRES step)")
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--n", type=int, default=1 << 20, help="points (default 2^20)")
    ap.add_argument("--op-eps", type=float, default=1e-12)
    ap.add_argument("--sampling", type=int, default=2000)
    ap.add_argument("--inner-eps", type=float, default=1e-7,
                    help="tolerance of the operator the mixed path applies in its FP32 inner iterations (0: the main operator)")
    ap.add_argument("--mixed-factor", choices=["bf16", "fp32", "bf16schur"], default="bf16",
                    help="ULV of the mixed path: FP32 factorization with bfloat16-stored or FP32-stored factors "
                         "(bf16schur: bfloat16-stored, and the Schur-complement products on bfloat16 inputs)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-n", type=int, default=16384, help="points of the CPU baseline sample")
    ap.add_argument("--workload", choices=["solve", "batch", "serve"], default="solve",
                    help="solve: the metric (default); batch: 64 variable-size problems factorized in shared launches; "
                         "serve: 16 resident problems under a Poisson request stream")
    ap.add_argument("--rate", type=float, default=10.0, help="serve: offered load, requests per second")
    ap.add_argument("--requests", type=int, default=160, help="serve: requests of the timed stream")
    ap.add_argument("--problem", choices=["sphere-yukawa", "cube-laplace"], default="sphere-yukawa",
                    help="sphere-yukawa: the metric's workload (points uniform on the unit sphere, exp(-r)/r); "
                         "cube-laplace: points uniform in the unit cube, 1/r (K(0) = 1e3 for both)")
    ap.add_argument("--cpu-phases", action="store_true",
                    help="only time the oracle's phases on the host cores at N=20000 and N=65536 (no GPU work)")
    return ap.parse_args()


def pct(a, q):
    return float(np.percentile(np.asarray(a), q))


def profile_class(N, ctx, cls):
    """Launches, summed device time and algorithmic work of one kernel class
    since the last profile reset."""
    import ctypes as C
    cnt, ms, work = C.c_longlong(), C.c_double(), C.c_double()
    N.check_gpu(N.gpu().mxg_profile_read(ctx.handle, cls, C.byref(cnt), C.byref(ms)))
    N.check_gpu(N.gpu().mxg_profile_work(ctx.handle, cls, C.byref(work)))
    return {"launches": cnt.value, "ms": ms.value, "work": work.value}


def measured_traffic(nr, kernel="k_pair_gemv<float, float, 8>"):
    """HBM bytes per launch of the FP32 near-product launch with nr right-hand
    sides, from the committed rocprofv3 FETCH_SIZE/WRITE_SIZE passes (None when
    absent, or measured for another kernel or batch width)."""
    path = os.path.join(ROOT, "profiles", "pair_gemv_matvec_fp32_traffic.md")
    try:
        text = open(path).read()
        rec = json.loads(text.split("```json", 1)[1].split("```", 1)[0])
        same = int(rec["rhs"]) == nr and rec.get("launch") == "leaf near products" and rec.get("kernel") == kernel
        return float(rec["traffic_bytes"]) if same else None
    except (OSError, IndexError, KeyError, ValueError):
        return None


def log(msg):
    """Progress on stderr; stdout carries only the JSON line."""
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def main():
    args = parse()
    if args.cpu_phases:
        cpu_phases([(20000, 0), (65536, 1)])
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py: this build targets one GPU (the H^2 solve path keeps the whole tree on one device); "
              f"--gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.workload == "batch":
        return bench_batch(args)
    if args.workload == "serve":
        return bench_serve(args)

    from mxc_amd import _native as N
    from mxc_amd.device import Context, kernel_desc
    from mxc_amd.solver import H2, Problem, generate_points, uniform_vector

    n, leaf, theta = args.n, 256, 2.0
    cube = args.problem == "cube-laplace"
    kernel, diag, alpha = ("laplace" if cube else "yukawa"), 1e3, 1.0
    tol, restart, max_cycles = 1e-10, 30, 20
    ctx = Context(0)
    t0 = time.time()
    pts = generate_points("cube" if cube else "sphere", n, 2)
    prob = Problem(ctx, pts, leaf)
    # the far-field sketches of the operator's construction are timed with
    # HIP events (one pair per level) and their flops counted
    N.check_gpu(N.gpu().mxg_profile_reset(ctx.handle))
    N.check_gpu(N.gpu().mxg_profile_enable(ctx.handle, 1))
    A = H2.operator(prob, args.op_eps, args.sampling, kernel, diag, alpha, theta)
    t_op = time.time() - t0
    sketch = profile_class(N, ctx, N.MXG_PROF_SKETCH)
    N.check_gpu(N.gpu().mxg_profile_enable(ctx.handle, 0))
    log(f"operator built in {t_op:.1f} s; sketch {sketch}")
    t0 = time.time()
    M = H2.preconditioner(prob, 64, args.sampling, kernel, diag, alpha, theta, eps=1e-6)
    t_pc = time.time() - t0
    log(f"preconditioner built in {t_pc:.1f} s")
    t0 = time.time()
    A_in = H2.operator(prob, args.inner_eps, args.sampling, kernel, diag, alpha, theta) if args.inner_eps > 0 else None
    t_in = time.time() - t0
    if A_in is not None:
        log(f"inner operator (eps {args.inner_eps:g}) built in {t_in:.1f} s")
    # launch shapes of the matvecs the solves run, picked by the autotuner
    tuning = {"fp64_operator": A.autotune(N.MXG_FP64, min(args.batch, 8)),
              "fp32_operator": (A_in or A).autotune(N.MXG_FP32, min(args.batch, 8))}
    log(f"autotuned matvec shapes: {tuning}")
    L = prob.levels

    bs = args.batch
    rhs = [ctx.array(np.stack([uniform_vector(n, -1.0, 1.0, 1000 + k * bs + q) for q in range(bs)], axis=1))
           for k in range(args.warmup + args.steps)]
    sol = {path: [ctx.zeros((n, bs)) for _ in range(args.steps)] for path in ("fp64", "mixed")}
    work = ctx.zeros((n, bs))

    def solve(B, X, inner_prec):
        X.zero()
        inner_op = A_in if inner_prec == N.MXG_FP32 else None
        res, it, inner = A.gmres_batch(M, B, X, tol, restart, max_cycles, N.GMRES_CGS2, inner_prec, inner_op)
        return [float(r[-1]) for r in res], [int(v) for v in it], [int(v) for v in inner]

    def run_path(factor_prec, inner_prec, xs):
        info = M.factorize(factor_prec)
        assert not info.any(), f"singular ULV blocks per level: {info}"
        fsec = M.factor_seconds()
        for k in range(args.warmup):
            solve(rhs[k], work, inner_prec)
        ctx.synchronize()
        lat, cycles, inners, res = [], [], [], []
        t_start = time.perf_counter()
        for k in range(args.steps):
            ts = time.perf_counter()
            r, it, inner = solve(rhs[args.warmup + k], xs[k], inner_prec)
            lat.append(time.perf_counter() - ts)
            res += r
            cycles += it
            inners += inner
        ctx.synchronize()
        elapsed = time.perf_counter() - t_start
        # kernel timing with HIP events, in a separate pass over the same
        # right-hand sides (the events stay out of the timed region)
        N.check_gpu(N.gpu().mxg_profile_reset(ctx.handle))
        N.check_gpu(N.gpu().mxg_profile_enable(ctx.handle, 1))
        for k in range(min(args.steps, 3)):
            solve(rhs[args.warmup + k], work, inner_prec)
        ctx.synchronize()
        N.check_gpu(N.gpu().mxg_profile_enable(ctx.handle, 0))
        kern = {}
        import ctypes as C
        for cls, name in ((0, "matvec_near_fp64"), (1, "matvec_near_fp32"), (2, "ulv_fwd_leaf"), (3, "ulv_bwd_leaf"),
                          (4, "matvec_upper"), (5, "matvec_far"), (6, "matvec_fp64"), (7, "matvec_fp32"),
                          (8, "ulv_solve")):
            cnt, ms = C.c_longlong(), C.c_double()
            N.check_gpu(N.gpu().mxg_profile_read(ctx.handle, cls, C.byref(cnt), C.byref(ms)))
            kern[name] = {"launches": cnt.value, "total_ms": ms.value,
                          "avg_ms": ms.value / cnt.value if cnt.value else None}
        log(f"factor {factor_prec} inner {inner_prec}: {args.steps} x {bs} solves in {elapsed:.2f} s")
        # dense FP64 check of every solution (outside the timed region)
        dense = []
        for k in range(args.steps):
            prob.dense_matvec(kernel, diag, alpha, xs[k], work)
            ax = work.download()
            b = rhs[args.warmup + k].download()
            for q in range(bs):
                dense.append(float(np.linalg.norm(b[:, q] - ax[:, q]) / np.linalg.norm(b[:, q])))
            log(f"dense check of step {k}: max {max(dense[-bs:]):.3e}")
        return dict(elapsed=elapsed, lat=lat, cycles=cycles, inner=inners, h2_resid=res, dense=dense, factor_s=fsec,
                    kernels=kern)

    fflops = factor_flops(M, pts, L, theta)
    fp64 = run_path(N.MXG_FP64, N.MXG_FP64, sol["fp64"])
    mixed_code = {"bf16": N.MXG_BF16, "fp32": N.MXG_FP32, "bf16schur": N.MXG_BF16_SCHUR}[args.mixed_factor]
    mixed = run_path(mixed_code, N.MXG_FP32, sol["mixed"])
    # error of every mixed-path solution against the FP64 path's for the same right-hand side
    mix_err = []
    for k in range(args.steps):
        xm, x64 = sol["mixed"][k].download(), sol["fp64"][k].download()
        mix_err += [float(np.linalg.norm(xm[:, q] - x64[:, q]) / np.linalg.norm(x64[:, q])) for q in range(bs)]

    # roofline of the dominant kernel of the mixed path: the FP32 launch of the
    # leaf near products of the matvec for the batch's right-hand sides.
    # Algorithmic bytes: every leaf near block once, and per right-hand side
    # the input x read once and the near sum of every leaf row written once.
    Afp32 = A_in if A_in is not None else A
    h = N.h2().mxc_solver_device(Afp32.handle)
    near_elems = int(N.gpu().mxg_h2_level_size(h, L, 0))
    nr = min(bs, 8)  # right-hand sides per launch
    bytes_prod_fp32 = 4 * (near_elems + 2 * nr * n)
    km = mixed["kernels"]["matvec_near_fp32"]
    achieved = bytes_prod_fp32 / (km["avg_ms"] * 1e-3) / 1e9 if km["avg_ms"] else None
    def level_sizes(op):
        out, hd = [], N.h2().mxc_solver_device(op.handle)
        for l in range(L + 1):
            d, r = op.dims(l)
            out.append({"level": l, "boxes": len(d), "sum_dims": int(d.sum()), "sum_ranks": int(r.sum()),
                        "near_elems": int(N.gpu().mxg_h2_level_size(hd, l, 0)),
                        "coupling_elems": int(N.gpu().mxg_h2_level_size(hd, l, 3)), "basis_elems": int(np.sum(d * r))})
        return out

    sizes = level_sizes(A)
    sizes_in = level_sizes(A_in) if A_in is not None else None

    value = args.steps * args.batch / mixed["elapsed"]
    value64 = args.steps * args.batch / fp64["elapsed"]
    out = {
        "metric": "solves/s at rel-res 1e-10, mixed vs FP64 ULV, N=2^20; p50/p99 solve latency",
        "value": round(value, 4),
        "unit": "solves/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(1e3 * mixed["elapsed"] / args.steps, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"bf16": "f32 ULV factorization with bf16-stored factors", "fp32": "f32 ULV",
                  "bf16schur": "f32 ULV factorization with bf16 Schur-complement inputs and bf16-stored factors"}
                 [args.mixed_factor] +
                 " + f32 inner operator, f64 Arnoldi/residual",
        "data": "synthetic (seeded points and right-hand sides)",
        "config": {
            "workload": ("H2 GMRES solve, unit-cube Laplace" if cube else "H2 GMRES solve, unit-sphere Yukawa") +
                        f" N={n}, leaf 256",
            "N": n, "leaf": leaf, "theta": theta, "kernel": "laplace 1/r, K(0)=1e3" if cube else "yukawa exp(-r)/r, K(0)=1e3",
            "operator_eps": args.op_eps, "inner_operator_eps": args.inner_eps, "sampling": args.sampling,
            "sampling_rule": "operators: the max(3, ceil(L/2)) nearest far shells of a box are kept whole, every "
                             "farther shell keeps at most `sampling` bodies; preconditioner: `sampling` bodies in "
                             "total over every outside body",
            "precond_eps": 1e-6, "precond_max_rank": 64,
            "restart": restart, "tol": tol, "levels": L, "rhs_per_step": args.batch,
        },
        "p50_ms": round(1e3 * pct(mixed["lat"], 50), 3),
        "p99_ms": round(1e3 * pct(mixed["lat"], 99), 3),
        "fp64_path": {
            "solves_per_s": round(value64, 4), "p50_ms": round(1e3 * pct(fp64["lat"], 50), 3),
            "p99_ms": round(1e3 * pct(fp64["lat"], 99), 3), "cycles": fp64["cycles"], "inner_iters": fp64["inner"],
            "dense_relres_max": max(fp64["dense"]), "factor_s": round(fp64["factor_s"], 4),
        },
        "mixed_over_fp64": round(value / value64, 3),
        "mixed_path": {
            "cycles": mixed["cycles"], "inner_iters": mixed["inner"], "h2_relres_max": max(mixed["h2_resid"]),
            "dense_relres_max": max(mixed["dense"]), "factor_s": round(mixed["factor_s"], 4),
            "rel_diff_vs_fp64_max": max(mix_err), "rel_diff_vs_fp64_median": float(np.median(mix_err)),
        },
        "setup_s": {"operator": round(t_op, 2), "preconditioner": round(t_pc, 2), "inner_operator": round(t_in, 2)},
        "factorization": {
            "algorithmic_flops": fflops,
            "fp64": {"s": round(fp64["factor_s"], 4), "tflops": round(fflops / fp64["factor_s"] / 1e12, 2),
                     "frac_of_fp64_mfma_peak": round(fflops / fp64["factor_s"] / 1e12 / 78.6, 3)},
            "fp32_bf16": {"s": round(mixed["factor_s"], 4), "tflops": round(fflops / mixed["factor_s"] / 1e12, 2),
                          "frac_of_fp32_mfma_peak": round(fflops / mixed["factor_s"] / 1e12 / 157.3, 3)},
        },
        "construction": {
            "operator_s": round(t_op, 2),
            "sketch": {"kernel": "k_sketch (FP64 MFMA) + k_sum_slices", "launches": sketch["launches"],
                       "s": round(sketch["ms"] * 1e-3, 3), "algorithmic_flops": sketch["work"],
                       "tflops": round(sketch["work"] / (sketch["ms"] * 1e-3) / 1e12, 2) if sketch["ms"] else None,
                       "frac_of_fp64_mfma_peak": round(sketch["work"] / (sketch["ms"] * 1e-3) / 1e12 / 78.6, 3)
                       if sketch["ms"] else None},
        },
        "kernels_mixed": mixed["kernels"],
        "kernels_fp64": fp64["kernels"],
        "matvec_autotune": {k: {"ms": round(v[0], 4), **v[1]} for k, v in tuning.items()},
        "operator_levels": sizes,
        "inner_operator_levels": sizes_in,
        "roofline": {
            "kernel": f"k_pair_gemv<float, float, {nr}>: the matvec's launch of the leaf near products (FP32)",
            "bound": "hbm",
            "achieved": round(achieved, 1) if achieved else None,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": round(achieved / HBM_PEAK_GBS, 4) if achieved else None,
            "algorithmic_bytes_per_launch": bytes_prod_fp32,
            "traffic": measured_traffic(nr),
        },
    }
    # release every device object before the CPU baseline and before exit, so
    # that no HIP handle is left to static destructors at interpreter shutdown
    ctx.close()
    if not args.no_cpu_baseline:
        log("cpu baseline")
        out["cpu_baseline"] = cpu_baseline(args.cpu_n, args.sampling, args.op_eps, n, cube)
    print(json.dumps(out))


def bench_batch(args):
    """64 independent problems (N log-uniform in [4096, 65536], theta in
    {1, 2, 3}, leaf 128/256, rank in {32, 64, 128}, operator eps 1e-8, cube
    Laplace): their FP32 ULV factorizations in shared launches against one
    after another, then one FP64 GMRES solve (CGS2, restart 30) to 1e-8 per
    problem, each checked with the dense FP64 product."""
    from mxc_amd import _native as N
    from mxc_amd.batch import Batch, log_uniform_specs
    from mxc_amd.device import Context

    ctx = Context(0)
    specs = log_uniform_specs(64, seed=4)
    t0 = time.time()
    b = Batch(ctx, specs)
    log(f"built {len(specs)} problems ({sum(s.n for s in specs)} points) in {time.time() - t0:.1f} s")
    Ms = [m.M for m in b.members]
    b.factorize(N.MXG_FP32)  # warm-up
    reps = max(1, args.steps // 2)
    batched = []
    for _ in range(reps):
        ctx.synchronize()
        ts = time.perf_counter()
        b.factorize(N.MXG_FP32)
        batched.append((time.perf_counter() - ts, b.factor_seconds()))
    single = []
    for _ in range(reps):
        ts = time.perf_counter()
        dev = 0.
        for M in Ms:
            M.factorize(N.MXG_FP32)
            dev += M.factor_seconds()
        single.append((time.perf_counter() - ts, dev))
    b.factorize(N.MXG_FP32)
    res_max, dense_max, solve_s, inner = 0., 0., 0., []
    for k, m in enumerate(b.members):
        n = m.spec.n
        rhs = np.random.default_rng(100 + k).uniform(-1, 1, n)
        db, dx, dy = ctx.array(rhs), ctx.zeros(n), ctx.empty(n)
        ts = time.perf_counter()
        res, cyc, inn = m.A.gmres(m.M, db, dx, 1e-8, restart=30, max_cycles=20, mode=N.GMRES_CGS2)
        solve_s += time.perf_counter() - ts
        inner.append(inn)
        m.problem.dense_matvec("laplace", 1e3, 0.0, dx, dy)
        res_max = max(res_max, float(res[-1]))
        dense_max = max(dense_max, float(np.linalg.norm(dy.download() - rhs) / np.linalg.norm(rhs)))
        for a in (db, dx, dy):
            a.free()
    wall_b = min(t for t, _ in batched)
    wall_s = min(t for t, _ in single)
    out = {
        "metric": "batched FP32 ULV factorization of 64 variable-size problems (shared launches)",
        "value": round(len(specs) / wall_b, 3), "unit": "factorizations/s", "n_gpus": 1, "steps": reps,
        "higher_is_better": True, "dtype": "f32 factorization, f64 GMRES", "data": "synthetic",
        "config": {"workload": "64 problems, N log-uniform [4096, 65536], theta {1,2,3}, leaf {128,256}, "
                               "rank {32,64,128}, operator eps 1e-8, cube Laplace K(0)=1e3",
                   "points": int(sum(s.n for s in specs))},
        "batched_wall_ms": round(1e3 * wall_b, 2), "batched_device_ms": round(1e3 * min(d for _, d in batched), 2),
        "one_by_one_wall_ms": round(1e3 * wall_s, 2), "one_by_one_device_ms": round(1e3 * min(d for _, d in single), 2),
        "speedup_wall": round(wall_s / wall_b, 3),
        "gmres": {"tol": 1e-8, "solves": len(specs), "seconds": round(solve_s, 3), "inner_iters_max": int(max(inner)),
                  "h2_relres_max": res_max, "dense_relres_max": dense_max},
    }
    b.close()
    ctx.close()
    print(json.dumps(out))


def bench_serve(args):
    """16 problems (N = 2^14, 2^15, 2^16 five times each and one 2^17; cube
    Laplace K(0)=1e3, theta 2, operator eps 1e-13, preconditioner eps 1e-6
    with rank <= 64, FP32 factorization with bfloat16-stored factors), made
    resident first; then a Poisson stream of requests, each a problem id and
    1..32 right-hand sides U(-1, 1), solved to 1e-10 by GMRES-IR with the
    inner steps replayed as HIP graphs; p50/p99 latency from arrival to
    solution at the offered load, and solves/s. A problem holds about 0.27 MB
    of HBM per point (N = 65536: 17 GB, most of it the FP64 operator at
    eps 1e-13), so 16 problems up to N = 2^19 (2.3 M points) would need
    about 630 GB: the sizes here keep all 16 within the HBM budget; larger
    sets are served by eviction and rebuilding (tests/test_gpu_serving.py)."""
    from mxc_amd.device import Context
    from mxc_amd.serving import ProblemSpec, Server, latency_stats, poisson_load

    sizes = [1 << 14, 1 << 15, 1 << 16] * 5 + [1 << 17]
    ctx = Context(0)
    free, total = ctx.mem_info()
    ctx.close()
    budget = int(0.8 * free)
    srv = Server(hbm_budget=budget, slots=4, max_rhs=32)
    try:
        for pid, n in enumerate(sizes):
            srv.register(pid, ProblemSpec(n=n, seed=500 + pid))
        t0 = time.time()
        warm = [srv.submit(pid, np.ones((n, 1))) for pid, n in enumerate(sizes)]
        for k, r in enumerate(warm):
            r.wait(1200)
            log(f"problem {k} (N={sizes[k]}) resident after {time.time() - t0:.1f} s")
        t_warm = time.time() - t0
        log(f"16 problems resident in {t_warm:.1f} s, {srv.resident_bytes() / 2**30:.1f} GiB, "
            f"evictions {srv.evictions}")
        ncalls = len(srv.calls)
        reqs = poisson_load(srv, list(range(16)), sizes, rate=args.rate, count=args.requests, seed=11, timeout=600)
        st = latency_stats(reqs, srv.calls[ncalls:])
        worst = max(max(r.resid) for r in reqs)
        out = {
            "metric": "serving: solves/s and p50/p99 request latency at a fixed offered load",
            "value": round(st["solves_per_s"], 3), "unit": "solves/s", "n_gpus": 1, "higher_is_better": True,
            "dtype": "f32 ULV (bf16-stored factors) + f32 inner operator, f64 Arnoldi/residual", "data": "synthetic",
            "config": {"workload": "16 resident problems N=2^14,2^15,2^16 (x5) and 2^17, cube Laplace, theta 2, "
                                   "tol 1e-10 (H2 residual 5e-11)",
                       "offered_load_req_per_s": args.rate, "requests": args.requests, "rhs_per_request": "1..32",
                       "slots": 4, "max_rhs_per_call": 32, "hbm_budget_gib": round(budget / 2**30, 1)},
            "p50_ms": round(st["p50_ms"], 2), "p99_ms": round(st["p99_ms"], 2), "mean_ms": round(st["mean_ms"], 2),
            "latency_stats": st, "h2_relres_max": worst, "resident_gib": round(srv.resident_bytes() / 2**30, 2),
            "evictions": srv.evictions, "warm_up_s": round(t_warm, 1),
        }
    finally:
        srv.close()
    print(json.dumps(out))


def factor_flops(M, points, levels, theta):
    """Algorithmic flops of one ULV factorization of M, from its level
    structure (ranks s_i, box sizes M_i, r_i = M_i - s_i, near pairs):
    per box 4 M^3 (two-sided rotation) + 2 r^3 (LU and inverse of the rr
    block) + 2 r^2 M (V) + 2 r^2 s + 2 s^2 r (rs and ss updates); per near
    pair j != i, 2 M_i M_j (M_i + M_j) (two-sided rotation) and the single
    Schur complement 2 s_i r_i s_j + 2 s_i r_j s_j + 2 s_i r_j^2 + 2 s_i^2 r_j."""
    from mxc_amd.solver import build_tree, csr

    _, child, body, R, Cc = build_tree(points, levels)
    rp, col = csr(False, theta, levels, child, body, R, Cc)
    total = 0.0
    for l in range(levels + 1):
        d, s = (a.astype(np.float64) for a in M.dims(l))
        r = d - s
        total += float(np.sum(4 * d**3 + 2 * r**3 + 2 * r**2 * d + 2 * r**2 * s + 2 * s**2 * r))
        c0 = (1 << l) - 1
        for i in range(1 << l):
            for p in range(rp[c0 + i], rp[c0 + i + 1]):
                j = col[p] - c0
                if j == i:
                    continue
                total += 2 * d[i] * d[j] * (d[i] + d[j])
                total += 2 * s[i] * r[i] * s[j] + 2 * s[i] * r[j] * s[j] + 2 * s[i] * r[j] ** 2 + 2 * s[i] ** 2 * r[j]
    return total


def host_info():
    """Host cores as the process sees them: nproc, OMP_NUM_THREADS, CPU model."""
    model = ""
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), "")
    except OSError:
        pass
    return {"nproc": len(os.sched_getaffinity(0)), "omp_num_threads": os.environ.get("OMP_NUM_THREADS"),
            "cpu_model": model}


def host_threads():
    """Threads of the CPU runs: OMP_NUM_THREADS when set (the host's share of
    cores), else every core the process may use."""
    env = os.environ.get("OMP_NUM_THREADS")
    return max(1, int(env)) if env and env.isdigit() else len(os.sched_getaffinity(0))


def timed_solves(A, M, O, ns, seconds, max_solves):
    times = []
    t_all = time.perf_counter()
    k = 0
    while True:
        b = O.uniform_vec(ns, -1, 1, 1000 + k)
        ts = time.perf_counter()
        A.gmres(M, b, 1e-10, 30, 20)
        times.append(time.perf_counter() - ts)
        k += 1
        if time.perf_counter() - t_all > seconds or k >= max_solves:
            break
    return float(np.mean(times)), k


def cpu_baseline(ns, sampling, op_eps, n_target, cube=False):
    """The CPU oracle (FP64 restatement of the reference) on the metric's
    workload at ns points (construction and factorization untimed): GMRES
    solves (reference mode, restart 30, tol 1e-10) timed for about 10 s on
    every host thread and for about 5 s on one thread. `value` is the
    all-threads rate scaled to N = n_target by the O(N) cost of a solve; the
    measured rates at ns are reported beside it."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    threads = host_threads()
    pts = O.cube_points(ns, 2) if cube else O.sphere_points(ns, 2)
    kern = "laplace" if cube else "yukawa"
    P = O.Problem(pts, 256)
    A = P.h2(kernel=kern, alpha=1.0, tol=op_eps, sampling=sampling, nthreads=threads)
    M = P.h2(kernel=kern, alpha=1.0, tol=1e-6, rank=64, orth=True, hss_far=True, sampling=sampling, nthreads=threads)
    M.factorize()
    per, k = timed_solves(A, M, O, ns, 10.0, 200)
    A.set_threads(1)
    M.set_threads(1)
    per1, k1 = timed_solves(A, M, O, ns, 5.0, 50)
    return {
        "value": round((1.0 / per) * ns / n_target, 6),
        "unit": "solves/s (scaled to N=2^20 from the measured rate at the sample size)",
        "cores": threads,
        "kind": "port",
        "sample": f"oracle GMRES solves (reference mode, CGS1, restart 30, tol 1e-10) at N={ns}, "
                  f"{'unit-cube Laplace' if cube else 'unit-sphere Yukawa'}, "
                  f"operator eps {op_eps:g}, preconditioner eps 1e-6 rank 64: {k} solves on {threads} threads "
                  f"({per * 1e3:.1f} ms each), {k1} on 1 thread ({per1 * 1e3:.1f} ms each)",
        "measured_solves_per_s_at_sample": round(1.0 / per, 4),
        "measured_solves_per_s_at_sample_1thread": round(1.0 / per1, 4),
        "host": host_info(),
    }


def cpu_phases(sizes):
    """The reference CPU path's phases timed with the oracle on the host cores
    (lorasp N 2 256 1e-10 100 2000: cube points, Laplace with K(0) = 1e3,
    b = A x_true): operator and preconditioner construction, factorization,
    one substitution, GMRES (restart 10, 50 cycles) to 1e-10 and the dense
    check, on every host thread; factorization, substitution and GMRES again
    on one thread. Prints one JSON line per size."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import oracle_lib as O

    threads = host_threads()
    for n, seed in sizes:
        ph = {}
        t = time.perf_counter()
        pts = O.cube_points(n, seed)
        P = O.Problem(pts, 256)
        x = O.uniform_vec(n, 0, 1, 5489)
        b = O.dense_matvec("laplace", 1e3, 0.0, P.bodies, x, threads)
        ph["dense_matvec_s"] = time.perf_counter() - t
        t = time.perf_counter()
        A = P.h2(tol=1e-10, sampling=2000, nthreads=threads)
        ph["operator_construct_s"] = time.perf_counter() - t
        t = time.perf_counter()
        M = P.h2(rank=100, orth=True, hss_far=True, sampling=2000, nthreads=threads)
        ph["preconditioner_construct_s"] = time.perf_counter() - t
        cerr = O.rel_err(A.matvec(x), b)
        res1 = {}
        for nt in (threads, 1):
            M1 = M if nt == threads else P.h2(rank=100, orth=True, hss_far=True, sampling=2000, nthreads=threads)
            M1.set_threads(nt)
            A.set_threads(nt)
            t = time.perf_counter()
            assert M1.factorize() == 0
            f = time.perf_counter() - t
            t = time.perf_counter()
            z = M1.precondition(b)
            sub = time.perf_counter() - t
            t = time.perf_counter()
            xs, res = A.gmres(M1, b, 1e-10, 10, 50)
            g = time.perf_counter() - t
            res1[nt] = dict(factor_s=f, substitution_s=sub, gmres_s=g, gmres_cycles=len(res) - 1,
                            gmres_residual=float(res[-1]), substitution_err=O.rel_err(z, x))
        t = time.perf_counter()
        rr = O.rel_err(O.dense_matvec("laplace", 1e3, 0.0, P.bodies, xs, threads), b)
        ph["dense_check_s"] = time.perf_counter() - t
        print(json.dumps({"N": n, "seed": seed, "threads": threads, "phases_all_threads": {**ph, **res1[threads]},
                          "phases_1_thread": res1[1], "construct_err": cerr, "dense_relres": rr,
                          "solution_err": O.rel_err(xs, x), "host": host_info()}), flush=True)


if __name__ == "__main__":
    main()
