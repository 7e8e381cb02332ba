#!/usr/bin/env python3
"""Benchmark of the electrostatic BEM hot path on one MI355X.

Headline workload (the 100,000-triangle case): 32 coaxial ring electrodes (short cylinders of
radius 1.9 m, axial width 5 cm, spread over 8 m) inside a grounded cylinder of radius 2 m,
exactly 100,000 triangles; ring voltages U[-20, 0] kV (seeded); 10,000,000 field points
uniform in the bore (r < 1.8 m, |z| < 3.5 m); 16 voltage sets; a 10,000-point axial adjoint.

A step is one mixed-precision pipeline on that workload with every input resident in HBM:
dense collocation matrix assembly (FP64 80 GB + FP32 40 GB), GMRES to 1e-10 relative residual
with FP32-matrix inner solves and FP64 refinement, then potential + field at the 10 M points
with the FP32 far field (FP64 near field and sums).  `value` = point-element pairs of the field
pass per second of the whole step (field evaluations per second with the solve time included),
valid when the step's field is within 1e-6 of the FP64 pipeline's (norm-wise, max / rms and
potential / voltage scale).  The FP64 pipeline (FP64 matrix, FP64 GMRES, FP64 field) is timed
the same way on the same inputs; `pipeline_speedup` is the ratio of the two.  Launch shapes
are tuned on the workload before they are timed (the far pass's element ranges on a 1 M-point
sample, the matrix-vector product's split target per precision before its own pipeline).

Also reported, as extra keys: the 20,000-triangle coax (1 M points in the gap; FP64 and mixed
field passes, GMRES, matrix-vector bandwidth), 64 voltage sets on it (lockstep GMRES on MFMA
block products) with 4 M points in uneven batches, and the 1,728-triangle cube.

Prints one JSON line.  --gpus N > 1 is refused: the library runs on one GPU.
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

# FLOPs per (point, element) pair for field + potential, by regime (add/sub/mul = 1, FMA = 2,
# sqrt/div/rsqrt/log/atan2 = 1): cubature 14 + 21 n_q; closed form 260 (convention);
# segments 14 + 21 n_nodes, analytic 100.  Census order: triangle 7/12/33/closed,
# rectangle 7/12/33/closed, segment 4/16/analytic, triangle 6-point (FP32 ladder).
FLOPS_PER_PAIR = np.array([161, 266, 707, 260, 161, 266, 707, 260, 98, 350, 100, 140], dtype=np.float64)
FAR = np.ones(12)
FAR[[3, 7, 10]] = 0.0  # the far kernels evaluate every pair but the closed-form ones
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector peak
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (published; it needs packed FP32)
PEAK_HBM_GBS = 8000.0
# the oracle's per-core field rate over the compiled reference's, measured on one core of the
# same host type: 3.68 / 2.95 M point-element pairs/s (1,728-triangle cube, 4,096 points)
CPU_ORACLE_OVER_REFERENCE = 3.68 / 2.95

# HBM bytes per launch of the mixed far kernel on the headline workload, by element-range size
# (which sets the dispatch: 640 -> 208,384 points x 160 ranges, 2,560 -> 818,176 x 40, 10,240 ->
# 1,973,760 x 16), from rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE runs scaled by the readings of
# a 1 GiB device copy (FETCH_SIZE x 1.99990, WRITE_SIZE x 1.0): 640: 0.847 GB read, 1.067 GB
# written; 2,560: 0.885 / 1.047 GB; 10,240: 0.996 / 1.011 GB (the writes are the partial-sum
# slices of the ranges).  Source: the profile named here; other sizes not measured (null).
TRAFFIC = {"bytes": {640: 1.914e9, 2560: 1.932e9, 10240: 2.006e9},
           "profile": "profiles/pmc_traffic_k_field_far32c_bore100k_10Mpts.md"}

METRIC = "point-element field evals/s + solve time, mixed vs FP64 at 1e-6 rel. error"


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def host_cores():
    """The cores this process may run on: its affinity set, within the CPU share the host
    grants the job (OMP_NUM_THREADS where the host sets it)."""
    n = len(os.sched_getaffinity(0))
    share = os.environ.get("OMP_NUM_THREADS", "")
    if share.isdigit() and int(share) > 0:
        n = min(n, int(share))
    return max(1, n)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3, help="timed mixed pipelines")
    ap.add_argument("--warmup", type=int, default=1, help="untimed mixed pipelines before them")
    ap.add_argument("--points", type=int, default=10_000_000)
    ap.add_argument("--elements", type=int, default=100_000)
    ap.add_argument("--tune-points", type=int, default=0,
                    help="points of the sample the far-pass element ranges are tuned on (0: all of them; "
                         "a 1 M-point sample preferred 10,240-element ranges, which are the slowest at 10 M)")
    ap.add_argument("--cpu-points", type=int, default=0,
                    help="points of the bounded CPU-baseline sample; 0: 1,600 per thread")
    ap.add_argument("--cpu-threads", type=int, default=0, help="OpenMP threads of the CPU baseline; 0: all cores")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch shapes instead of tuning them on the workload")
    ap.add_argument("--no-extras", action="store_true",
                    help="skip the error split, 16 voltage sets and adjoint of the headline workload")
    ap.add_argument("--no-coax", action="store_true", help="skip the 20,000-triangle coax, its batches and the cube")
    ap.add_argument("--field-only", choices=["mixed", "fp64"], default=None,
                    help="only field passes of the headline workload (provisional charge densities), "
                         "steps x, for profiling")
    ap.add_argument("--far-range", type=int, default=-1,
                    help="--field-only: elements per far-pass element range (-1: the library's default)")
    ap.add_argument("--block-products", action="store_true",
                    help="only block products on the 20,000-triangle coax matrix (64 right-hand sides on "
                         "FP32 and BF16 storage, 16 on FP32, 16 and 64 on FP64), steps x each, for profiling")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device (reference kernel for the rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        return copy_calibration()

    from kassiopeia_amd import kemfield as K

    ctx = K.Context(0)
    if args.field_only:
        return field_only(ctx, args)
    if args.block_products:
        return block_products(ctx, args)
    t_start = time.perf_counter()
    warm_up(ctx)
    line = run_bore(ctx, args)
    if not args.no_coax:
        line.update(run_coax(ctx, args))
    line["bench_wall_s"] = round(time.perf_counter() - t_start, 1)
    ctx.close()
    print(json.dumps(line), flush=True)


def block_products(ctx, args):
    """steps block products on the 20,000-triangle coax matrix at the library's default launch
    shapes: 64 right-hand sides on the FP32 and the BF16 storage (32 x 32 MFMA tiles), then 16 on
    the FP32 storage and 16 and 64 on the FP64 storage (16 x 16 MFMA tiles): the command the
    rocprofv3 summaries of the streamed block-product kernels run."""
    from kassiopeia_amd import geometry as G
    from kassiopeia_amd import kemfield as K

    mesh = G.coax_mesh()
    cont = K.KSurfaceContainer(ctx, mesh)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True, bf16=True)
    X = np.random.default_rng(9).normal(size=(mesh.n, 64))
    out = {"block_products": mesh.n}
    for prec, k in (("mixed", 64), ("bf16", 64), ("mixed", 16), ("fp64", 16), ("fp64", 64)):
        A.MultiplyBatch(X[:, :k], prec)
        _timing(ctx, True)
        for _ in range(args.steps):
            A.MultiplyBatch(X[:, :k], prec)
        kt, kn = _kernel_time(ctx)
        _timing(ctx, False)
        out[f"{prec}_{k}_ms"] = kt / max(kn, 1)
    A.close()
    cont.close()
    ctx.close()
    print(json.dumps(out), flush=True)


def warm_up(ctx):
    """Loads every kernel the pipelines launch (both precisions) on a small mesh, so that no
    timed region pays for module loading."""
    from kassiopeia_amd import geometry as G
    from kassiopeia_amd import kemfield as K

    cube = G.cube_mesh(6)
    cont = K.KSurfaceContainer(ctx, cube)
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True)
    x, _, _ = A.solve(cont.rhs(), tolerance=1e-10)
    A.solve(cont.rhs(), tolerance=1e-10, precision="mixed")
    A.solve_batch(np.stack([cont.rhs()] * 16, 1), tolerance=1e-10, precision="mixed")
    pts = G.points_shell(300_000, 0.9, 2.0, 3)
    for prec in ("fp64", "mixed"):
        K.KIntegratingElectrostaticFieldSolver(cont, x, prec).ElectricFieldAndPotential(pts)
    A.close()
    cont.close()


def _timing(ctx, on):
    ctx.check(ctx.lib.kem_ctx_set_kernel_timing(ctx.h, 1 if on else 0))


def _kernel_time(ctx):
    kt, kn = ctypes.c_double(), ctypes.c_int()
    ctx.check(ctx.lib.kem_ctx_kernel_time(ctx.h, ctypes.byref(kt), ctypes.byref(kn)))
    return kt.value, kn.value


def _plan(ctx, n, npts, range_elems, precision):
    out = (ctypes.c_int64 * 8)()
    ctx.check(ctx.lib.kem_field_plan(0, 1, precision, 0, n, 1, npts, range_elems, -1, out))
    return {"ranges": out[3], "slices": out[4], "chunk": out[5], "xcd": out[6]}


# FLOPs per pair of each path of the mixed cluster far pass (kem_field_census_paths order):
# 3-point cluster pairs, 6-point record pairs, per-pair path by the pair's own rule (6, 7, 12,
# 33 points; segments 4, 16 nodes); near pairs belong to the near pass
FLOPS_PER_PATH = np.array([14 + 21 * 3, 140, 140, 161, 266, 707, 98, 350, 0, 0], dtype=np.float64)


def _census_paths(ctx, cont, X, m):
    """pairs per path of the mixed cluster far pass, in the order the pass visits the points;
    returns (counts, algorithmic FLOPs = each pair at the rule its path evaluates (per-pair
    pairs at their own rule), executed FLOPs (the per-pair path at its wavefront rule))"""
    c = (ctypes.c_int64 * 10)()
    ctx.check(ctx.lib.kem_field_census_paths(ctx.h, cont.h, m, X[0].ptr, X[1].ptr, X[2].ptr, c))
    c = np.array(c[:], dtype=np.float64)
    alg = float((c * FLOPS_PER_PATH).sum())
    executed = float(c[0] * FLOPS_PER_PATH[0] + c[1] * FLOPS_PER_PATH[1] + 14.0 * c[2:8].sum() + 21.0 * c[9])
    return c, alg, executed


# FLOPs of one collocation entry (Dirichlet row: potential only) by regime, census order
# (triangle 7/12/33/closed, rectangle 7/12/33/closed, segment 4/16/analytic, unused)
ASM_FLOPS = np.array([95, 155, 407, 200, 95, 155, 407, 200, 59, 203, 100, 0], dtype=np.float64)
ASM_FAR = np.ones(12)
ASM_FAR[[3, 7, 10, 11]] = 0.0


def _assembly_flops(ctx, cont, mesh):
    """census of the collocation pairs (targets at the element centroids) by FP64 regime:
    (FLOPs of the cubature kernel's entries, FLOPs of the closed-form entries)"""
    from kassiopeia_amd import _lib
    from kassiopeia_amd import kemfield as K

    C = mesh.centroids()
    T = [K.DeviceArray(ctx, np.ascontiguousarray(C[:, k])) for k in range(3)]
    c = _census(ctx, cont, T, mesh.n, _lib.KEM_FP64)
    for a in T:
        a.close()
    return float((c * ASM_FLOPS * ASM_FAR).sum()), float((c * ASM_FLOPS * (1 - ASM_FAR)).sum())


def _assemble_timed(ctx, cont, **storage):
    """assemble with kernel timing: (matrix, wall seconds incl. allocation, kernel ms)"""
    from kassiopeia_amd import kemfield as K

    ctx.sync()
    _timing(ctx, True)
    t0 = time.perf_counter()
    A = K.KBoundaryIntegralMatrix(cont, **storage)
    ctx.sync()
    t = time.perf_counter() - t0
    kt, kn = _kernel_time(ctx)
    _timing(ctx, False)
    return A, t, kt


def _census(ctx, cont, X, m, precision):
    c = (ctypes.c_int64 * 12)()
    ctx.check(ctx.lib.kem_field_census(ctx.h, cont.h, m, X[0].ptr, X[1].ptr, X[2].ptr, precision, c))
    return np.array(c[:], dtype=np.float64)


def _errors(E, phi, Em, phim, vscale):
    """Error of (Em, phim) against (E, phi): norm-wise over the point set, the largest
    deviation relative to the RMS field, and the potential relative to the voltage scale; the
    per-point relative field error is amplified where |E| nearly vanishes and is not the bar."""
    dE = Em - E
    Ed = np.linalg.norm(dE, axis=1)
    En = np.linalg.norm(E, axis=1)
    with np.errstate(divide="ignore", invalid="ignore"):
        pp = Ed / En
    pp = pp[np.isfinite(pp)]
    return {"E_l2": float(np.linalg.norm(dE) / np.linalg.norm(E)),
            "E_max_over_rms": float(Ed.max() / np.sqrt((E ** 2).sum(1).mean())),
            "phi_max_over_V": float(np.abs(phim - phi).max() / vscale),
            "E_pointwise_p99": float(np.quantile(pp, 0.99)) if pp.size else 0.0}


def bore_workload(n_elements, m_points):
    from kassiopeia_amd import geometry as G

    rng = np.random.default_rng(20)
    volts = rng.uniform(-20000.0, 0.0, 32)
    mesh = G.ring_bore_mesh(volts=volts, n_target=n_elements)
    pts = G.points_annulus(m_points, 0.0, 1.8, 3.5, seed=55)
    return mesh, volts, pts, rng


def field_only(ctx, args):
    """steps field passes of the headline workload at one precision with provisional charge
    densities (the time per pair does not depend on their values): the command the rocprofv3
    traffic and kernel profiles of the far kernels run."""
    from kassiopeia_amd import _lib
    from kassiopeia_amd import autotune as T
    from kassiopeia_amd import kemfield as K

    mesh, volts, pts, _ = bore_workload(args.elements, args.points)
    N, M = mesh.n, pts.shape[0]
    cont = K.KSurfaceContainer(ctx, mesh)
    prec = {"fp64": _lib.KEM_FP64, "mixed": _lib.KEM_MIXED}[args.field_only]
    if args.far_range >= 0:
        # the element ranges the headline run tuned (a sweep here would add dispatches)
        T.set_knob(ctx, "far_range_" + args.field_only, args.far_range)
    X = [K.DeviceArray(ctx, pts[:, k]) for k in range(3)]
    S = K.DeviceArray(ctx, np.full(N, 1e-9))
    O = [K.DeviceArray.empty(ctx, M) for _ in range(4)]
    ctx.sync()
    times = []
    for _ in range(args.steps):
        t0 = time.perf_counter()
        ctx.check(ctx.lib.kem_field_eval_device(ctx.h, cont.h, S.ptr, M, X[0].ptr, X[1].ptr, X[2].ptr, prec,
                                                O[0].ptr, O[1].ptr, O[2].ptr, O[3].ptr))
        ctx.sync()
        times.append(time.perf_counter() - t0)
    rng_used = T.get(ctx, "far_range_" + args.field_only)
    plan = _plan(ctx, N, M, rng_used, prec)
    print(json.dumps({"field_only": args.field_only, "elements": N, "points": M, "pass_s": times,
                      "far_range": rng_used, "plan": plan}), flush=True)
    for a in X + O + [S]:
        a.close()
    cont.close()
    ctx.close()


def run_bore(ctx, args):
    from kassiopeia_amd import _lib
    from kassiopeia_amd import autotune as T
    from kassiopeia_amd import kemfield as K

    lib = ctx.lib
    mesh, volts, pts, rng = bore_workload(args.elements, args.points)
    N, M = mesh.n, pts.shape[0]
    vscale = float(np.abs(volts).max())
    log(f"bore: mesh {N} triangles, {M} points")
    cont = K.KSurfaceContainer(ctx, mesh)
    b = cont.rhs()
    out = {}

    # ------------------------------------------------------------ launch shapes (sampled)
    tuning = None
    if not args.no_autotune:
        t0 = time.perf_counter()
        sig0 = np.full(N, 1e-9)  # the time per pair does not depend on the charge densities
        sample = pts[:min(args.tune_points, M)] if args.tune_points > 0 else pts
        # the same candidates for both precisions, on the whole point set: the chunking of the
        # per-XCD element ranges depends on the point count (10 M points, s per pass: mixed
        # {640: 2.10, 1280: 2.13, 2560: 2.14, 5120: 2.19, 10240: 2.19}, FP64 {640: 9.67,
        # 1280: 9.70, 2560: 9.85, 5120: 9.98, 10240: 10.16})
        cands = (640, 2560, 10240)
        tuning = {"far_range_mixed": T.tune_field(cont, sig0, sample, "mixed", repeats=1, candidates=cands),
                  "far_range_fp64": T.tune_field(cont, sig0, sample, "fp64", repeats=1, candidates=cands)}
        log(f"autotune on {sample.shape[0]} points ({time.perf_counter() - t0:.1f} s): " +
            ", ".join(f"{k} {v['best']} {v['ms']}" for k, v in tuning.items()))

    X = [K.DeviceArray(ctx, pts[:, k]) for k in range(3)]
    O64 = [K.DeviceArray.empty(ctx, M) for _ in range(4)]
    Omx = [K.DeviceArray.empty(ctx, M) for _ in range(4)]

    def field(sigma, prec, outs, timed=False):
        S = K.DeviceArray(ctx, sigma)
        if timed:
            _timing(ctx, True)
        t0 = time.perf_counter()
        ctx.check(lib.kem_field_eval_device(ctx.h, cont.h, S.ptr, M, X[0].ptr, X[1].ptr, X[2].ptr, prec,
                                            outs[0].ptr, outs[1].ptr, outs[2].ptr, outs[3].ptr))
        ctx.sync()
        t = time.perf_counter() - t0
        kt = _kernel_time(ctx) if timed else (0.0, 0)
        if timed:
            _timing(ctx, False)
        S.close()
        return t, kt

    def pipeline(mixed, outs, keep_matrix=False):
        """assembly + GMRES 1e-10 + field pass, wall time between stream synchronisations"""
        t0 = time.perf_counter()
        A, t_asm, asm_ms = _assemble_timed(ctx, cont, fp64=True, fp32=mixed)
        x, st, _ = A.solve(b, tolerance=1e-10, precision="mixed" if mixed else "fp64")
        t_solve = time.perf_counter() - t0 - t_asm
        t_field, kt = field(x, _lib.KEM_MIXED if mixed else _lib.KEM_FP64, outs, timed=True)
        total = time.perf_counter() - t0
        if not keep_matrix:
            A.close()
            A = None
        return {"total": total, "assembly": t_asm, "asm_ms": asm_ms, "solve": t_solve, "field": t_field,
                "far_ms": kt[0],
                "far_regions": kt[1], "iterations": st.iterations, "refinements": st.refinements,
                "residual": st.true_residual, "sigma": x, "A": A}

    def tune_matvec(fp32):
        """the matrix-vector product's split target on this matrix, FP64 storage for the FP64
        pipeline, FP32 for the mixed one (93 of its ~95 products), each tuned before its own
        pipeline (the knob is shared)"""
        if args.no_autotune:
            return
        t0 = time.perf_counter()
        A = K.KBoundaryIntegralMatrix(cont, fp64=not fp32, fp32=fp32)
        r = T.tune_matvec(A, repeats=3, candidates=(256, 512, 768, 1024, 2048, 4096))
        A.close()
        key = "gemv_blocks_" + ("fp32" if fp32 else "fp64")
        tuning[key] = r
        log(f"{key} {r['best']} {r['ms']} ({time.perf_counter() - t0:.1f} s)")

    # ------------------------------------------------------------ FP64 pipeline (the plain path)
    tune_matvec(False)
    p64 = pipeline(False, O64)
    log(f"FP64 pipeline {p64['total']:.2f} s: assembly {p64['assembly']:.2f} s, GMRES {p64['iterations']} it "
        f"{p64['solve']:.2f} s (res {p64['residual']:.1e}), field {p64['field']:.2f} s")

    # ------------------------------------------------------------ mixed pipeline: warm-up + timed steps
    tune_matvec(True)
    for _ in range(args.warmup):
        pipeline(True, Omx)
    steps = [pipeline(True, Omx, keep_matrix=(k == args.steps - 1)) for k in range(args.steps)]
    pm = steps[-1]
    t_step = float(np.mean([s["total"] for s in steps]))
    log("mixed pipelines: " + ", ".join(f"{s['total']:.3f}" for s in steps) +
        f" s; last: assembly {pm['assembly']:.2f} s, GMRES {pm['iterations']} it / {pm['refinements']} ref "
        f"{pm['solve']:.2f} s (res {pm['residual']:.1e}), field {pm['field']:.2f} s")

    # ------------------------------------------------------------ errors against the FP64 pipeline
    phi64 = O64[0].get()
    E64 = np.stack([O64[k].get() for k in (1, 2, 3)], 1)
    phimx = Omx[0].get()
    Emx = np.stack([Omx[k].get() for k in (1, 2, 3)], 1)
    err = _errors(E64, phi64, Emx, phimx, vscale)
    mixed_ok = err["E_l2"] <= 1e-6 and err["E_max_over_rms"] <= 1e-6 and err["phi_max_over_V"] <= 1e-6
    speedup = p64["total"] / t_step
    log(f"mixed vs FP64 pipeline: |dE| l2 {err['E_l2']:.2e}, max|dE|/rms|E| {err['E_max_over_rms']:.2e}, "
        f"|dphi|/V {err['phi_max_over_V']:.2e}; pipeline {p64['total']:.2f} s -> {t_step:.2f} s = {speedup:.2f}x "
        f"(field pass {p64['field']:.2f} -> {pm['field']:.2f} s = {p64['field'] / pm['field']:.2f}x)")
    out.update(pipeline_s_fp64=p64["total"], pipeline_s_mixed=t_step, pipeline_speedup=speedup,
               pipeline_s_mixed_steps=[s["total"] for s in steps],
               assembly_s_fp64=p64["assembly"], assembly_s_mixed=pm["assembly"],
               solve_s_fp64=p64["solve"], solve_s_mixed=pm["solve"],
               gmres_iterations_fp64=p64["iterations"], gmres_iterations_mixed=pm["iterations"],
               refinements_mixed=pm["refinements"], residual_fp64=p64["residual"], residual_mixed=pm["residual"],
               field_s_fp64=p64["field"], field_s_mixed=pm["field"], field_speedup=p64["field"] / pm["field"],
               field_pairs_per_s_mixed=N * M / pm["field"], field_pairs_per_s_fp64=N * M / p64["field"],
               mixed_err_E_l2=err["E_l2"], mixed_err_E_max_over_rms=err["E_max_over_rms"],
               mixed_err_phi_max_over_V=err["phi_max_over_V"], mixed_err_E_pointwise_p99=err["E_pointwise_p99"],
               mixed_within_1e6=mixed_ok)

    # ------------------------------------------------------------ roofline of the mixed far kernel
    prec_mx = _lib.KEM_MIXED
    cmx, far_flops, far_flops_executed = _census_paths(ctx, cont, X, M)
    far_ms = float(np.mean([s["far_ms"] / max(s["far_regions"], 1) for s in steps]))
    rng_used = T.get(ctx, "far_range_mixed")
    plan = _plan(ctx, N, M, rng_used, prec_mx)
    launches = -(-M // plan["chunk"]) if plan["xcd"] else 1
    achieved = far_flops / (far_ms * 1e-3) / 1e12
    log(f"far kernel (mixed): {far_ms:.1f} ms per pass in {launches} launches, census {far_flops:.3e} FLOP "
        f"-> {achieved:.1f} TFLOP/s; pairs by regime {cmx.astype(int).tolist()}")

    # ------------------------------------------------------------ the assembly kernels' roofline
    asm_ms = pm["asm_ms"]
    asm_far_flops, asm_near_flops = _assembly_flops(ctx, cont, mesh)
    asm_tflops = (asm_far_flops + asm_near_flops) / (asm_ms * 1e-3) / 1e12
    log(f"assembly kernels (FP64 + FP32 storage): {asm_ms:.1f} ms for {asm_far_flops + asm_near_flops:.3e} FLOP "
        f"= {asm_tflops:.1f} TFLOP/s ({asm_tflops / PEAK_FP64_TFLOPS:.2f} of the FP64 peak)")
    out.update(assembly_kernel_ms=asm_ms, assembly_flops=asm_far_flops + asm_near_flops,
               assembly_tflops=asm_tflops, assembly_frac_fp64_peak=asm_tflops / PEAK_FP64_TFLOPS)

    # ------------------------------------------------------------ BF16 inner solves (FP64 refinement)
    ctx.sync()
    t0 = time.perf_counter()
    Ab = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=False, bf16=True)
    ctx.sync()
    tb_asm = time.perf_counter() - t0
    xb, stb16, _ = Ab.solve(b, tolerance=1e-10, precision="bf16")
    tb_solve = time.perf_counter() - t0 - tb_asm
    Ab.close()
    Otmp = [K.DeviceArray.empty(ctx, M) for _ in range(4)]
    tb_field, _ = field(xb, _lib.KEM_MIXED, Otmp)
    eb = _errors(E64, phi64, np.stack([Otmp[k].get() for k in (1, 2, 3)], 1), Otmp[0].get(), vscale)
    for a in Otmp:
        a.close()
    tb_total = tb_asm + tb_solve + tb_field
    log(f"BF16-inner pipeline: assembly {tb_asm:.2f} s, GMRES {stb16.iterations} it / {stb16.refinements} ref "
        f"{tb_solve:.2f} s (res {stb16.true_residual:.1e}), field {tb_field:.2f} s = {tb_total:.2f} s "
        f"({p64['total'] / tb_total:.2f}x FP64); error max/rms {eb['E_max_over_rms']:.2e}")
    out.update(bf16_pipeline_s=tb_total, bf16_assembly_s=tb_asm, bf16_solve_s=tb_solve,
               bf16_gmres_iterations=stb16.iterations, bf16_refinements=stb16.refinements,
               bf16_residual=stb16.true_residual, bf16_pipeline_speedup=p64["total"] / tb_total,
               bf16_err_E_max_over_rms=eb["E_max_over_rms"], bf16_err_E_l2=eb["E_l2"])

    # ------------------------------------------------------------ where the error comes from
    sig64 = p64["sigma"]
    A = pm["A"]
    if not args.no_extras:
        Otmp = [K.DeviceArray.empty(ctx, M) for _ in range(4)]
        field(sig64, _lib.KEM_MIXED, Otmp)
        ef = _errors(E64, phi64, np.stack([Otmp[k].get() for k in (1, 2, 3)], 1), Otmp[0].get(), vscale)
        field(pm["sigma"], _lib.KEM_FP64, Otmp)
        es = _errors(E64, phi64, np.stack([Otmp[k].get() for k in (1, 2, 3)], 1), Otmp[0].get(), vscale)
        for a in Otmp:
            a.close()
        log(f"  from the field pass (FP64 sigma): l2 {ef['E_l2']:.2e}, max/rms {ef['E_max_over_rms']:.2e}; "
            f"from the solve (FP64 field of the mixed sigma): l2 {es['E_l2']:.2e}, max/rms {es['E_max_over_rms']:.2e}")
        out.update(mixed_field_only_err_E_l2=ef["E_l2"], mixed_field_only_err_E_max_over_rms=ef["E_max_over_rms"],
                   mixed_solve_only_err_E_l2=es["E_l2"], mixed_solve_only_err_E_max_over_rms=es["E_max_over_rms"],
                   sigma_mixed_vs_fp64=float(np.linalg.norm(pm["sigma"] - sig64) / np.linalg.norm(sig64)))
        # 16 voltage sets in lockstep (mixed: FP32 MFMA block products, FP64 refinement)
        groups = mesh.group
        V16 = rng.uniform(-20000.0, 0.0, size=(33, 16))
        V16[32, :] = 0.0
        Xb, res, stb, _ = A.solve_batch(V16[groups, :], tolerance=1e-10, precision="mixed")
        log(f"16 voltage sets: {stb.seconds:.2f} s ({stb.iterations} it, {stb.refinements} ref, "
            f"worst {res.max():.1e})")
        out.update(batch16_solve_s=stb.seconds, batch16_iterations=stb.iterations,
                   batch16_worst_residual=float(res.max()))
        # its block products: 16 right-hand sides on the 40 GB FP32 and the 80 GB FP64 matrix
        for prec, key, bpe in (("mixed", "fp32", 4), ("fp64", "fp64", 8)):
            _timing(ctx, True)
            for _ in range(3):
                A.MultiplyBatch(Xb, prec)
            kt, kn = _kernel_time(ctx)
            _timing(ctx, False)
            ms = kt / max(kn, 1)
            gbs = bpe * N * N / (ms * 1e-3) / 1e9
            log(f"  block product of the 16 sets, {key} matrix: {ms:.2f} ms = {gbs:.0f} GB/s of matrix")
            out.update({f"batch16_bmm_ms_{key}": ms, f"batch16_bmm_gbs_{key}": gbs})
        # adjoint of L = mean (phi(axis) - target)^2 over 10k axial points
        zax = np.linspace(-3.5, 3.5, 10000)
        pax = np.ascontiguousarray(np.stack([0 * zax, 0 * zax, zax], 1))
        Eax, phax = K.KIntegratingElectrostaticFieldSolver(cont, sig64).ElectricFieldAndPotential(pax)
        g = 2.0 * (phax - (-10000.0)) / phax.size
        ctx.sync()
        t0 = time.perf_counter()
        dsig, dvf = np.zeros(N), np.zeros((N, 9))
        ctx.check(lib.kem_field_vjp(ctx.h, cont.h, K._ptr(sig64), pax.shape[0], K._ptr(pax), K._ptr(g),
                                    K._ptr(dsig), K._ptr(dvf)))
        t_fvjp = time.perf_counter() - t0
        o = _lib.SolveOpts()
        lib.kem_solve_opts_default(ctypes.byref(o))
        o.tolerance, o.precision = 1e-10, _lib.KEM_MIXED
        lam = np.zeros(N)
        sta = _lib.SolveStats()
        ctx.check(lib.kem_solve_transposed(A.h, ctypes.byref(o), K._ptr(dsig), K._ptr(lam), ctypes.byref(sta)))
        t0 = time.perf_counter()
        dvm = np.zeros((N, 9))
        ctx.check(lib.kem_assemble_vjp(ctx.h, cont.h, K._ptr(lam), K._ptr(sig64), K._ptr(dvm)))
        t_mvjp = time.perf_counter() - t0
        log(f"adjoint (10k axial points): field VJP {t_fvjp:.2f} s, transposed solve {sta.seconds:.2f} s "
            f"({sta.iterations} it), matrix VJP {t_mvjp:.2f} s")
        out.update(adjoint_field_vjp_s=t_fvjp, adjoint_solve_s=sta.seconds, adjoint_matrix_vjp_s=t_mvjp,
                   adjoint_solve_iterations=sta.iterations)
    A.close()

    # ------------------------------------------------------------ CPU baseline + parity vs the oracle
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as OR

        threads = args.cpu_threads or host_cores()
        k = min(args.cpu_points or 1600 * threads, M)
        t0 = time.perf_counter()
        phr, Er = OR.field(mesh.shapes, mesh.recs, sig64, pts[:k], threads=threads)
        tc = time.perf_counter() - t0
        rate = k * N / tc
        cpu = {"value": rate, "unit": "point-element pairs/s", "cores": threads, "kind": "port",
               "sample": f"{k} points x {N} triangles of the same workload (its FP64 charge densities), "
                         f"FP64 oracle (the reference's integrators and Kahan sum restated in C), "
                         f"{threads} OpenMP threads (all cores of this job), {tc:.1f} s; field pass only",
               "reference_equivalent_value": rate / CPU_ORACLE_OVER_REFERENCE}
        ep = float(np.max(np.abs(phi64[:k] - phr)) / np.max(np.abs(phr)))
        ee = float(np.max(np.linalg.norm(E64[:k] - Er, axis=1)) / np.max(np.linalg.norm(Er, axis=1)))
        cpu_field_s = N * M / rate
        out.update(parity_phi_fp64_vs_oracle=ep, parity_E_fp64_vs_oracle=ee, cpu_field_s_estimate=cpu_field_s,
                   mixed_pipeline_vs_cpu_field=cpu_field_s / t_step)
        log(f"cpu baseline {rate:.3e} pairs/s on {threads} threads: the CPU field pass alone would take "
            f"{cpu_field_s:.0f} s = {cpu_field_s / t_step:.0f}x the mixed GPU pipeline; FP64 GPU field vs "
            f"oracle on the sample: phi {ep:.2e}, E {ee:.2e}")
    for a in X + O64 + Omx:
        a.close()
    cont.close()

    value = N * M / t_step
    line = {
        "metric": METRIC,
        "value": value,
        "unit": "point-element pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": t_step * 1e3,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": f"32 ring electrodes + grounded r = 2 m cylinder, {N} triangles, {M} field points "
                               f"in the bore; step = mixed pipeline: FP64 + FP32 assembly, GMRES 1e-10 "
                               f"(FP32 inner solves, FP64 refinement), FP32-far-field pass; value = field "
                               f"pairs per second of the whole step",
                   "elements": N, "points": M, "gmres_tolerance": 1e-10},
        # the far kernel is bound by the vector ALU (no matrix-shaped work): achieved = census
        # FLOPs of its pairs (the pairs of each regime x 14 + 21 n_q) / its time from HIP events
        # on the library's stream over the timed steps; one pass is `launches` dispatches (chunks
        # of points), so this equals FLOPs per dispatch / average dispatch duration
        "roofline": {"bound": "valu", "compute": "vector ALU, fp32 (unpacked)", "achieved": achieved,
                     "peak": PEAK_FP32_TFLOPS, "unit": "TFLOP/s", "frac": achieved / PEAK_FP32_TFLOPS,
                     "traffic": TRAFFIC["bytes"].get(rng_used) if (N == 100_000 and M == 10_000_000) else None,
                     "traffic_unit": "HBM bytes per dispatch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "traffic_source": TRAFFIC["profile"], "kernel": "k_field_far32c", "kernel_ms_per_pass": far_ms,
                     "launches_per_pass": launches, "kernel_ms_per_launch": far_ms / launches,
                     "flops_per_pass": far_flops, "frac_of_unpacked_fp32_peak": achieved / PEAK_FP64_TFLOPS,
                     "flops_executed_per_pass": far_flops_executed,
                     "pairs_by_path": cmx.astype(int).tolist(),
                     "pairs_by_path_keys": ["3-point cluster", "6-point record", "per-pair 6", "per-pair 7",
                                            "per-pair 12", "per-pair 33", "segment 4", "segment 16", "near",
                                            "per-pair nodes executed"]},
        "cpu_baseline": cpu,
        "tuning": None if tuning is None else {k: v["best"] for k, v in tuning.items()},
    }
    line.update(out)
    return line


def run_coax(ctx, args):
    """The 20,000-triangle coax: solve, matrix-vector bandwidth, field passes over 1 M points,
    64 voltage sets with 4 M points in uneven batches, and the 1,728-triangle cube."""
    from kassiopeia_amd import _lib
    from kassiopeia_amd import autotune as T
    from kassiopeia_amd import geometry as G
    from kassiopeia_amd import kemfield as K

    lib = ctx.lib
    mesh = G.coax_mesh()
    pts = G.points_annulus(1_000_000, 0.5, 1.0, 2.0, seed=2024)
    N, M = mesh.n, pts.shape[0]
    cont = K.KSurfaceContainer(ctx, mesh)
    b = cont.rhs()
    out = {}
    # solve
    ctx.sync()
    t0 = time.perf_counter()
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True)
    ctx.sync()
    t_asm = time.perf_counter() - t0
    x64, st64, _ = A.solve(b, tolerance=1e-10)
    xmx, stmx, _ = A.solve(b, tolerance=1e-10, precision="mixed")
    log(f"coax20k: assembly {t_asm*1e3:.1f} ms; GMRES fp64 {st64.iterations} it {st64.seconds*1e3:.1f} ms; "
        f"mixed {stmx.iterations} it / {stmx.refinements} ref {stmx.seconds*1e3:.1f} ms")
    out.update(coax20k_assembly_ms=t_asm * 1e3, coax20k_solve_ms_fp64=st64.seconds * 1e3,
               coax20k_solve_ms_mixed=stmx.seconds * 1e3, coax20k_gmres_iterations_fp64=st64.iterations,
               coax20k_gmres_iterations_mixed=stmx.iterations)
    # single-RHS matrix-vector product (HBM)
    v = np.random.default_rng(5).normal(size=N)
    A.Multiply(v)
    _timing(ctx, True)
    for _ in range(10):
        A.Multiply(v)
    kt, kn = _kernel_time(ctx)
    _timing(ctx, False)
    ms_mv = kt / max(kn, 1)
    out.update(coax20k_matvec_ms_fp64=ms_mv, coax20k_matvec_gbs_fp64=8.0 * N * N / (ms_mv * 1e-3) / 1e9)
    # field passes (tuned element ranges of the coax's own points)
    if not args.no_autotune:
        T.tune_field(cont, x64, pts, "mixed", repeats=2, candidates=(1280, 2560, 5120, 10240))
        T.tune_field(cont, x64, pts, "fp64", repeats=1, candidates=(1280, 2560, 5120))
    X = [K.DeviceArray(ctx, pts[:, k]) for k in range(3)]
    S = K.DeviceArray(ctx, x64)
    O = [K.DeviceArray.empty(ctx, M) for _ in range(4)]
    Om = [K.DeviceArray.empty(ctx, M) for _ in range(4)]

    def timed(prec, outs, reps):
        ctx.check(lib.kem_field_eval_device(ctx.h, cont.h, S.ptr, M, X[0].ptr, X[1].ptr, X[2].ptr, prec,
                                            outs[0].ptr, outs[1].ptr, outs[2].ptr, outs[3].ptr))
        ctx.sync()
        _timing(ctx, True)
        t0 = time.perf_counter()
        for _ in range(reps):
            ctx.check(lib.kem_field_eval_device(ctx.h, cont.h, S.ptr, M, X[0].ptr, X[1].ptr, X[2].ptr, prec,
                                                outs[0].ptr, outs[1].ptr, outs[2].ptr, outs[3].ptr))
        ctx.sync()
        t = (time.perf_counter() - t0) * 1e3 / reps
        kt, kn = _kernel_time(ctx)
        _timing(ctx, False)
        return t, kt / max(kn, 1)

    ms64, kms64 = timed(_lib.KEM_FP64, O, 2)
    msmx, kmsmx = timed(_lib.KEM_MIXED, Om, 5)
    c64 = _census(ctx, cont, X, M, _lib.KEM_FP64)
    f64 = float((c64 * FLOPS_PER_PAIR * FAR).sum())
    _, fmx, _ = _census_paths(ctx, cont, X, M)
    phi = O[0].get()
    E = np.stack([O[k].get() for k in (1, 2, 3)], 1)
    err = _errors(E, phi, np.stack([Om[k].get() for k in (1, 2, 3)], 1), Om[0].get(), 18600.0)
    log(f"coax20k field: fp64 {ms64:.1f} ms (far {kms64:.1f} ms, {f64 / kms64 / 1e9:.1f} TFLOP/s), mixed "
        f"{msmx:.2f} ms (far {kmsmx:.2f} ms, {fmx / kmsmx / 1e9:.1f} TFLOP/s), {ms64 / msmx:.2f}x; error "
        f"l2 {err['E_l2']:.1e} max/rms {err['E_max_over_rms']:.1e} phi/V {err['phi_max_over_V']:.1e}")
    out.update(coax20k_field_ms_fp64=ms64, coax20k_field_ms_mixed=msmx,
               coax20k_pairs_per_s_mixed=N * M / msmx * 1e3, coax20k_pairs_per_s_fp64=N * M / ms64 * 1e3,
               coax20k_field_speedup=ms64 / msmx, coax20k_far_frac_fp64=f64 / kms64 / 1e9 / PEAK_FP64_TFLOPS,
               coax20k_far_frac_mixed=fmx / kmsmx / 1e9 / PEAK_FP32_TFLOPS, coax20k_mixed_err_E_l2=err["E_l2"],
               coax20k_mixed_err_E_max_over_rms=err["E_max_over_rms"],
               coax20k_mixed_err_phi_max_over_V=err["phi_max_over_V"],
               coax20k_pipeline_speedup=(t_asm + st64.seconds + ms64 * 1e-3) / (t_asm + stmx.seconds + msmx * 1e-3))
    for a in X + O + Om + [S]:
        a.close()

    # 64 voltage sets on 16 axial ring groups, and their field over 4 M points in uneven batches
    groups = G.coax_ring_groups(mesh, 16)
    V = np.random.default_rng(64).uniform(-1000.0, 0.0, size=(16, 64))
    B = V[groups, :]
    A.solve_batch(B, tolerance=1e-10, precision="mixed")  # warm
    Xb, res, stb, _ = A.solve_batch(B, tolerance=1e-10, precision="mixed")
    log(f"64 voltage sets: lockstep mixed GMRES {stb.seconds*1e3:.1f} ms ({stb.iterations} it, "
        f"{stb.refinements} ref, worst residual {res.max():.2e})")
    out.update(batch64_solve_ms=stb.seconds * 1e3, batch64_iterations=stb.iterations,
               batch64_refinements=stb.refinements, batch64_worst_residual=float(res.max()),
               batch64_speedup_vs_64_single=64 * stmx.seconds / stb.seconds)
    # block products of 64 right-hand sides on MFMA: FP32 matrix (k_bmm32) and BF16 matrix
    # (k_bmm16), kernel time from HIP events; and the 64 sets in lockstep on the BF16 matrix
    Ab = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=False, bf16=True)
    Xr = np.random.default_rng(9).normal(size=(N, 64))
    # 16 right-hand sides on the FP32 matrix and 16 / 64 on the FP64 matrix: the streamed
    # 16 x 16-tile products (k_bmmt), the matrix read once per call
    for prec, key, bpe, k in (("mixed", "fp32", 4, 16), ("fp64", "fp64", 8, 16), ("fp64", "fp64", 8, 64)):
        A.MultiplyBatch(Xr[:, :k], prec)
        _timing(ctx, True)
        for _ in range(5):
            A.MultiplyBatch(Xr[:, :k], prec)
        kt, kn = _kernel_time(ctx)
        _timing(ctx, False)
        ms = kt / max(kn, 1)
        gbs = bpe * N * N / (ms * 1e-3) / 1e9
        tfl = 2.0 * N * N * k / (ms * 1e-3) / 1e12
        log(f"block product, {k} right-hand sides, {key} matrix (16 x 16 tiles): {ms:.3f} ms = {gbs:.0f} GB/s of "
            f"matrix, {tfl:.1f} TFLOP/s")
        out.update({f"coax20k_bmm{k}_ms_{key}": ms, f"coax20k_bmm{k}_gbs_{key}": gbs,
                    f"coax20k_bmm{k}_tflops_{key}": tfl})
    for mat, prec, key, bpe in ((A, "mixed", "fp32", 4), (Ab, "bf16", "bf16", 2)):
        if not args.no_autotune:
            # the split-K target (and for FP32 the wavefront tile) tuned per matrix precision
            tb = T.tune_block_product(mat, nrhs=64, precision=prec)
            out[f"coax20k_bmm64_tuned_{key}"] = {tb["knob"]: tb["best"]}
        mat.MultiplyBatch(Xr, prec)
        _timing(ctx, True)
        for _ in range(5):
            mat.MultiplyBatch(Xr, prec)
        kt, kn = _kernel_time(ctx)
        _timing(ctx, False)
        ms = kt / max(kn, 1)
        gbs = bpe * N * N / (ms * 1e-3) / 1e9
        tfl = 2.0 * N * N * 64 / (ms * 1e-3) / 1e12
        log(f"block product, 64 right-hand sides, {key} matrix: {ms:.3f} ms = {gbs:.0f} GB/s of matrix, "
            f"{tfl:.1f} TFLOP/s")
        out.update({f"coax20k_bmm64_ms_{key}": ms, f"coax20k_bmm64_gbs_{key}": gbs,
                    f"coax20k_bmm64_tflops_{key}": tfl, f"coax20k_bmm64_frac_hbm_{key}": gbs / PEAK_HBM_GBS})
    Ab.solve_batch(B, tolerance=1e-10, precision="bf16")  # warm
    Xb16, res16, st16, _ = Ab.solve_batch(B, tolerance=1e-10, precision="bf16")
    xs16, sts16, _ = Ab.solve(b, tolerance=1e-10, precision="bf16")
    log(f"64 voltage sets, lockstep BF16 GMRES: {st16.seconds*1e3:.1f} ms ({st16.iterations} it, "
        f"{st16.refinements} ref, worst {res16.max():.2e}); single right-hand side BF16 {sts16.seconds*1e3:.1f} ms "
        f"({sts16.iterations} it, {sts16.refinements} ref)")
    out.update(batch64_solve_ms_bf16=st16.seconds * 1e3, batch64_iterations_bf16=st16.iterations,
               batch64_refinements_bf16=st16.refinements, batch64_worst_residual_bf16=float(res16.max()),
               coax20k_solve_ms_bf16=sts16.seconds * 1e3, coax20k_gmres_iterations_bf16=sts16.iterations)
    Ab.close()
    # Robin Hood (one correction per step), to 1e-8 or 3 s
    xr, strh, _ = A.solve(b, method="robin_hood", tolerance=1e-8, time_limit_s=3.0)
    log(f"Robin Hood: {strh.sub_iterations} corrections in {strh.seconds:.2f} s "
        f"({strh.sub_iterations / strh.seconds:.3e}/s), residual {strh.residual:.2e}")
    out.update(coax20k_robin_hood_s=strh.seconds, coax20k_robin_hood_corrections=strh.sub_iterations,
               coax20k_robin_hood_corrections_per_s=strh.sub_iterations / strh.seconds,
               coax20k_robin_hood_residual=strh.residual)
    npb = 4_000_000
    rng = np.random.default_rng(4096)
    sizes = []
    while sum(sizes) < npb:
        sizes.append(int(np.exp(rng.uniform(np.log(1000), np.log(256000)))))
    sizes[-1] -= sum(sizes) - npb
    Pb = G.points_annulus(npb, 0.5, 1.0, 2.0, seed=77)
    sidx = np.repeat(np.arange(len(sizes)) % 64, sizes).astype(np.int32)
    Sd = K.DeviceArray(ctx, np.asfortranarray(Xb).T)
    Id = K.DeviceArray(ctx, sidx)
    Xd = [K.DeviceArray(ctx, Pb[:, k]) for k in range(3)]
    Od = [K.DeviceArray.empty(ctx, npb) for _ in range(4)]

    def run_sets():
        ctx.check(lib.kem_field_eval_sets(ctx.h, cont.h, 64, Sd.ptr, Id.ptr, npb, Xd[0].ptr, Xd[1].ptr, Xd[2].ptr,
                                          _lib.KEM_MIXED, Od[0].ptr, Od[1].ptr, Od[2].ptr, Od[3].ptr))

    run_sets()
    ctx.sync()
    t0 = time.perf_counter()
    run_sets()
    ctx.sync()
    tb = time.perf_counter() - t0
    log(f"uneven batches: {len(sizes)} batches, 64 sets, {npb} points: mixed {tb*1e3:.1f} ms")
    out.update(batches_count=len(sizes), batches_field_ms_mixed=tb * 1e3, batches_pairs_per_s_mixed=N * npb / tb)
    for a in Xd + Od + [Sd, Id]:
        a.close()
    A.close()
    cont.close()

    # the 1,728-triangle cube (the reference's CPU-runnable case)
    cube = G.cube_mesh(12)
    ccont = K.KSurfaceContainer(ctx, cube)
    ctx.sync()
    t0 = time.perf_counter()
    Ac = K.KBoundaryIntegralMatrix(ccont)
    ctx.sync()
    t_casm = time.perf_counter() - t0
    xc, stc, _ = Ac.solve(ccont.rhs(), tolerance=1e-10)
    xrh, strc, _ = Ac.solve(ccont.rhs(), method="robin_hood", tolerance=1e-8)
    log(f"cube 1728 Robin Hood: {strc.sub_iterations} corrections in {strc.seconds*1e3:.1f} ms, residual "
        f"{strc.residual:.1e}")
    out.update(cube1728_robin_hood_ms=strc.seconds * 1e3, cube1728_robin_hood_corrections=strc.sub_iterations)
    Ac.close()
    cap = float((cube.areas() * xc).sum() / (4 * np.pi * 8.8541878128e-12))
    cpts = G.points_shell(4096, 0.9, 2.0, seed=7)
    fs = K.KIntegratingElectrostaticFieldSolver(ccont, xc)
    fs.ElectricFieldAndPotential(cpts)
    t0 = time.perf_counter()
    fs.ElectricFieldAndPotential(cpts)
    t_cfield = time.perf_counter() - t0
    ccont.close()
    log(f"cube 1728: assembly {t_casm*1e3:.2f} ms, GMRES {stc.iterations} it {stc.seconds*1e3:.2f} ms, "
        f"C/(4 pi eps0) {cap:.9f}, field 4096 points {t_cfield*1e3:.2f} ms")
    out.update(cube1728_assembly_ms=t_casm * 1e3, cube1728_gmres_ms=stc.seconds * 1e3,
               cube1728_gmres_iterations=stc.iterations, cube1728_capacitance=cap, cube1728_field_ms=t_cfield * 1e3)
    return out


def copy_calibration():
    """One device-to-device copy of 1 GiB (the HIP runtime's copy kernel): the reference
    dispatch whose rocprofv3 FETCH_SIZE / WRITE_SIZE readings scale those of the field kernels
    to bytes."""
    from kassiopeia_amd import kemfield as K

    ctx = K.Context(0)
    n = 1 << 30
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    ctx.check(ctx.lib.kem_dev_alloc(ctx.h, n, ctypes.byref(a)))
    ctx.check(ctx.lib.kem_dev_alloc(ctx.h, n, ctypes.byref(b)))
    hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime the library is bound to
    hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    hip.hipMemcpy.restype = ctypes.c_int
    rc = hip.hipMemcpy(b, a, n, 3)  # hipMemcpyDeviceToDevice
    ctx.check(ctx.lib.kem_dev_free(ctx.h, a))
    ctx.check(ctx.lib.kem_dev_free(ctx.h, b))
    ctx.close()
    if rc != 0:
        raise RuntimeError(f"hipMemcpy failed ({rc})")
    print(json.dumps({"copied_bytes": n}), flush=True)


if __name__ == "__main__":
    main()
