#!/usr/bin/env python3
"""Benchmark of the electrostatic BEM hot path on one MI355X.

Workload (the benchmark configuration with a 20,000-triangle mesh and 1 M field points):
two coaxial cylinders (r = 0.5 m and 1.0 m, length 4 m) with annular end caps, 20,000
triangles with 1 % seeded tangential jitter; inner electrode -18.6 kV, outer electrode and
caps 0 V.  Dense collocation matrix assembly (FP64 and FP32 storage), GMRES to 1e-10 relative
residual in FP64 and with FP32-matrix inner solves + FP64 refinement, then potential + field
at 1,000,000 points uniform in volume in the annular gap 0.5 < r < 1.0, |z| < 2.

A step is one field-evaluation pass over the 1 M points with the FP64 charge densities
(inputs resident in HBM); `value` is point-element pairs per second of that pass (the
mixed-precision pass when its error against the FP64 pass is within 1e-6).  The solve
(assembly + GMRES) is timed separately and reported beside it, as is the multi-right-hand-side
case on the same mesh: 16 axial ring groups, 64 voltage sets U[-1000, 0] V, lockstep GMRES
on the FP32 matrix (MFMA block products) with FP64 refinement, and the field of those 64
solutions over 4 M points in uneven batches (sizes log-uniform in 1k-256k) in one pass.

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
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 vector peak
PEAK_FP32_TFLOPS = 157.3  # MI355X FP32 vector peak (packed)
PEAK_HBM_GBS = 8000.0
# the oracle's per-core field rate over the compiled reference's, measured on one core of the
# same host type: 3.68 / 2.95 M point-element pairs/s (1,728-triangle cube, 4,096 points)
CPU_ORACLE_OVER_REFERENCE = 3.68 / 2.95

# HBM bytes per launch of the far-field kernels on the default workload (20,000 triangles,
# 1,000,000 points): rocprofv3 --pmc FETCH_SIZE and WRITE_SIZE of one dispatch, scaled by the
# readings of a 1 GiB device copy (bench.py --copy-calibration: FETCH_SIZE x 1.99990, WRITE_SIZE
# x 1.0), both element ranges at their default of 2,560 elements.
TRAFFIC_BYTES = {"k_field_far32c": (149889.4375 * 1.99990 + 250000.0) * 1024,
                 "k_field_far": (1223041.5 * 1.99990 + 250000.0) * 1024}


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=5)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--points", type=int, default=1_000_000)
    ap.add_argument("--cpu-points", type=int, default=0,
                    help="points of the bounded CPU-baseline sample (x 20,000 elements); 0: 4000 per thread")
    ap.add_argument("--cpu-threads", type=int, default=0,
                    help="OpenMP threads of the CPU baseline; 0: the host's cores, at most 16")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-solve", action="store_true")
    ap.add_argument("--no-batch", action="store_true", help="skip the multi-right-hand-side section")
    ap.add_argument("--batch-points", type=int, default=4_000_000)
    ap.add_argument("--workload", choices=["coax20k", "large"], default="coax20k",
                    help="large: 100,000-triangle ring bore, 16 voltage sets, 10 M points, adjoint")
    ap.add_argument("--no-autotune", action="store_true",
                    help="keep the library's default launch shapes instead of tuning them on the workload")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only copy 1 GiB device to device (reference kernel for the rocprofv3 byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py: the library runs on one GPU; --gpus {args.gpus} is not supported", file=sys.stderr)
        sys.exit(2)
    if args.copy_calibration:
        return copy_calibration()
    if args.workload == "large":
        return run_large(args)

    from kassiopeia_amd import _lib
    from kassiopeia_amd import geometry as G
    from kassiopeia_amd import kemfield as K

    mesh = G.coax_mesh()
    pts = G.points_annulus(args.points, 0.5, 1.0, 2.0, seed=2024)
    N, M = mesh.n, pts.shape[0]
    log(f"mesh {N} triangles, {M} points")

    ctx = K.Context(0)
    lib = ctx.lib
    cont = K.KSurfaceContainer(ctx, mesh)
    b = cont.rhs()

    def sync():
        ctx.check(lib.kem_ctx_sync(ctx.h))

    out = {}
    A = None
    # ---------------------------------------------------------------- solve
    if not args.no_solve:
        t0 = time.perf_counter()
        A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True)
        sync()
        t_asm = time.perf_counter() - t0
        x64, st64, _ = A.solve(b, tolerance=1e-10)
        xmx, stmx, _ = A.solve(b, tolerance=1e-10, precision="mixed")
        x64, st64, _ = A.solve(b, tolerance=1e-10)  # second run: warm
        xmx, stmx, _ = A.solve(b, tolerance=1e-10, precision="mixed")
        sig_err = float(np.linalg.norm(xmx - x64) / np.linalg.norm(x64))
        log(f"assembly {t_asm*1e3:.1f} ms; GMRES fp64 {st64.iterations} it {st64.seconds*1e3:.1f} ms "
            f"res {st64.true_residual:.2e}; mixed {stmx.iterations} it / {stmx.refinements} ref "
            f"{stmx.seconds*1e3:.1f} ms res {stmx.true_residual:.2e}; sigma rel diff {sig_err:.2e}")
        out.update(assembly_ms=round(t_asm * 1e3, 3), solve_ms_fp64=round(st64.seconds * 1e3, 3),
                   solve_ms_mixed=round(stmx.seconds * 1e3, 3), gmres_iterations_fp64=st64.iterations,
                   gmres_iterations_mixed=stmx.iterations, refinements_mixed=stmx.refinements,
                   residual_fp64=st64.true_residual, residual_mixed=stmx.true_residual,
                   sigma_mixed_vs_fp64=sig_err)
        sigma = x64
    else:
        sigma = np.random.default_rng(3).normal(size=N) * 1e-7

    # ---------------------------------------------------------------- single-RHS matvec (HBM)
    if A is not None:
        v = np.random.default_rng(5).normal(size=N)
        mv = {}
        for name, Am in (("fp64", A),):
            Am.Multiply(v)
            ctx.check(lib.kem_ctx_set_kernel_timing(ctx.h, 1))
            for _ in range(10):
                Am.Multiply(v)
            kt, kn = ctypes.c_double(), ctypes.c_int()
            ctx.check(lib.kem_ctx_kernel_time(ctx.h, ctypes.byref(kt), ctypes.byref(kn)))
            ctx.check(lib.kem_ctx_set_kernel_timing(ctx.h, 0))
            ms_mv = kt.value / max(kn.value, 1)
            mv[name] = (ms_mv, 8.0 * N * N / (ms_mv * 1e-3) / 1e9)
        A32 = K.KBoundaryIntegralMatrix(cont, fp64=False, fp32=True)
        A32.Multiply(v)
        ctx.check(lib.kem_ctx_set_kernel_timing(ctx.h, 1))
        for _ in range(10):
            A32.Multiply(v)
        kt, kn = ctypes.c_double(), ctypes.c_int()
        ctx.check(lib.kem_ctx_kernel_time(ctx.h, ctypes.byref(kt), ctypes.byref(kn)))
        ctx.check(lib.kem_ctx_set_kernel_timing(ctx.h, 0))
        A32.close()
        ms_mv = kt.value / max(kn.value, 1)
        mv["fp32"] = (ms_mv, 4.0 * N * N / (ms_mv * 1e-3) / 1e9)
        log("matvec (kernel time, HBM bytes of the matrix): " +
            ", ".join(f"{k} {t:.3f} ms = {g:.0f} GB/s ({g / PEAK_HBM_GBS * 100:.0f} % of 8 TB/s)"
                      for k, (t, g) in mv.items()))
        out.update(matvec_ms_fp64=mv["fp64"][0], matvec_gbs_fp64=mv["fp64"][1], matvec_ms_fp32=mv["fp32"][0],
                   matvec_gbs_fp32=mv["fp32"][1])

    # ---------------------------------------------------------------- launch shapes
    tuning = None
    if not args.no_autotune:
        from kassiopeia_amd import autotune as T

        t0 = time.perf_counter()
        tuning = {"far_range_mixed": T.tune_field(cont, sigma, pts, "mixed", repeats=2),
                  "far_range_fp64": T.tune_field(cont, sigma, pts, "fp64", repeats=1),
                  "near_blocks": T.tune_near(cont, sigma, pts, repeats=2)}
        if A is not None:
            tuning["gemv_blocks"] = T.tune_matvec(A, repeats=5)
            tuning["bmm_blocks"] = T.tune_block_product(A, nrhs=64, precision="mixed", repeats=3)
        log(f"autotune ({time.perf_counter() - t0:.1f} s): " +
            ", ".join(f"{k} {v['best']}" for k, v in tuning.items()))

    # ---------------------------------------------------------------- field pass
    X = [K.DeviceArray(ctx, pts[:, k]) for k in range(3)]
    S = K.DeviceArray(ctx, sigma)
    O = [K.DeviceArray.empty(ctx, M) for _ in range(4)]
    Om = [K.DeviceArray.empty(ctx, M) for _ in range(4)]

    def run(precision, outs):
        ctx.check(lib.kem_field_eval_device(ctx.h, cont.h, S.ptr, M, X[0].ptr, X[1].ptr, X[2].ptr, precision,
                                            outs[0].ptr, outs[1].ptr, outs[2].ptr, outs[3].ptr))

    def census(precision):
        c = (ctypes.c_int64 * 12)()
        ctx.check(lib.kem_field_census(ctx.h, cont.h, M, X[0].ptr, X[1].ptr, X[2].ptr, precision, c))
        return np.array(c[:], dtype=np.float64)

    def timed(precision, outs):
        """(pass ms, far-kernel ms): steps back-to-back on the library's stream between two
        stream synchronisations; the far kernel from HIP events the library records around it."""
        for _ in range(args.warmup):
            run(precision, outs)
        sync()
        ctx.check(lib.kem_ctx_set_kernel_timing(ctx.h, 1))
        t0 = time.perf_counter()
        for _ in range(args.steps):
            run(precision, outs)
        sync()
        t = time.perf_counter() - t0
        kt, kn = ctypes.c_double(), ctypes.c_int()
        ctx.check(lib.kem_ctx_kernel_time(ctx.h, ctypes.byref(kt), ctypes.byref(kn)))
        ctx.check(lib.kem_ctx_set_kernel_timing(ctx.h, 0))
        return t * 1e3 / args.steps, kt.value / max(kn.value, 1)

    pairs = float(N) * M
    c64 = census(_lib.KEM_FP64)
    cmx = census(_lib.KEM_MIXED)
    flops64 = float((c64 * FLOPS_PER_PAIR).sum())
    # the far-field kernel evaluates every pair but the closed-form ones
    far = np.ones(12)
    far[[3, 7, 10]] = 0.
    farflops64 = float((c64 * FLOPS_PER_PAIR * far).sum())
    farflopsmx = float((cmx * FLOPS_PER_PAIR * far).sum())
    log(f"regime census fp64 (tri 7/12/33/closed): {c64[:4].astype(int).tolist()}; "
        f"mixed: {cmx[:4].astype(int).tolist()} + 6-point {int(cmx[11])}")
    ms64, kms64 = timed(_lib.KEM_FP64, O)
    msmx, kmsmx = timed(_lib.KEM_MIXED, Om)
    # error of the mixed result against the FP64 path: norm-wise over the whole point set, the
    # largest deviation relative to the RMS field, and the potential relative to the electrode
    # voltage scale; the per-point relative field error is reported beside them (it is
    # amplified where |E| nearly vanishes)
    phi = O[0].get()
    E = np.stack([O[k].get() for k in (1, 2, 3)], 1)
    phim = Om[0].get()
    Emx = np.stack([Om[k].get() for k in (1, 2, 3)], 1)
    dE = Emx - E
    Ed, En = np.linalg.norm(dE, axis=1), np.linalg.norm(E, axis=1)
    err_E_l2 = float(np.linalg.norm(dE) / np.linalg.norm(E))
    err_E_rms = float(Ed.max() / np.sqrt((E ** 2).sum(1).mean()))
    errE = Ed / En
    vscale = float(np.abs(mesh.bc_value).max()) or 1.0
    err_p_max = float((np.abs(phim - phi) / vscale).max())
    err_E_max, err_E_p99 = float(errE.max()), float(np.quantile(errE, 0.99))
    log(f"field pass fp64 {ms64:.2f} ms (far kernel {kms64:.2f} ms, {pairs / ms64 * 1e3:.4e} pairs/s, "
        f"{flops64 / ms64 / 1e9:.2f} TFLOP/s); mixed {msmx:.2f} ms (far kernel {kmsmx:.2f} ms, "
        f"{pairs / msmx * 1e3:.4e} pairs/s), speed-up {ms64 / msmx:.2f}x; mixed vs fp64: "
        f"|dE| l2 {err_E_l2:.2e}, max|dE|/rms|E| {err_E_rms:.2e}, |dphi|/V max {err_p_max:.2e}; "
        f"per point |dE|/|E| max {err_E_max:.2e} p99 {err_E_p99:.2e}")
    mixed_ok = err_E_l2 <= 1e-6 and err_E_rms <= 1e-6 and err_p_max <= 1e-6
    ms = msmx if mixed_ok else ms64
    kms = kmsmx if mixed_ok else kms64
    kflops = farflopsmx if mixed_ok else farflops64
    value = pairs / (ms * 1e-3)
    achieved = kflops / (kms * 1e-3) / 1e12
    kname = "k_field_far32c" if mixed_ok else "k_field_far"
    # the measured traffic holds for the profiled launch: 20,000 triangles, 1,000,000 points,
    # one element range of 2,500 per XCD (element-range size >= 2,500)
    from kassiopeia_amd import autotune as T

    rng_used = T.get(ctx, "far_range_mixed" if mixed_ok else "far_range_fp64")
    default_workload = N == 20000 and M == 1_000_000 and rng_used >= 2500
    out.update(field_ms_fp64=ms64, field_ms_mixed=msmx, far_kernel_ms_fp64=kms64, far_kernel_ms_mixed=kmsmx,
               pairs_per_s_fp64=pairs / ms64 * 1e3,
               pairs_per_s_mixed=pairs / msmx * 1e3, speedup_mixed_vs_fp64=ms64 / msmx,
               mixed_err_E_l2=err_E_l2, mixed_err_E_max_over_rms=err_E_rms, mixed_err_phi_max=err_p_max,
               mixed_err_E_pointwise_max=err_E_max, mixed_err_E_pointwise_p99=err_E_p99,
               mixed_within_1e6=mixed_ok, fp64_pass_tflops=flops64 / ms64 / 1e9,
               fp64_far_kernel_frac=farflops64 / kms64 / 1e9 / PEAK_FP64_TFLOPS)
    for arr in X + O + Om + [S]:
        arr.close()

    # ---------------------------------------------------------------- multi-RHS (64 voltage sets)
    if not args.no_batch and A is not None:
        groups = G.coax_ring_groups(mesh, 16)
        V = np.random.default_rng(64).uniform(-1000.0, 0.0, size=(16, 64))
        B = V[groups, :]
        Xb, res, stb, rcb = A.solve_batch(B, tolerance=1e-10, precision="mixed")  # warm
        Xb, res, stb, rcb = A.solve_batch(B, tolerance=1e-10, precision="mixed")
        x1, st1, _ = A.solve(B[:, 0], tolerance=1e-10, precision="mixed")
        log(f"64 voltage sets: lockstep mixed GMRES {stb.seconds*1e3:.1f} ms ({stb.iterations} it, "
            f"{stb.refinements} ref, worst residual {res.max():.2e}); one set alone {st1.seconds*1e3:.1f} ms")
        out.update(batch64_solve_ms=round(stb.seconds * 1e3, 3), batch64_iterations=stb.iterations,
                   batch64_refinements=stb.refinements, batch64_worst_residual=float(res.max()),
                   single_rhs_mixed_solve_ms=round(st1.seconds * 1e3, 3),
                   batch64_speedup_vs_64_single=round(64 * st1.seconds / stb.seconds, 2))
        # uneven batches: sizes log-uniform in [1k, 256k] until 4 M points, batch b uses set b % 64
        rng = np.random.default_rng(4096)
        sizes = []
        while sum(sizes) < args.batch_points:
            sizes.append(int(np.exp(rng.uniform(np.log(1000), np.log(256000)))))
        sizes[-1] -= sum(sizes) - args.batch_points
        Pb = G.points_annulus(args.batch_points, 0.5, 1.0, 2.0, seed=77)
        sidx = np.repeat(np.arange(len(sizes)) % 64, sizes).astype(np.int32)
        Sd = K.DeviceArray(ctx, np.asfortranarray(Xb).T)
        Id = K.DeviceArray(ctx, sidx)
        Xd = [K.DeviceArray(ctx, Pb[:, k]) for k in range(3)]
        Od = [K.DeviceArray.empty(ctx, args.batch_points) for _ in range(4)]

        def run_sets(prec):
            ctx.check(lib.kem_field_eval_sets(ctx.h, cont.h, 64, Sd.ptr, Id.ptr, args.batch_points, Xd[0].ptr,
                                              Xd[1].ptr, Xd[2].ptr, prec, Od[0].ptr, Od[1].ptr, Od[2].ptr,
                                              Od[3].ptr))

        run_sets(_lib.KEM_MIXED)
        sync()
        t0 = time.perf_counter()
        run_sets(_lib.KEM_MIXED)
        sync()
        tb = time.perf_counter() - t0
        bpairs = float(N) * args.batch_points
        log(f"uneven batches: {len(sizes)} batches ({min(sizes)}..{max(sizes)} points), 64 sets, "
            f"{args.batch_points} points: mixed {tb*1e3:.1f} ms = {bpairs / tb:.4e} pairs/s")
        out.update(batches_count=len(sizes), batches_points=args.batch_points,
                   batches_field_ms_mixed=round(tb * 1e3, 3), batches_pairs_per_s_mixed=bpairs / tb)
        for arr in Xd + Od + [Sd, Id]:
            arr.close()
    if A is not None:
        A.close()

    # ---------------------------------------------------------------- the 1,728-triangle cube
    # (the reference's CPU-runnable case: FP64 assembly, GMRES 1e-10, capacitance, field at
    # 4,096 points in the shell 0.9 <= r <= 2 m)
    cube = G.cube_mesh(12)
    ccont = K.KSurfaceContainer(ctx, cube)
    t0 = time.perf_counter()
    Ac = K.KBoundaryIntegralMatrix(ccont)
    sync()
    t_casm = time.perf_counter() - t0
    xc, stc, _ = Ac.solve(ccont.rhs(), tolerance=1e-10)
    xc, stc, _ = Ac.solve(ccont.rhs(), tolerance=1e-10)  # warm
    Ac.close()
    cap = float((cube.areas() * xc).sum() / (4 * np.pi * 8.8541878128e-12))
    cpts = G.points_shell(4096, 0.9, 2.0, seed=7)
    fs = K.KIntegratingElectrostaticFieldSolver(ccont, xc)
    fs.ElectricFieldAndPotential(cpts)
    t0 = time.perf_counter()
    fs.ElectricFieldAndPotential(cpts)
    t_cfield = time.perf_counter() - t0
    ccont.close()
    log(f"cube 1728: assembly {t_casm*1e3:.2f} ms, GMRES {stc.iterations} it {stc.seconds*1e3:.2f} ms "
        f"(res {stc.true_residual:.2e}), C/(4 pi eps0) {cap:.9f}, field 4096 points {t_cfield*1e3:.2f} ms "
        f"(host arrays)")
    out.update(cube1728_assembly_ms=t_casm * 1e3, cube1728_gmres_ms=stc.seconds * 1e3,
               cube1728_gmres_iterations=stc.iterations, cube1728_capacitance=cap,
               cube1728_field_ms=t_cfield * 1e3)

    # ---------------------------------------------------------------- CPU baseline + parity
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as OR

        threads = args.cpu_threads or max(1, min(16, os.cpu_count() or 1))
        k = min(args.cpu_points or 4000 * threads, M)
        t0 = time.perf_counter()
        phr, Er = OR.field(mesh.shapes, mesh.recs, sigma, pts[:k], threads=threads)
        tc = time.perf_counter() - t0
        cpu = {"value": k * N / tc, "unit": "point-element pairs/s", "cores": threads, "kind": "port",
               "sample": f"{k} points x {N} triangles of the same workload, FP64 oracle (the reference's "
                         f"integrators restated in C), {threads} OpenMP threads, {tc:.1f} s",
               # per core the oracle's field sum runs 1.25x the compiled reference's (3.68 vs
               # 2.95 M pairs/s on the 1,728-triangle cube, one core of the same host type)
               "reference_equivalent_value": k * N / tc / CPU_ORACLE_OVER_REFERENCE}
        # normalised by the largest |phi| and |E| of the sample (pointwise ratios are amplified
        # where phi or E nearly vanish)
        ep = float(np.max(np.abs(phi[:k] - phr)) / np.max(np.abs(phr)))
        ee = float(np.max(np.linalg.norm(E[:k] - Er, axis=1)) / np.max(np.linalg.norm(Er, axis=1)))
        out.update(parity_phi_vs_oracle=ep, parity_E_vs_oracle=ee,
                   gpu_mixed_vs_cpu_speedup=(pairs / msmx * 1e3) / cpu["value"])
        log(f"cpu baseline {cpu['value']:.3e} pairs/s on {threads} threads; parity phi {ep:.2e} E {ee:.2e}")

    line = {
        "metric": "point-element field evals/s + solve time, mixed vs FP64 at 1e-6 rel. error",
        "value": value,
        "unit": "point-element pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64" if mixed_ok else "f64",
        "data": "synthetic",
        "config": {"workload": f"coaxial cylinders + annular caps, {N} triangles, GMRES 1e-10, "
                               f"{M} field points in the gap; value = "
                               + ("mixed-precision field pass (FP32 far field, FP64 near field and sums), "
                                  "error vs FP64 <= 1e-6" if mixed_ok else "FP64 field pass"),
                   "elements": N, "points": M},
        # the far-field kernel (the dominant one) is bound by the vector ALU (no matrix-shaped
        # work); the peak is MI355X's published figure of the arithmetic type: FP64 78.6 TFLOP/s,
        # FP32 157.3 TFLOP/s (the FP32 figure needs packed FP32, which the library's kernels do not use).
        # achieved = census FLOPs of the far pairs / far-kernel time from HIP events on the
        # library's stream over the timed steps
        "roofline": {"bound": "mfma", "compute": "vector ALU (" + ("fp32" if mixed_ok else "fp64") + ")",
                     "achieved": achieved, "peak": PEAK_FP32_TFLOPS if mixed_ok else PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s",
                     "frac": achieved / (PEAK_FP32_TFLOPS if mixed_ok else PEAK_FP64_TFLOPS),
                     "traffic": TRAFFIC_BYTES[kname] if default_workload else None,
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE + WRITE_SIZE, copy-calibrated)",
                     "kernel": kname,
                     "kernel_ms": kms, "flops_per_launch": kflops,
                     "pairs_by_regime": (cmx if mixed_ok else c64).astype(int).tolist()},
        "cpu_baseline": cpu,
        "tuning": None if tuning is None else {k: v["best"] for k, v in tuning.items()},
    }
    line.update(out)
    print(json.dumps(line), flush=True)


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


def run_large(args):
    """100,000-triangle case: 32 coaxial ring electrodes (r = 1.9 m) inside a grounded r = 2 m
    cylinder, voltages U[-20, 0] kV; FP64 pipeline (FP64 matrix, FP64 GMRES, FP64 field) against
    the mixed pipeline (FP32 + FP64 matrices, FP32 GMRES + FP64 refinement, FP32 far field) on
    10 M points uniform in the bore (r < 1.8 m, |z| < 3.5 m); 16 voltage sets in lockstep; the
    voltage + vertex adjoint of a 10k-point axial potential loss."""
    from kassiopeia_amd import _lib
    from kassiopeia_amd import geometry as G
    from kassiopeia_amd import kemfield as K

    rng = np.random.default_rng(20)
    volts = rng.uniform(-20000.0, 0.0, 32)
    mesh = G.ring_bore_mesh(volts=volts)
    N = mesh.n
    M = 10_000_000 if args.points == 1_000_000 else args.points
    pts = G.points_annulus(M, 0.0, 1.8, 3.5, seed=55)
    log(f"large: mesh {N} triangles, {M} points")
    ctx = K.Context(0)
    lib = ctx.lib
    cont = K.KSurfaceContainer(ctx, mesh)
    b = cont.rhs()

    def sync():
        ctx.check(lib.kem_ctx_sync(ctx.h))

    out = {}
    tuning = None
    if not args.no_autotune:
        # element-range sizes of the far pass, tuned on the whole point set with provisional
        # charge densities (the time per pair does not depend on their values)
        from kassiopeia_amd import autotune as T

        t0 = time.perf_counter()
        sig0 = np.ones(N) * 1e-9
        tuning = {p: T.tune_field(cont, sig0, pts, p, repeats=1, candidates=(0, 640, 2560, 10240))["best"]
                  for p in ("mixed", "fp64")}
        log(f"autotune ({time.perf_counter() - t0:.1f} s): far range mixed {tuning['mixed']}, fp64 {tuning['fp64']}")
    X = [K.DeviceArray(ctx, pts[:, k]) for k in range(3)]
    O = [K.DeviceArray.empty(ctx, M) for _ in range(4)]

    def field(sig, prec):
        S = K.DeviceArray(ctx, sig)
        sync()
        t0 = time.perf_counter()
        ctx.check(lib.kem_field_eval_device(ctx.h, cont.h, S.ptr, M, X[0].ptr, X[1].ptr, X[2].ptr, prec, O[0].ptr,
                                            O[1].ptr, O[2].ptr, O[3].ptr))
        sync()
        t = time.perf_counter() - t0
        S.close()
        return t, O[0].get(), np.stack([O[k].get() for k in (1, 2, 3)], 1)

    # FP64 pipeline
    t0 = time.perf_counter()
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=False)
    sync()
    t_asm64 = time.perf_counter() - t0
    x64, st64, _ = A.solve(b, tolerance=1e-10)
    A.close()
    log(f"FP64: assembly {t_asm64:.2f} s, GMRES {st64.iterations} it {st64.seconds:.2f} s")
    t_f64, phi64, E64 = field(x64, _lib.KEM_FP64)
    log(f"FP64 field {t_f64:.2f} s = {N * M / t_f64:.3e} pairs/s")
    # mixed pipeline
    t0 = time.perf_counter()
    A = K.KBoundaryIntegralMatrix(cont, fp64=True, fp32=True)
    sync()
    t_asm = time.perf_counter() - t0
    xmx, stmx, _ = A.solve(b, tolerance=1e-10, precision="mixed")
    log(f"mixed: assembly {t_asm:.2f} s, GMRES {stmx.iterations} it / {stmx.refinements} ref {stmx.seconds:.2f} s")
    t_fmx, phimx, Emx = field(xmx, _lib.KEM_MIXED)
    log(f"mixed field {t_fmx:.2f} s = {N * M / t_fmx:.3e} pairs/s")
    dE = Emx - E64
    err_E_l2 = float(np.linalg.norm(dE) / np.linalg.norm(E64))
    err_E_rms = float(np.linalg.norm(dE, axis=1).max() / np.sqrt((E64 ** 2).sum(1).mean()))
    err_p = float(np.abs(phimx - phi64).max() / np.abs(volts).max())
    t_plain = t_asm64 + st64.seconds + t_f64
    t_mixed = t_asm + stmx.seconds + t_fmx
    log(f"mixed vs FP64: |dE| l2 {err_E_l2:.2e}, max|dE|/rms|E| {err_E_rms:.2e}, |dphi|/V {err_p:.2e}; "
        f"pipeline {t_plain:.2f} s -> {t_mixed:.2f} s = {t_plain / t_mixed:.2f}x")
    # the two parts of that error: the mixed field pass with the FP64 charge densities, and the
    # FP64 field pass with the mixed solve's charge densities
    _, phif, Ef = field(x64, _lib.KEM_MIXED)
    dEf = np.linalg.norm(Ef - E64, axis=1)
    err_field_rms = float(dEf.max() / np.sqrt((E64 ** 2).sum(1).mean()))
    err_field_l2 = float(np.linalg.norm(Ef - E64) / np.linalg.norm(E64))
    _, phis, Es = field(xmx, _lib.KEM_FP64)
    err_solve_rms = float(np.linalg.norm(Es - E64, axis=1).max() / np.sqrt((E64 ** 2).sum(1).mean()))
    log(f"  from the field pass (same sigma): l2 {err_field_l2:.2e}, max/rms {err_field_rms:.2e}; from the "
        f"solve (FP64 field of the mixed sigma): max/rms {err_solve_rms:.2e}, sigma diff "
        f"{np.linalg.norm(xmx - x64) / np.linalg.norm(x64):.2e}")
    # 16 voltage sets
    groups = mesh.group
    V16 = rng.uniform(-20000.0, 0.0, size=(33, 16))
    V16[32, :] = 0.0
    Xb, res, stb, _ = A.solve_batch(V16[groups, :], tolerance=1e-10, precision="mixed")
    log(f"16 voltage sets: {stb.seconds:.2f} s ({stb.iterations} it, {stb.refinements} ref, worst {res.max():.1e})")
    # adjoint of L = mean (phi(axis) - target)^2 over 10k axial points (element-corner gradients)
    zax = np.linspace(-3.5, 3.5, 10000)
    pax = np.ascontiguousarray(np.stack([0 * zax, 0 * zax, zax], 1))
    ctx.set_integrator("numeric")
    Eax, phax = K.KIntegratingElectrostaticFieldSolver(cont, x64).ElectricFieldAndPotential(pax)
    g = 2.0 * (phax - (-10000.0)) / phax.size
    sync()
    t0 = time.perf_counter()
    dsig, dvf = np.zeros(N), np.zeros((N, 9))
    ctx.check(lib.kem_field_vjp(ctx.h, cont.h, K._ptr(x64), pax.shape[0], K._ptr(pax), K._ptr(g), K._ptr(dsig),
                                K._ptr(dvf)))
    t_fvjp = time.perf_counter() - t0
    o = _lib.SolveOpts()
    lib.kem_solve_opts_default(ctypes.byref(o))
    o.tolerance, o.precision = 1e-10, _lib.KEM_MIXED
    lam = np.zeros(N)
    sta = _lib.SolveStats()
    ctx.check(lib.kem_solve_transposed(A.h, ctypes.byref(o), K._ptr(dsig), K._ptr(lam), ctypes.byref(sta)))
    t0 = time.perf_counter()
    dvm = np.zeros((N, 9))
    ctx.check(lib.kem_assemble_vjp(ctx.h, cont.h, K._ptr(lam), K._ptr(x64), K._ptr(dvm)))
    t_mvjp = time.perf_counter() - t0
    dV = np.bincount(groups, weights=lam, minlength=33)
    log(f"adjoint: field VJP {t_fvjp:.2f} s, transposed solve {sta.seconds:.2f} s ({sta.iterations} it), "
        f"matrix VJP {t_mvjp:.2f} s; |dL/dV| max {np.abs(dV).max():.3e}")
    A.close()
    for arr in X + O:
        arr.close()
    # CPU baseline: the FP64 oracle on all host cores over a sample of the same points; the
    # CPU pipeline takes at least its field pass, so N M / rate bounds its time from below
    cpu = None
    if not args.no_cpu_baseline:
        from oracle import oracle as OR

        threads = args.cpu_threads or max(1, min(16, os.cpu_count() or 1))
        k = min(args.cpu_points or 64 * threads, M)
        t0 = time.perf_counter()
        OR.field(mesh.shapes, mesh.recs, x64, pts[:k], threads=threads)
        tc = time.perf_counter() - t0
        rate = k * N / tc
        cpu = {"value": rate, "unit": "point-element pairs/s", "cores": threads, "kind": "port",
               "sample": f"{k} points x {N} triangles of the same workload, FP64 oracle, {threads} OpenMP "
                         f"threads, {tc:.1f} s",
               "reference_equivalent_value": rate / CPU_ORACLE_OVER_REFERENCE}
        log(f"cpu baseline {rate:.3e} pairs/s on {threads} threads: field pass alone {N * M / rate:.0f} s, "
            f"{N * M / rate / t_mixed:.0f}x the mixed GPU pipeline")
    value = N * M / t_fmx
    line = {
        "metric": "point-element field evals/s + solve time, mixed vs FP64 at 1e-6 rel. error",
        "value": value, "unit": "point-element pairs/s", "n_gpus": 1, "steps": 1, "warmup": 0,
        "ms_per_step": t_fmx * 1e3, "higher_is_better": True, "scaling": "weak", "vs_baseline": None,
        "dtype": "f32+f64", "data": "synthetic",
        "config": {"workload": f"32 ring electrodes + grounded cylinder, {N} triangles, 16 voltage sets, "
                               f"{M} points in the bore; value = mixed-precision field pass", "elements": N,
                   "points": M},
        "assembly_s_fp64_only": t_asm64, "assembly_s_fp64_fp32": t_asm, "solve_s_fp64": st64.seconds,
        "solve_s_mixed": stmx.seconds, "gmres_iterations_fp64": st64.iterations,
        "gmres_iterations_mixed": stmx.iterations, "field_s_fp64": t_f64, "field_s_mixed": t_fmx,
        "pipeline_s_fp64": t_plain, "pipeline_s_mixed": t_mixed, "pipeline_speedup": t_plain / t_mixed,
        "mixed_err_E_l2": err_E_l2, "mixed_err_E_max_over_rms": err_E_rms, "mixed_err_phi_max": err_p,
        "mixed_field_only_err_E_l2": err_field_l2, "mixed_field_only_err_E_max_over_rms": err_field_rms,
        "mixed_solve_only_err_E_max_over_rms": err_solve_rms,
        "batch16_solve_s": stb.seconds, "batch16_worst_residual": float(res.max()),
        "adjoint_field_vjp_s": t_fvjp, "adjoint_solve_s": sta.seconds, "adjoint_matrix_vjp_s": t_mvjp,
        "tuning": tuning,
        "cpu_baseline": cpu,
        "cpu_field_s_estimate": None if cpu is None else N * M / cpu["value"],
        "mixed_pipeline_vs_cpu_field_lower_bound": None if cpu is None else N * M / cpu["value"] / t_mixed,
    }
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
