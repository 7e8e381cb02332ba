"""Autotuner of the library's launch shapes, measured on the device on the caller's own problem.

The C-ABI exposes these launch-shape knobs (``kem_ctx_set_tuning``):

* ``KEM_TUNE_FAR_RANGE_FP64`` / ``KEM_TUNE_FAR_RANGE_MIXED``: elements per element range of the
  far field pass over large point sets.  The elements split into 8 R ranges, R per XCD, so the
  range an XCD works on stays in its L2; 0 disables the split.
* ``KEM_TUNE_GEMV_BLOCKS``: target block count of the split-K matrix-vector product (GMRES,
  Robin Hood residuals, ``KSquareMatrix::Multiply``).
* ``KEM_TUNE_BMM_BLOCKS``: target block count of the multi-right-hand-side block products
  that load the matrix into registers (FP32 when n % 4 != 0, FP64 when n is odd).
* ``KEM_TUNE_BMM_STREAM_BLOCKS``: target workgroup count of the block products of 33-64
  right-hand sides that stream the matrix through LDS (BF16, and FP32 when n % 4 == 0).
* ``KEM_TUNE_BMM_TILE_BLOCKS``: largest workgroup count of the streamed block products of up
  to 32 right-hand sides (16 x 16 MFMA tiles: FP32 when n % 4 == 0, FP64 when n is even; one
  32-column tile: BF16).
* ``KEM_TUNE_NEAR_BLOCKS``: blocks of the near pass that runs beside the mixed far pass on a
  second stream (0: one per 256 points).  Its timed region is the whole field pass.
* ``KEM_TUNE_BMM_ROWS``: the wavefront tile of the FP32 MFMA block product, 64 rows (one float
  per lane and column) or 128 rows (a 16-byte load of four rows per lane feeding four MFMAs).
* ``KEM_TUNE_ROBIN_HOOD_GROUPS``: Robin Hood over several co-resident workgroups from 8,192
  unknowns on (0) or always one workgroup (1); both take the same corrections.

Each knob is swept over its candidates; every candidate runs once untimed and then ``repeats``
times, timed by the HIP events the library records around the kernels the knob shapes
(``kem_ctx_set_kernel_timing``).  The fastest candidate stays set on the context.  The results
are plain dicts that ``save`` / ``load`` keep as JSON and ``apply`` sets on another context.
"""
import ctypes
import json
import time

import numpy as np

from . import _lib
from . import kemfield as K

FAR_RANGE_CANDIDATES = (0, 640, 1280, 2560, 5120, 10240)
GEMV_CANDIDATES = (512, 1024, 2048, 4096, 8192)
BMM_CANDIDATES = (640, 1280, 2560, 5120, 10240)
STREAM_CANDIDATES = (160, 240, 320, 400, 480, 640, 960)
TILE_CANDIDATES = (256, 512, 768, 1024, 1536, 2048, 4096)
NEAR_CANDIDATES = (0, 128, 256, 384, 512)
BMM_ROW_CANDIDATES = (64, 128)

_KNOB = {"far_range_fp64": _lib.KEM_TUNE_FAR_RANGE_FP64, "far_range_mixed": _lib.KEM_TUNE_FAR_RANGE_MIXED,
         "gemv_blocks": _lib.KEM_TUNE_GEMV_BLOCKS, "bmm_blocks": _lib.KEM_TUNE_BMM_BLOCKS,
         "near_blocks": _lib.KEM_TUNE_NEAR_BLOCKS, "bmm_rows": _lib.KEM_TUNE_BMM_ROWS,
         "robin_hood_groups": _lib.KEM_TUNE_ROBIN_HOOD_GROUPS, "bmm_stream_blocks": _lib.KEM_TUNE_BMM_STREAM_BLOCKS,
         "bmm_tile_blocks": _lib.KEM_TUNE_BMM_TILE_BLOCKS}


def get(ctx: K.Context, name: str) -> int:
    v = ctypes.c_int64()
    ctx.check(ctx.lib.kem_ctx_get_tuning(ctx.h, _KNOB[name], ctypes.byref(v)))
    return int(v.value)


def set_knob(ctx: K.Context, name: str, value: int):
    ctx.check(ctx.lib.kem_ctx_set_tuning(ctx.h, _KNOB[name], int(value)))


def _sweep(ctx, name, candidates, run, repeats, wall=False):
    """ms per timed region for each candidate (wall: the whole call between two stream
    synchronisations); leaves the fastest set."""
    old = get(ctx, name)
    times = {}
    try:
        for v in candidates:
            set_knob(ctx, name, v)
            run()  # untimed: workspaces grow, caches warm
            ctx.check(ctx.lib.kem_ctx_sync(ctx.h))
            if wall:
                t0 = time.perf_counter()
                for _ in range(repeats):
                    run()
                ctx.check(ctx.lib.kem_ctx_sync(ctx.h))
                times[int(v)] = (time.perf_counter() - t0) * 1e3 / repeats
                continue
            ctx.check(ctx.lib.kem_ctx_set_kernel_timing(ctx.h, 1))
            for _ in range(repeats):
                run()
            t, n = ctypes.c_double(), ctypes.c_int()
            ctx.check(ctx.lib.kem_ctx_kernel_time(ctx.h, ctypes.byref(t), ctypes.byref(n)))
            ctx.check(ctx.lib.kem_ctx_set_kernel_timing(ctx.h, 0))
            times[int(v)] = t.value / max(n.value, 1)
    except BaseException:
        set_knob(ctx, name, old)
        raise
    best = min(times, key=times.get)
    set_knob(ctx, name, best)
    return {"best": best, "ms": times}


def tune_field(container: K.KSurfaceContainer, sigma, points, precision="mixed", repeats=2,
               candidates=FAR_RANGE_CANDIDATES):
    """Element-range size of the far pass for this container, charge densities and points
    (points: (m, 3) array; the split applies from 262,144 points on)."""
    ctx = container.ctx
    P = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    m = P.shape[0]
    X = [K.DeviceArray(ctx, np.ascontiguousarray(P[:, k])) for k in range(3)]
    S = K.DeviceArray(ctx, np.asarray(sigma, dtype=np.float64))
    O = [K.DeviceArray.empty(ctx, m) for _ in range(4)]
    prec = {"fp64": _lib.KEM_FP64, "mixed": _lib.KEM_MIXED}[precision]

    def run():
        ctx.check(ctx.lib.kem_field_eval_device(ctx.h, container.h, S.ptr, m, X[0].ptr, X[1].ptr, X[2].ptr, prec,
                                                O[0].ptr, O[1].ptr, O[2].ptr, O[3].ptr))

    try:
        return _sweep(ctx, "far_range_" + precision, candidates, run, repeats)
    finally:
        for a in X + O + [S]:
            a.close()


def tune_near(container: K.KSurfaceContainer, sigma, points, repeats=3, candidates=NEAR_CANDIDATES):
    """Grid of the near pass beside the mixed far pass, timed over the whole field pass."""
    ctx = container.ctx
    P = np.asarray(points, dtype=np.float64).reshape(-1, 3)
    m = P.shape[0]
    X = [K.DeviceArray(ctx, np.ascontiguousarray(P[:, k])) for k in range(3)]
    S = K.DeviceArray(ctx, np.asarray(sigma, dtype=np.float64))
    O = [K.DeviceArray.empty(ctx, m) for _ in range(4)]

    def run():
        ctx.check(ctx.lib.kem_field_eval_device(ctx.h, container.h, S.ptr, m, X[0].ptr, X[1].ptr, X[2].ptr,
                                                _lib.KEM_MIXED, O[0].ptr, O[1].ptr, O[2].ptr, O[3].ptr))

    try:
        return _sweep(ctx, "near_blocks", candidates, run, repeats, wall=True)
    finally:
        for a in X + O + [S]:
            a.close()


def tune_matvec(matrix: K.KBoundaryIntegralMatrix, repeats=5, candidates=GEMV_CANDIDATES):
    """Split-K target of the matrix-vector product (FP64 storage if present, else FP32)."""
    x = np.random.default_rng(0).normal(size=matrix.n)
    return _sweep(matrix.ctx, "gemv_blocks", candidates, lambda: matrix.Multiply(x), repeats)


def tune_block_product(matrix: K.KBoundaryIntegralMatrix, nrhs=64, precision="mixed", repeats=3,
                       candidates=None, rows=BMM_ROW_CANDIDATES):
    """Launch shape of the block product with nrhs right-hand sides ("fp64", "mixed" = FP32
    matrix, "bf16").  33-64 right-hand sides on a BF16 matrix, or on an FP32 matrix with
    n % 4 == 0, run the 32 x 32-tile kernels that stream the matrix through LDS: their workgroup
    target ``bmm_stream_blocks`` is swept (candidates default STREAM_CANDIDATES).  Up to 32
    right-hand sides on an FP32 matrix with n % 4 == 0, and an FP64 matrix with n even, run the
    16 x 16-tile streamed kernels, and 2-32 on a BF16 matrix the one-tile streamed kernel:
    ``bmm_tile_blocks`` is swept (TILE_CANDIDATES).  Otherwise
    (register-loaded kernels) the FP32 wavefront tile is swept first at the current split
    target, then the split target ``bmm_blocks`` with the chosen tile (candidates default
    BMM_CANDIDATES; the tile's sweep under "rows").  Returns the sweep of the split knob with
    its name under "knob"."""
    X = np.random.default_rng(1).normal(size=(matrix.n, nrhs))

    def run():
        matrix.MultiplyBatch(X, precision)

    streamed = nrhs > 32 and (precision == "bf16" or (precision == "mixed" and matrix.n % 4 == 0))
    if streamed:
        res = _sweep(matrix.ctx, "bmm_stream_blocks", candidates or STREAM_CANDIDATES, run, repeats)
        res["knob"] = "bmm_stream_blocks"
        return res
    tiled = ((precision == "mixed" and nrhs <= 32 and matrix.n % 4 == 0)
             or (precision == "fp64" and matrix.n % 2 == 0) or (precision == "bf16" and 2 <= nrhs <= 32))
    if tiled:
        res = _sweep(matrix.ctx, "bmm_tile_blocks", candidates or TILE_CANDIDATES, run, repeats)
        res["knob"] = "bmm_tile_blocks"
        return res
    tile = None
    if precision == "mixed":
        tile = _sweep(matrix.ctx, "bmm_rows", rows if matrix.n % 4 == 0 else (64,), run, repeats)
    res = _sweep(matrix.ctx, "bmm_blocks", candidates or BMM_CANDIDATES, run, repeats)
    res["knob"] = "bmm_blocks"
    if tile is not None:
        res["rows"] = tile
    return res


def current(ctx: K.Context) -> dict:
    return {name: get(ctx, name) for name in _KNOB}


def apply(ctx: K.Context, table: dict):
    """Sets every knob of table ({name: value} or the results of the tune_* functions)."""
    for name, v in table.items():
        set_knob(ctx, name, v["best"] if isinstance(v, dict) else v)


def save(table: dict, path: str):
    with open(path, "w") as f:
        json.dump({k: (v["best"] if isinstance(v, dict) else v) for k, v in table.items()}, f, indent=1)


def load(path: str) -> dict:
    with open(path) as f:
        return {k: int(v) for k, v in json.load(f).items() if k in _KNOB}
