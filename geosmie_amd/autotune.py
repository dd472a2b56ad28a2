"""Autotuner of the launch shapes of a gm_ctx (gm_ctx_tune).

Times the caller's own workload (a callable that enqueues one pass on the
current stream) under each candidate value of each knob -- coordinate
search, one knob at a time, the others held at their best so far -- and
leaves the fastest setting on the context.

Knobs (include/geosmie.h, enum gm_tune_param):
  mie_grid_mult  persistent Mie grid as a multiple of the resident wavefronts
  s12_waves      wavefronts (64 angles each) per S1/S2 block
  psd_chunk      points per split-K chunk of the few-PSD contraction
  s12_points     points per wavefront of the FP64 S1/S2 kernel
  mie_waves      register budget of the Mie recurrence kernels (0 unconstrained,
                 3 or 4 resident wavefronts per SIMD)
  s12_layout     FP64 S1/S2 blocks split angles (0) or points (1) among their waves,
                 or share LDS-staged angle tables among them (2)
  mie_ckpt       coated kernel: downward-chain states kept every 4 or 8 rows and
                 regenerated per block in LDS, or every row kept in scratch (0)
  mie_lanes      homogeneous kernel: lanes per point (1, or 16/32/64 with the serial
                 recurrences joined by parallel scans; 0 = chosen by batch)
(mlp_weight_terms, the BF16 terms per network weight, is a precision
setting reached through the same call; it is not tuned for speed.)
"""
import ctypes
import time

from . import _lib
from ._lib import check

KNOBS = {"mie_grid_mult": (0, (1, 2, 4)),
         "s12_waves": (1, (1, 2, 4)),
         "psd_chunk": (2, (64, 128, 256, 512)),
         "s12_points": (4, (8, 16)),
         "mie_waves": (5, (0, 3, 4)),
         "s12_layout": (6, (0, 1, 2)),
         "mie_ckpt": (7, (4, 8, 0)),
         "mie_lanes": (8, (0, 1, 16, 32, 64)),
         # a precision setting rather than a shape; set explicitly, never tuned for speed
         "mlp_weight_terms": (3, (1, 2)),
         # a behaviour setting: the integratePSD quirks reproduced (bit 0: the aliased
         # Qsca weighting of csca and cext); set explicitly, never tuned
         "psd_quirks": (9, (1, 0))}


def get(ctx, name):
    v = ctypes.c_int(0)
    check(_lib.lib().gm_ctx_get_tune(ctx, KNOBS[name][0], ctypes.byref(v)), "gm_ctx_get_tune")
    return v.value


def set_knob(ctx, name, value):
    check(_lib.lib().gm_ctx_tune(ctx, KNOBS[name][0], int(value)), "gm_ctx_tune")


def tune(ctx, workload, knobs=("mie_grid_mult", "s12_points", "s12_waves", "psd_chunk"), reps=3,
         warmup=1, candidates=None):
    """Returns {knob: {"best": value, "ms": {value: ms per pass}}}; the context
    keeps the best values."""
    t = _lib.torch()
    out = {}
    for name in knobs:
        cands = (candidates or {}).get(name, KNOBS[name][1])
        ms = {}
        for v in cands:
            set_knob(ctx, name, v)
            for _ in range(warmup):
                workload()
            t.cuda.synchronize()
            t0 = time.perf_counter()
            for _ in range(reps):
                workload()
            t.cuda.synchronize()
            ms[v] = (time.perf_counter() - t0) * 1e3 / reps
        best = min(ms, key=ms.get)
        set_knob(ctx, name, best)
        out[name] = {"best": best, "ms": ms}
    return out
