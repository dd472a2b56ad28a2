"""Autotuner of the launch shapes of a gm_ctx (gm_ctx_tune).

Times the caller's own workload (a callable that enqueues one pass on the
current stream) under each candidate value of each knob -- coordinate
search, one knob at a time, the others held at their best so far -- and
leaves the fastest setting on the context.  Choices can be kept in a JSON
cache per (device, workload name) and reused without timing.

Knobs (include/geosmie.h, enum gm_tune_param):
  mie_grid_mult  persistent Mie grid as a multiple of the resident wavefronts
  s12_waves      wavefronts (64 angles each) per S1/S2 block
  psd_chunk      points per split-K chunk of the few-PSD contraction
  s12_points     points per wavefront of the FP64 S1/S2 kernel
  mie_waves      register budget of the Mie recurrence kernels (0 unconstrained,
                 3 or 4 resident wavefronts per SIMD)
  s12_layout     FP64 S1/S2 blocks split angles (0) or points (1) among their waves,
                 or share LDS-staged angle tables among them (2)
  mie_ckpt       coated kernel: downward-chain states kept every 4, 5 or 8 rows and
                 regenerated per block in LDS, or every row kept in scratch (0);
                 non-zero also selects the checkpointed lane-per-point
                 homogeneous kernel (every 4 rows)
  mie_lanes      homogeneous kernel: lanes per point (1, or 16/32/64 with the serial
                 recurrences joined by parallel scans; 0 = chosen by batch)
  s12_row_tiles  fused S1/S2 row sums: 8-point tiles per work unit (0 = by batch)
  s12_share      split-BF16 fused S1/S2: angle-table tiles shared by four work units
                 through LDS (1) or fetched per wavefront (0)
(mlp_weight_terms, the BF16 terms per network weight, is a precision
setting reached through the same call; it is not tuned for speed.)
"""
import ctypes
import json
import os
import time

from . import _lib
from ._lib import check

KNOBS = {"mie_grid_mult": (0, (1, 2, 4)),
         "s12_waves": (1, (1, 2, 4)),
         "psd_chunk": (2, (64, 128, 256, 512)),
         "s12_points": (4, (8, 16)),
         "mie_waves": (5, (0, 3, 4)),
         "s12_layout": (6, (0, 1, 2)),
         "mie_ckpt": (7, (5, 4, 8, 0)),
         "mie_lanes": (8, (0, 1, 16, 32, 64)),
         "s12_row_tiles": (10, (0, 1, 2, 4)),
         "s12_share": (11, (1, 0)),
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


def cache_path():
    """The tuning cache file: $GM_AUTOTUNE_CACHE, else
    ~/.cache/geosmie_amd/autotune.json."""
    return os.environ.get("GM_AUTOTUNE_CACHE") or os.path.join(
        os.path.expanduser("~"), ".cache", "geosmie_amd", "autotune.json")


def _device_tag():
    t = _lib.torch()
    p = t.cuda.get_device_properties(t.cuda.current_device())
    return "%s/%s" % (getattr(p, "gcnArchName", "") or p.name, p.multi_processor_count)


def load_cache(path):
    try:
        with open(path) as f:
            d = json.load(f)
        return d if isinstance(d, dict) else {}
    except (OSError, ValueError):
        return {}


def store_cache(path, entry_key, choices):
    """Merge {knob: value} under entry_key into the JSON file at path (written
    to a temporary file, then renamed)."""
    d = load_cache(path)
    d.setdefault(entry_key, {}).update({k: int(v) for k, v in choices.items()})
    os.makedirs(os.path.dirname(os.path.abspath(path)), exist_ok=True)
    tmp = path + ".tmp%d" % os.getpid()
    with open(tmp, "w") as f:
        json.dump(d, f, indent=1, sort_keys=True)
    os.replace(tmp, path)


def tune(ctx, workload, knobs=("mie_grid_mult", "s12_points", "s12_waves", "psd_chunk"), reps=3,
         warmup=1, candidates=None, cache_key=None, cache=None, tie=0.0):
    """Returns {knob: {"best": value, "ms": {value: ms per pass}}}; the context
    keeps the best values.  With cache_key (a name for the workload), choices
    already in the cache file for this device and workload are applied without
    timing (reported with "cached": True), and new ones are saved there; cache
    is the file (default cache_path()).  tie: the first candidate, in list
    order, whose time is within this fraction of the fastest is kept (times
    that agree to within the timing noise go to the earlier-listed value)."""
    t = _lib.torch()
    out = {}
    path = cache or cache_path()
    entry = None
    if cache_key is not None:
        entry = "%s|%s" % (_device_tag(), cache_key)
        known = load_cache(path).get(entry, {})
        for name in list(knobs):
            if name in known and known[name] in KNOBS[name][1]:
                set_knob(ctx, name, known[name])
                out[name] = {"best": known[name], "ms": {}, "cached": True}
        knobs = [k for k in knobs if k not in out]
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
        fastest = min(ms.values())
        best = next(v for v in cands if ms[v] <= fastest * (1.0 + tie))
        set_knob(ctx, name, best)
        out[name] = {"best": best, "ms": ms}
    if entry is not None and knobs:
        store_cache(path, entry, {k: out[k]["best"] for k in knobs})
    return out
