"""Benchmark: coated-sphere Mie batch (FP64) on one MI355X.

Workload (a "step"): every point of the coated-sphere batch -- core
m = 1.95+0.79i, shells 1.33 + (1.45-1.33)/gf^3 for 12 growth factors
gf = 1.0..2.5, core/shell radius ratio 0.3..0.9 (7), 5000 shell radii
log-spaced 0.01-10 um, 20 wavelengths 0.3-3.0 um: 8.4 M points, size
parameter up to 209 -- through gm_mie_batch (plan + length-bucketed coated
kernel): a_n, b_n and Qext, Qsca, Qabs, Qb, g, qratio per point, FP64.
Inputs are resident in HBM before the timed region.

Prints one JSON line.  --gpus N > 1 exits: this build runs on one GPU.
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

METRIC = "(r,λ,m) Mie points/s + bulk LUT entries/s vs plain GPU at ≤1e-3 rel err"
PEAK_FP64_TFLOPS = 78.6  # MI355X FP64 (vector = matrix on gfx950)


def coated_workload():
    r = np.logspace(np.log10(0.01), np.log10(10.0), 5000)
    lam = np.linspace(0.3, 3.0, 20)
    ratio = np.linspace(0.3, 0.9, 7)
    gf = np.linspace(1.0, 2.5, 12)
    m_shell = 1.33 + (1.45 - 1.33) / gf ** 3
    Y = (2 * np.pi * r[None, :] / lam[:, None]).ravel()          # [lambda][r]
    n1 = Y.size
    y = np.tile(Y, 7 * 12)                                        # [shell][ratio][lambda][r]
    x = np.tile(np.repeat(ratio, n1), 12) * y
    m2 = np.repeat(m_shell, 7 * n1).astype(np.complex128)
    m1 = 1.95 + 0.79j
    return x, y, m1, m2


def coated_flops(x, y, m1, m2):
    """Algorithmic FP64 FLOP per coated point: 30 nmx + 270 nmax."""
    nmax = np.rint(2 + y + 4 * y ** (1.0 / 3.0))
    mx = np.maximum(np.abs(m1) * y, np.abs(m2) * y)
    nmx = np.rint(np.maximum(nmax, mx) + 16)
    return float(np.sum(30 * nmx + 270 * nmax)), float(nmax.sum()), float(nmx.sum())


def cpu_baseline(x, y, m1, m2, target_s=12.0):
    from oracle import cmie
    cores = len(os.sched_getaffinity(0))
    cores = min(cores, 16)
    probe = slice(0, None, 997)
    t = time.perf_counter()
    cmie.batch(x[probe], m1, y[probe], m2[probe], threads=cores)
    dt = max(time.perf_counter() - t, 1e-4)
    est_full = dt * x.size / x[probe].size
    stride = max(1, int(np.ceil(est_full / target_s)))
    sel = slice(0, None, stride)
    t = time.perf_counter()
    cmie.batch(x[sel], m1, y[sel], m2[sel], threads=cores)
    dt = time.perf_counter() - t
    n = x[sel].size
    return {"value": n / dt, "unit": "points/s", "cores": cores, "kind": "port",
            "sample": "every %d-th point of the 8.4M-point coated batch (%d points, %.1f s), "
                      "C OpenMP restatement oracle/cmie.c" % (stride, n, dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--copy-calib", action="store_true",
                    help="also run 3 device copies of 512 MiB (PMC byte-counter calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; --gpus %d is not supported "
              "(the path shards by independent table slabs, see DESIGN.md)" % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import torch
    from geosmie_amd import _lib
    from geosmie_amd._lib import check, ptr

    torch.cuda.set_device(0)
    x, y, m1, m2 = coated_workload()
    n = x.size
    flops, sum_nmax, sum_nmx = coated_flops(x, y, m1, m2)
    d_x = _lib.to_dev(x)
    d_y = _lib.to_dev(y)
    d_eps = _lib.cplx_dev(np.full(n, m1 * m1))
    d_eps2 = _lib.cplx_dev(m2 * m2)
    d_props = torch.empty(6 * n, dtype=torch.float64, device="cuda")
    L = _lib.lib()
    h = _lib.ctx()
    s = _lib.stream()
    size_max = float(y.max())

    def step():
        check(L.gm_mie_batch(h, n, ptr(d_x), ptr(d_y), ptr(d_eps), None, ptr(d_eps2), size_max,
                             None, None, None, ptr(d_props), None, _lib.GM_MIE_SORT, s),
              "gm_mie_batch")

    for _ in range(args.warmup):
        step()
    torch.cuda.synchronize()
    flags = _lib.check_flags()
    assert flags == 0, flags
    L.gm_ctx_set_timing(h, 1)
    ms_k = ctypes.c_double(0)
    cnt = ctypes.c_int64(0)
    L.gm_ctx_kernel_time(h, 1, ctypes.byref(ms_k), ctypes.byref(cnt))  # drop stale pairs
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    torch.cuda.synchronize()
    t1 = time.perf_counter()
    check(L.gm_ctx_kernel_time(h, 1, ctypes.byref(ms_k), ctypes.byref(cnt)), "kernel_time")
    L.gm_ctx_set_timing(h, 0)
    ms_step = (t1 - t0) * 1e3 / args.steps
    kern_ms = ms_k.value / max(cnt.value, 1)
    props = d_props.view(6, n)
    finite = bool(torch.isfinite(props).all().item())
    achieved = flops / (kern_ms * 1e-3) / 1e12
    out = {
        "metric": METRIC,
        "value": n / (ms_step * 1e-3),
        "unit": "points/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": ms_step,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f64",
        "data": "synthetic",
        "config": {"workload": "coated-sphere batch: core 1.95+0.79i, 12 water-mixed shells "
                               "(gf 1.0-2.5), 7 core/shell ratios 0.3-0.9, 5000 radii 0.01-10 um, "
                               "20 wavelengths 0.3-3.0 um; 8.4M points, y <= 209; "
                               "coefficients + Qext/Qsca/Qabs/Qb/g",
                   "points": n, "sum_nmax": sum_nmax, "sum_nmx": sum_nmx,
                   "outputs_finite": finite},
        "roofline": {"bound": "mfma", "pipe": "FP64 VALU (gfx950 FP64 vector peak = FP64 "
                                              "matrix peak = 78.6 TFLOP/s)",
                     "kernel": "k_mie_coated", "achieved": achieved, "peak": PEAK_FP64_TFLOPS,
                     "unit": "TFLOP/s", "frac": achieved / PEAK_FP64_TFLOPS, "traffic": None,
                     "kernel_ms": kern_ms, "flop_per_launch": flops},
    }
    if args.copy_calib:
        a = torch.ones(64 << 20, dtype=torch.float64, device="cuda")   # 512 MiB
        b = torch.empty_like(a)
        for _ in range(3):
            b.copy_(a)
        torch.cuda.synchronize()
        del a, b
    if not args.no_cpu:
        out["cpu_baseline"] = cpu_baseline(x, y, m1, m2)
    print(json.dumps(out))


if __name__ == "__main__":
    main()
