"""ctypes binding of libgeosmie_hip.so and device-buffer plumbing.

The library is the only compute path: if it is missing, or no ROCm device is
visible, every compute call raises (there is no CPU fallback).
"""
import ctypes
import os
import threading

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgeosmie_hip.so")

GM_OK = 0
GM_MIE_SORT = 1
GM_MIE_DN_CONVERGED = 2
PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}

_lib = None
_lock = threading.Lock()

c_i64 = ctypes.c_int64
c_i32 = ctypes.c_int32
c_u32 = ctypes.c_uint32
c_dbl = ctypes.c_double
c_vp = ctypes.c_void_p

# name -> (restype, argtypes); every symbol declared in include/geosmie.h
SIGNATURES = {
    "gm_version": (ctypes.c_char_p, []),
    "gm_last_error": (ctypes.c_char_p, []),
    "gm_ctx_create": (ctypes.c_int, [ctypes.c_int, ctypes.POINTER(c_vp)]),
    "gm_ctx_destroy": (ctypes.c_int, [c_vp]),
    "gm_ctx_check": (ctypes.c_int, [c_vp, c_vp, ctypes.POINTER(c_i32)]),
    "gm_ctx_set_timing": (ctypes.c_int, [c_vp, ctypes.c_int]),
    "gm_ctx_kernel_time": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.POINTER(c_dbl),
                                          ctypes.POINTER(c_i64)]),
    "gm_ctx_tune": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int]),
    "gm_ctx_get_tune": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]),
    "gm_nmax": (ctypes.c_int, [c_dbl]),
    "gm_coef_offsets": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp]),
    "gm_mie_batch": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp,
                                    c_vp, c_vp, c_vp, c_vp, c_u32, c_vp]),
    "gm_mie_batch_noalloc": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl,
                                            c_vp, c_vp, c_vp, c_vp, c_vp, c_u32, c_vp, c_vp]),
    "gm_shape_average": (ctypes.c_int, [c_vp, ctypes.c_int, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "gm_ri_bilinear": (ctypes.c_int, [c_vp, ctypes.c_int, ctypes.c_int, c_i64, c_vp, c_i64, c_vp,
                                      c_vp, c_vp, c_vp]),
    "gm_pmatrix_normalize6": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, c_vp, c_vp, c_vp]),
    "gm_gsf_matr": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, ctypes.c_int, c_vp, c_vp,
                                   c_vp]),
    "gm_angle_tables": (ctypes.c_int, [c_vp, ctypes.c_int, c_vp, ctypes.c_int, c_vp, c_vp,
                                       c_vp]),
    "gm_mie_s12": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_vp, c_vp,
                                  ctypes.c_int, c_vp, c_vp, ctypes.c_int, c_vp]),
    "gm_mie_props": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp]),
    "gm_s12_mueller": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_vp,
                                      c_vp, ctypes.c_int, c_vp, ctypes.c_int, c_vp]),
    "gm_psd_contract": (ctypes.c_int, [c_vp, ctypes.c_int, c_i64, ctypes.c_int, c_vp, c_vp,
                                       c_vp, c_vp]),
    "gm_integrate_psd": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, c_vp, c_vp, c_vp, c_vp,
                                        c_dbl, c_dbl, c_dbl, c_vp, c_vp]),
    "gm_psd_contract_rows": (ctypes.c_int, [c_vp, ctypes.c_int, c_vp, c_vp, c_i64, ctypes.c_int,
                                            c_vp, c_vp, c_vp, c_vp]),
    "gm_s12_contract_rows": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, ctypes.c_int,
                                            c_vp, c_vp, ctypes.c_int, ctypes.c_int, c_vp, c_vp,
                                            c_vp, c_vp, ctypes.c_int, c_vp, c_vp]),
    "gm_s12_split_coef": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, ctypes.c_int, c_vp, c_vp]),
    "gm_lognorm_rows": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, c_vp, c_vp, c_vp, c_vp,
                                       c_i64, c_vp]),
    "gm_integrate_psd_rows": (ctypes.c_int, [c_vp, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp,
                                             c_dbl, c_dbl, c_vp, c_vp, c_vp]),
    "gm_band_average": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, c_vp, ctypes.c_int,
                                       ctypes.c_int, c_vp, c_vp, c_vp]),
    "gm_pmatrix_normalize": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, c_vp, c_vp, c_vp]),
    "gm_gsf_expand": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, c_vp, ctypes.c_int,
                                     ctypes.c_int, c_vp, c_vp]),
    "gm_legendre_eval": (ctypes.c_int, [c_vp, c_i64, c_vp, ctypes.c_int, ctypes.c_int, c_vp,
                                        c_vp, c_vp]),
    "gm_s12_p11": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, ctypes.c_int, c_vp, c_vp,
                                  ctypes.c_int, c_vp, ctypes.c_int, c_vp]),
    "gm_lut_points": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, c_vp, c_vp, c_vp, c_dbl,
                                     c_dbl, c_vp, c_vp, c_vp]),
    "gm_lut_lognormal": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, c_vp, c_vp, c_vp, c_vp,
                                        c_vp]),
    "gm_lut_bulk": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, ctypes.c_int, ctypes.c_int, c_vp,
                                   ctypes.c_int, c_vp, c_vp, c_vp, c_vp, c_vp, ctypes.c_int,
                                   ctypes.c_int, c_vp, c_vp]),
    "gm_mlp_param_count": (ctypes.c_int64, [ctypes.c_int, c_vp]),
    "gm_mlp_forward": (ctypes.c_int, [c_vp, ctypes.c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp]),
    "gm_mlp_grad": (ctypes.c_int, [c_vp, ctypes.c_int, c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp,
                                   c_vp, c_vp]),
    "gm_mlp_adam": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, ctypes.c_float,
                                   ctypes.c_float, ctypes.c_float, ctypes.c_float, c_i64, c_vp]),
    "gm_surrogate_prep": (ctypes.c_int, [c_vp, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp,
                                         c_i64, c_dbl, c_vp, c_vp, c_vp]),
    "gm_surrogate_post": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, c_dbl, c_vp, c_vp,
                                         c_vp, c_vp, c_vp]),
    "gm_surrogate_targets": (ctypes.c_int, [c_vp, c_i64, ctypes.c_int, c_vp, c_dbl, c_vp, c_vp,
                                            c_vp, c_vp, c_vp]),
}


class GMError(RuntimeError):
    pass


def lib():
    """Load the shared library once; raise if it is absent (no fallback)."""
    global _lib
    with _lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise GMError("libgeosmie_hip.so not built (run python -m geosmie_amd.build): %s"
                              % LIB_PATH)
            # torch first: its HIP runtime is then the one this library binds
            # (same soname), so the process holds one runtime, not two that
            # both claim the GPU (the second one then finds no device)
            import torch  # noqa: F401
            L = ctypes.CDLL(LIB_PATH)
            missing = [name for name in SIGNATURES if not hasattr(L, name)]
            if missing:
                raise GMError("%s lacks declared entry points %s: rebuild it "
                              "(python -m geosmie_amd.build)" % (LIB_PATH, ", ".join(missing)))
            for name, (res, args) in SIGNATURES.items():
                f = getattr(L, name)
                f.restype = res
                f.argtypes = args
            _lib = L
    return _lib


def exported_symbols():
    L = lib()
    return [n for n in SIGNATURES if hasattr(L, n)]


def check(rc, what=""):
    if rc != GM_OK:
        msg = lib().gm_last_error().decode(errors="replace")
        if rc < 0:
            raise ValueError(msg)
        raise GMError("%s failed (status %d): %s" % (what, rc, msg))


# ---------------------------------------------------------------------------
# Device plumbing (torch-ROCm tensors as buffers, torch's current stream)
# ---------------------------------------------------------------------------
_torch = None


def torch():
    global _torch
    if _torch is None:
        import torch as t
        if not t.cuda.is_available():
            raise GMError("no ROCm GPU visible: geosmie_amd computes only on the GPU")
        _torch = t
    return _torch


def device():
    return torch().device("cuda", torch().cuda.current_device())


def to_dev(a, dtype=np.float64):
    """numpy (or scalar) -> contiguous device tensor."""
    t = torch()
    if isinstance(a, t.Tensor):
        return a.to(device=device()).contiguous()
    arr = np.ascontiguousarray(np.asarray(a, dtype=dtype))
    if not arr.flags.writeable:     # torch.from_numpy wants writable memory
        arr = arr.copy()
    return t.from_numpy(arr).to(device())


def cplx_dev(a):
    """complex array -> interleaved (re, im) float64 device tensor."""
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.complex128))
    return to_dev(arr.view(np.float64))


def ptr(t):
    return None if t is None else ctypes.c_void_p(t.data_ptr())


def stream():
    return ctypes.c_void_p(torch().cuda.current_stream().cuda_stream)


_ctx = {}


def ctx():
    """Per-(thread, device) gm_ctx handle."""
    t = torch()
    key = (threading.get_ident(), t.cuda.current_device())
    h = _ctx.get(key)
    if h is None:
        h = c_vp()
        check(lib().gm_ctx_create(key[1], ctypes.byref(h)), "gm_ctx_create")
        _ctx[key] = h
    return h


class PrivateCtx(object):
    """A gm_ctx owned by one object (destroyed with it); usable wherever a
    context handle is passed.  Work captured into a graph must use such a
    context: a shared one could grow (reallocate) a workspace the graph uses."""

    def __init__(self):
        t = torch()
        self._h = c_vp()
        check(lib().gm_ctx_create(t.cuda.current_device(), ctypes.byref(self._h)), "gm_ctx_create")
        self._as_parameter_ = self._h

    def __del__(self):
        h = getattr(self, "_h", None)
        if h is not None and h.value and _lib is not None:
            try:
                torch().cuda.synchronize()
                _lib.gm_ctx_destroy(h)
            except Exception:
                pass
            self._h = None


def private_ctx():
    return PrivateCtx()


def check_flags():
    """Synchronise and return the context's device status flags."""
    f = c_i32(0)
    check(lib().gm_ctx_check(ctx(), stream(), ctypes.byref(f)), "gm_ctx_check")
    return f.value
