"""S1/S2, angle tables and efficiencies from coefficients, on the GPU.

Host-side helpers for the pymiecoated surface (mie_props.py:72-149) that
operate on coefficient arrays the caller already holds.
"""
import numpy as np

from . import _lib
from ._lib import check, ptr, stream

PREC = {"fp64": 0, "fp32": 1, "bf16x3": 2}


def _prec(precision):
    """gm_precision code of a name in PREC (integers pass through to the library)."""
    return PREC[precision] if isinstance(precision, str) else int(precision)


def _coef_dev(an_list, bn_list):
    """Pack ragged coefficient lists -> (d_off, d_nmax, d_an, d_bn)."""
    t = _lib.torch()
    nmax = np.array([len(a) for a in an_list], dtype=np.int64)
    off = np.zeros(len(an_list) + 1, dtype=np.int64)
    np.cumsum(nmax, out=off[1:])
    an = np.concatenate([np.asarray(a, dtype=np.complex128) for a in an_list]) if len(an_list) else np.zeros(0, np.complex128)
    bn = np.concatenate([np.asarray(b, dtype=np.complex128) for b in bn_list]) if len(bn_list) else np.zeros(0, np.complex128)
    return (_lib.to_dev(off, np.int64), _lib.to_dev(nmax.astype(np.int32), np.int32),
            _lib.cplx_dev(an if an.size else np.zeros(1, np.complex128)),
            _lib.cplx_dev(bn if bn.size else np.zeros(1, np.complex128)))


def angle_tables_dev(cost, nmax_cap):
    """Device pi'/tau' tables [nmax_cap, n_ang] for cosines `cost`."""
    t = _lib.torch()
    cost = np.atleast_1d(np.asarray(cost, dtype=np.float64))
    d_c = _lib.to_dev(cost)
    dev = _lib.device()
    d_pi = t.empty((max(nmax_cap, 1), cost.size), dtype=t.float64, device=dev)
    d_tau = t.empty_like(d_pi)
    check(_lib.lib().gm_angle_tables(_lib.ctx(), int(cost.size), ptr(d_c), int(nmax_cap),
                                     ptr(d_pi), ptr(d_tau), stream()), "gm_angle_tables")
    return d_pi, d_tau


def s12_batch(an_list, bn_list, cost=None, pi_tab=None, tau_tab=None, precision="fp64"):
    """S1, S2 [n_sets, n_ang] for ragged coefficient lists.

    Either `cost` (cosines; tables built on the GPU) or device tables
    `pi_tab`, `tau_tab` of shape [nmax_cap, n_ang].
    """
    t = _lib.torch()
    n = len(an_list)
    d_off, d_nmax, d_an, d_bn = _coef_dev(an_list, bn_list)
    nmax_max = max((len(a) for a in an_list), default=1)
    if pi_tab is None:
        pi_tab, tau_tab = angle_tables_dev(cost, nmax_max)
    n_ang = pi_tab.shape[1]
    dev = _lib.device()
    d_s1 = t.empty(2 * n * n_ang, dtype=t.float64, device=dev)
    d_s2 = t.empty_like(d_s1)
    check(_lib.lib().gm_mie_s12(_lib.ctx(), n, ptr(d_off), ptr(d_nmax), ptr(d_an), ptr(d_bn),
                                int(n_ang), ptr(pi_tab), ptr(tau_tab), int(pi_tab.shape[0]),
                                ptr(d_s1), ptr(d_s2), _prec(precision), stream()), "gm_mie_s12")
    s1 = d_s1.cpu().numpy().view(np.complex128).reshape(n, n_ang)
    s2 = d_s2.cpu().numpy().view(np.complex128).reshape(n, n_ang)
    return s1, s2


def s12_single(an, bn, nmax, cost):
    """[(S1, S2)] of one coefficient set at each cosine in `cost`."""
    s1, s2 = s12_batch([an[:nmax]], [bn[:nmax]], cost=cost)
    return [(complex(a), complex(b)) for a, b in zip(s1[0], s2[0])]


def s12_pt(an, bn, nmax, pin, tin):
    """S1, S2 from caller-supplied premultiplied pi/tau (mie_props.py:116-117, 133-149).

    pin, tin: 1-D rows (one angle) or 2-D [n_ang, >= nmax].
    """
    pin = np.asarray(pin, dtype=np.float64)
    tin = np.asarray(tin, dtype=np.float64)
    single = pin.ndim == 1
    P = np.ascontiguousarray(np.atleast_2d(pin)[:, :nmax].T)
    T = np.ascontiguousarray(np.atleast_2d(tin)[:, :nmax].T)
    s1, s2 = s12_batch([an[:nmax]], [bn[:nmax]], pi_tab=_lib.to_dev(P), tau_tab=_lib.to_dev(T))
    if single:
        return complex(s1[0, 0]), complex(s2[0, 0])
    return s1[0], s2[0]


def props_from_coeffs(an, bn, nmax, size):
    """(qext, qsca, qabs, qb, asy, qratio) of one coefficient set (mie_props.py:28-70)."""
    t = _lib.torch()
    d_off, d_nmax, d_an, d_bn = _coef_dev([an[:nmax]], [bn[:nmax]])
    d_size = _lib.to_dev(np.array([size], dtype=np.float64))
    d_props = t.empty(6, dtype=t.float64, device=_lib.device())
    check(_lib.lib().gm_mie_props(_lib.ctx(), 1, ptr(d_off), ptr(d_an), ptr(d_bn), ptr(d_size),
                                  ptr(d_props), stream()), "gm_mie_props")
    return tuple(float(v) for v in d_props.cpu().numpy())
