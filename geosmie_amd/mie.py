"""pymiecoated call surface on the gfx950 kernels.

Mirrors src/pymiecoated/pymiecoated/mie_coated.py (``Mie``, ``MieScatterProps``),
mie_coeffs.py (``mie_coeffs``, ``MieCoeffs``) and mie_props.py (``mie_props``,
``mie_S12``): same names, argument meaning and ValueError messages.  Every
number is computed by libgeosmie_hip.so on the GPU.
"""
import numpy as np

from . import _lib
from ._lib import check, ptr, stream

_PROP_KEYS = ("qext", "qsca", "qabs", "qb", "asy", "qratio")


def _batch(eps, x, mu=None, y=None, eps2=None, want_coeffs=True, flags=_lib.GM_MIE_SORT):
    """Run gm_mie_batch on host arrays; returns (an_list, bn_list, nmax, props[6,n])."""
    t = _lib.torch()
    x = np.atleast_1d(np.asarray(x, dtype=np.float64))
    n = x.size
    size = x if y is None else np.atleast_1d(np.asarray(y, dtype=np.float64))
    size_max = float(size.max()) if n else 0.0
    d_x = _lib.to_dev(x)
    d_y = _lib.to_dev(size) if y is not None else None
    dev = _lib.device()

    def per_point(v):
        # a scalar is filled on the device (no n-element host array or copy)
        v = np.asarray(v, dtype=np.complex128)
        if v.ndim == 0:
            d = t.empty((n, 2), dtype=t.float64, device=dev)
            d[:, 0] = float(v.real)
            d[:, 1] = float(v.imag)
            return d.view(-1)
        return _lib.cplx_dev(np.broadcast_to(v, (n,)))
    d_eps = per_point(eps)
    d_mu = None if mu is None else per_point(mu)
    d_eps2 = None if eps2 is None else per_point(eps2)
    d_props = t.empty(6 * n, dtype=t.float64, device=dev)
    d_nmax = t.empty(n, dtype=t.int32, device=dev)
    L = _lib.lib()
    h = _lib.ctx()
    s = stream()
    d_off = d_an = d_bn = None
    if want_coeffs:
        d_off = t.empty(n + 1, dtype=t.int64, device=dev)
        check(L.gm_coef_offsets(h, n, ptr(d_y if d_y is not None else d_x), ptr(d_off), s),
              "gm_coef_offsets")
        total = int(d_off[-1].item())
        d_an = t.empty(2 * max(total, 1), dtype=t.float64, device=dev)
        d_bn = t.empty(2 * max(total, 1), dtype=t.float64, device=dev)
    check(L.gm_mie_batch(h, n, ptr(d_x), ptr(d_y), ptr(d_eps), ptr(d_mu), ptr(d_eps2), size_max,
                         ptr(d_off), ptr(d_an), ptr(d_bn), ptr(d_props), ptr(d_nmax), flags, s),
          "gm_mie_batch")
    flags_out = _lib.check_flags()
    if flags_out & 2:
        raise ValueError("Multilayer calculations for magnetic particles "
                         "are not currently supported.")
    if flags_out & 1:
        raise _lib.GMError("a point exceeded the size bound")
    props = d_props.cpu().numpy().reshape(6, n)
    nmax = d_nmax.cpu().numpy().astype(np.int64)
    an_l = bn_l = None
    if want_coeffs:
        off = d_off.cpu().numpy()
        an = d_an.cpu().numpy().view(np.complex128)
        bn = d_bn.cpu().numpy().view(np.complex128)
        an_l = [an[off[i]:off[i + 1]] for i in range(n)]
        bn_l = [bn[off[i]:off[i + 1]] for i in range(n)]
    return an_l, bn_l, nmax, props


def mie_batch(x, m=None, y=None, m2=None, eps=None, eps2=None, mu=None, coeffs=False,
              converged_dn=False, sort=True):
    """Batched Mie on the GPU: dict of efficiency arrays (+ coefficient lists).

    x: size parameters (core sizes when y is given); m / eps: core index or
    permittivity (scalar or per point); y, m2 / eps2: shell; mu: permeability.
    Returns {'qext','qsca','qabs','qb','asy','qratio': [n], 'nmax': [n]} and,
    with coeffs=True, 'an', 'bn' lists of complex arrays.
    """
    if eps is None:
        if m is None:
            raise ValueError("Must specify x and either eps or m.")
        eps = np.asarray(m, dtype=np.complex128) ** 2
    if y is not None and eps2 is None:
        if m2 is None:
            raise ValueError("Must specify both y and m2 for coated particles.")
        eps2 = np.asarray(m2, dtype=np.complex128) ** 2
    flags = (_lib.GM_MIE_SORT if sort else 0) | (_lib.GM_MIE_DN_CONVERGED if converged_dn else 0)
    an, bn, nmax, props = _batch(eps, x, mu=mu, y=y, eps2=eps2, want_coeffs=coeffs, flags=flags)
    out = dict(zip(_PROP_KEYS, props))
    out["nmax"] = nmax
    if coeffs:
        out["an"], out["bn"] = an, bn
    return out


class MieCoeffs(object):
    """Wrapper for the Mie coefficients (mie_coeffs.py:30-34)."""

    def __init__(self, par):
        (self.an, self.bn, self.nmax) = mie_coeffs(par)


def mie_coeffs(params):
    """Input validation and dispatch (mie_coeffs.py:36-73) -> (an, bn, nmax)."""
    eps = complex(params["eps"]) if params.get("eps") is not None else None
    x = float(params["x"]) if params.get("x") is not None else None
    if x is None or eps is None:
        raise ValueError("Must specify x and either eps or m.")
    mu = complex(params["mu"]) if params.get("mu") is not None else complex(1.0)
    y = float(params["y"]) if params.get("y") is not None else None
    eps2 = complex(params["eps2"]) if params.get("eps2") is not None else None
    coated = (y is not None)
    if coated == (eps2 is None):
        raise ValueError("Must specify both y and m2 for coated particles.")
    if coated and mu != complex(1.0):
        raise ValueError("Multilayer calculations for magnetic particles " +
                         "are not currently supported.")
    an, bn, nmax, _ = _batch(eps, x, mu=None if mu == 1 else mu, y=y, eps2=eps2)
    return an[0], bn[0], int(nmax[0])


def mie_props(coeffs, y):
    """Efficiencies from coefficients (mie_props.py:72-75).

    The sums run on the GPU in the same kernel that produced the
    coefficients; this entry point recomputes them from an arbitrary
    (an, bn) pair with the device S12/props path.
    """
    from .s12 import props_from_coeffs
    vals = props_from_coeffs(coeffs.an, coeffs.bn, coeffs.nmax, y)
    return dict(zip(_PROP_KEYS, vals))


def mie_S12(coeffs, u):
    """S1, S2 at one cos(theta) (mie_props.py:108-114)."""
    from .s12 import s12_single
    return s12_single(coeffs.an, coeffs.bn, coeffs.nmax, [u])[0]


def mie_S12_pt(coeffs, pin, tin):
    """S1, S2 from caller-supplied premultiplied pi/tau rows (mie_props.py:116-117)."""
    from .s12 import s12_pt
    return s12_pt(coeffs.an, coeffs.bn, coeffs.nmax, pin, tin)


class Cache(dict):
    """FIFO cache of 10 entries (src/pymiecoated/pymiecoated/mie_aux.py:22-33)."""

    def __init__(self, size=10):
        super(Cache, self).__init__()
        self.size = size
        self.log = []

    def __setitem__(self, key, value):
        super(Cache, self).__setitem__(key, value)
        self.log.append(key)
        if len(self.log) > self.size:
            del self[self.log[0]]
            self.log.pop(0)


class MieScatterProps(object):
    """Coefficients, efficiencies and S12 of one parameter set (mie_coated.py:181-211)."""

    def __init__(self, params, ajv=(), ayv=()):
        par = dict(zip(("eps", "mu", "x", "y", "eps2"), params[:5]))
        self.size = par["x"] if par["y"] is None else par["y"]
        eps, mu, x, y, eps2 = (par[k] for k in ("eps", "mu", "x", "y", "eps2"))
        if x is None or eps is None:
            raise ValueError("Must specify x and either eps or m.")
        mu = complex(mu) if mu is not None else complex(1.0)
        coated = y is not None
        if coated == (eps2 is None):
            raise ValueError("Must specify both y and m2 for coated particles.")
        if coated and mu != complex(1.0):
            raise ValueError("Multilayer calculations for magnetic particles " +
                             "are not currently supported.")
        an, bn, nmax, props = _batch(complex(eps), float(x), mu=None if mu == 1 else mu,
                                     y=None if y is None else float(y),
                                     eps2=None if eps2 is None else complex(eps2))
        self.an, self.bn, self.nmax = an[0], bn[0], int(nmax[0])
        self._props = dict(zip(_PROP_KEYS, (float(v) for v in props[:, 0])))

    def prop(self, prop_name):
        return self._props[prop_name]

    def S12(self, u):
        from .s12 import s12_single
        return s12_single(self.an, self.bn, self.nmax, [u])[0]

    def S12_pt(self, pin, tin):
        from .s12 import s12_pt
        return s12_pt(self.an, self.bn, self.nmax, pin, tin)


class Mie(object):
    """Mie scattering from homogeneous and coated spheres (mie_coated.py:214-396).

    Same attributes, keyword arguments, properties and errors as the
    reference class; all computation on the GPU.
    """

    def __init__(self, **kwargs):
        self._cache = Cache()
        self.eps = None
        self.mu = 1.0
        self._x = None
        self._y = None
        self.eps2 = None
        self.ajv = ()
        self.ayv = ()
        for k in ["eps", "mu", "eps2"]:
            if k in kwargs:
                self.__dict__[k] = kwargs[k]
        if "m" in kwargs:
            self.m = kwargs["m"]
        if "m2" in kwargs:
            self.m2 = kwargs["m2"]
        if "x" in kwargs:
            self.x = kwargs["x"]
        if "y" in kwargs:
            self.y = kwargs["y"]
        if "ajv" in kwargs:
            self.ajv = tuple(kwargs["ajv"])
        if "ayv" in kwargs:
            self.ayv = tuple(kwargs["ayv"])

    def _params_signature(self):
        return (self.eps, self.mu, self.x, self.y, self.eps2)

    def _sp(self):
        sig = self._params_signature()
        if sig not in self._cache:
            self._cache[sig] = MieScatterProps(sig)
        return self._cache[sig]

    def qext(self):
        return self._sp().prop("qext")

    def qsca(self):
        return self._sp().prop("qsca")

    def qabs(self):
        return self._sp().prop("qabs")

    def qb(self):
        return self._sp().prop("qb")

    def asy(self):
        return self._sp().prop("asy")

    def qratio(self):
        return self._sp().prop("qratio")

    def S12(self, u):
        if abs(u) > 1:
            raise ValueError("The cosine u must be between -1 and 1.")
        return self._sp().S12(u)

    def S12_pt(self, pin, tin):
        return self._sp().S12_pt(pin, tin)

    def _get_m(self):
        return np.sqrt(self.eps / self.mu)

    def _set_m(self, m):
        self.mu = 1.0
        self.eps = m ** 2

    m = property(_get_m, _set_m)

    def _get_m2(self):
        # The reference's getter recurses (mie_coated.py:366-367); the
        # intended value is the shell refractive index.
        return np.sqrt(self.eps2)

    def _set_m2(self, m2):
        self.eps2 = m2 ** 2

    m2 = property(_get_m2, _set_m2)

    def _get_x(self):
        return self._x

    def _set_x(self, x):
        if x >= 0.0:
            self._x = x
        else:
            raise ValueError("The size x cannot be smaller than 0.")

    x = property(_get_x, _set_x)

    def _get_y(self):
        return self._y

    def _set_y(self, y):
        if y >= self.x:
            self._y = y
        else:
            raise ValueError("The size y cannot be smaller than x.")

    y = property(_get_y, _set_y)
