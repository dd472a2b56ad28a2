"""GSF expansion and Legendre evaluation on the GPU.

Drop-in for the rungsf / convertncdf 'pygeos' path (src/gsf/convertncdf.py:
170-189, 319-401), which runs the Fortran program spher_expan.x once per
(bin, wavelength, rh) through a text file.  Here all entries of a table go
to the device in one call.
"""
import numpy as np

from . import _lib
from ._lib import check, ptr, stream

NSPHER = 129        # src/gsf/params.h:14
TABLE_P = [11, 12, 33, 34, 22, 44]   # element order of the runoptics table p axis


def pmom_batch(ang_deg, pmat, ng=NSPHER, quantize=True):
    """pmom[n, 6, ng] from phase matrices pmat[n, 6, n_ang] (table order p11, p12,
    p33, p34, p22, p44), as convertncdf.py:369-394 fills pmom(bin, wavelength,
    rh, p, m)."""
    t = _lib.torch()
    ang = np.ascontiguousarray(ang_deg, dtype=np.float64)
    pmat = np.ascontiguousarray(pmat, dtype=np.float64)
    if pmat.ndim == 2:
        pmat = pmat[None]
    n = pmat.shape[0]
    d_f = _lib.to_dev(pmat)
    d_o = t.empty(n * 6 * ng, dtype=t.float64, device=_lib.device())
    check(_lib.lib().gm_gsf_expand(_lib.ctx(), n, ang.size, ang.ctypes.data_as(_lib.c_vp),
                                   ptr(d_f), int(ng), int(quantize), ptr(d_o), stream()),
          "gm_gsf_expand")
    return d_o.cpu().numpy().reshape(n, 6, ng)


def spher_expan(ang_deg, f11, f22, f33, f44, f12, f34, ng=NSPHER, quantize=True):
    """One phase matrix in the program's column order (spher_expan.f:218-220);
    returns rows alpha1, alpha2, alpha3, alpha4, beta1, beta2 [6, ng] as in
    '<file>.expan_coeff' (spher_expan.f:93-107)."""
    pm = np.stack([f11, f12, f33, f34, f22, f44])
    out = pmom_batch(ang_deg, pm, ng=ng, quantize=quantize)[0]
    return np.stack([out[0], out[4], out[2], out[5], out[1], out[3]])


GM_GSF_RAW = 2      # gm_gsf_expand flag: coefficients before the 1/alpha1(0) normalisation


def matr_batch(coef, ang_deg):
    """The phase matrices [n, 6, n_ang] (table order p11, p12, p33, p34, p22,
    p44) back from expansion coefficients coef [n, 6, l1max] in
    gm_gsf_expand's order, at angles ang_deg (MATR, spher_expan.f:419-517)."""
    t = _lib.torch()
    c = np.ascontiguousarray(np.atleast_3d(coef) if np.ndim(coef) == 3 else coef[None],
                             dtype=np.float64)
    ang = np.ascontiguousarray(ang_deg, dtype=np.float64)
    d_c, d_th = _lib.to_dev(c), _lib.to_dev(np.radians(ang))
    d_o = t.empty(c.shape[0] * 6 * ang.size, dtype=t.float64, device=_lib.device())
    check(_lib.lib().gm_gsf_matr(_lib.ctx(), c.shape[0], c.shape[2], ptr(d_c), ang.size,
                                 ptr(d_th), ptr(d_o), stream()), "gm_gsf_matr")
    return d_o.cpu().numpy().reshape(c.shape[0], 6, ang.size)


def alt_angles(ang_deg):
    """The program's midpoint check grid (spher_expan.f:240-247)."""
    a = np.asarray(ang_deg, dtype=np.float64)
    n = a.size
    alt = a.copy()
    h = n // 2
    alt[1:h] = 0.5 * (a[0:h - 1] + a[1:h])
    alt[h:n - 1] = 0.5 * (a[h + 1:n] + a[h:n - 1])
    return alt


def expan_matr(ang_deg, f11, f22, f33, f44, f12, f34, ng=NSPHER):
    """The program's fit check for one matrix (one_calc, spher_expan.f:157-177):
    returns (matr [6, n_ang] in the program's column order F11, F22, F33, F44,
    F12, F34 -- the '<file>.expan_matr' values, :85-91 --, fiterr), the fit
    error being the largest |F11 - F11 fitted| at the input angles and at the
    midpoint grid (ERREVAL with MAXABS, :635-687).  The fitted matrix is MATR
    of the unnormalised coefficients on the device."""
    ang = np.asarray(ang_deg, dtype=np.float64)
    pm = np.stack([f11, f12, f33, f34, f22, f44])[None]
    raw = pmom_batch(ang, pm, ng=ng, quantize=GM_GSF_RAW)
    alt = alt_angles(ang)
    both = matr_batch(np.concatenate([raw, raw]), ang)
    m_alt = matr_batch(raw, alt)[0]
    m = both[0]
    f11alt = np.interp(alt, ang, np.asarray(f11, np.float64))
    err = max(np.max(np.abs(np.asarray(f11) - m[0])), np.max(np.abs(f11alt - m_alt[0])))
    return np.stack([m[0], m[4], m[2], m[5], m[1], m[3]]), float(err)


def hbleg(x, pmoms, leg=None):
    """sum_l pmoms[l] P_l(x) (convertncdf.py:32-43).  `leg` (the reference's
    scratch array) is accepted and ignored."""
    t = _lib.torch()
    x = np.ascontiguousarray(np.atleast_1d(x), dtype=np.float64)
    c = np.ascontiguousarray(np.atleast_2d(pmoms), dtype=np.float64)
    d_o = t.empty(c.shape[0] * x.size, dtype=t.float64, device=_lib.device())
    d_x, d_c = _lib.to_dev(x), _lib.to_dev(c)   # keep alive until the kernel has read them
    check(_lib.lib().gm_legendre_eval(_lib.ctx(), x.size, ptr(d_x), c.shape[0], c.shape[1],
                                      ptr(d_c), ptr(d_o), stream()), "gm_legendre_eval")
    out = d_o.cpu().numpy().reshape(c.shape[0], x.size)
    return out[0] if np.ndim(pmoms) == 1 else out


def add_pmom(table, ng=NSPHER, quantize=True):
    """Fill table['pmom'][bin, wavelength, rh, p, m] from the p11..p44 variables
    of a runoptics table dict (the pygeos path of convertncdf.processFileRaw)."""
    p = np.stack([table[k] for k in ("p11", "p12", "p33", "p34", "p22", "p44")], axis=-2)
    shp = p.shape[:-2]
    flat = p.reshape(-1, 6, p.shape[-1])
    pm = pmom_batch(table["ang"], flat, ng=ng, quantize=quantize)
    table["pmom"] = pm.reshape(shp + (6, ng))
    return table
