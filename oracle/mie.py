"""FP64 numpy/scipy restatement of pymiecoated (the reference Mie core).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows, formula by formula:
  src/pymiecoated/pymiecoated/mie_coeffs.py:36-251
  src/pymiecoated/pymiecoated/mie_props.py:28-231
  src/pymiecoated/pymiecoated/mie_coated.py:31-89, 129-179

Riccati-Bessel functions come from ``scipy.special.jv/yv`` exactly as the
reference calls them (mie_coeffs.py:156-161, 223-224; mie_coated.py:136-138).

Pinned by (tests/test_oracle_mie.py):
  * the six efficiency and two S12 golden values of each of the three cases of
    src/pymiecoated/pymiecoated/test/test_mie.py:45-109 (tolerance 2.2e-13);
  * the independent arbitrary-precision Mie of ``oracle/mpmie.py`` at sizes
    0.01 .. 2000 (homogeneous) and coated cases;
  * closed forms: the Rayleigh limit of Qsca, the optical theorem
    Re S(0) = x^2 Qext / 4, qb = 4|S1(pi)|^2 / x^2, S1(0) = S2(0),
    S1(pi) = -S2(pi), Qabs = 0 for real m, Qsca = (1/x^2) int (|S1|^2+|S2|^2),
    g = <cos> of the phase function; coated reductions to homogeneous
    (m1 = m2, x = 0, x = y);
  * scipy.special.lpmv for pi_n, tau_n.
"""
import math

import numpy as np
from scipy.special import jv, yv


def py_round(v):
    """Python's round(): half to even (mie_coeffs.py:99,101 use builtin round)."""
    return int(round(float(v)))


def nmax_of(size):
    """Series length, mie_coeffs.py:99 (and :203 with y)."""
    return py_round(2.0 + size + 4.0 * size ** (1.0 / 3.0))


def nmx_homog(x, eps, mu=1.0):
    """D_n start index for the homogeneous sphere, mie_coeffs.py:96,101."""
    z = np.sqrt(complex(eps) * complex(mu)) * x
    return py_round(max(nmax_of(x), abs(z)) + 16)


def nmx_coated(x, y, eps1, eps2):
    """D_n start index for the coated sphere, mie_coeffs.py:196-205."""
    m1 = np.sqrt(complex(eps1))
    m2 = np.sqrt(complex(eps2))
    mx = max(abs(m1 * y), abs(m2 * y))
    return py_round(max(nmax_of(y), mx) + 16)


def _dn_down(z, nmx, nmax):
    """Downward log-derivative recurrence, mie_coeffs.py:118-122.

    dnx[j-1] = r - 1/(dnx[j] + r), r = (j+1)/z, from dnx[nmx-1] = 0; returns
    D_1..D_nmax (dnx[:nmax]).
    """
    dnx = np.zeros(nmx, dtype=np.complex128)
    for j in range(nmx - 1, 0, -1):
        r = (j + 1.0) / z
        dnx[j - 1] = r - 1.0 / (dnx[j] + r)
    return dnx[:nmax].copy()


def bessel_tables(x, nmax):
    """jv(n+1.5, x), yv(n+1.5, x) for n = 0..nmax-1 (mie_coated.py:129-147)."""
    nu = np.arange(nmax) + 1.5
    return jv(nu, x), yv(nu, x)


def single_mie_coeff(eps, mu, x, ajv=None, ayv=None):
    """Homogeneous-sphere a_n, b_n: mie_coeffs.py:132-180 (numba twin :83-130).

    Returns (an, bn, nmax) with an, bn complex128[nmax].
    """
    eps = complex(eps)
    mu = complex(mu)
    z = np.sqrt(eps * mu) * x
    m = np.sqrt(eps / mu)
    nmax = nmax_of(x)
    nmx = py_round(max(nmax, abs(z)) + 16)
    n = np.arange(nmax)
    if ajv is None or len(ajv) == 0:
        ajv, ayv = bessel_tables(x, nmax)
    sx = math.sqrt(0.5 * math.pi * x)
    px = sx * np.asarray(ajv, dtype=np.float64)
    p1x = np.concatenate(([math.sin(x)], px[:nmax - 1]))
    chx = -sx * np.asarray(ayv, dtype=np.float64)
    ch1x = np.concatenate(([math.cos(x)], chx[:nmax - 1]))
    gsx = px - 1j * chx
    gs1x = p1x - 1j * ch1x
    dn = _dn_down(z, nmx, nmax)
    n1 = n + 1
    da = dn / m + n1 / x
    db = dn * m + n1 / x
    an = (da * px - p1x) / (da * gsx - gs1x)
    bn = (db * px - p1x) / (db * gsx - gs1x)
    return an, bn, nmax


def coated_mie_coeff(eps1, eps2, x, y):
    """Coated-sphere a_n, b_n: mie_coeffs.py:183-251 (scipy jv/yv at complex v, w)."""
    m1 = np.sqrt(complex(eps1))
    m2 = np.sqrt(complex(eps2))
    m = m2 / m1
    u = m1 * x
    v = m2 * x
    w = m2 * y
    nmax = nmax_of(y)
    mx = max(abs(m1 * y), abs(w))
    nmx = py_round(max(nmax, mx) + 16)
    n = np.arange(nmax)
    dnu = _dn_down(u, nmx, nmax)
    dnv = _dn_down(v, nmx, nmax)
    dnw = _dn_down(w, nmx, nmax)
    nu = n + 1.5
    sv = np.sqrt(0.5 * np.pi * v)
    sw = np.sqrt(0.5 * np.pi * w)
    sy = math.sqrt(0.5 * math.pi * y)
    pv = sv * jv(nu, v)
    pw = sw * jv(nu, w)
    py = sy * jv(nu, y)
    chv = -sv * yv(nu, v)
    chw = -sw * yv(nu, w)
    chy = -sy * yv(nu, y)
    p1y = np.concatenate(([math.sin(y)], py[:nmax - 1]))
    ch1y = np.concatenate(([math.cos(y)], chy[:nmax - 1]))
    gsy = py - 1j * chy
    gs1y = p1y - 1j * ch1y
    uu = m * dnu - dnv
    vv = dnu / m - dnv
    fv = pv / chv
    ku1 = uu * fv / pw
    kv1 = vv * fv / pw
    pt = pw - chw * fv
    prat = pw / pv / chv
    ku2 = uu * pt + prat
    kv2 = vv * pt + prat
    dns = ku1 / ku2 + dnw
    gns = kv1 / kv2 + dnw
    nrat = (n + 1) / y
    a1 = dns / m2 + nrat
    b1 = m2 * gns + nrat
    an = (py * a1 - p1y) / (gsy * a1 - gs1y)
    bn = (py * b1 - p1y) / (gsy * b1 - gs1y)
    return an, bn, nmax


def mie_coeffs(eps, x, mu=1.0, y=None, eps2=None):
    """Validation and dispatch, mie_coeffs.py:36-73.

    Raises ValueError with the reference messages.  Returns (an, bn, nmax).
    """
    if x is None or eps is None:
        raise ValueError("Must specify x and either eps or m.")
    eps = complex(eps)
    x = float(x)
    mu = complex(mu) if mu is not None else complex(1.0)
    y = float(y) if y is not None else None
    eps2 = complex(eps2) if eps2 is not None else None
    coated = y is not None
    if coated == (eps2 is None):
        raise ValueError("Must specify both y and m2 for coated particles.")
    if coated and mu != complex(1.0):
        raise ValueError("Multilayer calculations for magnetic particles "
                         "are not currently supported.")
    if not coated:
        y = x
        eps2 = eps
    if x == y or eps == eps2:
        return single_mie_coeff(eps, mu, y)
    if x == 0:
        return single_mie_coeff(eps2, mu, y)
    return coated_mie_coeff(eps, eps2, x, y)


def mie_props_raw(an, bn, nmax, y):
    """Efficiencies, mie_props.py:28-70.

    Returns (qext, qsca, qabs, qb, asy, qratio).
    """
    anp, anpp = an.real, an.imag
    bnp, bnpp = bn.real, bn.imag
    n = np.arange(1.0, nmax + 1.0)
    cn = 2 * n + 1
    c1n = n * (n + 2) / (n + 1)
    c2n = cn / (n * (n + 1))
    y2 = y ** 2
    qext = 2 * np.sum(cn * (anp + bnp)) / y2
    qsca = 2 * np.sum(cn * (anp ** 2 + anpp ** 2 + bnp ** 2 + bnpp ** 2)) / y2
    qabs = qext - qsca
    q = np.sum((an - bn) * cn * (-1.0) ** n)
    qb = (q.real ** 2 + q.imag ** 2) / y2
    g1 = np.zeros((4, nmax))
    g1[:, :nmax - 1] = np.vstack((anp[1:], anpp[1:], bnp[1:], bnpp[1:]))
    asy1 = c1n * (anp * g1[0] + anpp * g1[1] + bnp * g1[2] + bnpp * g1[3])
    asy2 = c2n * (anp * bnp + anpp * bnpp)
    asy = 4 / y2 * np.sum(asy1 + asy2) / qsca
    return qext, qsca, qabs, qb, asy, qb / qsca


PROP_KEYS = ("qext", "qsca", "qabs", "qb", "asy", "qratio")


def mie_pt(u, nmax):
    """pi_n, tau_n premultiplied by (2n+1)/(n(n+1)), mie_props.py:166-231.

    pi_1 = 1, pi_2 = 3u, pi_k = (2k-1)/(k-1) u pi_{k-1} - k/(k-1) pi_{k-2};
    tau_1 = u, tau_2 = 6u^2 - 3, tau_k = k u pi_k - (k+1) pi_{k-1}.
    Handles nmax >= 1 (the reference indexes p[1] unconditionally).  u may be
    a scalar (rows [nmax]) or an array of cosines (rows [len(u), nmax]); the
    recurrence is the same, carried out for all cosines at once.
    """
    scalar = np.ndim(u) == 0
    u = np.atleast_1d(np.asarray(u, dtype=np.float64))
    p = np.zeros((nmax, u.size))
    t = np.zeros((nmax, u.size))
    p[0] = 1.0
    t[0] = u
    if nmax > 1:
        p[1] = 3 * u
        t[1] = 6 * u ** 2 - 3
    for ni in range(2, nmax):
        nf = float(ni)
        p[ni] = (2 * nf + 1) / nf * p[ni - 1] * u - (nf + 1) / nf * p[ni - 2]
        t[ni] = (nf + 1) * u * p[ni] - (nf + 2) * p[ni - 1]
    k = np.arange(1, nmax + 1, dtype=np.float64)
    n2 = ((2 * k + 1) / (k * (k + 1)))[:, None]
    p = (p * n2).T
    t = (t * n2).T
    if scalar:
        return p[0].copy(), t[0].copy()
    return np.ascontiguousarray(p), np.ascontiguousarray(t)


def pt_table(costarr, nmax):
    """[N_ang, nmax] pi'/tau' tables, mie_coated.py:160-179."""
    if len(costarr) == 0:
        return np.zeros((0, nmax)), np.zeros((0, nmax))
    return mie_pt(np.asarray(costarr, dtype=np.float64), nmax)


def s12_from_pt(an, bn, pin, tin):
    """S1, S2 for prefix-length pi/tau rows, mie_props.py:133-149."""
    nmax = len(an)
    p = pin[..., :nmax]
    t = tin[..., :nmax]
    s1 = p @ an + t @ bn
    s2 = t @ an + p @ bn
    return s1, s2


def mie_S12(an, bn, nmax, u):
    """Single-angle S1, S2, mie_props.py:108-131."""
    if abs(u) > 1:
        raise ValueError("The cosine u must be between -1 and 1.")
    p, t = mie_pt(u, nmax)
    return s12_from_pt(an, bn, p, t)


def size_range(xarr, costarr, mr, mi):
    """MultipleMie.calculateS12SizeRange, mie_coated.py:61-89.

    Returns dict: 's1','s2' complex [N_x, N_ang]; 'qext','qsca','qabs','asy',
    'qb','qratio' float [N_x]; plus 'nmax' int [N_x].
    """
    eps = complex(mr, mi) ** 2
    nx = len(xarr)
    nang = len(costarr)
    out = {k: np.zeros(nx) for k in ("qext", "qsca", "qabs", "asy", "qb", "qratio")}
    out["s1"] = np.zeros((nx, nang), dtype=np.complex128)
    out["s2"] = np.zeros((nx, nang), dtype=np.complex128)
    out["nmax"] = np.zeros(nx, dtype=np.int64)
    nmaxall = max(nmax_of(float(v)) for v in xarr) if nx else 1
    P, T = pt_table(costarr, nmaxall)
    for i, xx in enumerate(xarr):
        xx = float(xx)
        an, bn, nmax = single_mie_coeff(eps, 1.0, xx)
        s1, s2 = s12_from_pt(an, bn, P, T)
        out["s1"][i] = s1
        out["s2"][i] = s2
        qext, qsca, qabs, qb, asy, qratio = mie_props_raw(an, bn, nmax, xx)
        out["qext"][i] = qext
        out["qsca"][i] = qsca
        out["qabs"][i] = qabs
        out["qb"][i] = qb
        out["asy"][i] = asy
        out["qratio"][i] = qratio
        out["nmax"][i] = nmax
    return out
