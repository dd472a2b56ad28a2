"""numpy restatement of the GSF (generalised spherical function) expansion.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows src/gsf/spher_expan.f (Mishchenko's expansion program) as driven by
src/gsf/convertncdf.py:170-189, 319-394 in 'pygeos' mode:
  GAUSS        spher_expan.f:520-579   Gauss-Legendre nodes/weights (Newton)
  LINTERPOL    spher_expan.f:593-623   linear interpolation, extrapolating ends
  one_calc     spher_expan.f:157-169   F(acos x_i) at NG = NSPHER = 129 nodes
  GENER        spher_expan.f:363-407   P^{00}, P^{22}, P^{2-2}, P^{02} recurrences
  SPHER_EXPAN  spher_expan.f:269-358   alpha1..4, beta1..2
  MATR         spher_expan.f:419-517   matrix from coefficients (round trip)
  output       spher_expan.f:93-107    normalised by CNORM = 1/alpha1(0), F17.10
  hbleg        convertncdf.py:32-43    Legendre series evaluator

Pinned by (tests/test_oracle_gsf.py): vectors from the reference program
itself (oracle/_ref/spher_expan.x built from src/gsf/spher_expan.f by
oracle/Makefile; committed in tests/golden/gsf_vectors.npz), numpy's
leggauss, closed forms (Rayleigh: alpha1 = (1, 0, 1/2); Henyey-Greenstein:
alpha1_l = (2l+1) g^l) and MATR round trips.
"""
import math

import numpy as np

NSPHER = 129                      # params.h:14
D2R = math.acos(-1.0) / 180.0     # params.h:3-4


def gauss(n):
    """Nodes z (ascending) and weights w on [-1, 1] (GAUSS with IND1=0, IND2=0)."""
    z = np.zeros(n)
    w = np.zeros(n)
    ind = n % 2
    k = n // 2 + ind
    f = float(n)
    x = 0.0
    for i in range(1, k + 1):
        m = n + 1 - i                      # Fortran 1-based index of the node
        if i == 1:
            x = 1.0 - 2.0 / ((f + 1.0) * f)
        if i == 2:
            x = (z[n - 1] - 1.0) * 4.0 + z[n - 1]
        if i == 3:
            x = (z[n - 2] - z[n - 1]) * 1.6 + z[n - 2]
        if i > 3:
            x = (z[m] - z[m + 1]) * 3.0 + z[m + 2]
        if i == k and ind == 1:
            x = 0.0
        niter = 0
        check = 1e-16
        while True:
            pb = 1.0
            niter += 1
            if niter > 100:
                check *= 10.0
            pc = x
            dj = 1.0
            for _ in range(2, n + 1):
                dj += 1.0
                pa = pb
                pb = pc
                pc = x * pb + (x * pb - pa) * (dj - 1.0) / dj
            pa = 1.0 / ((pb - x * pc) * f)
            pb = pa * pc * (1.0 - x * x)
            x = x - pb
            if not abs(pb) > check * abs(x):
                break
        z[m - 1] = x
        w[m - 1] = 2.0 * pa * pa * (1.0 - x * x)
        if i == k and ind == 1:
            continue
        z[i - 1] = -z[m - 1]
        w[i - 1] = w[m - 1]
    return z, w


def linterpol(xx, yy, x):
    """LINTERPOL with IEX = 1 (extrapolate outside the grid)."""
    nn = len(xx)
    if x < xx[0]:
        return (yy[0] - yy[1]) / (xx[0] - xx[1]) * (x - xx[0]) + yy[0]
    if x > xx[nn - 1]:
        return (yy[nn - 1] - yy[nn - 2]) / (xx[nn - 1] - xx[nn - 2]) * (x - xx[nn - 1]) + yy[nn - 1]
    i = 1
    while i < nn and not (xx[i] > x):
        i += 1
    if i == nn:
        # x equal to the last node: the program's loop runs past the grid and
        # reads the zeros beyond it in its COMMON arrays, so the slope term is
        # multiplied by x - xx[nn-1] = 0 and the last value comes back
        return yy[nn - 1]
    return (yy[i] - yy[i - 1]) / (xx[i] - xx[i - 1]) * (x - xx[i - 1]) + yy[i - 1]


def gener(u, l1max):
    """P1 (P^00_l), P2 (P^22_l), P3 (P^2-2_l), P4 (P^02_l) for l = 0..l1max-1."""
    p1 = np.zeros(l1max)
    p2 = np.zeros(l1max)
    p3 = np.zeros(l1max)
    p4 = np.zeros(l1max)
    dup, dum, du = 1.0 + u, 1.0 - u, u * u
    d6 = 0.25 * math.sqrt(6.0)
    p1[0], p1[1], p1[2] = 1.0, u, 0.5 * (3.0 * du - 1.0)
    p2[2] = 0.25 * dup * dup
    p3[2] = 0.25 * dum * dum
    p4[2] = d6 * (du - 1.0)
    for L1 in range(3, l1max):             # Fortran L1 = 3..LMAX, fills index L1+1
        L = L1 - 1
        c1 = 1.0 / (L + 1)
        c2 = float(2 * L + 1)
        c3 = 1.0 / math.sqrt(float((L + 1) * (L + 1) - 4))
        c4 = math.sqrt(float(L * L - 4))
        c5 = 1.0 / (float(L) * float((L + 1) * (L + 1) - 4))
        c6 = float(2 * L + 1) * float(L * (L + 1))
        c7 = float((2 * L + 1) * 4)
        c8 = float(L + 1) * float(L * L - 4)
        cu1, cu2, dl = c2 * u, c6 * u, float(L1 - 1)
        # 0-based: new index L1, current L1-1, previous L1-2
        p1[L1] = c1 * (cu1 * p1[L1 - 1] - dl * p1[L1 - 2])
        p2[L1] = c5 * ((cu2 - c7) * p2[L1 - 1] - c8 * p2[L1 - 2])
        p3[L1] = c5 * ((cu2 + c7) * p3[L1 - 1] - c8 * p3[L1 - 2])
        p4[L1] = c3 * (cu1 * p4[L1 - 1] - c4 * p4[L1 - 2])
    return p1, p2, p3, p4


def spher_expan(x, w, f11, f22, f33, f44, f12, f34, l1max):
    """Expansion coefficients (unnormalised), SPHER_EXPAN."""
    al = np.zeros((6, l1max))
    for i in range(len(x)):
        p1, p2, p3, p4 = gener(x[i], l1max)
        wi = w[i]
        ff11, ff22, ff33 = f11[i] * wi, f22[i] * wi, f33[i] * wi
        ff44, ff12, ff34 = f44[i] * wi, f12[i] * wi, f34[i] * wi
        al[0] += ff11 * p1
        al[3] += ff44 * p1
        al[1] += (ff22 + ff33) * p2
        al[2] += (ff22 - ff33) * p3
        al[4] += ff12 * p4
        al[5] += ff34 * p4
    cl = np.arange(l1max) + 0.5
    out = np.empty_like(al)
    out[0] = al[0] * cl
    a2 = al[1] * cl * 0.5
    a3 = al[2] * cl * 0.5
    out[1] = a2 + a3
    out[2] = a2 - a3
    out[3] = al[3] * cl
    out[4] = al[4] * cl
    out[5] = al[5] * cl
    return out


def expand(ang_deg, f11, f22, f33, f44, f12, f34, ng=NSPHER, quantize=True):
    """spher_expan.x on one phase matrix; rows alpha1..4, beta1..2 [6, ng].

    Inputs in the program's column order (angle, F11, F22, F33, F44, F12,
    F34; convertncdf.py:177).  Normalised by 1/alpha1(0) and, with
    quantize=True, rounded to 10 decimals as the F17.10 text boundary does.
    Returns (coeffs, cnorm).
    """
    angl = np.asarray(ang_deg, dtype=np.float64) * D2R
    x, w = gauss(ng)
    F = [np.asarray(v, dtype=np.float64) for v in (f11, f22, f33, f44, f12, f34)]
    Fn = [np.array([linterpol(angl, Fi, math.acos(xi)) for xi in x]) for Fi in F]
    al = spher_expan(x, w, *Fn, l1max=ng)
    cnorm = 1.0 / al[0, 0]
    al = al * cnorm
    if quantize:
        al = np.vectorize(lambda v: float("%.10f" % v))(al)
    return al, cnorm


def matr(al, ang_rad):
    """F11, F22, F33, F44, F12, F34 at angles from coefficients (MATR)."""
    a1, a2, a3, a4, b1, b2 = al
    l1max = al.shape[1]
    lmax = l1max - 1
    d6 = math.sqrt(6.0) * 0.25
    out = np.zeros((6, len(ang_rad)))
    for j, t in enumerate(ang_rad):
        u = math.cos(t)
        f11 = f2 = f3 = f44 = f12 = f34 = 0.0
        p1 = p2 = p3 = p4 = 0.0
        pp1 = 1.0
        pp2 = 0.25 * (1 + u) * (1 + u)
        pp3 = 0.25 * (1 - u) * (1 - u)
        pp4 = d6 * (u * u - 1)
        for L1 in range(1, l1max + 1):
            L = L1 - 1
            dl, dl1 = float(L), float(L1)
            f11 += a1[L1 - 1] * pp1
            f44 += a4[L1 - 1] * pp1
            if L != lmax:
                pl1 = float(2 * L + 1)
                p = (pl1 * u * pp1 - dl * p1) / dl1
                p1, pp1 = pp1, p
            if L < 2:
                continue
            f2 += (a2[L1 - 1] + a3[L1 - 1]) * pp2
            f3 += (a2[L1 - 1] - a3[L1 - 1]) * pp3
            f12 += b1[L1 - 1] * pp4
            f34 += b2[L1 - 1] * pp4
            if L == lmax:
                continue
            pl1 = float(2 * L + 1)
            pl2 = dl * dl1 * u
            pl3 = dl1 * (dl * dl - 4.0)
            pl4 = 1.0 / (dl * (dl1 * dl1 - 4.0))
            p = (pl1 * (pl2 - 4.0) * pp2 - pl3 * p2) * pl4
            p2, pp2 = pp2, p
            p = (pl1 * (pl2 + 4.0) * pp3 - pl3 * p3) * pl4
            p3, pp3 = pp3, p
            p = (pl1 * u * pp4 - math.sqrt(dl * dl - 4.0) * p4) / math.sqrt(dl1 * dl1 - 4.0)
            p4, pp4 = pp4, p
        out[:, j] = [f11, (f2 + f3) * 0.5, (f2 - f3) * 0.5, f44, f12, f34]
    return out


def alt_angles(ang_deg):
    """The program's second check grid (READMATRIX, spher_expan.f:240-247):
    first and last angles kept, midpoints of each angle with its lower
    neighbour in the first half and its upper neighbour in the second."""
    a = np.asarray(ang_deg, dtype=np.float64)
    n = a.size
    alt = a.copy()
    for i in range(2, n // 2 + 1):            # Fortran i = 2 .. NANG/2
        alt[i - 1] = 0.5 * (a[i - 2] + a[i - 1])
    for i in range(n // 2 + 1, n):            # Fortran i = NANG/2+1 .. NANG-1
        alt[i - 1] = 0.5 * (a[i] + a[i - 1])
    return alt


def fit(ang_deg, f11, f22, f33, f44, f12, f34, ng=NSPHER):
    """one_calc's fit check (spher_expan.f:157-177) for one matrix: the
    unnormalised coefficients' MATR at the input angles (the '.expan_matr'
    rows F11..F34) and the fit error max(MAXABS at the input angles, MAXABS at
    the midpoint grid against the linearly interpolated F11) (ERREVAL,
    :635-687, ERRTYP = MAXABS, 0-180 degrees).  Returns (matr [6, N], fiterr)."""
    angl = np.asarray(ang_deg, dtype=np.float64) * D2R
    x, w = gauss(ng)
    F = [np.asarray(v, dtype=np.float64) for v in (f11, f22, f33, f44, f12, f34)]
    Fn = [np.array([linterpol(angl, Fi, math.acos(xi)) for xi in x]) for Fi in F]
    al = spher_expan(x, w, *Fn, l1max=ng)
    alt = alt_angles(ang_deg) * D2R
    f11alt = np.array([linterpol(angl, F[0], t) for t in alt])
    m_alt = matr(al, alt)
    m = matr(al, angl)
    err = max(np.max(np.abs(F[0] - m[0])), np.max(np.abs(f11alt - m_alt[0])))
    return m, err


def hbleg(x, pmoms):
    """Legendre series sum_l pmoms[l] P_l(x) (convertncdf.py:32-43)."""
    x = np.asarray(x, dtype=np.float64)
    L = len(pmoms)
    leg = np.zeros((max(L, 2),) + x.shape)
    leg[0] = 1.0
    leg[1] = x
    for m in range(2, L):
        leg[m] = 1.0 / m * ((2.0 * m - 1.0) * x * leg[m - 1] - (m - 1.0) * leg[m - 2])
    ret = np.zeros_like(x)
    for m in range(L):
        ret = ret + pmoms[m] * leg[m]
    return ret


# convertncdf.py:381-394: program row (alpha1, alpha2, alpha3, alpha4, beta1,
# beta2) -> table p index in the order p = [11, 12, 33, 34, 22, 44]
PMOM_ORDER = [0, 4, 2, 5, 1, 3]


def pmom_from_table_slab(ang_deg, p11, p12, p33, p34, p22, p44, ng=NSPHER):
    """pmom[p, m] for one (bin, wavelength, rh) slab, convertncdf.py:173-189, 381-394."""
    al, _ = expand(ang_deg, p11, p22, p33, p44, p12, p34, ng=ng)
    out = np.zeros((6, ng))
    for ii, npol in enumerate(PMOM_ORDER):
        out[npol] = al[ii]
    return out
