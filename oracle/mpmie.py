"""Arbitrary-precision Mie coefficients (mpmath), an independent truth source.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Written from Bohren & Huffman (1983), ch. 4 and 8: Riccati-Bessel functions
psi_n(z) = z j_n(z), chi_n(z) = -z y_n(z), xi_n = psi_n - i chi_n by the
three-term recurrence psi_{n+1} = (2n+1)/z psi_n - psi_{n-1} carried out at
``dps`` decimal digits (enough that the upward recurrence's loss of digits
beyond n > |z| is irrelevant), and the logarithmic derivative
D_n = psi_n'/psi_n by the downward recurrence started far above n_max.

The homogeneous a_n, b_n use the same algebraic form as
src/pymiecoated/pymiecoated/mie_coeffs.py:124-128; the coated ones the same
combination as mie_coeffs.py:230-249, which is algebraically exact, so at
high precision both give the mathematical coefficients.  Returned as
complex128 numpy arrays.
"""
import mpmath as mp
import numpy as np


def _riccati(z, nmax):
    """psi_0..psi_nmax, chi_0..chi_nmax of (complex or real) z at current dps."""
    psi = [mp.sin(z), mp.sin(z) / z - mp.cos(z)]
    chi = [mp.cos(z), mp.cos(z) / z + mp.sin(z)]
    for n in range(1, nmax):
        psi.append((2 * n + 1) / z * psi[n] - psi[n - 1])
        chi.append((2 * n + 1) / z * chi[n] - chi[n - 1])
    return psi, chi


def _dn(z, nmax, nstart=None):
    """D_1..D_nmax of z by downward recurrence from D_nstart = 0.

    ``nstart=None`` starts far above both nmax and |z| (converged D_n); an
    integer reproduces a given start index such as the reference's
    nmx = round(max(nmax, |z|) + 16) (mie_coeffs.py:101, 118-121).
    """
    if nstart is None:
        nstart = int(nmax + 200 + abs(complex(z)))
    d = mp.mpc(0)
    out = [None] * (nmax + 1)
    for n in range(nstart, 0, -1):
        r = mp.mpf(n) / z
        d = r - 1 / (d + r)          # d now holds D_{n-1}
        if n - 1 <= nmax:
            out[n - 1] = d
    return out


def _dps_for(zabs, nmax):
    # upward recurrence of psi loses ~ (nmax - |z|) * a few digits at most
    return 40 + int(0.6 * max(0, nmax - zabs)) + int(zabs ** (1.0 / 3.0) * 6)


def homog(m, x, nmax=None, mu=1.0, start="converged"):
    """a_n, b_n (n = 1..nmax) of a homogeneous sphere.

    m is the complex refractive index sqrt(eps/mu); the D_n argument is
    z = sqrt(eps mu) x = m mu x (mie_coeffs.py:96-97).  ``start`` is
    "converged" (mathematical coefficients) or "reference" (D_n started at the
    reference's nmx with D = 0).
    """
    x = float(x)
    if nmax is None:
        nmax = int(round(2 + x + 4 * x ** (1.0 / 3.0)))
    zc = complex(m) * complex(mu) * x
    nstart = None
    if start == "reference":
        nstart = int(round(max(nmax, abs(zc)) + 16))
    with mp.workdps(_dps_for(abs(zc), nmax)):
        mm = mp.mpc(m)
        xx = mp.mpf(x)
        psi, chi = _riccati(xx, nmax)
        xi = [p - 1j * c for p, c in zip(psi, chi)]
        dn = _dn(mm * mp.mpc(mu) * xx, nmax, nstart)
        an = np.zeros(nmax, dtype=np.complex128)
        bn = np.zeros(nmax, dtype=np.complex128)
        for n in range(1, nmax + 1):
            da = dn[n] / mm + n / xx
            db = dn[n] * mm + n / xx
            an[n - 1] = complex((da * psi[n] - psi[n - 1]) / (da * xi[n] - xi[n - 1]))
            bn[n - 1] = complex((db * psi[n] - psi[n - 1]) / (db * xi[n] - xi[n - 1]))
    return an, bn, nmax


def coated(m1, m2, x, y, nmax=None, start="converged"):
    """a_n, b_n of a coated sphere (core index m1, radius x; shell m2, radius y).

    ``start="reference"`` starts all three D_n recurrences at the reference's
    nmx = round(max(nmax, |m1 y|, |m2 y|) + 16) (mie_coeffs.py:203-218).
    """
    x = float(x)
    y = float(y)
    if nmax is None:
        nmax = int(round(2 + y + 4 * y ** (1.0 / 3.0)))
    zabs = max(abs(complex(m1)), abs(complex(m2))) * y
    nstart = None
    if start == "reference":
        nstart = int(round(max(nmax, zabs) + 16))
    with mp.workdps(_dps_for(zabs, nmax) + int(abs(complex(m2).imag) * y / 2.3)):
        M1 = mp.mpc(m1)
        M2 = mp.mpc(m2)
        X = mp.mpf(x)
        Y = mp.mpf(y)
        m = M2 / M1
        u, v, w = M1 * X, M2 * X, M2 * Y
        dnu, dnv, dnw = (_dn(u, nmax, nstart), _dn(v, nmax, nstart),
                         _dn(w, nmax, nstart))
        pv, chv = _riccati(v, nmax)
        pw, chw = _riccati(w, nmax)
        py, chy = _riccati(Y, nmax)
        an = np.zeros(nmax, dtype=np.complex128)
        bn = np.zeros(nmax, dtype=np.complex128)
        for n in range(1, nmax + 1):
            uu = m * dnu[n] - dnv[n]
            vv = dnu[n] / m - dnv[n]
            fv = pv[n] / chv[n]
            pt = pw[n] - chw[n] * fv
            prat = pw[n] / pv[n] / chv[n]
            dns = (uu * fv / pw[n]) / (uu * pt + prat) + dnw[n]
            gns = (vv * fv / pw[n]) / (vv * pt + prat) + dnw[n]
            a1 = dns / M2 + n / Y
            b1 = M2 * gns + n / Y
            gsy = py[n] - 1j * chy[n]
            gs1y = py[n - 1] - 1j * chy[n - 1]
            an[n - 1] = complex((py[n] * a1 - py[n - 1]) / (gsy * a1 - gs1y))
            bn[n - 1] = complex((py[n] * b1 - py[n - 1]) / (gsy * b1 - gs1y))
    return an, bn, nmax


def s12(an, bn, cost, dps=40):
    """S1, S2 at each cosine, summed in mpmath (Bohren & Huffman eq. 4.74).

    pi_n, tau_n by the upward recurrences of B&H eq. 4.47 (pi_0 = 0, pi_1 = 1,
    pi_n = (2n-1)/(n-1) mu pi_{n-1} - n/(n-1) pi_{n-2}, tau_n = n mu pi_n -
    (n+1) pi_{n-1}), stable for every |mu| <= 1; the sums
    S1 = sum (2n+1)/(n(n+1)) (a_n pi_n + b_n tau_n), S2 likewise with pi and
    tau exchanged, the definition src/pymiecoated/pymiecoated/mie_props.py:133-149
    evaluates.  an, bn are complex128 (or mpc) arrays; returns complex128
    arrays [len(cost)].
    """
    s1 = np.zeros(len(cost), dtype=np.complex128)
    s2 = np.zeros(len(cost), dtype=np.complex128)
    with mp.workdps(dps):
        A = [mp.mpc(complex(a)) for a in an]
        B = [mp.mpc(complex(b)) for b in bn]
        for i, c in enumerate(cost):
            mu = mp.mpf(float(c))
            pm1, p = mp.mpf(0), mp.mpf(1)
            t1, t2 = mp.mpc(0), mp.mpc(0)
            for n in range(1, len(A) + 1):
                if n > 1:
                    pm1, p = p, ((2 * n - 1) * mu * p - n * pm1) / (n - 1)
                tau = n * mu * p - (n + 1) * pm1
                f = mp.mpf(2 * n + 1) / (n * (n + 1))
                t1 += f * (A[n - 1] * p + B[n - 1] * tau)
                t2 += f * (A[n - 1] * tau + B[n - 1] * p)
            s1[i] = complex(t1)
            s2[i] = complex(t2)
    return s1, s2
