"""numpy restatement of the runoptics size grid, PSD and integration.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows src/geosmie/dointegration.py and src/geosmie/particleparams.py:
  getDR                 dointegration.py:93-101
  getXArrCarma          dointegration.py:103-153
  getLogNormPSD         particleparams.py:113-127
  lognormal PSD         dointegration.py:549-583 (calculatePSD, 'lognorm')
  calculateScatVals     dointegration.py:1044-1050
  integratePSD          dointegration.py:1064-1209 (mixing = 'dumix')
  post-integration      dointegration.py:950-1001
  angle grid            dointegration.py:739-742

Pinned by (tests/test_oracle_lut.py): closed-form lognormal moments, the
geometric grid ratio, a one-bin (delta) PSD that must return that bin's
values, the trapezoid normalisation int P11 sin(theta) = 2 checked against
scipy.integrate.trapezoid, the qb back-out identity, the Qsca-weight alias
of csca/cext (dointegration.py:1111-1112, 1154-1157) checked against an
independent weighted-mean computation, and golden vectors written by
running the reference's own integratePSD (tests/golden/make_lut_vectors.py).
"""
import numpy as np

SCATKEYS = ['p11', 'p12', 'p22', 'p33', 'p34', 'p44']          # dointegration.py:16
SCALARKEYS = ['qext', 'qsca', 'qabs', 'qb', 'g', 'csca', 'cext']  # dointegration.py:18


def angle_grid():
    """The 371-angle output grid in degrees (dointegration.py:739-742)."""
    a1 = np.linspace(0., 1., 100, endpoint=False)
    a2 = np.linspace(1., 10., 100, endpoint=False)
    a3 = np.linspace(10., 180., 171, endpoint=True)
    return np.concatenate([a1, a2, a3])


def getDR(arr):
    """Centred bin widths (dointegration.py:93-101)."""
    arr = np.asarray(arr, dtype=np.float64)
    d = np.empty_like(arr)
    d[0] = arr[1] - arr[0]
    d[1:-1] = ((arr[1:-1] - arr[:-2]) + (arr[2:] - arr[1:-1])) / 2.
    d[-1] = arr[-1] - arr[-2]
    return d


def getXArrCarma(minx, maxx, nbinperdecade):
    """CARMA volume-ratio grid and widths (dointegration.py:103-153)."""
    rat = maxx / minx
    nbin = np.log10(maxx / minx) * nbinperdecade
    rmRat = (rat ** 3) ** (1. / nbin)
    rMin = minx * ((1. + rmRat) / 2.) ** (1. / 3.)
    cpi = 4. / 3. * np.pi
    rvolmin = cpi * rMin ** 3.
    vrfact = ((3. / 2. / np.pi / (rmRat + 1)) ** (1. / 3.)) * (rmRat ** (1. / 3.) - 1.)
    rvol = [float(rvolmin) * float(rmRat) ** i for i in range(int(nbin))]
    return (np.array([(v / cpi) ** (1. / 3.) for v in rvol]),
            np.array([vrfact * v ** (1. / 3.) for v in rvol]))


def getLogNormPSD(rmode, sigma, xxArr, lambd, rmax, rmin):
    """Lognormal dN/dx in size-parameter space (particleparams.py:113-127)."""
    xconv = 2 * np.pi / lambd
    xmode, xmax, xmin = rmode * xconv, rmax * xconv, rmin * xconv
    xxArr = np.asarray(xxArr, dtype=np.float64)
    dNdx = 1. / (xxArr * (2 * np.pi) ** 0.5 * np.log(sigma)) * \
        np.exp(-(np.log(xxArr / xmode) ** 2) / (2. * np.log(sigma) ** 2))
    dNdx[xxArr >= xmax] = 0.
    dNdx[xxArr <= xmin] = 0.
    return dNdx


def lognorm_psd(rmode, sigma, xxarr, drarr, lam, rmax, rmin):
    """One lognormal mode as runoptics builds it (dointegration.py:569-583).

    Returns (psd normalised to sum 1, mass-effective radius).
    """
    thispsd = getLogNormPSD(rmode, sigma, xxarr, lam, rmax, rmin)
    thispsd = thispsd * drarr
    thispsd = thispsd / np.sum(thispsd)
    rr = xxarr * lam / 2. / np.pi
    return thispsd, np.sum(rr ** 4. * thispsd) / np.sum(rr ** 3. * thispsd)


def scat_vals(s1, s2):
    """Six Mueller planes [6, N_x, N_ang] (dointegration.py:1044-1050)."""
    a1 = np.abs(s1) ** 2
    a2 = np.abs(s2) ** 2
    out = np.empty((6,) + s1.shape)
    out[0] = 0.5 * (a1 + a2)
    out[1] = 0.5 * (a2 - a1)
    out[2] = 0.5 * (a1 + a2)
    out[3] = (s1 * np.conj(s2)).real
    out[4] = -(np.conj(s1) * s2).imag
    out[5] = (s1 * np.conj(s2)).real
    return out


def raw_mie(xxarr, lam, size_range_out):
    """rawMie (dointegration.py:1211-1254) from a size_range() result."""
    rr = np.asarray(xxarr) * lam / (2. * np.pi)
    v = size_range_out
    ret = {}
    sv = scat_vals(v['s1'], v['s2'])
    for i, k in enumerate(SCATKEYS):
        ret[k] = sv[i]
    ret['qext'] = np.asarray(v['qext'])
    ret['qsca'] = np.asarray(v['qsca'])
    ret['qabs'] = np.asarray(v['qabs'])
    ret['qb'] = np.asarray(v['qb'])
    ret['g'] = np.asarray(v['asy'])
    ret['csca'] = ret['qsca'] * np.pi * rr ** 2
    ret['cext'] = ret['qext'] * np.pi * rr ** 2
    return ret


def integrate_psd(xxarr, rawret, psd, fracs, lam, reff0, rhop0, rhop, alias=True):
    """integratePSD with mixing = 'dumix' (dointegration.py:1064-1209).

    rawret: list (one per mode) of dicts with SCATKEYS planes [N_x, N_ang]
    and SCALARKEYS arrays [N_x], keys in SCATKEYS + SCALARKEYS order (the
    order rawMie builds, :1245, which fixes where the Qsca weighting of key
    'g' starts to leak into later keys).  alias=False weights csca and cext
    by area only (the switch the device path offers).
    """
    rrarr = np.asarray(xxarr) * lam / (2. * np.pi)
    retkeys = SCATKEYS + SCALARKEYS
    ret = {k: 0. for k in ['num', 'area', 'volume', 'mass', 'rEff', 'rMass',
                            'bsca', 'bext', 'bbck', 'lidar_ratio']}
    for k in retkeys:
        ret[k] = np.zeros(rawret[0][k].shape[1]) if k in SCATKEYS else 0.
    for fi, frac in enumerate(fracs):
        rr = rawret[fi]
        w = psd[fi]
        th = {}
        num = np.sum(w)
        rarr2 = rrarr ** 2. * w
        rarr3 = rrarr ** 3. * w
        rarr4 = rrarr ** 4. * w
        th['num'] = num
        th['area'] = np.pi * np.sum(rarr2) / num
        th['volume'] = 4. / 3. * np.pi * np.sum(rarr3) / num
        th['mass'] = th['volume'] * rhop
        th['rEff'] = np.sum(rarr3) / np.sum(rarr2)
        reff_mass = np.sum(rarr4) / np.sum(rarr3)
        th['rMass'] = 4. / 3. * np.pi * rhop * reff_mass ** 3.
        rMass0 = 4. / 3. * np.pi * rhop0 * reff0[fi] ** 3
        weight = np.pi * rrarr ** 2.        # thisarea (aliased as thisweight)
        p11back = None
        for k in retkeys:
            if k == 'qb':
                beta = np.sum(np.asarray(rr['qsca']) * rarr2) * np.pi
                thisp11 = np.asarray(rr['qb']) * weight * (4. * np.pi)
                p11back = np.dot(thisp11.T, w) / beta
            if k == 'g':
                weight = weight * np.asarray(rr['qsca'])   # persists for later keys
            elif not alias and k in ('csca', 'cext'):
                # the weighting without the alias: area only, as for qext, qsca
                weight = np.pi * rrarr ** 2.
            sumarea = np.sum(w * weight)
            if k in SCATKEYS:
                th[k] = np.dot(rr[k].T, w)
            elif k == 'qb':
                th[k] = p11back * th['qsca'] / (4. * np.pi)
            else:
                th[k] = np.dot(np.asarray(rr[k]), w * weight) / sumarea
        mc = 1. / rhop / th['rEff'] * th['rMass'] / rMass0
        th['bsca'] = 3. / 4. * th['qsca'] * mc
        th['bext'] = 3. / 4. * th['qext'] * mc
        th['bbck'] = 3. / 4. * th['qb'] * mc / (4 * np.pi)
        th['lidar_ratio'] = 0.
        for k in th:
            ret[k] = ret[k] + frac * th[k]
    return ret


def post_integration(ret, ang_deg, mass0=None, rhop=None):
    """lidar ratio, ssa, P normalisation, pback (dointegration.py:950-988).

    Returns a new dict; `ang_deg` is the angle grid of the P planes.
    """
    out = dict(ret)
    qext, qsca, qb = float(ret['qext']), float(ret['qsca']), float(ret['qb'])
    out['lidar_ratio'] = qext / qb * 4 * np.pi
    out['ssa'] = qsca / qext
    theta = np.radians(ang_deg)
    p11 = np.asarray(ret['p11'])
    p11n = 2. * p11 / np.trapezoid(p11 * np.sin(theta), theta)
    for k in ['p12', 'p22', 'p33', 'p34', 'p44']:
        out[k] = np.asarray(ret[k]) * p11n / p11
    out['p11'] = p11n
    out['pback'] = np.array([out[k][-1] for k in ['p11', 'p12', 'p33', 'p34', 'p22', 'p44']])
    return out
