"""Particle description front end of runoptics (host side, before the GPU path).

Drop-ins for the readers and humidity rules runoptics applies before the Mie
batch of every (bin, wavelength, rh) entry:

  getPPJSON, getM, getMSep, getWaterM, humidityGrowth
                          src/geosmie/particleparams.py:18-108
  wtpct, dens, grow_v75   src/geosmie/carma_utils.py:136-314 (CARMA sulfate
                          growth, rhDep type 'su')
  initializeXarr          src/geosmie/dointegration.py:425-447 (lognorm, ss)
  getHumidRefractiveIndex src/geosmie/dointegration.py:493-536
  calculatePSD            src/geosmie/dointegration.py:539-622 (lognorm, ss)

The size grid helpers getXArrCarma, getDR and getLogNormPSD live in
geosmie_amd.lut.  Relative paths in a particle file (refractive-index
tables, the water table) resolve against `data_dir` (runoptics resolves them
against its working directory).
"""
import json
import os

import numpy as np

from .lut import getDR, getLogNormPSD, getXArrCarma


# ---------------------------------------------------------------------------
# readers
# ---------------------------------------------------------------------------
def _resolve(path, data_dir):
    if data_dir is None or os.path.isabs(path):
        return path
    return os.path.join(data_dir, path)


def getM(fn):
    """OPAC 'gads' table: 61 rows after 17 header lines; columns wavelength
    [um], ref.real, ref.imag (negative); wavelength returned in m
    (particleparams.py:67-70)."""
    lam, mr, mi = np.loadtxt(fn, skiprows=17, max_rows=61, comments=None, unpack=True,
                             usecols=[1, 8, 9])
    return np.array([lam * 1e-6, mr, mi])


def getMSep(fn, sep):
    """Separated table 'wavelength[um], refreal, refimag' (particleparams.py:73-76)."""
    d = np.loadtxt(fn, unpack=True, delimiter=sep)
    d[0] = d[0] * 1e-6
    return d


def getWaterM(fn="data/refrac.water.txt", data_dir=None):
    """Water refractive index (columns 0-2 after 13 header lines; particleparams.py:79-83)."""
    lam, mr, mi = np.loadtxt(_resolve(fn, data_dir), skiprows=13, unpack=True, usecols=[0, 1, 2])
    return np.array([lam * 1e-6, mr, mi])


def getPPJSON(partid, data_dir=None):
    """Particle JSON plus its refractive-index tables as 'mList' (particleparams.py:18-56)."""
    with open(partid) as fp:
        data = json.load(fp)
    ri = data["ri"]
    readers = {"gads": getM,
               "csv": lambda p: getMSep(p, ","),
               "wsv": lambda p: getMSep(p, None)}
    if ri["format"] not in readers:
        raise ValueError("refractive index format %s not yet supported" % ri["format"])
    data["mList"] = [readers[ri["format"]](_resolve(p, data_dir)) for p in ri["path"]]
    psd = data["psd"]
    if psd["type"] == "du":   # promote scalar sub-bin limits to one-element lists (:39-55)
        pp = psd["params"]
        pairs = [(a, b) if isinstance(a, list) else ([a], [b])
                 for a, b in zip(pp["rMinMaj"], pp["rMaxMaj"])]
        pp["rMinMaj"] = [a for a, _ in pairs]
        pp["rMaxMaj"] = [b for _, b in pairs]
    return data


# ---------------------------------------------------------------------------
# humidity growth
# ---------------------------------------------------------------------------
# Tabazadeh H2SO4 weight-percent fit: (a1, b1, c1, d1, a2, b2, c2, d2) for the
# low (< 0.05), middle and high (> 0.85) activity ranges (carma_utils.py:146-173)
_WTP_FIT = (
    (12.37208932, -0.16125516114, -30.490657554, -2.1133114241,
     13.455394705, -0.1921312255, -34.285174607, -1.7620073078),
    (11.820654354, -0.20786404244, -4.807306373, -5.1727540348,
     12.891938068, -0.23233847708, -6.4261237757, -4.9005471319),
    (-180.06541028, -0.38601102592, -93.317846778, 273.88132245,
     -176.95814097, -0.36257048154, -90.469744201, 267.45509988),
)


def wtpct(relhum, temp=220.):
    """H2SO4 weight percent of sulfate aerosol in equilibrium with relhum at temp
    (Tabazadeh fit, carma_utils.py:136-182)."""
    a = relhum
    if a < 0.05:
        a = max(a, 1e-6)
        c = _WTP_FIT[0]
    elif a <= 0.85:
        c = _WTP_FIT[1]
    else:
        a = min(a, 1.0)
        c = _WTP_FIT[2]
    lo = c[0] * a ** c[1] + c[2] * a + c[3]
    hi = c[4] * a ** c[5] + c[6] * a + c[7]
    t = lo + (hi - lo) * ((temp - 190.) / 70.)
    w = 100. * t * 98. / (t * 98. + 1000.)
    return min(max(w, 1.), 100.)


# density of H2SO4 solution: rho = c0 + c1 * T at these weight percents (carma_utils.py:199-224)
_DEN_WTP = np.array([0., 1., 5., 10., 20., 25., 30., 35., 40., 41., 45., 50., 53., 55., 56., 60.,
                     65., 66., 70., 72., 73., 74., 75., 76., 78., 79., 80., 81., 82., 83., 84.,
                     85., 86., 87., 88., 89., 90., 91., 92., 93., 94., 95., 96., 97., 98., 100.])
_DEN_C0 = np.array([1., 1.13185, 1.17171, 1.22164, 1.3219, 1.37209, 1.42185, 1.4705, 1.51767,
                    1.52731, 1.56584, 1.61834, 1.65191, 1.6752, 1.68708, 1.7356, 1.7997, 1.81271,
                    1.86696, 1.89491, 1.9092, 1.92395, 1.93904, 1.95438, 1.98574, 2.00151,
                    2.01703, 2.03234, 2.04716, 2.06082, 2.07363, 2.08461, 2.09386, 2.10143,
                    2.10764, 2.11283, 2.11671, 2.11938, 2.12125, 2.1219, 2.12723, 2.12654,
                    2.12621, 2.12561, 2.12494, 2.12093])
_DEN_C1 = np.array([0., -0.000435022, -0.000479481, -0.000531558, -0.000622448, -0.000660866,
                    -0.000693492, -0.000718251, -0.000732869, -0.000735755, -0.000744294,
                    -0.000761493, -0.000774238, -0.00078392, -0.000788939, -0.00080946,
                    -0.000839848, -0.000845825, -0.000874337, -0.000890074, -0.00089873,
                    -0.000908778, -0.000920012, -0.000932184, -0.000959514, -0.000974043,
                    -0.000988264, -0.00100258, -0.00101634, -0.00102762, -0.00103757,
                    -0.00104337, -0.00104563, -0.00104458, -0.00104144, -0.00103719,
                    -0.00103089, -0.00102262, -0.00101355, -0.00100249, -0.00100934,
                    -0.000998299, -0.000990961, -0.000985845, -0.000984529, -0.000989315])


def dens(relhum, temp=220.):
    """Sulfate solution density [g cm-3]: linear in weight percent between the
    table points bracketing wtpct(relhum, temp) (carma_utils.py:186-237)."""
    w = wtpct(relhum, temp=temp)
    i = 0
    while w > _DEN_WTP[i]:
        i += 1
    d_hi = _DEN_C0[i] + _DEN_C1[i] * temp
    if i == 0 or w == _DEN_WTP[i]:
        return d_hi
    d_lo = _DEN_C0[i - 1] + _DEN_C1[i - 1] * temp
    f = (_DEN_WTP[i] - w) / (_DEN_WTP[i] - _DEN_WTP[i - 1])
    return d_lo * f + d_hi * (1.0 - f)


def grow_v75(relhum, rd, temp=220.):
    """Wet/dry radius ratio of sulfate with the Kelvin correction of water
    activity (carma_utils.py:241-314).  rd dry radius [m]."""
    rdry = rd * 100.                       # cm
    rho_dry, mw_acid, rgas = 1.923, 98., 8.31447e7
    # saturation vapour pressure over water (Pa) and water mass concentration (g cm-3)
    es = 611. * np.exp(2.501e6 / 461. * (1. / 273.16 - 1. / temp))
    n_v = relhum * es / (1.38e-23 * temp)
    h2o = n_v / 6.022e23 * 0.018 * 1000. / 1.e6
    # Kelvin term at an assumed 80 wt % solution
    w_k = 80.
    d79 = 2.00151 - 0.000974043 * temp
    d80 = 2.01703 - 0.000988264 * temp
    ddw = d80 - d79
    s79 = 79.3556 - 0.0267212 * temp
    s86 = 75.608 - 0.0269204 * temp
    dsw = (s86 - s79) / (85.9195 - 79.432)
    sig = s79 + dsw * (80.0 - 79.432)
    rwet = rdry * (100. * rho_dry / w_k / d80) ** (1. / 3.)
    kb = 1. + w_k * ddw / d80 - 3. * w_k * dsw / (2. * sig)
    ka = 2. * mw_acid * sig / (d79 * rgas * temp * rwet)
    h2o_k = h2o / np.exp(ka * kb)
    # effective humidity after the Kelvin correction; the final weight percent
    # is taken at the default 220 K, the density at temp (as the source does)
    rh_eff = h2o_k * 6.022e23 / 18. * 1.3807e-16 * temp / (es * 10.)
    rho_wet = dens(rh_eff, temp=temp)
    rwet = rdry * (100. * rho_dry / wtpct(rh_eff) / rho_wet) ** (1. / 3.)
    return rwet / rdry


def humidityGrowth(params, siz0, rh, allrh):
    """Humidified size of siz0 at relative humidity rh (particleparams.py:85-108).
    params is the particle file's 'rhDep' block; the growth factor of 'simple'
    is looked up at the first position of rh in allrh."""
    rhi = list(allrh).index(rh)
    kind, p = params["type"], params["params"]
    if kind in ("simple", "trivial"):
        return siz0 * p["gf"][rhi]
    if kind == "ss":      # Gerber (1985), cm inside
        if rh == 0.0:
            return siz0
        s = siz0 * 100
        return (p["c1"] * s ** p["c2"] / (p["c3"] * s ** p["c4"] - np.log10(rh)) + s ** 3.) \
            ** (1. / 3.) / 100.
    if kind == "su":
        if rh == 0.0:
            return siz0
        return siz0 * float(grow_v75(rh, siz0, temp=p["temp"]))
    raise ValueError("unknown rhDep type %s" % kind)


# ---------------------------------------------------------------------------
# size grid, humid refractive index, PSD
# ---------------------------------------------------------------------------
def initializeXarr(params, radind, minlam, maxlam):
    """Size-parameter grid and widths of one bin (dointegration.py:425-447)."""
    pp, kind = params["psd"]["params"], params["psd"]["type"]
    if kind == "lognorm":
        lo = pp["rmin0"][radind][0] * 2 * np.pi / maxlam
        hi = pp["rmax0"][radind][-1] * 2 * np.pi / minlam * 3.0
    elif kind == "ss":
        lo = pp["rMinMaj"][radind] * 2 * np.pi / maxlam
        hi = pp["rMaxMaj"][radind] * 2 * np.pi / minlam * 10
    elif kind == "du":
        # the reference raises the log10 values to the tenth power (not ten to
        # them), so the 1000-point grid is not monotonic: it falls from
        # log10(minx)**10 to ~0 where the logarithm changes sign and rises to
        # log10(maxx)**10; kept as written (:456-462).  No widths: the du PSD
        # takes them from the grid itself.
        lo = pp["rMinMaj"][radind][0] * 2 * np.pi / maxlam
        hi = pp["rMaxMaj"][radind][-1] * 2 * np.pi / minlam
        return (np.linspace(np.log10(lo), np.log10(hi), 1000)) ** 10., None
    else:
        raise ValueError("psd type %s is not supported by the Mie table path" % kind)
    if "numperdec" not in pp:
        raise ValueError("numperdec missing from dict")
    return getXArrCarma(lo, hi, pp["numperdec"][radind])


def getHumidRefractiveIndex(params, radind, rhi, rh, nref0, nrefwater):
    """(mr, mi, gf, rrat) of the wet particle: volume mixing with water,
    n = n_w + (n_0 - n_w) rrat^3, rrat = dry / wet radius (dointegration.py:493-536)."""
    kind = params["rhDep"]["type"]
    if kind == "simple" or (kind == "trivial" and rhi == 0):
        gf = params["rhDep"]["params"]["gf"][rhi]
        rrat = 1. / gf
    elif kind in ("ss", "su"):
        pp = params["psd"]["params"]
        if kind == "ss" and "rMinMaj" in pp:
            r0 = pp["rMinMaj"][radind]
        else:
            r0 = pp["rmin0"][radind][0]
        rrat = r0 / humidityGrowth(params["rhDep"], r0, rh[rhi], rh)
        gf = 1. / rrat
    else:
        raise ValueError("PROBLEM! NO RHDEP DEFINED!")
    n = [nrefwater + (n0 - nrefwater) * rrat ** 3. for n0 in nref0]
    return [v.real for v in n], [v.imag for v in n], gf, rrat


def lognorm_psd_rows(params, radind, rhs, rh, xxarr, drarr, lam):
    """calculatePSD's lognormal weights for many rh at once: psd[mode][i] is
    calculatePSD(params, radind, rhs[i], ...)[0][mode], bit for bit (the same
    element-wise numpy operations and one row-wise sum per rh), for the table
    loop, which needs all rh of a wavelength (dointegration.py:549-583)."""
    pp, rd = params["psd"]["params"], params["rhDep"]
    if params["psd"]["type"] != "lognorm":
        raise ValueError("lognorm_psd_rows: psd type must be lognorm")
    xx = np.asarray(xxarr, dtype=np.float64)
    k = 2 * np.pi / lam
    out = []
    for m, r0 in enumerate(pp["r0"][radind]):
        rmode = np.array([humidityGrowth(rd, r0, r, rh) for r in rhs], np.float64)
        rmax = np.array([humidityGrowth(rd, pp["rmax0"][radind][m], r, rh) for r in rhs],
                        np.float64)
        rmin = pp["rmin0"][radind][m]
        ls = np.log(pp["sigma"][radind][m])
        # getLogNormPSD (particleparams.py:113-127), one row per rh
        d = 1. / (xx * (2 * np.pi) ** 0.5 * ls) * np.exp(
            -(np.log(xx / (rmode * k)[:, None]) ** 2) / (2. * ls ** 2))
        d = np.where(xx >= (rmax * k)[:, None], 0., d)
        d = np.where(xx <= rmin * k, 0., d)
        w = d * drarr
        out.append(w / np.sum(w, axis=1, keepdims=True))
    return out


def calculatePSD(params, radind, onerh, rh, xxarr, drarr, rrat, lam, want_ref=True):
    """Number weights per mode on the size grid, mass-weighted effective radius
    per mode, and the bin edges (rLow, rUp) (dointegration.py:539-622).
    want_ref=False skips the effective radii (returned as None), which the
    table loop only needs at the dry state."""
    pp, kind, rd = params["psd"]["params"], params["psd"]["type"], params["rhDep"]
    rarr = xxarr / (2 * np.pi / lam)
    rr = xxarr * lam / 2. / np.pi
    psd, ref = [], []
    if kind == "lognorm":
        rmodes = [humidityGrowth(rd, r0, onerh, rh) for r0 in pp["r0"][radind]]
        rmax0 = list(pp["rmax0"][radind])
        rmaxs = [humidityGrowth(rd, r, onerh, rh) for r in rmax0]
        rmins = list(pp["rmin0"][radind])        # rmin does not grow
        rLow, rUp = rmins[0], rmax0[0]
        for k in range(len(rmodes)):
            w = getLogNormPSD(rmodes[k], pp["sigma"][radind][k], xxarr, lam, rmaxs[k], rmins[k])
            w = w * drarr
            w = w / np.sum(w)
            psd.append(w)
            if want_ref:
                ref.append(np.sum(rr ** 4. * w) / np.sum(rr ** 3. * w))
    elif kind == "ss":
        rmin_d, rmax_d = pp["rMinMaj"][radind], pp["rMaxMaj"][radind]
        rLow, rUp = rmin_d, rmax_d
        rmin_w = humidityGrowth(rd, rmin_d, onerh, rh)
        rmax_w = humidityGrowth(rd, rmax_d, onerh, rh)
        rrat = rmin_d / rmin_w
        dr = getDR(rarr)
        # sea-salt number distribution in r80 space (dointegration.py:600-613)
        k80 = 1.65 * rrat
        r80 = rarr * k80 * 1e6
        a = 4.7 * (1. + 30. * r80) ** (-0.017 * r80 ** (-1.44))
        b = (0.433 - np.log10(r80)) / 0.433
        dndr = 1.373 * r80 ** (-a) * (1. + 0.057 * r80 ** 3.45) * 10. ** (1.607 * np.exp(-b ** 2.))
        dndr = dndr * k80
        dndr[rarr < rmin_w] = 0.
        dndr[rarr > rmax_w] = 0.
        dndr = dndr / np.sum(dndr)
        w = dndr * dr
        w = w / np.sum(w)
        psd.append(w)
        if want_ref:
            ref.append(np.sum(rr ** 4 * w) / np.sum(rr ** 3 * w))
    elif kind == "du":
        # r^-4 number distribution per minor sub-bin, with the grid's own
        # widths; no humidity growth (dointegration.py:626-662)
        dr = getDR(xxarr)
        for rmin, rmax in zip(pp["rMinMaj"][radind], pp["rMaxMaj"][radind]):
            rLow, rUp = rmin, rmax
            dndr = rarr ** (-4.)
            dndr[rarr < rmin] = 0.
            dndr[rarr > rmax] = 0.
            if np.max(dndr) == 0.:
                raise ValueError("du sub-bin [%g, %g] m holds no grid radius at wavelength %g m"
                                 % (rmin, rmax, lam))
            w = np.copy(dndr) * dr
            w /= np.sum(w)
            psd.append(w)
            if want_ref:
                ref.append(np.sum(rr ** 4. * w) / np.sum(rr ** 3. * w))
    else:
        raise ValueError("psd type %s is not supported by the Mie table path" % kind)
    return psd, (ref if want_ref else None), rLow, rUp
