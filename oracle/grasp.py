"""CPU restatement of runoptics' GRASP kernel mode.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows src/geosmie/dointegration.py:
  bilinear_interpolation      :29-38   (corner order f1 (i,j), f2 (i+1,j),
                                        f3 (i+1,j+1), f4 (i,j+1))
  get_interpolated            :40-72   (bisect_left - 1, both indices clamped
                                        to [0, len-2], the index itself clipped
                                        to the grid's range)
  integrateShapes             :74-91   (fractions normalised, tensordot over
                                        the shape axis)
  kernel size grid and widths :763-775 (r/dr constant on a geometric grid)
  the kernel branch of fun    :891-947, 969-976 (interpolated efficiencies and
                                        phase-matrix kernels, csca/cext from the
                                        grid radii, integratePSD, the effective-
                                        radius rescaling of the mass efficiencies,
                                        and the cubic spline in cos(theta) to the
                                        Mie angle grid)
Pinned by tests/golden/grasp_vectors.npz, written by running the reference's
own three helpers (tests/golden/make_grasp_vectors.py), and by exactness on
bilinear functions (tests/test_oracle_grasp.py).
"""
import bisect

import numpy as np
from scipy.interpolate import CubicSpline

from . import lut as OL

KERNEL_SCALAR_KEYS = ("ext", "abs", "sca", "qext", "qabs", "qsca", "qb", "g", "cext", "csca",
                      "cabs")                                  # dointegration.py:896
KERNEL_ELEMS = ("p11", "p22", "p33", "p44", "p12", "p34")      # :899, element axis order


def bilinear_interpolation(x1, x2, y1, y2, f1, f2, f3, f4, x, y):
    t = (x - x1) / (x2 - x1)
    u = (y - y1) / (y2 - y1)
    return ((1.0 - t) * (1.0 - u) * f1 + t * (1.0 - u) * f2 + (1.0 - t) * u * f4 +
            t * u * f3)


def corner_weights(mr, mi, allmr, allmi):
    """(mri, mii, w1, w2, w3, w4): the bracketing indices and the four corner
    weights get_interpolated applies (f1 (mri, mii), f2 (mri+1, mii),
    f3 (mri+1, mii+1), f4 (mri, mii+1)), each weight formed as the reference's
    expression forms it."""
    mri = bisect.bisect_left(allmr, mr) - 1
    mri = min(len(allmr) - 2, max(0, mri))
    mii = bisect.bisect_left(allmi, mi) - 1
    mii = min(len(allmi) - 2, max(0, mii))
    mr = np.clip(mr, allmr[0], allmr[-1])
    mi = np.clip(mi, allmi[0], allmi[-1])
    t = (mr - allmr[mri]) / (allmr[mri + 1] - allmr[mri])
    u = (mi - allmi[mii]) / (allmi[mii + 1] - allmi[mii])
    return mri, mii, (1.0 - t) * (1.0 - u), t * (1.0 - u), t * u, (1.0 - t) * u


def get_interpolated(data, mr, mi, allmr, allmi):
    """data [n_mr, n_mi, ...] at (mr, mi)."""
    mri = bisect.bisect_left(allmr, mr) - 1
    mri = min(len(allmr) - 2, max(0, mri))
    mii = bisect.bisect_left(allmi, mi) - 1
    mii = min(len(allmi) - 2, max(0, mii))
    v1, v2 = data[mri, mii], data[mri + 1, mii]
    v3, v4 = data[mri + 1, mii + 1], data[mri, mii + 1]
    mr = np.clip(mr, allmr[0], allmr[-1])
    mi = np.clip(mi, allmi[0], allmi[-1])
    return bilinear_interpolation(allmr[mri], allmr[mri + 1], allmi[mii], allmi[mii + 1],
                                  v1, v2, v3, v4, mr, mi)


def integrate_shapes(variables, fracs0):
    """{key: array}: keys with a leading shape axis (4-D and 6-D) averaged with
    the normalised fractions, 1-D keys passed through."""
    fracs = np.array(fracs0) / np.sum(fracs0)
    out = {}
    for k, v in variables.items():
        v = np.asarray(v)
        if v.ndim in (4, 6):
            out[k] = np.tensordot(fracs, v, axes=(0, 0))
        elif v.ndim == 1:
            out[k] = v
    return out


def kernel_grid_widths(x):
    """drarr of a geometric kernel size grid (dointegration.py:770-775)."""
    x = np.asarray(x, np.float64)
    rmrat = (x[1] / x[0]) ** 3
    vrfact = ((3. / 2. / np.pi / (rmrat + 1)) ** (1. / 3.)) * (rmrat ** (1. / 3.) - 1.)
    return vrfact * (4. / 3. * np.pi * x ** 3.) ** (1. / 3.)


def kernel_entry(avg, x, mr, mi, psd, fracs, lam, reff0, rhop0, rhop, ang_k, ang_out,
                 psd_nominal=None, xx_nominal=None):
    """One (bin, wavelength, rh) entry in kernel mode, up to the P
    normalisation: the interpolated kernels (:894-914), csca/cext (:916-919),
    integratePSD (:920-927), the mass-efficiency rescaling when psd_nominal
    is given (:929-944), lidar ratio and ssa (:953-961) and the P elements
    spline-interpolated from ang_k to ang_out (:969-976) and normalised
    (:978-988).  avg: shape-averaged variables; mr, mi scalars."""
    allmr = np.asarray(avg["mr"])
    allmi = -np.asarray(avg["mi"])
    raw = {}
    for k in ("qext", "qabs", "qsca", "qb", "g"):
        raw[k] = get_interpolated(np.asarray(avg[k]), mr, mi, allmr, allmi)
    for e, name in enumerate(KERNEL_ELEMS):
        raw[name] = get_interpolated(np.asarray(avg["scama"])[:, :, :, e, :], mr, mi, allmr,
                                     allmi)
    rarr = np.asarray(x) * lam / (2 * np.pi)
    raw["csca"] = raw["qsca"] * np.pi * rarr ** 2
    raw["cext"] = raw["qext"] * np.pi * rarr ** 2
    ret = OL.integrate_psd(x, [raw] * len(psd), psd, fracs, lam, reff0, rhop0, rhop)
    if psd_nominal is not None:
        rr = np.asarray(xx_nominal) * lam / (2. * np.pi)
        for k in range(len(fracs)):
            reff = np.sum(rr ** 3. * psd_nominal[k]) / np.sum(rr ** 2. * psd_nominal[k])
            for key in ("bsca", "bext", "bbck"):
                ret[key] = ret[key] * ret["rEff"] / reff
            ret["rEff"] = reff
    out = dict(ret)
    out["lidar_ratio"] = ret["qext"] / ret["qb"] * 4 * np.pi
    out["ssa"] = ret["qsca"] / ret["qext"]
    ck = np.cos(np.radians(ang_k))
    co = np.cos(np.radians(ang_out))
    for key in KERNEL_ELEMS:
        cs = CubicSpline(ck[::-1], np.asarray(ret[key])[::-1])
        out[key] = cs(co[::-1])[::-1]
    theta = np.radians(ang_out)
    p11n = 2. * out["p11"] / np.trapezoid(out["p11"] * np.sin(theta), theta)
    for k in ("p12", "p22", "p33", "p34", "p44"):
        out[k] = out[k] * p11n / out["p11"]
    out["p11"] = p11n
    out["pback"] = np.array([out[k][-1] for k in ("p11", "p12", "p33", "p34", "p22", "p44")])
    return out


def table(params, variables, fracs0, water=None, ang_out=None):
    """Every entry of a kernel-mode table (dointegration.py:fun, 'mode':
    'kernel'), keyed (bin, wavelength index, rh index), plus the per-(bin, rh)
    values.  variables: the kernel file's arrays; fracs0: the shape
    distribution."""
    from scipy.interpolate import interp1d
    from geosmie_amd import particle as PF
    ang_out = OL.angle_grid() if ang_out is None else np.asarray(ang_out, np.float64)
    avg = integrate_shapes(variables, fracs0)
    x = np.asarray(variables["x"], np.float64)
    dr = kernel_grid_widths(x)
    ang_k = np.linspace(0., 180., np.asarray(variables["scama"]).shape[-1])
    pp = params["psd"]["params"]
    kind = params["psd"]["type"]
    n_bin = len(pp["r0"]) if kind == "lognorm" else len(pp["rMinMaj"])
    mlist = params["mList"]
    lams = mlist[0][0]
    fmr, fmi = interp1d(mlist[0][0], mlist[0][1]), interp1d(mlist[0][0], mlist[0][2])
    triv = params["rhDep"]["type"] == "trivial"
    if not triv:
        wmr, wmi = interp1d(water[0], water[1]), interp1d(water[0], water[2])
    out, nl = {}, {}
    for radind in range(n_bin):
        rhop00 = params["rhop0"]
        rhop0 = rhop00[radind] if isinstance(rhop00, list) else rhop00
        xx_n, dr_n = PF.initializeXarr(params, radind, lams[0], lams[-1])
        fracs = pp["fracs"][0] if len(pp["fracs"]) == 1 else pp["fracs"][radind]
        rescale = kind == "lognorm" or (kind == "du" and len(pp["fracs"][radind]) == 1)
        for li, lam in enumerate(lams):
            nref0 = [complex(fmr(lam), -fmi(lam))]
            nw = complex(1, 0) if triv else complex(wmr(lam), wmi(lam))
            rh = list(params["rh"])
            if "maxrh" in params:
                a = np.array(rh)
                a[a > params["maxrh"]] = params["maxrh"]
                rh = list(a)
            _, _, _, rrat0 = PF.getHumidRefractiveIndex(params, radind, 0, rh, nref0, nw)
            _, reff0, _, _ = PF.calculatePSD(params, radind, 0., rh, x, dr, rrat0, lam)
            mass0 = None
            for rhi, onerh in enumerate(rh):
                if triv and rhi > 0:
                    out[radind, li, rhi] = out[radind, li, 0]
                    nl[radind, rhi] = nl[radind, 0]
                    continue
                mr, mi, gf, rrat = PF.getHumidRefractiveIndex(params, radind, rhi, rh, nref0, nw)
                psd, _, rlo, rup = PF.calculatePSD(params, radind, onerh, rh, x, dr, rrat, lam)
                psd_n = None
                if rescale:
                    psd_n, _, _, _ = PF.calculatePSD(params, radind, onerh, rh, xx_n, dr_n, rrat,
                                                     lam)
                rhop = rrat ** 3. * rhop0 + (1. - rrat ** 3.) * 1000.
                ret = kernel_entry(avg, x, float(mr[0]), float(mi[0]), psd, fracs, lam, reff0,
                                   rhop0, rhop, ang_k, ang_out, psd_nominal=psd_n,
                                   xx_nominal=xx_n)
                ret["rLow"], ret["rUp"] = rlo, rup
                ret["growth_factor"] = gf
                if rhi == 0:
                    mass0 = ret["volume"] * rhop0
                ret["rhop"] = rhop
                ret["area"] = ret["area"] / mass0
                ret["volume"] = ret["volume"] / mass0
                ret["refreal"] = mr[0]
                ret["refimag"] = -np.abs(mi[0])
                out[radind, li, rhi] = ret
                nl[radind, rhi] = {k: ret[k] for k in ("mass", "volume", "area", "rEff", "rMass",
                                                       "rhop", "growth_factor", "rLow", "rUp")}
    return out, nl
