"""CPU restatement of runoptics' Mie table loop (one table, all entries).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows src/geosmie/dointegration.py:fun (672-1036), Mie mode: per bin the
size grid (initializeXarr), per wavelength the particle and water indices
(interp1d of the tables, particle imaginary part negated, :813-822), rh
capped at maxrh (:824-828), the dry state's effective radii (:835-836); per
rh the humid index and PSD, the wet density (:866-868), rawMie per
refractive index (oracle.mie.size_range + oracle.lut.raw_mie), integratePSD
(oracle.lut.integrate_psd), the post-integration block (:950-1001: lidar
ratio, ssa, rLow/rUp, P normalisation, pback, growth factor, mass0 at the
first rh, area and volume per dry mass, refreal/refimag) and, for rhDep
'trivial', the copy of the dry entry to every rh (copyDryValues, :462-490).

The humidity rules and PSDs are geosmie_amd.particle, which is pinned bit for
bit against vectors written by the reference's own functions
(tests/golden/make_front_vectors.py, tests/test_particle_cpu.py); the Mie,
integration and post-integration pieces are the oracle's own, pinned as
their modules state.  Returns entries keyed (bin, wavelength index, rh index).
"""
import numpy as np
from scipy.interpolate import interp1d

from geosmie_amd import particle as PF

from . import lut as OL
from . import mie as OM


def table(params, water=None, ang=None):
    ang = OL.angle_grid() if ang is None else np.asarray(ang, np.float64)
    cost = np.cos(np.radians(ang))
    pp = params["psd"]["params"]
    n_bin = len(pp["r0"]) if params["psd"]["type"] == "lognorm" else len(pp["rMinMaj"])
    # (du: one bin per rMinMaj list, its sub-bins being the PSD's modes)
    mlist = params["mList"]
    lams = mlist[0][0]
    fmr = [interp1d(m[0], m[1]) for m in mlist]
    fmi = [interp1d(m[0], m[2]) for m in mlist]
    triv = params["rhDep"]["type"] == "trivial"
    if not triv:
        wmr, wmi = interp1d(water[0], water[1]), interp1d(water[0], water[2])
    rh = list(params["rh"])
    out, nl = {}, {}
    for radind in range(n_bin):
        rhop00 = params["rhop0"]
        rhop0 = rhop00[radind] if isinstance(rhop00, list) else rhop00
        xx, dr = PF.initializeXarr(params, radind, lams[0], lams[-1])
        for li, lam in enumerate(lams):
            nref0 = [complex(fmr[i](lam), -fmi[i](lam)) for i in range(len(mlist))]
            nw = complex(1, 0) if triv else complex(wmr(lam), wmi(lam))
            if "maxrh" in params:
                a = np.array(rh)
                a[a > params["maxrh"]] = params["maxrh"]
                rh = list(a)
            _, _, _, rrat0 = PF.getHumidRefractiveIndex(params, radind, 0, rh, nref0, nw)
            _, reff0, _, _ = PF.calculatePSD(params, radind, 0., rh, xx, dr, rrat0, lam)
            mass0 = None
            for rhi, onerh in enumerate(rh):
                if triv and rhi > 0:
                    out[radind, li, rhi] = out[radind, li, 0]
                    nl[radind, rhi] = nl[radind, 0]
                    continue
                mr, mi, gf, rrat = PF.getHumidRefractiveIndex(params, radind, rhi, rh, nref0, nw)
                psd, _, rlo, rup = PF.calculatePSD(params, radind, onerh, rh, xx, dr, rrat, lam)
                rhop = rrat ** 3. * rhop0 + (1. - rrat ** 3.) * 1000.
                raws = [OL.raw_mie(xx, lam, OM.size_range(xx, cost, mr[k], mi[k]))
                        for k in range(len(mr))]
                if len(raws) == 1:
                    raws = [raws[0]] * len(psd)
                ret = OL.integrate_psd(xx, raws, psd, pp["fracs"][radind], lam, reff0, rhop0,
                                       rhop)
                ret = OL.post_integration(ret, ang)
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
