"""Band averaging of optics tables on the GPU (runbands).

Drop-ins for src/geosmie/bandaverage.py and runbands.py:

  getBands(mode)        band edges of GEOS5, RRTMG, RRTMGP and PURDUE
                        (bandaverage.py:73-124)
  doAverage(...)        one spectrum's band mean (bandaverage.py:18-50)
  average_table(...)    every averaged variable of a table at once: one
                        device launch for all (variable, bin, rh, band)
                        means (bandaverage.py:159-180), then the GEOS5
                        merge of bands 0 and 2 and the noIR option
                        (:183-215)
  fun(...)              read a table file, write the band table
                        (bandaverage.py:126-294)

The band means are bit-identical to the reference: the sample wavenumbers
are formed on the host with the reference's own numpy expressions, and the
kernel (gm_band_average) applies numpy.interp's rule and numpy's pairwise
summation order.  The reference's solar weighting is an unused stub
(doAverage ignores it); use_solar is accepted and likewise has no effect.
"""
import numpy as np

from . import _lib
from ._lib import check, ptr, stream

VARS = ["qsca", "qext", "bsca", "bext", "g", "bbck", "refreal", "refimag"]   # :157
N_SUB = 100                                                                   # :33


def getBands(mode):
    """(lBandLow, lBandUp, bandMeanM, useWavenum, numBands) of a band set.
    GEOS5 and PURDUE edges are wavelengths in m (numpy arrays); RRTMG and
    RRTMGP edges are wavenumbers in cm-1 (lists); bandMeanM is in m."""
    use_wn, mod = True, 0
    if mode == "GEOS5":
        use_wn, mod = False, -1    # bands 0 and 2 are merged into one
        lo = np.array([.175, .225, .285, .300, .325, .400, .690, 1.220, 2.270, 29.412, 18.519,
                       12.5, 10.204, 9.091, 8.230, 7.246, 5.263, 3.333, 16.129]) * 1.e-6
        up = np.array([.225, .285, .300, .325, .400, .690, 1.220, 2.270, 3.850, 40., 29.412,
                       18.519, 12.5, 10.204, 9.091, 8.230, 7.246, 5.263, 18.519]) * 1.e-6
    elif mode == "RRTMG":       # shortwave then longwave bands, wavenumbers in cm-1
        lo = [2600., 3250., 4000., 4650., 5150., 6150., 7700., 8050., 12850., 16000., 22650.,
              29000., 38000., 820.] + \
             [10., 350., 500., 630., 700., 820., 980., 1080., 1180., 1390., 1480., 1800., 2080.,
              2250., 2380., 2600.]
        up = [3250., 4000., 4650., 5150., 6150., 7700., 8050., 12850., 16000., 22650., 29000.,
              38000., 50000., 2600.] + \
             [350., 500., 630., 700., 820., 980., 1080., 1180., 1390., 1480., 1800., 2080.,
              2250., 2380., 2600., 3250.]
    elif mode == "RRTMGP":
        lo = [820., 2680., 3250., 4000., 4650., 5150., 6150., 7700., 8050., 12850., 16000.,
              22650., 29000., 38000.] + \
             [10., 250., 500., 630., 700., 820., 980., 1080., 1180., 1390., 1480., 1800., 2080.,
              2250., 2390., 2680.]
        up = [2680., 3250., 4000., 4650., 5150., 6150., 7700., 8050., 12850., 16000., 22650.,
              29000., 38000., 50000.] + \
             [250., 500., 630., 700., 820., 980., 1080., 1180., 1390., 1480., 1800., 2080.,
              2250., 2390., 2680., 3250.]
    elif mode == "PURDUE":
        use_wn = False
        lo = np.array([.175, .225, .245, .280, .295, .310, .325, .400, .700, 1.220, 2.270, 3.33,
                       5.26, 7.25, 8.23, 9.09, 10.2, 12.5, 16.13, 18.52, 29.41]) * 1.e-6
        up = np.array([.225, .280, .260, .295, .310, .320, .400, .700, 1.220, 2.270, 10.00, 5.26,
                       7.25, 8.23, 9.09, 10.2, 12.5, 18.52, 18.52, 29.41, 40.0]) * 1.e-6
    else:
        raise ValueError("Band type must be one of: ['GEOS5', 'RRTMG', 'RRTMGP', 'PURDUE']")
    mean = [(a + b) / 2. for a, b in zip(lo, up)]
    if use_wn:
        mean = [(v * 100.) ** (-1) for v in mean]
    return lo, up, mean, use_wn, len(lo) + mod


def band_samples(lam, lo, up, use_wn, n_sub=N_SUB):
    """Sorted table wavenumbers, the sort order, and the clipped sample
    wavenumbers [n_band][n_sub] (bandaverage.py:26-41, interp1d's sort)."""
    wnum = (100. * np.asarray(lam, np.float64)) ** (-1)
    order = np.argsort(wnum, kind="mergesort")
    samples = np.empty((len(lo), n_sub))
    for b in range(len(lo)):
        bl, br = lo[b], up[b]
        if not use_wn:
            bl = (100. * bl) ** (-1)
            br = (100. * br) ** (-1)
        beg, end = (bl, br) if bl < br else (br, bl)
        samples[b] = np.clip(np.linspace(beg, end, n_sub), np.min(wnum), np.max(wnum))
    return wnum[order], order, samples


def band_means(lam, y, lo, up, use_wn, n_sub=N_SUB):
    """out[s][b]: band means of the spectra y [n_series][n_lam] (one launch)."""
    t = _lib.torch()
    y = np.ascontiguousarray(np.atleast_2d(y), dtype=np.float64)
    xp, order, samples = band_samples(lam, lo, up, use_wn, n_sub)
    d_x = _lib.to_dev(xp)
    d_y = _lib.to_dev(np.ascontiguousarray(y[:, order]))
    d_s = _lib.to_dev(samples)
    d_o = t.empty(y.shape[0] * len(lo), dtype=t.float64, device=_lib.device())
    check(_lib.lib().gm_band_average(_lib.ctx(), y.shape[0], xp.size, ptr(d_x), ptr(d_y), len(lo),
                                     int(n_sub), ptr(d_s), ptr(d_o), stream()), "gm_band_average")
    return d_o.cpu().numpy().reshape(y.shape[0], len(lo))


def doAverage(lam, varIn, bandl, bandr, useWavenum, solardata=None):
    """Mean of one spectrum over one band (bandaverage.py:18-50)."""
    return float(band_means(lam, varIn, [bandl], [bandr], useWavenum)[0, 0])


def average_table(variables, lam, mode, classic=False, noIR=False):
    """Band tables of VARS from table arrays variables[name] with dims
    (bin, wavelength, rh) (classic: (radius, rh, lambda)).  Returns
    (output dict, bandMeanM, lBandLow, lBandUp, numBands) after the GEOS5
    merge (bandaverage.py:159-215)."""
    lo, up, mean, use_wn, nb = getBands(mode)
    first = np.asarray(variables[VARS[0]])
    if classic:
        nbin, nrh, nlam = first.shape
        series = np.stack([np.asarray(variables[v]) for v in VARS])        # v, bin, rh, lam
    else:
        nbin, nlam, nrh = first.shape
        series = np.stack([np.asarray(variables[v]).transpose(0, 2, 1) for v in VARS])
    means = band_means(lam, series.reshape(-1, nlam), lo, up, use_wn)
    means = means.reshape(len(VARS), nbin, nrh, len(lo))
    out = {}
    for i, v in enumerate(VARS):
        out[v] = means[i] if classic else np.ascontiguousarray(means[i].transpose(0, 2, 1))
    ax = 2 if classic else 1
    if noIR and mode == "GEOS5":
        ind = np.where(lo > 3e-6)[0]
        for v, val in (("qsca", 0), ("bsca", 0), ("qext", 1e-32), ("bext", 1e-32)):
            idx = [slice(None)] * 3
            idx[ax] = ind
            out[v][tuple(idx)] = val
    if mode == "GEOS5":
        bw1, bw2 = up[0] - lo[0], up[2] - lo[2]
        for v in VARS:
            a = np.take(out[v], 0, axis=ax)
            c = np.take(out[v], 2, axis=ax)
            avg = (bw1 * a + bw2 * c) / (bw1 + bw2)
            idx = [slice(None)] * 3
            idx[ax] = 2
            out[v][tuple(idx)] = avg
            idx[ax] = slice(1, None)
            out[v] = np.ascontiguousarray(out[v][tuple(idx)])
        mean = mean[1:]
    return out, mean, lo, up, nb


def fun(path, part, opfn, mode, useSolar=False, noIR=False, format="NETCDF4",
        band_up_quirk=True):
    """Band-average one table file written by geosmie_amd.table (either
    layout, NetCDF-4 or NetCDF-3) into opfn (bandaverage.py:126-294).

    The output is NetCDF-4 as the reference writes it (:222-290: the averaged
    variables compressed with zlib, the others not), or NetCDF-3 with
    format="NETCDF3_64BIT_OFFSET".  band_up_quirk keeps the reference's
    bandUp, which holds the lower band edges (:266-268 assign lBandLow to
    it); False writes the upper edges."""
    from .table import read
    v = read(path)
    classic = "wavelength" not in v
    lam_nm, rad_nm = ("lambda", "radius") if classic else ("wavelength", "bin")
    lam = v[lam_nm][0]
    out, mean, lo, up, nb = average_table({k: v[k][0] for k in VARS}, lam, mode, classic, noIR)
    rh = v["rh"][0]
    radius = v[rad_nm][0]
    nbin = radius.size
    if use_wavenumber_edges(mode):
        lo_m = np.array(lo) ** (-1) * 0.01
        up_m = np.array(up) ** (-1) * 0.01
        lo, up = up_m[::-1], lo_m[::-1]
    lo = np.asarray(lo) * 1e6
    up = np.asarray(up) * 1e6
    names = np.zeros((nbin, 80), "S1")
    for i in range(nbin):
        s = ("%s%03d" % (part, i + 1)).encode()
        names[i, :len(s)] = [bytes([ch]) for ch in s]
    edge_lo = lo[:nb]
    edge_up = edge_lo if band_up_quirk else up[:nb]
    scal = (rad_nm, "rh", lam_nm) if classic else (rad_nm, lam_nm, "rh")
    # (name, type, dimensions, compressed, data, attributes) in the reference's order
    plan = [(k, "f8", scal, True, out[k], {}) for k in VARS]
    plan += [(lam_nm, "f8", (lam_nm,), False, np.asarray(mean), {}),
             ("rh", "f8", ("rh",), False, rh, {}),
             ("qname", "c", (rad_nm, "nchar"), False, names, {}),
             ("bandLow", "f8", (lam_nm,), False, edge_lo,
              {"long_name": "Lower edges of the bands", "units": "micrometers"}),
             ("bandUp", "f8", (lam_nm,), False, edge_up,
              {"long_name": "Upper edges of the bands", "units": "micrometers"}),
             (rad_nm, "f8", (rad_nm,), False, radius.astype(np.float64), {})]
    plan += [(k, "f8", (rad_nm,), False, v[k][0][:, 0], {}) for k in ("rLow", "rUp")]
    plan += [(k, "f8", (rad_nm, "rh"), False, v[k][0], {}) for k in ("rEff", "rMass")]
    dims = [("rh", rh.size), (lam_nm, nb), (rad_nm, nbin), ("nchar", 80)]
    if format == "NETCDF4":
        from . import nc4
        with nc4.Writer(opfn) as w:
            for d, n in dims:
                w.dim(d, n)
            for name, typ, dd, z, _, at in plan:
                w.var(name, typ, dd, zlib=z, attrs=at)
            for name, _, _, _, data, _ in plan:
                w.put(name, data)
        return opfn
    if format != "NETCDF3_64BIT_OFFSET":
        raise ValueError("format must be NETCDF4 or NETCDF3_64BIT_OFFSET, not %r" % (format,))
    from scipy.io import netcdf_file
    f = netcdf_file(opfn, "w", version=2)
    try:
        for d, n in dims:
            f.createDimension(d, n)
        for name, typ, dd, _, data, at in plan:
            var = f.createVariable(name, "c" if typ == "c" else "d", dd)
            var[:] = data
            for k, t in at.items():
                setattr(var, k, t)
    finally:
        f.close()
    return opfn


def use_wavenumber_edges(mode):
    return mode in ("RRTMG", "RRTMGP")
