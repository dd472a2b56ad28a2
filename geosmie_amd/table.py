"""Optics table writer in runoptics' layout.

Follows createNCDF and the per-(bin, wavelength, rh) writes of
src/geosmie/dointegration.py:205-422, 1006-1031, and the pmom variable that
the GSF step adds (src/gsf/convertncdf.py:70-79, 381-394, zero-padded to m):

  dimensions   rh, wavelength [m], bin (1-indexed), p = [11,12,33,34,22,44], ang
               (legacy layout: rh, lambda, radius, nPol, ang)
  variables    p11..p44 (bin, wavelength, rh, ang); the scalars
               (bin, wavelength, rh); pback (bin, wavelength, rh, p); the
               per-bin-RH scalars (bin, rh); pmom (bin, wavelength, rh, p, m)
               (legacy: radius, rh, lambda order; pback (nPol, radius, rh, lambda);
               pmom (nPol, nMom, radius, rh, lambda))

The file is NetCDF-4 (HDF5) as the reference writes it: netCDF's HDF5
conventions, every variable compressed with shuffle + deflate level 4
(netCDF4-python's compression='zlib') and int64 bin and p indices, written by
libgeosmie_nc4.so (geosmie_amd.nc4).  format="NETCDF3_64BIT_OFFSET" writes
the NetCDF-3 form instead (scipy.io.netcdf_file: no compression, int32
indices, since NetCDF-3 has no int64).  Names, dimension orders, units and
long names are the reference's.
"""
import numpy as np

P_ORDER = [11, 12, 33, 34, 22, 44]
SCAT_KEYS = ["p11", "p12", "p22", "p33", "p34", "p44"]
SCALAR_KEYS = ["qsca", "qabs", "qext", "g", "ssa", "qb", "bsca", "bext", "csca", "cext", "bbck",
               "lidar_ratio", "refreal", "refimag"]
NL_SCALAR_KEYS = ["mass", "volume", "area", "rEff", "rMass", "rUp", "rLow", "rhop",
                  "growth_factor"]

# (units, long_name) of every variable (dointegration.py:224-334)
META = {
    "p11": ("dimensionless", "P11 element of the normalized scattering matrix"),
    "p12": ("dimensionless", "P12 element of the normalized scattering matrix"),
    "p22": ("dimensionless", "P22 element of the normalized scattering matrix"),
    "p33": ("dimensionless", "P33 element of the normalized scattering matrix"),
    "p34": ("dimensionless", "P34 element of the normalized scattering matrix"),
    "p44": ("dimensionless", "P44 element of the normalized scattering matrix"),
    "qsca": ("dimensionless", "scattering efficiency"),
    "qabs": ("dimensionless", "absorption efficiency"),
    "qext": ("dimensionless", "extinction efficiency"),
    "g": ("dimensionless", "asymmetry factor"),
    "ssa": ("dimensionless", "single-scattering albedo"),
    "bsca": ("m2 (kg dry mass)-1", "mass scattering efficiency"),
    "bext": ("m2 (kg dry mass)-1", "mass extinction efficiency"),
    "csca": ("m2", "mass scattering cross-section"),
    "cext": ("m2", "mass extinction cross-section"),
    "bbck": ("m2 (kg dry mass)-1 sr-1", "mass backscatter efficiency"),
    "lidar_ratio": ("", "lidar ratio"),
    "mass": ("kg", "particle mass"),
    "volume": ("m3 kg-1", "particle volume per kg dry mass"),
    "area": ("m2 kg-1", "particle cross sectional area per kg dry mass"),
    "rEff": ("m", "effective radius of bin"),
    "rMass": ("kg", "effective mass of wet particle"),
    "rUp": ("m", "upper edge of radius bin"),
    "rLow": ("m", "lower edge of radius bin"),
    "pback": ("dimensionless", "phase function in backscatter direction, ordered as P11, P12, "
                               "P33, P34, P22, P44"),
    "qb": ("dimensionless", "backscattering efficiency"),
    "rhop": ("kg m-3", "wet particle density"),
    "growth_factor": ("fraction", "growth factor = ratio of wet to dry particle radius"),
    "refreal": ("dimensionless", "real refractive index of wet particle"),
    "refimag": ("dimensionless", "imaginary refractive index of wet particle"),
    "pmom": (None, "Moments of the generalized spherical functions order over nPol as the 11, 12, "
                   "33, 34, 22, 44 elements of the scattering law matrix for a Fourier "
                   "decomposition of the phase matrix such as those used in discrete ordinate "
                   "solutions to the RTE"),
}


class OpticsTable(object):
    """In-memory optics table; fill with set_entry / set_bin_rh, write with write()."""

    def __init__(self, n_bin, rh, wavelength_m, ang_deg, classic=False, n_mom=0):
        self.classic = bool(classic)
        self.rh = np.asarray(rh, np.float64)
        self.lam = np.asarray(wavelength_m, np.float64)
        self.ang = np.asarray(ang_deg, np.float64)
        self.n_bin, self.n_rh, self.n_lam, self.n_ang = int(n_bin), self.rh.size, self.lam.size, \
            self.ang.size
        self.n_mom = int(n_mom)
        self.radius = "radius" if classic else "bin"
        self.lamname = "lambda" if classic else "wavelength"
        self.pname = "nPol" if classic else "p"
        self.mname = "nMom" if classic else "m"
        r, lm, p, mo = self.radius, self.lamname, self.pname, self.mname
        if classic:   # dointegration.py:356-363
            scal = (r, "rh", lm)
            fun = (r, "rh", lm, "ang")
            ele = (p, r, "rh", lm)
        else:         # :364-371
            scal = (r, lm, "rh")
            fun = (r, lm, "rh", "ang")
            ele = (r, lm, "rh", p)
        self.dims = {}
        for k in SCAT_KEYS:
            self.dims[k] = fun
        for k in SCALAR_KEYS:
            self.dims[k] = scal
        for k in NL_SCALAR_KEYS:
            self.dims[k] = (r, "rh")
        self.dims["pback"] = ele
        if self.n_mom:   # convertncdf.py:70-79
            self.dims["pmom"] = (p, mo, r, "rh", lm) if classic else (r, lm, "rh", p, mo)
        size = {r: self.n_bin, "rh": self.n_rh, lm: self.n_lam, "ang": self.n_ang, p: 6,
                mo: self.n_mom}
        self.size = size
        self.data = {k: np.zeros([size[d] for d in dims]) for k, dims in self.dims.items()}

    def set_entry(self, radind, li, rhi, ret, pmom=None):
        """Write one (bin, wavelength, rh) result dict as the reference does
        (dointegration.py:1006-1031): p11..p44 [ang], the scalars, pback [6]
        in p order; pmom [6, <= n_mom] optional, zero-padded to n_mom."""
        for k in SCAT_KEYS:
            if k in ret:
                self._put(k, radind, li, rhi, np.asarray(ret[k]))
        for k in SCALAR_KEYS:
            if k in ret:
                self._put(k, radind, li, rhi, float(ret[k]))
        if "pback" in ret:
            pb = np.asarray(ret["pback"], np.float64)
            if self.classic:
                self.data["pback"][:, radind, rhi, li] = pb
            else:
                self.data["pback"][radind, li, rhi, :] = pb
        if pmom is not None:
            src = np.asarray(pmom, np.float64)[:, :self.n_mom]
            pm = np.zeros((6, self.n_mom))
            pm[:, :src.shape[1]] = src
            if self.classic:
                self.data["pmom"][:, :, radind, rhi, li] = pm
            else:
                self.data["pmom"][radind, li, rhi] = pm

    def set_entries(self, radind, li, rhis, ret, pmom=None):
        """set_entry for several rh of one (bin, wavelength) at once: ret maps
        each key to an array with a leading axis over rhis (p11.. [k, ang],
        scalars [k], pback [k, 6]); pmom [k, 6, <= n_mom]."""
        rhis = np.asarray(rhis, dtype=np.intp)
        for k in SCAT_KEYS + SCALAR_KEYS:
            if k in ret:
                v = np.asarray(ret[k], np.float64)
                if self.classic:
                    self.data[k][radind, rhis, li] = v
                else:
                    self.data[k][radind, li, rhis] = v
        if "pback" in ret:
            pb = np.asarray(ret["pback"], np.float64)
            if self.classic:
                self.data["pback"][:, radind, rhis, li] = pb.T
            else:
                self.data["pback"][radind, li, rhis, :] = pb
        if pmom is not None:
            src = np.asarray(pmom, np.float64)[:, :, :self.n_mom]
            pm = np.zeros((rhis.size, 6, self.n_mom))
            pm[:, :, :src.shape[2]] = src
            if self.classic:
                self.data["pmom"][:, :, radind, rhis, li] = pm.transpose(1, 2, 0)
            else:
                self.data["pmom"][radind, li, rhis] = pm

    def set_bins_rh(self, radind, rhis, ret):
        """set_bin_rh for several rh at once (ret values are arrays over rhis)."""
        rhis = np.asarray(rhis, dtype=np.intp)
        for k in NL_SCALAR_KEYS:
            if k in ret:
                self.data[k][radind, rhis] = np.asarray(ret[k], np.float64)

    def set_bin_rh(self, radind, rhi, ret):
        """Per-(bin, rh) scalars (mass, volume, area, rEff, rMass, rUp, rLow,
        rhop, growth_factor)."""
        for k in NL_SCALAR_KEYS:
            if k in ret:
                self.data[k][radind, rhi] = float(ret[k])

    def copy_rh(self, radind, li, src, dst):
        """Entry (radind, li, dst) and the (radind, dst) scalars := those at rh
        index src (copyDryValues, dointegration.py:462-490)."""
        for k, dims in self.dims.items():
            a = self.data[k]
            idx_s, idx_d = [], []
            for d in dims:
                if d == self.radius:
                    idx_s.append(radind), idx_d.append(radind)
                elif d == "rh":
                    idx_s.append(src), idx_d.append(dst)
                elif d == self.lamname:
                    idx_s.append(li), idx_d.append(li)
                else:
                    idx_s.append(slice(None)), idx_d.append(slice(None))
            a[tuple(idx_d)] = a[tuple(idx_s)]

    def _put(self, key, radind, li, rhi, value):
        if self.classic:
            self.data[key][radind, rhi, li] = value
        else:
            self.data[key][radind, li, rhi] = value

    def _coords(self):
        r, lm, p = self.radius, self.lamname, self.pname
        # dointegration.py:336-354: rh, wavelength, bin (i8), p (i8), ang
        return [("rh", self.rh, "f8", ("fraction", "relative humidity")),
                (lm, self.lam, "f8", ("m", "wavelength")),
                (r, np.arange(1, self.n_bin + 1), "i8",
                 ("dimensionless", "radius bin index (1-indexed)")),
                (p, np.array(P_ORDER), "i8",
                 ("dimensionless", "Scattering matrix element index, ordered as P11, P12, P33, "
                                   "P34, P22, P44")),
                ("ang", self.ang, "f8", ("degrees", "scattering angle"))]

    def write(self, path, format="NETCDF4"):
        """Write the table: NetCDF-4 with zlib as the reference does
        (dointegration.py:205-354, 418), or format="NETCDF3_64BIT_OFFSET"."""
        if format == "NETCDF4":
            return self._write_nc4(path)
        if format != "NETCDF3_64BIT_OFFSET":
            raise ValueError("format must be NETCDF4 or NETCDF3_64BIT_OFFSET, not %r" % (format,))
        from scipy.io import netcdf_file
        f = netcdf_file(path, "w", version=2)
        try:
            r, lm, p = self.radius, self.lamname, self.pname
            for d in ("rh", lm, r, p, "ang"):
                f.createDimension(d, self.size[d])
            if self.n_mom:
                f.createDimension(self.mname, self.n_mom)
            for name, vals, typ, (units, lname) in self._coords():
                v = f.createVariable(name, "i" if typ == "i8" else "d", (name,))
                v[:] = np.asarray(vals, np.int32 if typ == "i8" else np.float64)
                v.units = units
                v.long_name = lname
            for k, dims in self.dims.items():
                v = f.createVariable(k, "d", dims)
                v[:] = self.data[k]
                units, v.long_name = META[k]
                if units is not None:
                    v.units = units
        finally:
            f.close()

    def _write_nc4(self, path):
        from . import nc4
        r, lm, p = self.radius, self.lamname, self.pname
        with nc4.Writer(path) as w:
            for d in ("rh", lm, r, p, "ang"):
                w.dim(d, self.size[d])
            if self.n_mom:
                w.dim(self.mname, self.n_mom)   # convertncdf.py:70-79
            coords = self._coords()
            for name, _, typ, (units, lname) in coords:
                w.var(name, typ, (name,), attrs={"long_name": lname, "units": units})
            for k, dims in self.dims.items():
                units, lname = META[k]
                at = {"long_name": lname}
                if units is not None:
                    at["units"] = units
                w.var(k, "f8", dims, attrs=at)
            for name, vals, _, _ in coords:
                w.put(name, vals)
            for k in self.dims:
                w.put(k, self.data[k])


def read(path):
    """Variables of a written table as numpy arrays, {name: (array, dimension
    names, units)}; NetCDF-4 or NetCDF-3 by the file's signature."""
    from . import nc4
    if nc4.is_hdf5(path):
        return {k: (a, dims, at.get("units", "")) for k, (a, dims, at) in nc4.read(path).items()}
    from scipy.io import netcdf_file
    f = netcdf_file(path, "r", mmap=False)
    try:
        return {k: (v[:].copy(), v.dimensions, getattr(v, "units", b"").decode()
                    if isinstance(getattr(v, "units", b""), bytes) else getattr(v, "units", ""))
                for k, v in f.variables.items()}
    finally:
        f.close()
