"""GRASP kernel mode of runoptics on the GPU ('mode': 'kernel').

Drop-in for the kernel branch of src/geosmie/dointegration.py:fun
(:755-776, :891-947, :969-976) with its helpers integrateShapes and
get_interpolated (:28-91).  Instead of Mie theory, the per-size efficiencies
and phase-matrix kernels come from precomputed tables on a (shape, m_r, m_i,
size) grid (GRASP spheroid kernels, converted by convertkernels.py):

  * the shape axis is averaged once with the particle's shape distribution
    (gm_shape_average);
  * per wavelength, every rh's wet index is located in the (m_r, m_i) grid
    on the host (bisect, clamped indices, clipped index, exactly as
    get_interpolated does) and the tables are blended on the device
    (gm_ri_bilinear, bit-identical to numpy's evaluation of
    bilinear_interpolation);
  * the phase-matrix integrals over the size grid, and the integratePSD
    scalars, are the Mie path's row contractions (gm_psd_contract_rows,
    gm_integrate_psd_rows), since integratePSD treats kernel and Mie inputs
    alike;
  * the six elements (p22 and p44 are independent for spheroids) are
    carried from the kernels' 181 angles to the table's 371 by the cubic
    spline in cos(theta) the reference fits (CubicSpline, not-a-knot), applied
    on the device as one contraction with the spline's matrix, then
    normalised (gm_pmatrix_normalize6).

Kernel tables hold, per the reference's file: m_r [n_mr], m_i [n_mi] (stored
negative; the reference flips the sign, :893), x [n_x] (a geometric grid), the
scalar efficiencies qext, qabs, qsca, qb, g [n_shape, n_mr, n_mi, n_x] and
scama [n_shape, n_mr, n_mi, n_x, 6, n_ang] with elements in the order p11,
p22, p33, p44, p12, p34 (:899).  The kernel files of GRASP are NetCDF-4 and
are not shipped with the reference; KernelTables.read takes NetCDF-4 files
(geosmie_amd.nc4), NetCDF-3 files (scipy) or .npz files with the same
variable names.
"""
import bisect

import numpy as np

from . import _lib
from ._lib import check, ptr, stream

SCALAR_KEYS = ("qext", "qsca", "qabs", "qb", "g")     # gm_integrate_psd's prop planes 0-4
KERNEL_ELEMS = ("p11", "p22", "p33", "p44", "p12", "p34")


def kernel_grid_widths(x):
    """drarr of a geometric kernel size grid: r/dr constant
    (dointegration.py:770-775)."""
    x = np.asarray(x, np.float64)
    rmrat = (x[1] / x[0]) ** 3
    vrfact = ((3. / 2. / np.pi / (rmrat + 1)) ** (1. / 3.)) * (rmrat ** (1. / 3.) - 1.)
    return vrfact * (4. / 3. * np.pi * x ** 3.) ** (1. / 3.)


def corner_weights(mr, mi, allmr, allmi):
    """Bracketing indices and the weights of the corners (i, j), (i+1, j),
    (i+1, j+1), (i, j+1) for one refractive index, by get_interpolated's
    rules (dointegration.py:40-72), each weight formed as its expression is."""
    mri = bisect.bisect_left(allmr, mr) - 1
    mri = min(len(allmr) - 2, max(0, mri))
    mii = bisect.bisect_left(allmi, mi) - 1
    mii = min(len(allmi) - 2, max(0, mii))
    mr = np.clip(mr, allmr[0], allmr[-1])
    mi = np.clip(mi, allmi[0], allmi[-1])
    t = (mr - allmr[mri]) / (allmr[mri + 1] - allmr[mri])
    u = (mi - allmi[mii]) / (allmi[mii + 1] - allmi[mii])
    return mri, mii, ((1.0 - t) * (1.0 - u), t * (1.0 - u), t * u, (1.0 - t) * u)


def spline_matrix(ang_in, ang_out):
    """[len(ang_out), len(ang_in)]: the not-a-knot cubic spline in cos(theta)
    through values at ang_in, evaluated at ang_out, as a linear map (the
    spline of each unit vector).  dointegration.py:969-976 fits it with
    CubicSpline on ascending cosines."""
    from scipy.interpolate import CubicSpline
    ci = np.cos(np.radians(np.asarray(ang_in, np.float64)))[::-1]
    co = np.cos(np.radians(np.asarray(ang_out, np.float64)))[::-1]
    n = ci.size
    S = CubicSpline(ci, np.eye(n), axis=0)(co)        # [n_out (ascending cos), n_in (ascending)]
    return np.ascontiguousarray(S[::-1, ::-1])


class KernelTables(object):
    """Shape-averaged kernel tables resident on the device."""

    def __init__(self, variables, fracs0):
        """variables: the kernel file's arrays (see the module docstring);
        fracs0: the shape distribution's fractions (normalised here as
        integrateShapes does, dointegration.py:79)."""
        t = _lib.torch()
        L, h, s = _lib.lib(), _lib.ctx(), stream()
        v = {k: np.asarray(variables[k], np.float64) for k in ("mr", "mi", "x") + SCALAR_KEYS +
             ("scama",)}
        self.allmr = v["mr"]
        self.allmi = -v["mi"]                      # stored negative (:893)
        self.x = v["x"]
        n_shape, n_mr, n_mi, n_x = v["qext"].shape
        sc = v["scama"]
        if sc.shape[:4] != (n_shape, n_mr, n_mi, n_x) or sc.shape[4] != 6:
            raise ValueError("scama must be [shape, m_r, m_i, x, 6, angle]")
        if (n_mr, n_mi, n_x) != (self.allmr.size, self.allmi.size, self.x.size):
            raise ValueError("kernel tables do not match their m_r, m_i, x axes")
        if n_mr < 2 or n_mi < 2:
            raise ValueError("kernel tables need at least 2 x 2 refractive indices")
        fr = np.asarray(fracs0, np.float64).reshape(-1)
        if fr.size != n_shape:
            raise ValueError("%d shape fractions for %d kernel shapes" % (fr.size, n_shape))
        w = fr / np.sum(fr)
        self.n_mr, self.n_mi, self.n_x, self.n_ang = n_mr, n_mi, n_x, sc.shape[5]
        d_w = _lib.to_dev(w)
        dev = _lib.device()
        self.d_scalar = {}
        for k in SCALAR_KEYS:
            d_in = _lib.to_dev(v[k])
            out = t.empty(n_mr * n_mi * n_x, dtype=t.float64, device=dev)
            check(L.gm_shape_average(h, n_shape, out.numel(), ptr(d_w), ptr(d_in), ptr(out), s),
                  "gm_shape_average")
            self.d_scalar[k] = out
        d_in = _lib.to_dev(sc)
        self.d_scama = t.empty(n_mr * n_mi * n_x * 6 * self.n_ang, dtype=t.float64, device=dev)
        check(L.gm_shape_average(h, n_shape, self.d_scama.numel(), ptr(d_w), ptr(d_in),
                                 ptr(self.d_scama), s), "gm_shape_average")
        t.cuda.synchronize()

    @classmethod
    def read(cls, path, shape_dist):
        """From a kernel file (NetCDF-4, NetCDF-3 or .npz with the file's
        variable names) and a shape distribution file (first column, as
        np.loadtxt reads it at dointegration.py:765)."""
        fracs = np.loadtxt(shape_dist, usecols=[0], unpack=True, ndmin=1)
        names = ("mr", "mi", "x", "scama") + SCALAR_KEYS
        if str(path).endswith(".npz"):
            with np.load(path, allow_pickle=False) as f:
                var = {k: np.array(f[k]) for k in names}
        else:
            from . import nc4
            if nc4.is_hdf5(path):
                v = nc4.read(path, attrs=())
                var = {k: np.array(v[k][0]) for k in names}
            else:
                from scipy.io import netcdf_file
                with netcdf_file(path, "r", mmap=False) as f:
                    var = {k: np.array(f.variables[k][:]) for k in names}
        return cls(var, fracs)

    def queries(self, mr, mi):
        """Device (idx [n_q, 2] int32, w [n_q, 4]) for refractive indices mr, mi."""
        idx = np.empty((len(mr), 2), np.int32)
        w = np.empty((len(mr), 4))
        for q, (a, b) in enumerate(zip(mr, mi)):
            i, j, ww = corner_weights(np.array([a]), np.array([b]), self.allmr, self.allmi)
            idx[q] = i, j
            w[q] = [float(np.asarray(c).reshape(-1)[0]) for c in ww]
        return _lib.to_dev(idx, np.int32), _lib.to_dev(w)

    def interpolate(self, mr, mi, props_out=None):
        """Tables at n_q refractive indices: (props [6, n_q * n_x] with planes
        qext, qsca, qabs, qb, g, 0 -- the layout gm_integrate_psd_rows reads --
        and scama [n_q, n_x, 6, n_ang]) as device tensors."""
        t = _lib.torch()
        L, h, s = _lib.lib(), _lib.ctx(), stream()
        d_idx, d_w = self.queries(mr, mi)
        nq, nx = len(mr), self.n_x
        dev = _lib.device()
        props = t.zeros(6 * nq * nx, dtype=t.float64, device=dev)
        for p, k in enumerate(SCALAR_KEYS):
            check(L.gm_ri_bilinear(h, self.n_mr, self.n_mi, nx, ptr(self.d_scalar[k]), nq,
                                   ptr(d_idx), ptr(d_w), ptr(props[p * nq * nx:]), s),
                  "gm_ri_bilinear")
        n_el = nx * 6 * self.n_ang
        scama = t.empty(nq * n_el, dtype=t.float64, device=dev)
        check(L.gm_ri_bilinear(h, self.n_mr, self.n_mi, n_el, ptr(self.d_scama), nq, ptr(d_idx),
                               ptr(d_w), ptr(scama), s), "gm_ri_bilinear")
        self._keep = (d_idx, d_w)
        return props, scama


RHOW = 1000.


def build_kernel_table(params, water=None, classic=False, n_mom=0, kernels=None, ang=None):
    """An OpticsTable in kernel mode (dointegration.py:fun with 'mode':
    'kernel').  kernels: a KernelTables (default: read from
    params['kernel_params']['path'] and ['shape_dist'])."""
    from scipy.interpolate import interp1d
    from . import particle
    from .lut import SCALAR_OUT, angle_grid
    from .table import OpticsTable
    t = _lib.torch()
    kp = params.get("kernel_params")
    if kernels is None:
        if not kp or "path" not in kp or "shape_dist" not in kp:
            raise ValueError("kernel mode needs kernel_params with 'path' and 'shape_dist'")
        kernels = KernelTables.read(kp["path"], kp["shape_dist"])
    kind = params["psd"]["type"]
    pp = params["psd"]["params"]
    if kind == "lognorm":
        radii = list(pp["r0"])
    elif kind == "ss":
        radii = list(pp["rMinMaj"])
    elif kind == "du":
        radii = [r[0] for r in pp["rMinMaj"]]
    else:
        raise ValueError("psd type %s is not supported" % kind)
    mlist = params["mList"]
    if len(mlist) != 1:
        # get_interpolated bisects one refractive index (dointegration.py:41)
        raise ValueError("kernel mode takes one refractive-index component")
    lams = np.asarray(mlist[0][0])
    f_mr, f_mi = interp1d(mlist[0][0], mlist[0][1]), interp1d(mlist[0][0], mlist[0][2])
    trivial = params["rhDep"]["type"] == "trivial"
    if not trivial:
        if water is None:
            raise ValueError("the water refractive index table is required for rhDep %s"
                             % params["rhDep"]["type"])
        w_mr, w_mi = interp1d(water[0], water[1]), interp1d(water[0], water[2])
    ang_out = angle_grid() if ang is None else np.asarray(ang, np.float64)
    ang_k = np.linspace(0., 180., kernels.n_ang)              # :749-750
    rh_table = list(params["rh"])
    table = OpticsTable(len(radii), rh_table, lams, ang_out, classic=classic, n_mom=n_mom)
    xx = kernels.x
    dr = kernel_grid_widths(xx)
    nx = xx.size
    L, h, s = _lib.lib(), _lib.ctx(), stream()
    dev = _lib.device()
    d_S = _lib.to_dev(spline_matrix(ang_k, ang_out).T)       # [n_ang_k][n_out]
    d_th = _lib.to_dev(np.radians(ang_out))
    n_out = ang_out.size
    rhop00 = params["rhop0"]
    for radind in range(len(radii)):
        rhop0 = rhop00[radind] if isinstance(rhop00, list) else rhop00
        xx_n, dr_n = particle.initializeXarr(params, radind, lams[0], lams[-1])  # nominal (:782)
        fracs = pp["fracs"][0] if len(pp["fracs"]) == 1 else pp["fracs"][radind]   # :922-925
        rescale = kind == "lognorm" or (kind == "du" and len(pp["fracs"][radind]) == 1)
        for li, lam in enumerate(lams):
            nref0 = [complex(f_mr(lam), -f_mi(lam))]
            nw = complex(1, 0) if trivial else complex(w_mr(lam), w_mi(lam))
            rh = list(params["rh"])
            if "maxrh" in params:
                rh = list(np.where(np.array(rh) > params["maxrh"], params["maxrh"],
                                   np.array(rh)))
            _, _, _, rrat0 = particle.getHumidRefractiveIndex(params, radind, 0, rh, nref0, nw)
            _, reff0, _, _ = particle.calculatePSD(params, radind, 0., rh, xx, dr, rrat0, lam)
            todo = [0] if trivial else list(range(len(rh)))
            meta, rows, q_mr, q_mi = [], [], [], []
            for q, rhi in enumerate(todo):
                mr, mi, gf, rrat = particle.getHumidRefractiveIndex(params, radind, rhi, rh,
                                                                    nref0, nw)
                psd, _, rlo, rup = particle.calculatePSD(params, radind, rh[rhi], rh, xx, dr,
                                                         rrat, lam, want_ref=False)
                psd_n = None
                if rescale:
                    psd_n, _, _, _ = particle.calculatePSD(params, radind, rh[rhi], rh, xx_n,
                                                           dr_n, rrat, lam, want_ref=False)
                rhop = rrat ** 3. * rhop0 + (1. - rrat ** 3.) * RHOW
                q_mr.append(float(np.asarray(mr).reshape(-1)[0]))
                q_mi.append(float(np.asarray(mi).reshape(-1)[0]))
                mode_rows = []
                for k in range(len(fracs)):
                    mode_rows.append(len(rows))
                    rows.append((q, psd[k], reff0[k], rhop))
                meta.append((rhi, mr, mi, gf, rhop, rlo, rup, mode_rows, psd_n))
            nq, nr = len(todo), len(rows)
            props, scama = kernels.interpolate(q_mr, q_mi)
            n = nq * nx
            # dense [rows][points] weights: row (rh, mode) over its rh's block
            wts = np.zeros((nr, n))
            for r, (q, psd, _, _) in enumerate(rows):
                wts[r, q * nx:(q + 1) * nx] = psd
            d_w = _lib.to_dev(wts)
            n_cols = 6 * kernels.n_ang
            d_pk = t.empty(nr * n_cols, dtype=t.float64, device=dev)
            import ctypes
            for r0 in range(0, nr, 64):
                k = min(64, nr - r0)
                st = (ctypes.c_int64 * k)(*[rows[r0 + i][0] * nx for i in range(k)])
                ct = (ctypes.c_int64 * k)(*([nx] * k))
                check(L.gm_psd_contract_rows(h, k, st, ct, n, n_cols, ptr(d_w[r0:r0 + k]),
                                             ptr(scama), ptr(d_pk[r0 * n_cols:]), s),
                      "gm_psd_contract_rows")
            d_x = _lib.to_dev(np.tile(xx, nq))
            d_lp = t.full((n,), float(lam), dtype=t.float64, device=dev)
            d_r0 = _lib.to_dev(np.array([r[2] for r in rows], np.float64))
            d_rho = _lib.to_dev(np.array([r[3] for r in rows], np.float64))
            d_sc = t.empty(nr * len(SCALAR_OUT), dtype=t.float64, device=dev)
            check(L.gm_integrate_psd_rows(h, nr, n, ptr(d_x), ptr(d_lp), ptr(props), ptr(d_w),
                                          ptr(d_r0), float(rhop0), 0.0, ptr(d_rho), ptr(d_sc), s),
                  "gm_integrate_psd_rows")
            pk = d_pk.cpu().numpy().reshape(nr, 6, kernels.n_ang)
            sc = d_sc.cpu().numpy().reshape(nr, len(SCALAR_OUT))
            _kernel_wavelength(table, radind, li, lam, fracs, rhop0, meta, pk, sc, xx_n, rescale,
                               d_S, d_th, n_out, ang_out, n_mom, trivial, len(rh))
    return table


def _kernel_wavelength(table, radind, li, lam, fracs, rhop0, meta, pk, sc, xx_n, rescale, d_S,
                       d_th, n_out, ang_out, n_mom, trivial, n_rh):
    """Mode mixing, the mass-efficiency rescaling, the spline to the table's
    angles, normalisation and the table writes of one (bin, wavelength)."""
    from .lut import SCALAR_OUT
    t = _lib.torch()
    L, h, s = _lib.lib(), _lib.ctx(), stream()
    ne = len(meta)
    mode_rows = np.array([m[7] for m in meta], dtype=np.intp)
    mixed_p = np.zeros((ne,) + pk.shape[1:])
    mixed_s = np.zeros((ne, len(SCALAR_OUT)))
    for k in range(len(fracs)):
        mixed_p += fracs[k] * pk[mode_rows[:, k]]
        mixed_s += fracs[k] * sc[mode_rows[:, k]]
    S = {k: mixed_s[:, i].copy() for i, k in enumerate(SCALAR_OUT)}
    if rescale:
        # limited kernel-grid resolution: mass efficiencies referred to the
        # effective radius of the nominal grid, once per mode (:929-944)
        rr = np.asarray(xx_n) * lam / (2. * np.pi)
        for e, m in enumerate(meta):
            for k in range(len(fracs)):
                psd_n = m[8][k]
                reff = np.sum(rr ** 3. * psd_n) / np.sum(rr ** 2. * psd_n)
                for key in ("bsca", "bext", "bbck"):
                    S[key][e] = S[key][e] * S["rEff"][e] / reff
                S["rEff"][e] = reff
    # elements p11, p22, p33, p44, p12, p34 -> table order p11, p12, p22, p33, p34, p44,
    # then the spline to the output angles as one contraction (rows x n_ang_k)(n_ang_k x n_out)
    p6 = np.ascontiguousarray(mixed_p[:, [0, 4, 1, 2, 5, 3], :])
    d_p = _lib.to_dev(p6)
    d_pn = t.empty(ne * 6 * n_out, dtype=t.float64, device=_lib.device())
    check(L.gm_psd_contract(h, ne * 6, p6.shape[2], n_out, ptr(d_p), ptr(d_S), ptr(d_pn), s),
          "gm_psd_contract")
    d_b = t.empty(ne * 6, dtype=t.float64, device=_lib.device())
    check(L.gm_pmatrix_normalize6(h, ne, n_out, ptr(d_th), ptr(d_pn), ptr(d_b), s),
          "gm_pmatrix_normalize6")
    p_norm = d_pn.cpu().numpy().reshape(ne, 6, n_out)
    pback = d_b.cpu().numpy().reshape(ne, 6)
    pmom = None
    if n_mom:
        from .gsf import pmom_batch
        # table order p11, p12, p33, p34, p22, p44
        pmom = pmom_batch(ang_out, np.ascontiguousarray(p_norm[:, [0, 1, 3, 4, 2, 5], :]),
                          ng=n_mom)
    ret = {k: S[k] for k in ("qext", "qsca", "qabs", "qb", "g", "csca", "cext", "bsca", "bext",
                             "bbck")}
    ret["lidar_ratio"] = S["qext"] / S["qb"] * 4 * np.pi
    ret["ssa"] = S["qsca"] / S["qext"]
    ret.update(p11=p_norm[:, 0], p12=p_norm[:, 1], p22=p_norm[:, 2], p33=p_norm[:, 3],
               p34=p_norm[:, 4], p44=p_norm[:, 5], pback=pback)
    ret["refreal"] = np.array([np.asarray(m[1]).reshape(-1)[0] for m in meta])
    ret["refimag"] = -np.abs(np.array([np.asarray(m[2]).reshape(-1)[0] for m in meta]))
    mass0 = S["volume"][0] * rhop0
    nl = {"mass": S["mass"], "rEff": S["rEff"], "rMass": S["rMass"],
          "rLow": np.array([m[5] for m in meta]), "rUp": np.array([m[6] for m in meta]),
          "growth_factor": np.array([m[3] for m in meta]),
          "rhop": np.array([m[4] for m in meta]), "area": S["area"] / mass0,
          "volume": S["volume"] / mass0}
    rhis = np.array([m[0] for m in meta], dtype=np.intp)
    table.set_entries(radind, li, rhis, ret, pmom=pmom)
    table.set_bins_rh(radind, rhis, nl)
    if trivial:
        for rhi in range(1, n_rh):
            table.copy_rh(radind, li, 0, rhi)
