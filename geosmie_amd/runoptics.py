"""runoptics on the GPU: a whole optics table from a particle description.

Drop-in for the Mie branch of src/geosmie/dointegration.py:fun (672-1036)
and the runoptics command (src/geosmie/runoptics.py):

  * the front end is geosmie_amd.particle (readers, humidity rules, PSDs);
  * per bin the size-parameter grid is fixed, so the angle tables are built
    once per bin; per wavelength ONE device pass covers every rh of that
    wavelength: the points are the grid repeated once per (rh, refractive
    index), each with its own wet index, so the Mie batch, S1/S2 and Mueller
    planes of all rh run as one launch sequence;
  * every (rh, mode) is one weight row over its own block of points: the
    phase-matrix integrals come from gm_psd_contract_rows and the
    integratePSD scalars from gm_integrate_psd_rows with each row's wet
    density; modes are mixed with their fractions in the order integratePSD
    adds them (dointegration.py:1199-1200), then P is normalised on the
    device (gm_pmatrix_normalize) and the post-integration values follow
    (:950-1001);
  * the result goes to an OpticsTable (geosmie_amd.table) in either layout.

Mie mode ('mode': 'mie') with psd types lognorm, ss and du (the du grid is
the reference's non-monotonic one, see particle.initializeXarr); GRASP
kernel mode ('mode': 'kernel') is geosmie_amd.grasp.build_kernel_table, which
build_table dispatches to.
"""
import ctypes
import os

import numpy as np

from . import _lib, particle
from ._lib import check, ptr, stream
from .lut import SCALAR_OUT, _DevMie, angle_grid
from .table import OpticsTable

RHOW = 1000.          # density of water (dointegration.py:866)
ROWS_PER_CALL = 64    # gm_psd_contract_rows row limit


def _interp(xp, fp):
    from scipy.interpolate import interp1d
    return interp1d(xp, fp)


class _WavelengthPass(object):
    """Device pass over every rh of one (bin, wavelength)."""

    def __init__(self, xx, pi_t, tau_t, n_ang, precision, dr=None):
        self.xx, self.pi_t, self.tau_t, self.dr = xx, pi_t, tau_t, dr
        self.n_ang, self.precision = n_ang, precision

    def launch(self, lam, m_blocks, rows, rhop0):
        """Enqueue the pass; returns a handle for collect().  m_blocks:
        complex index per point block (each block = the whole grid); rows:
        list of (block, psd [N_x], reff0, rhop, spec), spec = (mode radius,
        ln sigma, lower and upper cut-off, in size-parameter units) of a
        lognormal row whose weights are formed on the device (psd None), or
        None when psd holds the weights.  Every host-to-device copy
        comes first and nothing here waits for a kernel, so the caller can
        prepare the next wavelength while these kernels run."""
        t = _lib.torch()
        nx, nb = self.xx.size, len(m_blocks)
        n = nx * nb
        nr = len(rows)
        dev = _lib.device()
        L, h, s = _lib.lib(), _lib.ctx(), stream()
        # host -> device first (a pageable copy waits for the stream's work)
        eps = np.asarray(m_blocks, np.complex128) ** 2
        d_eps = _lib.to_dev(eps.view(np.float64).reshape(nb, 2))
        on_dev = rows[0][4] is not None     # lognormal rows: weights formed on the device
        if on_dev:
            d_psd = _lib.to_dev(np.array([r[4] for r in rows], np.float64))
        else:
            d_psd = _lib.to_dev(np.stack([r[1] for r in rows]))
        d_blk = _lib.to_dev(np.array([r[0] for r in rows], np.int64), dtype=np.int64)
        d_r0 = _lib.to_dev(np.array([r[2] for r in rows], np.float64))
        d_rho = _lib.to_dev(np.array([r[3] for r in rows], np.float64))
        # the grid repeated once per block, and its coefficient offsets: the
        # same for every wavelength of the bin, so built once per block count
        if getattr(self, "_grid", None) is None or self._grid[0] != nb:
            d_x = _lib.to_dev(self.xx).repeat(nb)
            d_off = t.empty(n + 1, dtype=t.int64, device=dev)
            check(L.gm_coef_offsets(h, n, ptr(d_x), ptr(d_off), s), "gm_coef_offsets")
            self._grid = (nb, d_x, d_off, int(d_off[-1].item()))
            if self.dr is not None:      # one grid and its widths, for gm_lognorm_rows
                self._d_grid1 = (_lib.to_dev(self.xx), _lib.to_dev(self.dr))
        _, d_x, d_off, total = self._grid
        d_eps = d_eps.repeat_interleave(nx, dim=0).reshape(-1)
        dm = _DevMie(d_x, None, d_eps=d_eps, x_max=float(self.xx.max()), offsets=(d_off, total))
        # dense [rows][points] weights, zero outside each row's block: lognormal
        # rows formed on the device, others scattered from the host PSDs
        d_w = t.zeros((nr, nb, nx), dtype=t.float64, device=dev)
        d_col = None
        if on_dev:
            d_col = d_blk * nx
            check(L.gm_lognorm_rows(h, nr, nx, ptr(self._d_grid1[0]), ptr(self._d_grid1[1]),
                                    ptr(d_psd), ptr(d_col), ptr(d_w), n, s), "gm_lognorm_rows")
        else:
            d_w[t.arange(nr, device=dev), d_blk] = d_psd
        d_w = d_w.view(nr, n)
        n_cols = 4 * self.n_ang
        d_pm = t.empty(nr * n_cols, dtype=t.float64, device=dev)
        # FP64: S1/S2, the Mueller elements and their PSD sums in one pass
        # (no Mueller planes in memory); reduced precisions go through them
        d_mu = None if self.precision == 0 else dm.mueller(self.pi_t, self.tau_t, self.precision)
        cap = int(self.pi_t.shape[0])
        for r0 in range(0, nr, ROWS_PER_CALL):
            k = min(ROWS_PER_CALL, nr - r0)
            starts = (ctypes.c_int64 * k)(*[rows[r0 + i][0] * nx for i in range(k)])
            counts = (ctypes.c_int64 * k)(*([nx] * k))
            out = ptr(d_pm[r0 * n_cols:(r0 + k) * n_cols])
            if d_mu is None:
                check(L.gm_s12_contract_rows(h, n, ptr(dm.d_off), ptr(dm.d_nmax), ptr(dm.d_an),
                                             ptr(dm.d_bn), self.n_ang, ptr(self.pi_t),
                                             ptr(self.tau_t), cap, k, starts, counts,
                                             ptr(d_w[r0:r0 + k]), out, 0, None, s),
                      "gm_s12_contract_rows")
            else:
                check(L.gm_psd_contract_rows(h, k, starts, counts, n, n_cols,
                                             ptr(d_w[r0:r0 + k]), ptr(d_mu), out, s),
                      "gm_psd_contract_rows")
        d_lp = t.full((n,), float(lam), dtype=t.float64, device=dev)
        d_sc = t.empty(nr * len(SCALAR_OUT), dtype=t.float64, device=dev)
        check(L.gm_integrate_psd_rows(h, nr, n, ptr(dm.d_x), ptr(d_lp), ptr(dm.d_props), ptr(d_w),
                                      ptr(d_r0), float(rhop0), 0.0, ptr(d_rho), ptr(d_sc), s),
              "gm_integrate_psd_rows")
        # keep every buffer the enqueued kernels use alive until collect()
        return (nr, d_pm, d_sc, (dm, d_mu, d_w, d_lp, d_r0, d_rho, d_psd, d_blk, d_col))

    def collect(self, handle):
        """The un-normalised phase-matrix integrals [rows][4][n_ang] and the
        scalars [rows][16] (gm_integrate_psd order) of a launched pass."""
        nr, d_pm, d_sc, _ = handle
        pm = d_pm.cpu().numpy().reshape(nr, 4, self.n_ang)
        sc = d_sc.cpu().numpy().reshape(nr, len(SCALAR_OUT))
        return pm, sc

    def run(self, lam, m_blocks, rows, rhop0):
        return self.collect(self.launch(lam, m_blocks, rows, rhop0))


def _normalize(pm, ang):
    """P normalisation and pback of entries pm [e][4][n_ang] on the device."""
    t = _lib.torch()
    d_p = _lib.to_dev(np.ascontiguousarray(pm))
    d_b = t.empty(pm.shape[0] * 6, dtype=t.float64, device=_lib.device())
    d_th = _lib.to_dev(np.radians(ang))
    check(_lib.lib().gm_pmatrix_normalize(_lib.ctx(), pm.shape[0], pm.shape[2], ptr(d_th),
                                          ptr(d_p), ptr(d_b), stream()), "gm_pmatrix_normalize")
    return d_p.cpu().numpy().reshape(pm.shape), d_b.cpu().numpy().reshape(-1, 6)


def build_table(params, water=None, classic=False, precision=0, ang=None, n_mom=0,
                kernels=None):
    """An OpticsTable for a particle description (getPPJSON's dict with
    'mList'); water = getWaterM() array (needed unless rhDep is 'trivial').
    n_mom > 0 adds pmom, the GSF moments of every entry's normalised phase
    matrix (the rungsf step, src/gsf/convertncdf.py:170-189, 369-394; 129 in
    the reference), computed on the device for all rh of a wavelength at once.
    'mode': 'kernel' builds the table from GRASP kernel tables instead
    (geosmie_amd.grasp; kernels = a grasp.KernelTables, or read from
    params['kernel_params'])."""
    mode = params.get("mode", "mie")
    if mode == "kernel":
        from .grasp import build_kernel_table
        return build_kernel_table(params, water=water, classic=classic, n_mom=n_mom,
                                  kernels=kernels, ang=ang)
    if mode != "mie":
        raise ValueError("mode %s is not supported (mie or kernel)" % mode)
    kind = params["psd"]["type"]
    pp = params["psd"]["params"]
    if kind == "lognorm":
        radii = list(pp["r0"])
    elif kind == "ss":
        radii = list(pp["rMinMaj"])
    elif kind == "du":       # the first sub-bin's lower edge names the bin (:722-724)
        radii = [r[0] for r in pp["rMinMaj"]]
    else:
        raise ValueError("psd type %s is not supported by the Mie table path" % kind)
    mlist = params["mList"]
    lams = np.asarray(mlist[0][0])
    part_mr = [_interp(m[0], m[1]) for m in mlist]
    part_mi = [_interp(m[0], m[2]) for m in mlist]
    rhdep = params["rhDep"]["type"]
    if rhdep != "trivial":
        if water is None:
            raise ValueError("the water refractive index table is required for rhDep %s" % rhdep)
        w_mr, w_mi = _interp(water[0], water[1]), _interp(water[0], water[2])
    ang = angle_grid() if ang is None else np.asarray(ang, np.float64)
    rh_table = list(params["rh"])
    table = OpticsTable(len(radii), rh_table, lams, ang, classic=classic, n_mom=n_mom)
    rhop00 = params["rhop0"]
    cost = np.cos(np.radians(ang))
    from .s12 import angle_tables_dev
    for radind in range(len(radii)):
        rhop0 = rhop00[radind] if isinstance(rhop00, list) else rhop00
        xx, dr = particle.initializeXarr(params, radind, lams[0], lams[-1])
        nmax_cap = int(np.rint(2 + xx.max() + 4 * xx.max() ** (1. / 3.)))
        pi_t, tau_t = angle_tables_dev(cost, nmax_cap)
        dev_pass = _WavelengthPass(xx, pi_t, tau_t, ang.size, precision, dr)
        fracs = pp["fracs"][radind]
        water = (w_mr, w_mi) if rhdep != "trivial" else None

        def prepare(lam):
            return _prepare(params, radind, lam, xx, dr, fracs, rhop0, part_mr, part_mi, water)
        # the host front end of wavelength li+1 (humid indices, PSDs) runs
        # while the kernels of wavelength li execute
        prep = prepare(lams[0])
        for li, lam in enumerate(lams):
            handle = dev_pass.launch(lam, prep[0], prep[1], rhop0)
            nxt = prepare(lams[li + 1]) if li + 1 < len(lams) else None
            _wavelength(table, radind, li, fracs, rhop0, dev_pass, ang, prep, handle, n_mom)
            prep = nxt
    return table


def _prepare(params, radind, lam, xx, dr, fracs, rhop0, part_mr, part_mi, water):
    """Host front end of one (bin, wavelength): the point blocks (one wet index
    each), the weight rows and the per-rh metadata (dointegration.py:811-868)."""
    nref0 = [complex(f(lam), -g(lam)) for f, g in zip(part_mr, part_mi)]
    nw = complex(water[0](lam), water[1](lam)) if water is not None else complex(1, 0)
    rh = list(params["rh"])
    if "maxrh" in params:     # rh above maxrh computed at maxrh (:824-828)
        rh = list(np.where(np.array(rh) > params["maxrh"], params["maxrh"], np.array(rh)))
    trivial = params["rhDep"]["type"] == "trivial"
    # dry state: effective radii the mass efficiencies are referred to (:835-836)
    _, _, _, rrat0 = particle.getHumidRefractiveIndex(params, radind, 0, rh, nref0, nw)
    _, reff0, _, _ = particle.calculatePSD(params, radind, 0., rh, xx, dr, rrat0, lam)
    todo = [0] if trivial else list(range(len(rh)))
    lognorm = params["psd"]["type"] == "lognorm"
    pp, rd = params["psd"]["params"], params["rhDep"]
    kx = 2 * np.pi / lam
    blocks, rows, meta = [], [], []
    for rhi in todo:
        mr, mi, gf, rrat = particle.getHumidRefractiveIndex(params, radind, rhi, rh, nref0, nw)
        if lognorm:
            # the weights are formed on the device (gm_lognorm_rows): per mode the
            # grown mode radius and upper cut-off, in size-parameter units, as
            # calculatePSD passes them to getLogNormPSD (dointegration.py:551-576)
            spec = [(particle.humidityGrowth(rd, r0, rh[rhi], rh) * kx,
                     np.log(pp["sigma"][radind][k]), pp["rmin0"][radind][k] * kx,
                     particle.humidityGrowth(rd, pp["rmax0"][radind][k], rh[rhi], rh) * kx)
                    for k, r0 in enumerate(pp["r0"][radind])]
            psd = [None] * len(spec)
            rlo, rup = pp["rmin0"][radind][0], pp["rmax0"][radind][0]
        else:
            psd, _, rlo, rup = particle.calculatePSD(params, radind, rh[rhi], rh, xx, dr, rrat,
                                                     lam, want_ref=False)
            spec = [None] * len(psd)
        rhop = rrat ** 3. * rhop0 + (1. - rrat ** 3.) * RHOW
        b0 = len(blocks)
        blocks.extend(complex(a, b) for a, b in zip(mr, mi))
        n_ri = len(mr)
        mode_rows = []
        for k in range(len(fracs)):
            mode_rows.append(len(rows))
            rows.append((b0 + (k if n_ri > 1 else 0), psd[k], reff0[k], rhop, spec[k]))
        meta.append((rhi, mr, mi, gf, rhop, rlo, rup, mode_rows))
    return blocks, rows, meta, trivial, len(rh)


def _wavelength(table, radind, li, fracs, rhop0, dev_pass, ang, prep, handle, n_mom=0):
    """All rh entries of one (bin, wavelength) (dointegration.py:869-1031):
    the launched device pass's results, mode mixing, normalisation and the
    table writes."""
    blocks, rows, meta, trivial, n_rh = prep
    pm, sc = dev_pass.collect(handle)
    # mix modes in integratePSD's order (ret += frac * thisret, :1199-1200),
    # all entries of the wavelength at once, then normalise on the device
    rhis = np.array([m[0] for m in meta], dtype=np.intp)
    mode_rows = np.array([m[7] for m in meta], dtype=np.intp)     # [entry][mode]
    mixed_p = np.zeros((len(meta), 4, ang.size))
    mixed_s = np.zeros((len(meta), len(SCALAR_OUT)))
    for k in range(len(fracs)):
        mixed_p += fracs[k] * pm[mode_rows[:, k]]
        mixed_s += fracs[k] * sc[mode_rows[:, k]]
    p_norm, pback = _normalize(mixed_p, ang)
    pmom = None
    if n_mom:
        from .gsf import pmom_batch
        # table order p11, p12, p33, p34, p22 (= p11), p44 (= p33)
        p6 = p_norm[:, [0, 1, 2, 3, 0, 2], :]
        pmom = pmom_batch(ang, p6, ng=n_mom)
    S = {k: mixed_s[:, i] for i, k in enumerate(SCALAR_OUT)}
    ret = {k: S[k] for k in ("qext", "qsca", "qabs", "qb", "g", "csca", "cext", "bsca", "bext",
                             "bbck")}
    ret["lidar_ratio"] = S["qext"] / S["qb"] * 4 * np.pi          # :957-961
    ret["ssa"] = S["qsca"] / S["qext"]
    ret.update(p11=p_norm[:, 0], p12=p_norm[:, 1], p22=p_norm[:, 0], p33=p_norm[:, 2],
               p34=p_norm[:, 3], p44=p_norm[:, 2], pback=pback)
    ret["refreal"] = np.array([m[1][0] for m in meta])            # :995-996
    ret["refimag"] = -np.abs(np.array([m[2][0] for m in meta]))
    # area and volume per dry mass: mass0 from the first rh (:983-988)
    mass0 = S["volume"][0] * rhop0
    nl = {"mass": S["mass"], "rEff": S["rEff"], "rMass": S["rMass"],
          "rLow": np.array([m[5] for m in meta]), "rUp": np.array([m[6] for m in meta]),
          "growth_factor": np.array([m[3] for m in meta]),
          "rhop": np.array([m[4] for m in meta]), "area": S["area"] / mass0,
          "volume": S["volume"] / mass0}
    table.set_entries(radind, li, rhis, ret, pmom=pmom)
    table.set_bins_rh(radind, rhis, nl)
    if trivial:   # copyDryValues (:462-490): every rh repeats the dry entry
        for rhi in range(1, n_rh):
            table.copy_rh(radind, li, 0, rhi)


def fun(partID0, oppfx=".", oppclassic=False, data_dir=None, precision=0, n_mom=0):
    """runoptics for one particle file: builds the table and writes
    optics_<name>.nomom.nc4 (or .nomom.legacy.nc4) under oppfx
    (dointegration.py:672-1036, 205-213).  n_mom > 0 also runs the GSF step
    and writes optics_<name>.nc4 with pmom instead, the name rungsf gives its
    output (convertncdf.py:328)."""
    part = os.path.basename(partID0).replace(".json", "")
    src = partID0.replace("-orig", "") if "-orig" in part else partID0
    params = particle.getPPJSON(src, data_dir)
    water = None
    if params["rhDep"]["type"] != "trivial":
        water = particle.getWaterM(data_dir=data_dir)
    table = build_table(params, water=water, classic=oppclassic, precision=precision, n_mom=n_mom)
    name = "optics_%s.nomom.legacy.nc4" % part if oppclassic else "optics_%s.nomom.nc4" % part
    if n_mom:
        name = name.replace("nomom.", "")
    path = os.path.join(oppfx, name)
    table.write(path)
    return path
