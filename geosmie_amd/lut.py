"""runoptics LUT engine surface on the GPU.

Drop-ins for the size-range Mie driver and the integration step of
src/geosmie/dointegration.py and src/pymiecoated/pymiecoated/mie_coated.py:

  MultipleMie           mie_coated.py:41-179 (preCalculate, calculateS12SizeRange)
  rawMie                dointegration.py:1211-1254
  calculateScatVals     dointegration.py:1044-1050
  integratePSD          dointegration.py:1064-1209 (mixing 'dumix')
  post_integration      dointegration.py:950-1001 (lidar ratio, ssa, P
                        normalisation, pback)
  lut_entries           the whole per-(wavelength, rh) chain kept on the
                        device: Mie batch -> S1/S2 -> Mueller -> PSD
                        contraction -> scalars -> normalisation.
plus the host-side size grid / lognormal helpers runoptics calls before the
hot path (getDR, getXArrCarma, getLogNormPSD, angle_grid).
"""
import numpy as np

from . import _lib
from ._lib import check, ptr, stream

scatkeys = ['p11', 'p12', 'p22', 'p33', 'p34', 'p44']         # dointegration.py:16
scalarkeys = ['qext', 'qsca', 'qabs', 'qb', 'g', 'csca', 'cext']  # dointegration.py:18
SCALAR_OUT = ['num', 'area', 'volume', 'mass', 'rEff', 'rMass', 'qext', 'qsca', 'qabs', 'qb',
              'g', 'csca', 'cext', 'bsca', 'bext', 'bbck']
PROP_KEYS = ("qext", "qsca", "qabs", "qb", "asy", "qratio")


# ---------------------------------------------------------------------------
# host-side grid and PSD helpers (the step before the hot path)
# ---------------------------------------------------------------------------
def angle_grid():
    """371 output angles in degrees: 0-1 (100), 1-10 (100), 10-180 (171) (dointegration.py:739-742)."""
    return np.concatenate([np.linspace(0., 1., 100, endpoint=False),
                           np.linspace(1., 10., 100, endpoint=False),
                           np.linspace(10., 180., 171)])


def getDR(arr):
    """Centred bin widths (dointegration.py:93-101)."""
    arr = np.asarray(arr, dtype=np.float64)
    d = np.empty_like(arr)
    d[0] = arr[1] - arr[0]
    d[1:-1] = ((arr[1:-1] - arr[:-2]) + (arr[2:] - arr[1:-1])) / 2.
    d[-1] = arr[-1] - arr[-2]
    return d


def getXArrCarma(minx, maxx, nbinperdecade):
    """CARMA volume-ratio size grid and widths (dointegration.py:103-153)."""
    nbin = np.log10(maxx / minx) * nbinperdecade
    rmrat = ((maxx / minx) ** 3) ** (1. / nbin)
    rmin = minx * ((1. + rmrat) / 2.) ** (1. / 3.)
    cpi = 4. / 3. * np.pi
    vrfact = ((3. / 2. / np.pi / (rmrat + 1)) ** (1. / 3.)) * (rmrat ** (1. / 3.) - 1.)
    rvolmin = cpi * rmin ** 3.
    # scalar powers (libm), not numpy's vector pow: the grid matches runoptics' bit for bit
    rvol = [float(rvolmin) * float(rmrat) ** i for i in range(int(nbin))]
    r = np.array([(v / cpi) ** (1. / 3.) for v in rvol])
    dr = np.array([vrfact * v ** (1. / 3.) for v in rvol])
    return r, dr


def getLogNormPSD(rmode, sigma, xxArr, lambd, rmax, rmin):
    """Lognormal dN/dx in size-parameter space (particleparams.py:113-127)."""
    k = 2 * np.pi / lambd
    xx = np.asarray(xxArr, dtype=np.float64)
    ls = np.log(sigma)
    d = 1. / (xx * (2 * np.pi) ** 0.5 * ls) * np.exp(-(np.log(xx / (rmode * k)) ** 2) / (2. * ls ** 2))
    d[xx >= rmax * k] = 0.
    d[xx <= rmin * k] = 0.
    return d


def lognorm_psd(rmode, sigma, xxarr, drarr, lam, rmax, rmin):
    """PSD weights as calculatePSD builds them (dointegration.py:569-583): (psd, reff_mass)."""
    p = getLogNormPSD(rmode, sigma, xxarr, lam, rmax, rmin) * drarr
    p = p / np.sum(p)
    rr = np.asarray(xxarr) * lam / 2. / np.pi
    return p, np.sum(rr ** 4. * p) / np.sum(rr ** 3. * p)


# ---------------------------------------------------------------------------
# device chain
# ---------------------------------------------------------------------------
class _DevMie(object):
    """Device coefficients of one homogeneous batch (kept resident)."""

    def __init__(self, xx, m, d_eps=None, x_max=None, offsets=None, yy=None, m2=None):
        """xx, m host arrays; or xx a device tensor of size parameters with
        d_eps (device, interleaved re/im of m^2) and x_max given.  offsets =
        (device coefficient offsets of xx, their total) reuses the offsets of
        an earlier batch on the same size parameters (no host wait).  yy, m2
        (host arrays): coated spheres with core sizes xx, shell sizes yy and
        shell index m2; the series length and efficiencies follow yy."""
        t = _lib.torch()
        self.d_y = self.d_eps2 = None
        if hasattr(xx, "data_ptr"):
            self.d_x, self.d_eps = xx, d_eps
            n = int(xx.numel())
            xmax = float(x_max)
        else:
            xx = np.ascontiguousarray(xx, dtype=np.float64)
            n = xx.size
            self.d_x = _lib.to_dev(xx)
            self.d_eps = _lib.cplx_dev(np.broadcast_to(np.asarray(m, np.complex128) ** 2, (n,)))
            xmax = float(xx.max())
            if yy is not None:
                yy = np.ascontiguousarray(yy, dtype=np.float64)
                if yy.shape != xx.shape:
                    raise ValueError("core and shell size arrays differ in length")
                self.d_y = _lib.to_dev(yy)
                self.d_eps2 = _lib.cplx_dev(
                    np.broadcast_to(np.asarray(m2, np.complex128) ** 2, (n,)))
                xmax = float(yy.max())
        self.n = n
        dev = _lib.device()
        L, h, s = _lib.lib(), _lib.ctx(), stream()
        size = self.d_y if self.d_y is not None else self.d_x
        if offsets is None:
            self.d_off = t.empty(n + 1, dtype=t.int64, device=dev)
            check(L.gm_coef_offsets(h, n, ptr(size), ptr(self.d_off), s), "gm_coef_offsets")
            total = int(self.d_off[-1].item())
        else:
            self.d_off, total = offsets
        self.d_an = t.empty(2 * max(total, 1), dtype=t.float64, device=dev)
        self.d_bn = t.empty_like(self.d_an)
        self.d_props = t.empty(6 * n, dtype=t.float64, device=dev)
        self.d_nmax = t.empty(n, dtype=t.int32, device=dev)
        check(L.gm_mie_batch(h, n, ptr(self.d_x), ptr(self.d_y), ptr(self.d_eps), None,
                             ptr(self.d_eps2), xmax, ptr(self.d_off), ptr(self.d_an),
                             ptr(self.d_bn), ptr(self.d_props), ptr(self.d_nmax),
                             _lib.GM_MIE_SORT, s), "gm_mie_batch")
        if self.d_y is not None and _lib.check_flags() & 2:
            raise ValueError("Multilayer calculations for magnetic particles "
                             "are not currently supported.")
        self.nmax_max = int(np.rint(2 + xmax + 4 * xmax ** (1. / 3.)))

    def mueller(self, pi_t, tau_t, precision=0):
        t = _lib.torch()
        n_ang = pi_t.shape[1]
        d_mu = t.empty(self.n * 4 * n_ang, dtype=t.float64, device=_lib.device())
        check(_lib.lib().gm_s12_mueller(_lib.ctx(), self.n, ptr(self.d_off), ptr(self.d_nmax),
                                        ptr(self.d_an), ptr(self.d_bn), int(n_ang), ptr(pi_t),
                                        ptr(tau_t), int(pi_t.shape[0]), ptr(d_mu), precision,
                                        stream()), "gm_s12_mueller")
        return d_mu

    def s12(self, pi_t, tau_t, precision=0):
        t = _lib.torch()
        n_ang = pi_t.shape[1]
        d1 = t.empty(2 * self.n * n_ang, dtype=t.float64, device=_lib.device())
        d2 = t.empty_like(d1)
        check(_lib.lib().gm_mie_s12(_lib.ctx(), self.n, ptr(self.d_off), ptr(self.d_nmax),
                                    ptr(self.d_an), ptr(self.d_bn), int(n_ang), ptr(pi_t),
                                    ptr(tau_t), int(pi_t.shape[0]), ptr(d1), ptr(d2), precision,
                                    stream()), "gm_mie_s12")
        return d1, d2


class MultipleMie(object):
    """Batched size-range Mie (mie_coated.py:41-179) on the GPU.

    The reference precomputes jv/yv per size and pi/tau per n_max on the
    host; here the Riccati-Bessel functions are computed inside the
    coefficient kernel and one device pi/tau table serves every size.

    Coated size ranges are not reachable in the reference: its branch calls an
    undefined name (mie_coated.py:80) and has no argument for the shell's
    index.  Here a coated range is xArr (core sizes) with yArr (shell sizes);
    calculateS12SizeRange(mr, mi, mr2, mi2) then takes the core index
    (mr, mi) and the shell index (mr2, mi2), and the efficiencies are
    normalised by the shell size y, as Mie(x=, y=) does (mie_coated.py:196).
    """

    def __init__(self, xArr, yArr, costarr):
        self.xArr = np.asarray(xArr, dtype=np.float64)
        self.yArr = None if yArr is None else np.asarray(yArr, dtype=np.float64)
        self.costarr = np.asarray(costarr, dtype=np.float64)
        self._tab = None

    def preCalculate(self):
        from .s12 import angle_tables_dev
        size = self.xArr if self.yArr is None else self.yArr
        nmax = int(np.rint(2 + size.max() + 4 * size.max() ** (1. / 3.)))
        self._tab = angle_tables_dev(self.costarr, nmax)

    def size_range_arrays(self, mr, mi, want="s12", mr2=None, mi2=None):
        """Arrays: 's1','s2' [N_x, N_ang] (or 'mueller' [N_x, 4, N_ang]) and
        efficiencies.  (mr2, mi2): shell index of a coated range (yArr given)."""
        if self._tab is None:
            self.preCalculate()
        m = complex(mr, mi)
        if self.yArr is not None:
            if mr2 is None:
                raise ValueError("Must specify both y and m2 for coated particles.")
            dm = _DevMie(self.xArr, m, yy=self.yArr, m2=complex(mr2, 0.0 if mi2 is None else mi2))
        else:
            dm = _DevMie(self.xArr, m)
        out = {}
        n_ang = self.costarr.size
        if want == "mueller":
            out["mueller"] = dm.mueller(*self._tab).cpu().numpy().reshape(dm.n, 4, n_ang)
        else:
            d1, d2 = dm.s12(*self._tab)
            out["s1"] = d1.cpu().numpy().view(np.complex128).reshape(dm.n, n_ang)
            out["s2"] = d2.cpu().numpy().view(np.complex128).reshape(dm.n, n_ang)
        props = dm.d_props.cpu().numpy().reshape(6, dm.n)
        for i, k in enumerate(PROP_KEYS):
            out[k] = props[i]
        return out

    def calculateS12SizeRange(self, mr, mi, mr2=None, mi2=None):
        """Same return structure as the reference (mie_coated.py:61-89);
        (mr2, mi2) is the shell index of a coated range."""
        a = self.size_range_arrays(mr, mi, mr2=mr2, mi2=mi2)
        ret = {"s12": [list(zip(a["s1"][i].tolist(), a["s2"][i].tolist()))
                       for i in range(self.xArr.size)]}
        for k in ("qext", "qsca", "qabs", "asy", "qb", "qratio"):
            ret[k] = a[k].tolist()
        return ret


def calculateScatVals(vals, scatkeys_, ret, numsizes, s1arr, s2arr):
    """Six Mueller planes into ret[6, N_x, N_ang] (dointegration.py:1044-1050)."""
    a1 = np.abs(s1arr) ** 2
    a2 = np.abs(s2arr) ** 2
    ret[0] = 0.5 * (a1 + a2)
    ret[1] = 0.5 * (a2 - a1)
    ret[2] = ret[0]
    c = s1arr * np.conj(s2arr)
    ret[3] = c.real
    ret[4] = c.imag
    ret[5] = ret[3]


def rawMie(mm, scatkeys_, scalarkeys_, lam, mr, mi, psd, costarr):
    """Per-size Mueller planes and efficiencies (dointegration.py:1211-1254)."""
    a = mm.size_range_arrays(mr, mi, want="mueller")
    rr = mm.xArr * lam / (2. * np.pi)
    mu = a["mueller"]
    ret = {"p11": mu[:, 0], "p12": mu[:, 1], "p22": mu[:, 0].copy(), "p33": mu[:, 2],
           "p34": mu[:, 3], "p44": mu[:, 2].copy()}
    ret["qext"] = a["qext"]
    ret["qsca"] = a["qsca"]
    ret["qabs"] = a["qabs"]
    ret["qb"] = a["qb"]
    ret["g"] = a["asy"]
    ret["csca"] = a["qsca"] * np.pi * rr ** 2
    ret["cext"] = a["qext"] * np.pi * rr ** 2
    return ret


def _contract(w, v):
    """out[k][c] = sum_p w[k][p] v[p][c] on the device (numpy in/out)."""
    t = _lib.torch()
    w = np.ascontiguousarray(np.atleast_2d(w), dtype=np.float64)
    v = np.ascontiguousarray(v, dtype=np.float64).reshape(w.shape[1], -1)
    d_w, d_v = _lib.to_dev(w), _lib.to_dev(v)
    d_o = t.empty(w.shape[0] * v.shape[1], dtype=t.float64, device=_lib.device())
    check(_lib.lib().gm_psd_contract(_lib.ctx(), w.shape[0], w.shape[1], v.shape[1], ptr(d_w),
                                     ptr(d_v), ptr(d_o), stream()), "gm_psd_contract")
    return d_o.cpu().numpy().reshape(w.shape[0], v.shape[1])


def _integrate_scalars(xxarr, props, psd, fracs, lam, reff0, rhop0, rhop, alias_quirk=True):
    t = _lib.torch()
    psd = np.ascontiguousarray(np.atleast_2d(psd), dtype=np.float64)
    d_out = t.empty(len(SCALAR_OUT), dtype=t.float64, device=_lib.device())
    # device copies held in locals until the kernels have consumed them
    d_x, d_p, d_w = _lib.to_dev(xxarr), _lib.to_dev(props), _lib.to_dev(psd)
    d_f = _lib.to_dev(np.asarray(fracs, np.float64))
    d_r = _lib.to_dev(np.asarray(reff0, np.float64))
    h = _lib.ctx()
    from .autotune import get, set_knob
    old = get(h, "psd_quirks")
    set_knob(h, "psd_quirks", 1 if alias_quirk else 0)
    try:
        check(_lib.lib().gm_integrate_psd(
            h, psd.shape[1], psd.shape[0], ptr(d_x), ptr(d_p), ptr(d_w), ptr(d_f), ptr(d_r),
            float(lam), float(rhop0), float(rhop), ptr(d_out), stream()), "gm_integrate_psd")
    finally:
        set_knob(h, "psd_quirks", old)
    return dict(zip(SCALAR_OUT, d_out.cpu().numpy().tolist()))


def integratePSD(xxarr, rawret, psd, fracs, lam, reff0, rhop0, rhop, alias_quirk=True):
    """integratePSD (dointegration.py:1064-1209, mixing 'dumix') on the GPU.

    rawret: list (one per mode) of rawMie dicts; psd: list of weight arrays.
    Each mode's phase-matrix planes and scalars are integrated with its own
    rawret entry and weighted by fracs, as the reference loops over modes.
    alias_quirk (default True, the reference's behaviour): csca and cext
    carry the Qsca weight the aliased thisweight gives every key after 'g'
    (:1111-1112, 1154-1157); False weights them by area only.
    """
    fracs = list(fracs)
    ret = {}
    for fi, frac in enumerate(fracs):
        rr = rawret[fi]
        planes = np.stack([rr["p11"], rr["p12"], rr["p33"], rr["p34"]], axis=1)  # [N_x, 4, N_ang]
        integ = _contract(psd[fi], planes).reshape(4, -1)
        props = np.stack([rr["qext"], rr["qsca"], rr["qabs"], rr["qb"], rr["g"],
                          np.zeros_like(rr["qext"])])
        sc = _integrate_scalars(xxarr, props, [psd[fi]], [1.0], lam, [reff0[fi]], rhop0, rhop,
                                alias_quirk)
        th = dict(sc)
        th["p11"], th["p12"], th["p33"], th["p34"] = integ
        th["p22"], th["p44"] = integ[0].copy(), integ[2].copy()
        th["lidar_ratio"] = 0.0
        for k, v in th.items():
            ret[k] = ret.get(k, 0.0) + frac * np.asarray(v)
    return ret


def post_integration(ret, ang_deg):
    """lidar ratio, ssa, normalised P, pback (dointegration.py:950-988), P on the GPU."""
    t = _lib.torch()
    out = dict(ret)
    qext, qsca, qb = float(ret["qext"]), float(ret["qsca"]), float(ret["qb"])
    out["lidar_ratio"] = qext / qb * 4 * np.pi
    out["ssa"] = qsca / qext
    n_ang = len(ang_deg)
    pm = np.stack([ret["p11"], ret["p12"], ret["p33"], ret["p34"]])[None]
    d_p = _lib.to_dev(pm.copy())
    d_b = t.empty(6, dtype=t.float64, device=_lib.device())
    d_th = _lib.to_dev(np.radians(np.asarray(ang_deg, dtype=np.float64)))
    check(_lib.lib().gm_pmatrix_normalize(_lib.ctx(), 1, n_ang, ptr(d_th), ptr(d_p), ptr(d_b),
                                          stream()), "gm_pmatrix_normalize")
    p = d_p.cpu().numpy().reshape(4, n_ang)
    out["p11"], out["p12"], out["p33"], out["p34"] = p
    out["p22"], out["p44"] = p[0].copy(), p[2].copy()
    out["pback"] = d_b.cpu().numpy()
    return out


def lut_entries(xx, lam, m, psd, ang_deg, fracs=(1.0,), reff0=None, rhop0=1000.0, rhop=1000.0,
                precision=0, normalize=True):
    """One device pass for a set of PSD rows sharing one size grid and index.

    xx [N_x] size parameters at wavelength `lam`, complex index m, psd
    [n_rows][N_x] number weights (each row one table entry's mode).  Returns
    dict of numpy arrays: 'pmat' [n_rows, 4, N_ang] (normalised if asked),
    'pback' [n_rows, 6], scalars per row (SCALAR_OUT) and 'ssa', 'lidar_ratio'.
    """
    from .s12 import angle_tables_dev
    t = _lib.torch()
    psd = np.ascontiguousarray(np.atleast_2d(psd), dtype=np.float64)
    n_rows, n_x = psd.shape
    ang_deg = np.asarray(ang_deg, dtype=np.float64)
    cost = np.cos(np.radians(ang_deg))
    dm = _DevMie(xx, m)
    pi_t, tau_t = angle_tables_dev(cost, dm.nmax_max)
    d_mu = dm.mueller(pi_t, tau_t, precision)
    n_ang = cost.size
    d_w = _lib.to_dev(psd)
    dev = _lib.device()
    d_pm = t.empty(n_rows * 4 * n_ang, dtype=t.float64, device=dev)
    L, h, s = _lib.lib(), _lib.ctx(), stream()
    check(L.gm_psd_contract(h, n_rows, n_x, 4 * n_ang, ptr(d_w), ptr(d_mu), ptr(d_pm), s),
          "gm_psd_contract")
    if reff0 is None:
        rr = np.asarray(xx) * lam / (2 * np.pi)
        reff0 = np.sum(rr[None] ** 4 * psd, 1) / np.sum(rr[None] ** 3 * psd, 1)
    reff0 = np.broadcast_to(np.asarray(reff0, np.float64), (n_rows,))
    scal = np.zeros((n_rows, len(SCALAR_OUT)))
    d_sc = t.empty(len(SCALAR_OUT), dtype=t.float64, device=dev)
    d_one = _lib.to_dev(np.array([1.0]))
    for r in range(n_rows):
        d_r0 = _lib.to_dev(reff0[r:r + 1].copy())
        check(L.gm_integrate_psd(h, n_x, 1, ptr(dm.d_x), ptr(dm.d_props), ptr(d_w[r]),
                                 ptr(d_one), ptr(d_r0), float(lam), float(rhop0), float(rhop),
                                 ptr(d_sc), s), "gm_integrate_psd")
        scal[r] = d_sc.cpu().numpy()
    d_b = t.empty(n_rows * 6, dtype=t.float64, device=dev)
    if normalize:
        d_th = _lib.to_dev(np.radians(ang_deg))
        check(L.gm_pmatrix_normalize(h, n_rows, n_ang, ptr(d_th), ptr(d_pm), ptr(d_b), s),
              "gm_pmatrix_normalize")
    out = {"pmat": d_pm.cpu().numpy().reshape(n_rows, 4, n_ang),
           "pback": d_b.cpu().numpy().reshape(n_rows, 6) if normalize else None}
    for i, k in enumerate(SCALAR_OUT):
        out[k] = scal[:, i]
    out["ssa"] = out["qsca"] / out["qext"]
    out["lidar_ratio"] = out["qext"] / out["qb"] * 4 * np.pi
    return out
