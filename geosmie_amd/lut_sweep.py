"""LUT sweep on the GPU: bulk optics over (m_r, m_i, wavelength, RH, PSD) grids.

One slab is one (m_dry, wavelength, growth factor) combination over a dry
radius grid: an iteration of runoptics' (lambda, rh) loop
(src/geosmie/dointegration.py:787-889) with the 'simple' humidity rule
(:497-507: wet radius gf r, index m_w + (m_dry - m_w) / gf^3).  For every
slab and every lognormal (r_mode, sigma_g) pair the sweep returns float32
rows  Qext, SSA, g, alpha1_0 .. alpha1_{n_mom-1}:  integratePSD's
area-weighted efficiencies and Qsca-weighted g (dointegration.py:1102-1190,
961) and the normalised Legendre moments of P11 that spher_expan.f forms
(SPHER_EXPAN :313-347, CNORM :95-103).

Device chain per batch of slabs (nothing returns to the host in between):
gm_lut_points -> gm_coef_offsets -> gm_mie_batch (sorted, FP64) ->
gm_angle_tables (cached) -> gm_s12_p11 (MFMA) -> gm_lut_bulk.
"""
import numpy as np

from . import _lib
from ._lib import check, ptr, stream
from .lut import angle_grid, getDR

M_WATER = 1.33 + 0j
NSPHER = 129     # Gauss nodes of spher_expan.f (params.h:14)
N_MOM = 32

# Growth factors of the 36-point RH grid (src/config/geosparticles/su.json:5-14).
SU_RH = np.array([0.00, 0.05, 0.10, 0.15, 0.20, 0.25, 0.30, 0.35, 0.40, 0.45, 0.50,
                  0.55, 0.60, 0.65, 0.70, 0.75, 0.80, 0.81, 0.82, 0.83, 0.84, 0.85,
                  0.86, 0.87, 0.88, 0.89, 0.90, 0.91, 0.92, 0.93, 0.94, 0.95, 0.96,
                  0.97, 0.98, 0.99])
SU_GF = np.array([1.00, 1.04, 1.08, 1.12, 1.16, 1.20, 1.23, 1.27, 1.31, 1.35, 1.39,
                  1.43, 1.46, 1.50, 1.54, 1.59, 1.64, 1.65, 1.66, 1.67, 1.68, 1.69,
                  1.71, 1.72, 1.74, 1.75, 1.77, 1.79, 1.82, 1.84, 1.87, 1.91, 1.94,
                  1.99, 2.05, 2.16])


def full_grid():
    """The full sweep: m_r 1.30-1.80 (26), m_i 1e-8-1 log (30), 3000 dry radii
    0.005-50 um log, 50 wavelengths 0.2-40 um log, the 36 growth factors
    above, 8 r_mode 0.01-5 um log x 5 sigma_g 1.4-2.6 (40 lognormal PSDs)."""
    return {"mr": np.linspace(1.30, 1.80, 26), "mi": np.logspace(-8.0, 0.0, 30),
            "r": np.logspace(np.log10(0.005), np.log10(50.0), 3000),
            "lam": np.logspace(np.log10(0.2), np.log10(40.0), 50), "gf": SU_GF.copy(),
            "rmode": np.logspace(-2.0, np.log10(5.0), 8), "sigma": np.linspace(1.4, 2.6, 5)}


def nmax_rule(s):
    """n_max = round(2 + s + 4 s^(1/3)), half to even (mie_coeffs.py:99)."""
    s = np.asarray(s, dtype=np.float64)
    return np.rint(2 + s + 4 * s ** (1.0 / 3.0)).astype(np.int64)


def lognormal_weights(r, rmode, sigma):
    """Number weights n(r) dr on the radius grid r, normalised to sum 1: the
    lognormal of particleparams.py:113-127 times getDR widths, as
    calculatePSD builds a mode (dointegration.py:569-577)."""
    r = np.asarray(r, dtype=np.float64)
    ls = np.log(sigma)
    d = np.exp(-np.log(r / rmode) ** 2 / (2 * ls * ls)) / (r * np.sqrt(2 * np.pi) * ls)
    w = d * getDR(r)
    return w / w.sum()


def effective_radius(r, w):
    """sum r^3 w / sum r^2 w (integratePSD rEff, dointegration.py:1114)."""
    return float(np.sum(r ** 3 * w) / np.sum(r ** 2 * w))


class LutSweep(object):
    """Bulk optics of a fixed set of lognormal PSDs over arbitrary slabs."""

    def __init__(self, r_dry, rmode, sigma, ang_deg=None, ng=NSPHER, n_mom=N_MOM,
                 precision="fp64", m_water=M_WATER, pairs=None):
        """PSDs: every (r_mode, sigma_g) of rmode x sigma, or the explicit
        list `pairs` of (r_mode, sigma_g) when given."""
        t = _lib.torch()
        self.r = np.ascontiguousarray(r_dry, dtype=np.float64)
        self.n_r = self.r.size
        if pairs is None:
            pairs = [(a, b) for a in np.atleast_1d(rmode) for b in np.atleast_1d(sigma)]
        self.pairs = [(float(a), float(b)) for a, b in pairs]
        w = np.stack([lognormal_weights(self.r, a, b) for a, b in self.pairs])
        self.w = w
        self.n_psd = w.shape[0]
        self.reff_dry = np.array([effective_radius(self.r, wk) for wk in w])
        self.ang = np.ascontiguousarray(angle_grid() if ang_deg is None else ang_deg,
                                        dtype=np.float64)
        self.n_ang = self.ang.size
        self.ng, self.n_mom = int(ng), int(n_mom)
        self.n_out = 3 + self.n_mom
        from .s12 import PREC
        self.precision = PREC[precision] if isinstance(precision, str) else int(precision)
        self.m_water = complex(m_water)
        self.d_r = _lib.to_dev(self.r)
        self.d_dr = _lib.to_dev(getDR(self.r))
        self.d_w = _lib.to_dev(w)
        self.d_cost = _lib.to_dev(np.cos(np.radians(self.ang)))
        self.nmax_cap = 0
        self.d_pi = self.d_tau = None
        self._t = t

    def _tables(self, nmax_cap):
        if nmax_cap <= self.nmax_cap:
            return
        t = self._t
        cap = int(nmax_cap * 1.1) + 8        # headroom: later batches reuse the table
        dev = _lib.device()
        self.d_pi = t.empty(cap * self.n_ang, dtype=t.float64, device=dev)
        self.d_tau = t.empty_like(self.d_pi)
        check(_lib.lib().gm_angle_tables(_lib.ctx(), self.n_ang, ptr(self.d_cost), cap,
                                         ptr(self.d_pi), ptr(self.d_tau), stream()),
              "gm_angle_tables")
        self.nmax_cap = cap

    def weights_dev(self, rmode, sigma):
        """Lognormal weights [n, n_r] on the device (gm_lut_lognormal)."""
        t = self._t
        rmode = np.ascontiguousarray(np.atleast_1d(rmode), dtype=np.float64)
        sigma = np.ascontiguousarray(np.atleast_1d(sigma), dtype=np.float64)
        n = rmode.size
        d_w = t.empty((n, self.n_r), dtype=t.float64, device=_lib.device())
        d_rm, d_sg = _lib.to_dev(rmode), _lib.to_dev(sigma)
        check(_lib.lib().gm_lut_lognormal(_lib.ctx(), n, self.n_r, ptr(self.d_r), ptr(self.d_dr),
                                          ptr(d_rm), ptr(d_sg), ptr(d_w), stream()),
              "gm_lut_lognormal")
        return d_w

    def slab_cost(self, lam, gf):
        """Sum of n_max over the radius grid for each (lam, gf) (host estimate)."""
        lam = np.atleast_1d(lam)
        gf = np.atleast_1d(gf)
        x = 2 * np.pi * gf[:, None] * self.r[None, :] / lam[:, None]
        return nmax_rule(x).sum(axis=1)

    def run(self, lam, m_dry, gf, w=None, out=None):
        """Bulk rows of n_slab slabs: float32 device tensor [n_slab, n_psd, 3 + n_mom].

        lam, gf: [n_slab]; m_dry: complex [n_slab] (positive imaginary part);
        w: optional per-slab weights [n_slab, n_psd, n_r] (device or host)
        instead of the sweep's shared PSDs; out: optional float32 device view
        of n_slab * n_psd * (3 + n_mom) elements to write into.
        """
        t = self._t
        lam = np.ascontiguousarray(np.atleast_1d(lam), dtype=np.float64)
        gf = np.ascontiguousarray(np.atleast_1d(gf), dtype=np.float64)
        m_dry = np.ascontiguousarray(np.atleast_1d(m_dry), dtype=np.complex128)
        n_slab = lam.size
        if gf.size != n_slab or m_dry.size != n_slab:
            raise ValueError("lam, m_dry and gf must have one entry per slab")
        if np.any(lam <= 0) or np.any(gf < 1.0) or np.any(m_dry.imag < 0):
            raise ValueError("need lam > 0, gf >= 1 and Im(m_dry) >= 0")
        n_psd = self.n_psd
        if w is not None:
            if not hasattr(w, "data_ptr"):
                w = _lib.to_dev(np.ascontiguousarray(w, dtype=np.float64))
            n_psd = w.numel() // (n_slab * self.n_r)
            if n_psd * n_slab * self.n_r != w.numel():
                raise ValueError("w must be [n_slab, n_psd, n_r]")
        L, h, s = _lib.lib(), _lib.ctx(), stream()
        dev = _lib.device()
        N = n_slab * self.n_r
        d_lam = _lib.to_dev(lam)
        d_gf = _lib.to_dev(gf)
        d_m = _lib.cplx_dev(m_dry)
        d_x = t.empty(N, dtype=t.float64, device=dev)
        d_eps = t.empty(2 * N, dtype=t.float64, device=dev)
        mw = self.m_water
        check(L.gm_lut_points(h, n_slab, self.n_r, ptr(self.d_r), ptr(d_lam), ptr(d_m), ptr(d_gf),
                              mw.real, mw.imag, ptr(d_x), ptr(d_eps), s), "gm_lut_points")
        size_max = float(np.max(2 * np.pi * gf / lam) * self.r.max()) * (1 + 1e-12)
        d_off = t.empty(N + 1, dtype=t.int64, device=dev)
        check(L.gm_coef_offsets(h, N, ptr(d_x), ptr(d_off), s), "gm_coef_offsets")
        total = int(d_off[-1].item())
        d_an = t.empty(2 * max(total, 1), dtype=t.float64, device=dev)
        d_bn = t.empty_like(d_an)
        d_props = t.empty(6 * N, dtype=t.float64, device=dev)
        d_nmax = t.empty(N, dtype=t.int32, device=dev)
        check(L.gm_mie_batch(h, N, ptr(d_x), None, ptr(d_eps), None, None, size_max, ptr(d_off),
                             ptr(d_an), ptr(d_bn), ptr(d_props), ptr(d_nmax), _lib.GM_MIE_SORT, s),
              "gm_mie_batch")
        self._tables(int(nmax_rule(size_max)))
        d_p11 = t.empty(N * self.n_ang, dtype=t.float64, device=dev)
        check(L.gm_s12_p11(h, N, ptr(d_off), ptr(d_nmax), ptr(d_an), ptr(d_bn), self.n_ang,
                           ptr(self.d_pi), ptr(self.d_tau), self.nmax_cap, ptr(d_p11),
                           self.precision, s), "gm_s12_p11")
        if out is None:
            out = t.empty(n_slab * n_psd * self.n_out, dtype=t.float32, device=dev)
        elif out.numel() != n_slab * n_psd * self.n_out or out.dtype != t.float32:
            raise ValueError("out must hold n_slab * n_psd * (3 + n_mom) float32 values")
        check(L.gm_lut_bulk(h, n_slab, self.n_r, self.n_ang, n_psd,
                            ptr(self.d_w if w is None else w), 0 if w is None else 1, ptr(d_x),
                            ptr(d_lam), ptr(d_props), ptr(d_p11),
                            self.ang.ctypes.data_as(_lib.c_vp), self.ng, self.n_mom, ptr(out), s),
              "gm_lut_bulk")
        return out.view(n_slab, n_psd, self.n_out)

    def sweep(self, mr, mi, lam, gf, max_coeffs=1.5e8, max_points=6e6, progress=None):
        """The full product grid: float32 device tensor
        [n_mr, n_mi, n_lam, n_gf, n_psd, 3 + n_mom] (slab order mr, mi, lam, gf).
        Slabs are cut into consecutive batches of bounded coefficient count."""
        t = self._t
        mr, mi = np.atleast_1d(mr), np.atleast_1d(mi)
        lam, gf = np.atleast_1d(lam), np.atleast_1d(gf)
        shape = (mr.size, mi.size, lam.size, gf.size)
        n_slab = int(np.prod(shape))
        out = t.empty(n_slab * self.n_psd * self.n_out, dtype=t.float32, device=_lib.device())
        cost_lg = self.slab_cost(np.repeat(lam, gf.size), np.tile(gf, lam.size)).reshape(
            lam.size, gf.size)
        per_slab = np.broadcast_to(cost_lg[None, None], shape).ravel()
        max_slabs = max(1, int(max_points // self.n_r))
        idx = np.arange(n_slab)
        i_mr, i_mi, i_lam, i_gf = np.unravel_index(idx, shape)
        m_all = mr[i_mr] + 1j * mi[i_mi]
        start = 0
        row = self.n_psd * self.n_out
        while start < n_slab:
            csum = np.cumsum(per_slab[start:start + max_slabs])
            cnt = max(1, int(np.searchsorted(csum, max_coeffs, side="right")))
            stop = min(n_slab, start + cnt)
            self.run(lam[i_lam[start:stop]], m_all[start:stop], gf[i_gf[start:stop]],
                     out=out[start * row:stop * row])
            if progress is not None:
                progress(stop, n_slab)
            start = stop
        return out.view(*shape, self.n_psd, self.n_out)
