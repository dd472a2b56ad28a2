"""The runoptics chain for one size grid, kept on the device and replayable
as a HIP graph.

One step computes, for a homogeneous batch of points (size parameters x at
refractive index m, grouped into rows that share a wavelength) what
runoptics' per-(wavelength, rh) iteration computes
(src/geosmie/dointegration.py:787-1001):

  gm_coef_offsets, gm_mie_batch          a_n, b_n and efficiencies (FP64)
  (gm_angle_tables                        pi'_n, tau'_n: once, at construction)
  (gm_s12_split_coef                       FP32 / split-BF16: U = a + b, V = a - b
                                          converted once)
  gm_s12_contract_rows                    S1/S2 on MFMA -> s11, s12, s33, s34 ->
                                          phase-matrix integrals of integratePSD,
                                          in one pass (fused=False: gm_s12_mueller,
                                          then gm_psd_contract_rows)
  gm_integrate_psd_rows                   the integratePSD scalars of every row
  gm_pmatrix_normalize                    P normalisation and pback
  gm_gsf_expand (optional)                GSF moments of every row

Every buffer is allocated up front and the chain owns a private gm_ctx, so
the launches of one step can be captured once (hipStreamBeginCapture via
torch.cuda.graph) and replayed: the graph's kernel arguments stay valid
because nothing else can grow or free that context's workspaces.
"""
import ctypes

import numpy as np

from . import _lib
from ._lib import check, ptr


class RunopticsChain(object):

    def __init__(self, x, m, ang, w, lam_of_row, ranges, precision=0, gsf_ng=0, rhop0=1000.0,
                 rhop=1000.0, fused=True):
        """x [n] size parameters, m complex index (scalar or [n]), ang [n_ang]
        degrees, w [rows, n] number weights (row r nonzero only on its points
        ranges[r] = (start, count)), lam_of_row [rows] wavelengths.  fused:
        S1/S2 and the row integrals in one pass, without the [n][4][n_ang]
        Mueller planes (at FP32 / split-BF16 on coefficients converted once
        per step by gm_s12_split_coef)."""
        t = _lib.torch()
        self.t = t
        x = np.ascontiguousarray(x, dtype=np.float64)
        w = np.ascontiguousarray(np.atleast_2d(w), dtype=np.float64)
        self.n = x.size
        self.size_max = float(x.max())
        self.nmax_cap = int(np.rint(2 + self.size_max + 4 * self.size_max ** (1. / 3.)))
        self.ang = np.ascontiguousarray(ang, dtype=np.float64)
        self.n_ang = self.ang.size
        self.rows = w.shape[0]
        self.lam = np.asarray(lam_of_row, dtype=np.float64)
        self.ranges = [(int(a), int(c)) for a, c in ranges]
        if len(self.ranges) != w.shape[0] or not 1 <= w.shape[0] <= 64:
            raise ValueError("one (start, count) range per weight row, 1 to 64 rows")
        self.precision = int(precision)
        self.fused = bool(fused)
        self.gsf_ng = int(gsf_ng)
        self.rhop0, self.rhop = float(rhop0), float(rhop)
        self.ctx = _lib.private_ctx()
        dev = _lib.device()
        self.d_x = _lib.to_dev(x)
        self.d_eps = _lib.cplx_dev(np.broadcast_to(np.asarray(m, np.complex128) ** 2, (self.n,)))
        self.d_cost = _lib.to_dev(np.cos(np.radians(self.ang)))
        self.d_theta = _lib.to_dev(np.radians(self.ang))
        self.d_w = _lib.to_dev(w)
        nm = np.rint(2 + x + 4 * x ** (1. / 3.)).astype(np.int64)
        self.sum_nmax = int(nm.sum())
        f64 = t.float64
        self.d_off = t.empty(self.n + 1, dtype=t.int64, device=dev)
        self.d_an = t.empty(2 * self.sum_nmax, dtype=f64, device=dev)
        self.d_bn = t.empty_like(self.d_an)
        self.d_props = t.empty(6 * self.n, dtype=f64, device=dev)
        self.d_nmax = t.empty(self.n, dtype=t.int32, device=dev)
        self.d_pi = t.empty(self.nmax_cap * self.n_ang, dtype=f64, device=dev)
        self.d_tau = t.empty_like(self.d_pi)
        self.d_mu = None   # Mueller planes, allocated when a step needs them
        self.d_cl = None   # converted coefficients (16 B per pair), reduced precision
        self.d_pm = t.empty(self.rows * 4 * self.n_ang, dtype=f64, device=dev)
        self.d_back = t.empty(self.rows * 6, dtype=f64, device=dev)
        self.d_sc = t.empty(self.rows * 16, dtype=f64, device=dev)
        rr = x[None, :] * self.lam[:, None] / (2 * np.pi)
        self.d_reff0 = _lib.to_dev((w * rr ** 4).sum(1) / (w * rr ** 3).sum(1))
        lam_pts = np.ones(self.n)
        for r, (a, c) in enumerate(self.ranges):
            lam_pts[a:a + c] = self.lam[r]
        self.d_lam_pts = _lib.to_dev(lam_pts)
        self._seg_start = (ctypes.c_int64 * self.rows)(*[a for a, _ in self.ranges])
        self._seg_count = (ctypes.c_int64 * self.rows)(*[c for _, c in self.ranges])
        if self.gsf_ng:
            self.d_F = t.empty(self.rows * 6 * self.n_ang, dtype=f64, device=dev)
            self.d_pmom = t.empty(self.rows * 6 * self.gsf_ng, dtype=f64, device=dev)
        self.graph = None
        # pi'/tau' depend only on the angles and nmax_cap: built once, as
        # MultipleMie.preCalculate does (mie_coated.py:149-151, 160-179)
        check(_lib.lib().gm_angle_tables(self.ctx, self.n_ang, ptr(self.d_cost), self.nmax_cap,
                                         ptr(self.d_pi), ptr(self.d_tau), _lib.stream()),
              "gm_angle_tables")

    def step(self):
        """Enqueue one pass on torch's current stream."""
        L, h, s, P = _lib.lib(), self.ctx, _lib.stream(), ptr
        check(L.gm_coef_offsets(h, self.n, P(self.d_x), P(self.d_off), s), "gm_coef_offsets")
        check(L.gm_mie_batch(h, self.n, P(self.d_x), None, P(self.d_eps), None, None,
                             self.size_max, P(self.d_off), P(self.d_an), P(self.d_bn),
                             P(self.d_props), P(self.d_nmax), _lib.GM_MIE_SORT, s), "gm_mie_batch")
        if self.fused:
            cl = None
            if self.precision:
                if self.d_cl is None:
                    self.d_cl = self.t.empty(2 * self.sum_nmax, dtype=self.t.float64,
                                             device=_lib.device())
                check(L.gm_s12_split_coef(h, self.sum_nmax, P(self.d_an), P(self.d_bn),
                                          self.precision, P(self.d_cl), s), "gm_s12_split_coef")
                cl = P(self.d_cl)
            check(L.gm_s12_contract_rows(h, self.n, P(self.d_off), P(self.d_nmax), P(self.d_an),
                                         P(self.d_bn), self.n_ang, P(self.d_pi), P(self.d_tau),
                                         self.nmax_cap, self.rows, self._seg_start,
                                         self._seg_count, P(self.d_w), P(self.d_pm),
                                         self.precision, cl, s), "gm_s12_contract_rows")
        else:
            if self.d_mu is None:
                self.d_mu = self.t.empty(self.n * 4 * self.n_ang, dtype=self.t.float64,
                                         device=_lib.device())
            check(L.gm_s12_mueller(h, self.n, P(self.d_off), P(self.d_nmax), P(self.d_an),
                                   P(self.d_bn), self.n_ang, P(self.d_pi), P(self.d_tau),
                                   self.nmax_cap, P(self.d_mu), self.precision, s),
                  "gm_s12_mueller")
            # one launch integrates every row over its own points
            check(L.gm_psd_contract_rows(h, self.rows, self._seg_start, self._seg_count, self.n,
                                         4 * self.n_ang, P(self.d_w), P(self.d_mu), P(self.d_pm),
                                         s), "gm_psd_contract_rows")
        check(L.gm_integrate_psd_rows(h, self.rows, self.n, P(self.d_x), P(self.d_lam_pts),
                                      P(self.d_props), P(self.d_w), P(self.d_reff0), self.rhop0,
                                      self.rhop, None, P(self.d_sc), s), "gm_integrate_psd_rows")
        check(L.gm_pmatrix_normalize(h, self.rows, self.n_ang, P(self.d_theta), P(self.d_pm),
                                     P(self.d_back), s), "gm_pmatrix_normalize")
        if self.gsf_ng:
            pm = self.d_pm.view(self.rows, 4, self.n_ang)
            F = self.d_F.view(self.rows, 6, self.n_ang)
            # table order p11, p12, p33, p34, p22 (= p11), p44 (= p33)
            for dst, src in enumerate((0, 1, 2, 3, 0, 2)):
                F[:, dst] = pm[:, src]
            check(L.gm_gsf_expand(h, self.rows, self.n_ang,
                                  self.ang.ctypes.data_as(ctypes.c_void_p), P(self.d_F),
                                  self.gsf_ng, 0, P(self.d_pmom), s), "gm_gsf_expand")

    def capture(self):
        """Run one step eagerly on a side stream (sizes every workspace), then
        capture one step into a HIP graph; replay() runs it."""
        t = self.t
        side = t.cuda.Stream()
        side.wait_stream(t.cuda.current_stream())
        with t.cuda.stream(side):
            self.step()
        t.cuda.current_stream().wait_stream(side)
        t.cuda.synchronize()
        g = t.cuda.CUDAGraph()
        with t.cuda.graph(g):
            self.step()
        self.graph = g
        return g

    def replay(self):
        if self.graph is None:
            self.capture()
        self.graph.replay()

    def results(self):
        """Host copies: pmat [rows, 4, n_ang] (normalised), pback [rows, 6],
        scalars [rows, 16] (gm_integrate_psd order), pmom [rows, 6, ng] or None."""
        out = {"pmat": self.d_pm.view(self.rows, 4, self.n_ang).cpu().numpy(),
               "pback": self.d_back.view(self.rows, 6).cpu().numpy(),
               "scalars": self.d_sc.view(self.rows, 16).cpu().numpy(),
               "props": self.d_props.view(6, self.n).cpu().numpy()}
        out["pmom"] = (self.d_pmom.view(self.rows, 6, self.gsf_ng).cpu().numpy()
                       if self.gsf_ng else None)
        return out
