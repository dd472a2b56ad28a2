"""MLP surrogate of the bulk LUT, with the plain GPU Mie path as fallback and judge.

No counterpart in the reference: the reference builds every table entry with
Mie theory.  Here a network maps a query

    q = (ln r_eff [um], sigma_g, m_r, ln m_i, lambda [um], f_w)

(wet effective radius of a lognormal number distribution, its width, the dry
refractive index, the wavelength, the water volume fraction 1 - gf^-3 of the
humidity growth) to the table row  Qext, SSA, g, alpha1_0 .. alpha1_31
that the LUT sweep (geosmie_amd.lut_sweep) computes with plain GPU Mie.
The network sees the query through the variables Mie theory depends on:
ln(2 pi r_eff / lambda), sigma_g, the volume-mixed index m_w + (m - m_w)(1 - f_w)
(real part, log imaginary part), ln(gf / lambda) (where the radius grid ends
in size-parameter space) and f_w.

  * Network: 6 -> 256 -> 256 -> 256 -> 256 -> 35, SiLU, BF16 weights, FP32
    accumulation; forward, backward and Adam are the HIP kernels of
    gm_mlp.hip; query transforms and the gate are gm_surrogate.hip.
  * Training data: the sweep's own tables (GPU Mie), split into training
    slabs on the sweep grid and validation slabs drawn at random inside it.
  * Gate: a query is answered by the network only inside the training box
    (in ln r_mode, sigma_g, m_r, ln m_i, ln lambda, f_w) and in a cell of a
    uniform grid over that box whose validation entries all met the error
    bound (relative error of Qext, SSA and g <= bound / safety).  Every other
    query is re-solved by plain GPU Mie (LutSweep.run with its own PSD),
    its S1/S2 at FP32 by default: the precision of the tables the network was
    fitted and validated on (Q and g come from the FP64 series either way;
    the P11 moments move by ~1e-7); the re-solve then runs 1.4x faster than
    with FP64 S1/S2.
"""
import ctypes
import time

import numpy as np

from . import _lib
from ._lib import check, ptr, stream
from .lut_sweep import M_WATER, N_MOM, LutSweep, full_grid

DIMS = (6, 256, 256, 256, 256, 35)
N_OUT = 3 + N_MOM
GATE_CELLS = (8, 6, 6, 10, 10, 6)
G_FLOOR = 1e-8        # floor of g in its log target (gm_surrogate.hip)
FEATURES = 3          # version of the feature / target transforms new models are trained with
# 2: Qext target ln Qext; 3: ln(Qext / B), B the bulk Rayleigh extinction saturated at 2


def sigma_factor(sigma):
    """r_eff / r_mode of a lognormal number distribution: exp(2.5 ln^2 sigma_g)."""
    return np.exp(2.5 * np.log(sigma) ** 2)


def make_queries(rmode, sigma, mr, mi, lam, gf):
    """Queries [n, 6] float32 from dry mode radius, width, dry index, wavelength, growth."""
    rmode, sigma, mr, mi, lam, gf = np.broadcast_arrays(*[np.asarray(v, dtype=np.float64) for v in
                                                         (rmode, sigma, mr, mi, lam, gf)])
    reff = gf * rmode * sigma_factor(sigma)
    return np.stack([np.log(reff), sigma, mr, np.log(mi), lam, 1.0 - gf ** -3.0],
                    axis=-1).astype(np.float32)


def decode_queries(q):
    """(rmode, sigma, m_dry, lam, gf) of queries [n, 6] (float64)."""
    q = np.asarray(q, dtype=np.float64)
    gf = (1.0 - q[:, 5]) ** (-1.0 / 3.0)
    rmode = np.exp(q[:, 0]) / (gf * sigma_factor(q[:, 1]))
    return rmode, q[:, 1], q[:, 2] + 1j * np.exp(q[:, 3]), q[:, 4], gf


def gate_coords(q):
    """Gate coordinates (ln r_mode, sigma_g, m_r, ln m_i, ln lambda, f_w) of queries."""
    rmode, sigma, m, lam, gf = decode_queries(q)
    return np.stack([np.log(rmode), sigma, m.real, np.log(m.imag), np.log(lam),
                     np.asarray(q, np.float64)[:, 5]], axis=-1)


def init_params(dims, seed=0):
    """Uniform init, variance 2/fan_in for hidden layers and 1/fan_in for the last."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(len(dims) - 1):
        n_in, n_out = dims[i], dims[i + 1]
        gain = 2.0 if i < len(dims) - 2 else 1.0
        a = np.sqrt(3.0 * gain / n_in)
        out.append(rng.uniform(-a, a, n_in * n_out))
        out.append(np.zeros(n_out))
    return np.concatenate(out).astype(np.float32)


class MLP(object):
    """FP32 master parameters and Adam moments on the device; HIP kernels."""

    def __init__(self, dims=DIMS, seed=0, params=None, weight_terms=1):
        t = _lib.torch()
        self.t = t
        self.weight_terms = int(weight_terms)
        self.dims = tuple(int(d) for d in dims)
        self.n_layers = len(self.dims) - 1
        self._cd = (ctypes.c_int * len(self.dims))(*self.dims)
        L = _lib.lib()
        self.n_params = int(L.gm_mlp_param_count(self.n_layers, self._cd))
        if self.n_params < 0:
            raise ValueError("bad network dims %r" % (self.dims,))
        if params is None:
            params = init_params(self.dims, seed)
        params = np.asarray(params, dtype=np.float32)
        if params.size != self.n_params:
            raise ValueError("expected %d parameters, got %d" % (self.n_params, params.size))
        dev = _lib.device()
        self.p = _lib.to_dev(params, np.float32)
        self.g = t.zeros(self.n_params, dtype=t.float32, device=dev)
        self.m = t.zeros_like(self.g)
        self.v = t.zeros_like(self.g)
        self.loss = t.zeros(1, dtype=t.float32, device=dev)
        self.step = 0

    def _precision(self):
        # BF16 terms per weight on the matrix cores (a setting of the context)
        check(_lib.lib().gm_ctx_tune(_lib.ctx(), 3, self.weight_terms), "gm_ctx_tune")

    def forward(self, x, out=None):
        self._precision()
        n = x.shape[0]
        y = out if out is not None else self.t.empty((n, self.dims[-1]), dtype=self.t.float32,
                                                     device=_lib.device())
        check(_lib.lib().gm_mlp_forward(_lib.ctx(), self.n_layers, self._cd, n, ptr(self.p),
                                        ptr(x), ptr(y), stream()), "gm_mlp_forward")
        return y

    def grad(self, x, target, cw):
        self._precision()
        check(_lib.lib().gm_mlp_grad(_lib.ctx(), self.n_layers, self._cd, x.shape[0], ptr(self.p),
                                     ptr(x), ptr(target), ptr(cw), ptr(self.g), ptr(self.loss),
                                     stream()), "gm_mlp_grad")
        return self.loss

    def adam(self, lr, beta1=0.9, beta2=0.999, eps=1e-8):
        self.step += 1
        check(_lib.lib().gm_mlp_adam(_lib.ctx(), self.n_params, ptr(self.p), ptr(self.g),
                                     ptr(self.m), ptr(self.v), float(lr), float(beta1),
                                     float(beta2), float(eps), self.step, stream()), "gm_mlp_adam")

    def params_host(self):
        return self.p.cpu().numpy()


def _f32dev(a):
    return _lib.to_dev(np.ascontiguousarray(a, dtype=np.float32), np.float32)


class Surrogate(object):
    """Trained network + standardisation + gate, with a plain-Mie fallback."""

    # S1/S2 precision of the plain re-solve of flagged queries
    fallback_precision = "fp32"

    def __init__(self, params, feat_mean, feat_std, tgt_mean, tgt_std, box_lo, box_hi,
                 cell_ok, cells=GATE_CELLS, r_dry=None, bound=1e-3, dims=DIMS, info=None,
                 weight_terms=1, features=FEATURES):
        t = _lib.torch()
        self.t = t
        if features not in (2, 3):
            raise ValueError("unknown feature transform %d" % features)
        self.features = int(features)
        self.dims = tuple(int(d) for d in dims)
        self.mlp = MLP(self.dims, params=params, weight_terms=weight_terms)
        self.feat_mean = np.asarray(feat_mean, np.float32)
        self.feat_std = np.asarray(feat_std, np.float32)
        self.tgt_mean = np.asarray(tgt_mean, np.float32)
        self.tgt_std = np.asarray(tgt_std, np.float32)
        self.box_lo = np.asarray(box_lo, np.float32)
        self.box_hi = np.asarray(box_hi, np.float32)
        self.cells = tuple(int(c) for c in cells)
        self.cell_ok = np.asarray(cell_ok, np.uint8).ravel()
        if self.cell_ok.size != int(np.prod(self.cells)):
            raise ValueError("cell_ok must have prod(cells) entries")
        self.bound = float(bound)
        self.info = dict(info or {})
        self.r_dry = np.asarray(full_grid()["r"] if r_dry is None else r_dry, np.float64)
        self._dev = {k: _f32dev(getattr(self, k)) for k in
                     ("feat_mean", "feat_std", "tgt_mean", "tgt_std", "box_lo", "box_hi")}
        self._cell_ok = _lib.to_dev(self.cell_ok, np.uint8)
        self._ccells = (ctypes.c_int * 6)(*self.cells)
        self._fallback = {}

    def set_cells(self, cells, cell_ok=None):
        """Replace the gate's cell grid (and its trusted-cell table, all
        trusted when not given) on the host and the device together."""
        cells = tuple(int(c) for c in cells)
        n_cells = int(np.prod(cells))
        ok = np.ones(n_cells, np.uint8) if cell_ok is None else np.asarray(cell_ok, np.uint8).ravel()
        if len(cells) != 6 or ok.size != n_cells:
            raise ValueError("cell_ok must have prod(cells) entries for 6 cell counts")
        self.cells = cells
        self.cell_ok = ok
        self._cell_ok = _lib.to_dev(ok, np.uint8)
        self._ccells = (ctypes.c_int * 6)(*cells)

    # -- network path ------------------------------------------------------
    def predict(self, q):
        """Network rows [n, 35] (float32, device) and gate flags [n] (int32, device)."""
        t = self.t
        q = q if hasattr(q, "data_ptr") else _f32dev(q)
        n = q.shape[0]
        dev = _lib.device()
        x = t.empty((n, 6), dtype=t.float32, device=dev)
        flags = t.empty(n, dtype=t.int32, device=dev)
        d = self._dev
        L, h, s = _lib.lib(), _lib.ctx(), stream()
        check(L.gm_surrogate_prep(h, n, ptr(q), ptr(d["feat_mean"]), ptr(d["feat_std"]),
                                  ptr(d["box_lo"]), ptr(d["box_hi"]), self._ccells,
                                  ptr(self._cell_ok), self._cell_ok.numel(), M_WATER.real,
                                  ptr(x), ptr(flags), s), "gm_surrogate_prep")
        y = self.mlp.forward(x)
        rows = t.empty_like(y)
        check(L.gm_surrogate_post(h, n, self.dims[-1], self._baseline_q(q), M_WATER.real, ptr(y),
                                  ptr(d["tgt_mean"]), ptr(d["tgt_std"]), ptr(rows), s),
              "gm_surrogate_post")
        return rows, flags

    def _baseline_q(self, q):
        """The queries the Qext baseline is formed from (None: no baseline)."""
        return ptr(q) if self.features >= 3 else None

    # -- plain GPU Mie -----------------------------------------------------
    def solver(self, precision=None):
        """The plain GPU Mie solver at S1/S2 precision `precision` (default
        fallback_precision), built once per precision."""
        p = precision or self.fallback_precision
        if p not in self._fallback:
            self._fallback[p] = LutSweep(self.r_dry, rmode=[1.0], sigma=[2.0], precision=p)
        return self._fallback[p]

    def solve_plain(self, q, chunk=1024, precision=None):
        """Rows [n, 35] of queries by plain GPU Mie (one slab and one PSD each),
        S1/S2 at `precision` ("fp64", "fp32"; default fallback_precision)."""
        t = self.t
        q = np.asarray(q.cpu().numpy() if hasattr(q, "data_ptr") else q, dtype=np.float64)
        n = q.shape[0]
        out = t.empty((n, N_OUT), dtype=t.float32, device=_lib.device())
        sw = self.solver(precision)
        rmode, sigma, m, lam, gf = decode_queries(q)
        for a in range(0, n, chunk):
            b = min(n, a + chunk)
            w = sw.weights_dev(rmode[a:b], sigma[a:b])
            sw.run(lam[a:b], m[a:b], gf[a:b], w=w, out=out[a:b].view(-1))
        return out

    def query(self, q):
        """Rows for every query: network where the gate allows, plain Mie elsewhere.
        Returns (rows [n, 35] float32 device, number re-solved by plain Mie)."""
        t = self.t
        qd = q if hasattr(q, "data_ptr") else _f32dev(q)
        rows, flags = self.predict(qd)
        idx = t.nonzero(flags != 0).flatten()
        nf = int(idx.numel())
        if nf:
            rows[idx] = self.solve_plain(qd[idx])
        return rows, nf

    # -- persistence ---------------------------------------------------------
    def save(self, path):
        np.savez_compressed(path, params=self.mlp.params_host(), feat_mean=self.feat_mean,
                            feat_std=self.feat_std, tgt_mean=self.tgt_mean, tgt_std=self.tgt_std,
                            box_lo=self.box_lo, box_hi=self.box_hi, cell_ok=self.cell_ok,
                            cells=np.array(self.cells), r_dry=self.r_dry, bound=self.bound,
                            dims=np.array(self.dims), weight_terms=self.mlp.weight_terms,
                            features=self.features,
                            info_keys=np.array(list(self.info.keys()), dtype=str),
                            info_vals=np.array([float(v) for v in self.info.values()]))

    @classmethod
    def load(cls, path):
        z = np.load(path, allow_pickle=False)
        ver = int(z["features"]) if "features" in z.files else 1
        if ver not in (2, 3):
            raise ValueError("%s was trained with feature transform %d, this library reads 2 and "
                             "3: retrain with python -m geosmie_amd.train_surrogate" % (path, ver))
        info = dict(zip([str(k) for k in z["info_keys"]], z["info_vals"].tolist()))
        return cls(z["params"], z["feat_mean"], z["feat_std"], z["tgt_mean"], z["tgt_std"],
                   z["box_lo"], z["box_hi"], z["cell_ok"], cells=tuple(z["cells"]),
                   r_dry=z["r_dry"], bound=float(z["bound"]), dims=tuple(z["dims"]), info=info,
                   weight_terms=int(z["weight_terms"]) if "weight_terms" in z.files else 1,
                   features=ver)


def rel_errors(rows, ref):
    """Relative errors of Qext, SSA, g: [n, 3] (numpy)."""
    a = np.asarray(rows, np.float64)[:, :3]
    b = np.asarray(ref, np.float64)[:, :3]
    return np.abs(a - b) / np.abs(b)


def _features(q):
    """Network features of queries, as gm_surrogate_prep forms them (before
    standardisation): ln x_eff, sigma_g, Re m_mix, ln Im m_mix, ln(gf/lambda), f_w."""
    q = np.asarray(q, np.float64)
    llam = np.log(q[:, 4])
    dry = 1.0 - q[:, 5]
    lgf = -np.log(dry) / 3.0
    mw = M_WATER.real
    return np.stack([q[:, 0] - llam + np.log(2 * np.pi), q[:, 1], mw + (q[:, 2] - mw) * dry,
                     q[:, 3] + np.log(dry), lgf - llam, q[:, 5]], axis=-1)


def rayleigh_log_b(q):
    """ln B of queries (FP64), B = Q_R / (1 + Q_R / 2) with Q_R the bulk
    Rayleigh extinction 8/3 |K|^2 x_eff^4 exp(6 s^2) + 4 Im(K) x_eff at the
    volume-mixed index (gm_surrogate_post forms the same in FP32)."""
    q = np.asarray(q, np.float64)
    dry = 1.0 - q[:, 5]
    lx = q[:, 0] - np.log(q[:, 4]) + np.log(2 * np.pi)
    ls = np.log(q[:, 1])
    mw = M_WATER.real
    m = (mw + (q[:, 2] - mw) * dry) + 1j * np.exp(q[:, 3]) * dry
    K = (m * m - 1) / (m * m + 2)
    ln_sca = np.log(8. / 3.) + np.log(np.abs(K) ** 2) + 4 * lx + 6 * ls * ls
    ln_abs = np.log(4.) + np.log(K.imag) + lx
    lq = np.logaddexp(ln_sca, ln_abs)
    return lq - np.logaddexp(0.0, lq - np.log(2.0))


def _targets_host(R, Q=None):
    """Targets (before standardisation) of table rows, as gm_surrogate_targets
    forms them (Q given: with the Rayleigh baseline of transform 3)."""
    R = np.asarray(R, np.float64)
    T = np.empty_like(R)
    T[:, 0] = np.log(R[:, 0]) - (rayleigh_log_b(Q) if Q is not None else 0.0)
    T[:, 1] = np.log(R[:, 1])
    T[:, 2] = np.log(np.maximum(R[:, 2], G_FLOOR))
    T[:, 3:] = R[:, 3:] / (2 * np.arange(R.shape[1] - 3) + 1)
    return T


def sweep_tables(g, mr, mi, lam, gf, pairs=None, precision="fp32", log=print):
    """Tables of the product grid (mr, mi, lam, gf) x PSDs with the GPU sweep:
    (queries [N, 6] host float32, rows [N, 35] device float32, seconds).
    PSDs: the grid's r_mode x sigma_g, or the explicit (r_mode, sigma_g) pairs."""
    t = _lib.torch()
    sweep = LutSweep(g["r"], g["rmode"], g["sigma"], precision=precision, pairs=pairs)
    t0 = time.perf_counter()
    tab = sweep.sweep(mr, mi, lam, gf)
    t.cuda.synchronize()
    dt = time.perf_counter() - t0
    pr = np.array([p[0] for p in sweep.pairs])
    ps = np.array([p[1] for p in sweep.pairs])
    A, B, C, D, P = np.meshgrid(np.arange(mr.size), np.arange(mi.size), np.arange(lam.size),
                                np.arange(gf.size), np.arange(sweep.n_psd), indexing="ij")
    Q = make_queries(pr[P.ravel()], ps[P.ravel()], mr[A.ravel()], mi[B.ravel()], lam[C.ravel()],
                     gf[D.ravel()])
    log("sweep: %d slabs x %d PSDs = %d entries in %.1f s" % (
        Q.shape[0] // sweep.n_psd, sweep.n_psd, Q.shape[0], dt))
    return Q, tab.view(-1, N_OUT), dt


def new_surrogate(g, Q, rows, seed=0, cells=GATE_CELLS, bound=1e-3):
    """Untrained surrogate with standardisation from (a subsample of) the data
    and the training box of the grid g; every cell trusted until fit_gate."""
    t = _lib.torch()
    n = rows.shape[0]
    rng = np.random.default_rng(seed)
    sub = np.sort(rng.choice(n, size=min(n, 2_000_000), replace=False))
    F = _features(Q[sub])
    feat_mean, feat_std = F.mean(0), F.std(0)
    feat_std[feat_std == 0] = 1.0
    T = _targets_host(rows[t.from_numpy(sub).to(rows.device)].cpu().numpy(), Q[sub])
    tgt_mean, tgt_std = T.mean(0), T.std(0)
    tgt_std[tgt_std < 1e-6] = 1.0
    box_lo = np.array([np.log(g["rmode"].min()), g["sigma"].min(), g["mr"].min(),
                       np.log(g["mi"].min()), np.log(g["lam"].min()), 1 - g["gf"].min() ** -3.0])
    box_hi = np.array([np.log(g["rmode"].max()), g["sigma"].max(), g["mr"].max(),
                       np.log(g["mi"].max()), np.log(g["lam"].max()), 1 - g["gf"].max() ** -3.0])
    return Surrogate(init_params(DIMS, seed), feat_mean, feat_std, tgt_mean, tgt_std,
                     box_lo - 1e-6, box_hi + 1e-6, np.ones(int(np.prod(cells)), np.uint8),
                     cells=cells, r_dry=g["r"], bound=bound)


def training_arrays(sur, Q, rows):
    """Standardised features X [N, 6] and targets Y [N, 35] on the device (HIP kernels)."""
    t = _lib.torch()
    n = rows.shape[0]
    dev = _lib.device()
    d_q = _f32dev(Q)
    X = t.empty((n, 6), dtype=t.float32, device=dev)
    Y = t.empty((n, N_OUT), dtype=t.float32, device=dev)
    flags = t.empty(n, dtype=t.int32, device=dev)
    L, h, s = _lib.lib(), _lib.ctx(), stream()
    d = sur._dev
    check(L.gm_surrogate_prep(h, n, ptr(d_q), ptr(d["feat_mean"]), ptr(d["feat_std"]),
                              ptr(d["box_lo"]), ptr(d["box_hi"]), sur._ccells, ptr(sur._cell_ok),
                              sur._cell_ok.numel(), M_WATER.real, ptr(X), ptr(flags), s),
          "gm_surrogate_prep")
    check(L.gm_surrogate_targets(h, n, N_OUT, sur._baseline_q(d_q), M_WATER.real, ptr(rows),
                                 ptr(d["tgt_mean"]), ptr(d["tgt_std"]), ptr(Y), s),
          "gm_surrogate_targets")
    return X, Y


def fit(mlp, X, Y, steps, batch=16384, lr=2e-3, lr_final=1e-5, warmup=500, cw_head=1000.0,
        seed=0, log=print):
    """Adam on shuffled minibatches, linear warmup then cosine decay; returns
    [(step, loss)] sampled every 1000 steps."""
    t = _lib.torch()
    n = X.shape[0]
    dev = _lib.device()
    cw = np.ones(N_OUT, np.float32)
    cw[:3] = cw_head
    cw[3] = 0.0                                   # alpha1_0 = 1 identically
    cw = _f32dev(cw / cw.sum() * N_OUT)
    gen = t.Generator(device=dev)
    gen.manual_seed(seed)
    perm = t.randperm(n, device=dev, generator=gen)
    pos = 0
    hist = []
    t0 = time.perf_counter()
    for it in range(steps):
        if pos + batch > n:
            perm = t.randperm(n, device=dev, generator=gen)
            pos = 0
        idx = perm[pos:pos + batch]
        pos += batch
        if it < warmup:
            lr_t = lr * (it + 1) / warmup
        else:
            c = (it - warmup) / max(1, steps - warmup)
            lr_t = lr_final + 0.5 * (lr - lr_final) * (1 + np.cos(np.pi * c))
        mlp.grad(X[idx], Y[idx], cw)
        mlp.adam(lr_t)
        if it % 10000 == 0 or it == steps - 1:
            hist.append((it, float(mlp.loss.item())))
            log("step %6d  loss %.3e  lr %.2e  %.1f s" % (it, hist[-1][1], lr_t,
                                                         time.perf_counter() - t0))
    t.cuda.synchronize()
    return hist, time.perf_counter() - t0


def validation_tables(g, n_slabs, n_psd=40, seed=1, precision="fp32", chunk=1000):
    """Random slabs inside the grid's box, each with its own n_psd random
    (r_mode, sigma_g) pairs (so every part of the box is sampled):
    (queries, rows) on the host."""
    t = _lib.torch()
    r = np.random.default_rng(seed)
    ns = int(n_slabs)
    v_mr = r.uniform(g["mr"].min(), g["mr"].max(), ns)
    v_mi = np.exp(r.uniform(np.log(g["mi"].min()), np.log(g["mi"].max()), ns))
    v_lam = np.exp(r.uniform(np.log(g["lam"].min()), np.log(g["lam"].max()), ns))
    v_gf = (1 - r.uniform(0, 1 - g["gf"].max() ** -3.0, ns)) ** (-1.0 / 3.0)
    v_rm = np.exp(r.uniform(np.log(g["rmode"].min()), np.log(g["rmode"].max()), (ns, n_psd)))
    v_sg = r.uniform(g["sigma"].min(), g["sigma"].max(), (ns, n_psd))
    vsw = LutSweep(g["r"], [1.0], [2.0], precision=precision)
    vq, vrows = [], []
    for a in range(0, ns, chunk):
        b = min(ns, a + chunk)
        w = vsw.weights_dev(v_rm[a:b].ravel(), v_sg[a:b].ravel())
        out = vsw.run(v_lam[a:b], v_mr[a:b] + 1j * v_mi[a:b], v_gf[a:b], w=w)
        vrows.append(out.view(-1, N_OUT).cpu().numpy())
        del w
        k = np.arange(a, b).repeat(n_psd)
        vq.append(make_queries(v_rm[a:b].ravel(), v_sg[a:b].ravel(), v_mr[k], v_mi[k], v_lam[k],
                               v_gf[k]))
    t.cuda.synchronize()
    return np.concatenate(vq), np.concatenate(vrows)


def regate(sur, g=None, n_val_slabs=60000, seed=11, safety=1.0, min_count=3, precision="fp32",
           cells=None, log=print):
    """Re-fit the trust gate of a trained surrogate on fresh validation slabs
    (optionally on a new cell grid)."""
    g = dict(full_grid() if g is None else g)
    if cells is not None:
        sur.set_cells(cells)
    t0 = time.perf_counter()
    vq, vrows = validation_tables(g, n_val_slabs, seed=seed, precision=precision)
    rep, err = evaluate(sur, vq, vrows)
    rep["cells_trusted"], rep["val_frac_trusted"] = fit_gate(sur, vq, err, safety, min_count)
    rep["val_s"] = time.perf_counter() - t0
    log("gate: median err %.2e, %.1f%% within bound; %.1f%% of cells trusted, %.1f%% of "
        "validation entries in trusted cells" % (rep["val_err_median"],
                                                 100 * rep["val_frac_within_bound"],
                                                 100 * rep["cells_trusted"],
                                                 100 * rep["val_frac_trusted"]))
    info = dict(sur.info)
    info.update(rep)
    sur.info = info
    return rep


def evaluate(sur, vq, vrows):
    """Validation report (per-quantity relative errors) and the per-entry max error."""
    rep = {}
    pred, _ = sur.predict(vq)
    e3 = rel_errors(pred.cpu().numpy(), vrows)
    err = e3.max(axis=1)
    for j, k in enumerate(("qext", "ssa", "g")):
        rep["val_%s_median" % k] = float(np.median(e3[:, j]))
        rep["val_%s_p90" % k] = float(np.quantile(e3[:, j], 0.9))
    xe = np.exp(vq[:, 0].astype(np.float64)) * 2 * np.pi / vq[:, 4]
    for lo_, hi_ in ((0, 0.3), (0.3, 3), (3, 30), (30, 1e9)):
        sel = (xe >= lo_) & (xe < hi_)
        if sel.any():
            rep["val_err_median_x%g" % lo_] = float(np.median(err[sel]))
    rep["val_entries"] = int(vq.shape[0])
    rep["val_err_median"] = float(np.median(err))
    rep["val_err_p90"] = float(np.quantile(err, 0.9))
    rep["val_frac_within_bound"] = float(np.mean(err <= sur.bound))
    return rep, err


def fit_gate(sur, vq, err, safety=1.0, min_count=3):
    """Trust a cell when it holds >= min_count validation entries and all are
    within bound / safety; returns (fraction of cells, fraction of entries) trusted."""
    gc = gate_coords(vq)
    lo, hi = sur.box_lo.astype(np.float64), sur.box_hi.astype(np.float64)
    ci = np.zeros(vq.shape[0], np.int64)
    for j, nc in enumerate(sur.cells):
        c = np.floor((gc[:, j] - lo[j]) / (hi[j] - lo[j]) * nc).astype(np.int64)
        ci = ci * nc + np.clip(c, 0, nc - 1)
    n_cells = int(np.prod(sur.cells))
    cmax = np.zeros(n_cells)
    np.maximum.at(cmax, ci, err)
    cnt = np.bincount(ci, minlength=n_cells)
    ok = (cnt >= min_count) & (cmax <= sur.bound / safety)
    sur.set_cells(sur.cells, ok)
    return float(ok.mean()), float(ok[ci].mean())


def train(grid=None, train_stride=(1, 1, 1, 1), n_val_slabs=20000, steps=20000, batch=16384,
          lr=2e-3, lr_final=1e-5, warmup=500, cw_head=1000.0, seed=0, bound=1e-3, safety=1.0,
          min_count=3, cells=GATE_CELLS, log=print, precision="fp32", n_aug_psd=40,
          weight_terms=2, init=None):
    """Build training tables with the GPU sweep, train the network with the HIP
    kernels, fit the gate on validation slabs; returns (Surrogate, report).

    train_stride subsamples the (m_r, m_i, lambda, gf) axes of the sweep grid
    (1 = every value).  n_aug_psd > 0 adds a second pass over the same slabs
    with that many random (r_mode, sigma_g) pairs inside the grid's ranges.
    init: a Surrogate to continue training (fresh Adam state)."""
    t = _lib.torch()
    g = dict(full_grid() if grid is None else grid)
    mr, mi, lam, gf = (g[k][::s] for k, s in zip(("mr", "mi", "lam", "gf"), train_stride))
    rep = {}
    Q, rows, dt = sweep_tables(g, mr, mi, lam, gf, precision=precision, log=log)
    rep["sweep_s"] = dt
    if n_aug_psd:
        r = np.random.default_rng(seed + 7)
        pairs = list(zip(np.exp(r.uniform(np.log(g["rmode"].min()), np.log(g["rmode"].max()),
                                          n_aug_psd)),
                         r.uniform(g["sigma"].min(), g["sigma"].max(), n_aug_psd)))
        Q2, rows2, dt2 = sweep_tables(g, mr, mi, lam, gf, pairs=pairs, precision=precision,
                                      log=log)
        Q = np.concatenate([Q, Q2])
        rows = t.cat([rows, rows2])
        del Q2, rows2
        rep["sweep_s"] += dt2
    rep["train_entries"] = int(rows.shape[0])
    if init is not None:
        # continue from a trained model (its standardisation and weight terms)
        sur = init
        prev_steps = float(sur.info.get("train_steps", 0.0))
    else:
        sur = new_surrogate(g, Q, rows, seed=seed, cells=cells, bound=bound)
        sur.mlp.weight_terms = int(weight_terms)
        prev_steps = 0.0
    X, Y = training_arrays(sur, Q, rows)
    del Q, rows
    hist, rep["train_s"] = fit(sur.mlp, X, Y, steps, batch=batch, lr=lr, lr_final=lr_final,
                               warmup=warmup, cw_head=cw_head, seed=seed, log=log)
    rep["train_steps"], rep["batch"], rep["final_loss"] = prev_steps + steps, batch, hist[-1][1]
    rep["weight_terms"] = sur.mlp.weight_terms
    del X, Y
    t0 = time.perf_counter()
    vq, vrows = validation_tables(g, n_val_slabs, seed=seed + 1, precision=precision)
    r2, err = evaluate(sur, vq, vrows)
    rep.update(r2)
    rep["cells_trusted"], rep["val_frac_trusted"] = fit_gate(sur, vq, err, safety, min_count)
    rep["val_s"] = time.perf_counter() - t0
    log("validation per quantity: " + ", ".join("%s %.2e/%.2e" % (k, rep["val_%s_median" % k],
                                                                  rep["val_%s_p90" % k])
                                                 for k in ("qext", "ssa", "g")))
    log("validation median by effective size parameter: " + ", ".join(
        "%s %.2e" % (k[16:], v) for k, v in rep.items() if k.startswith("val_err_median_x")))
    log("validation: median err %.2e, p90 %.2e, %.1f%% within bound; %.1f%% of cells trusted, "
        "%.1f%% of validation entries in trusted cells" % (
            rep["val_err_median"], rep["val_err_p90"], 100 * rep["val_frac_within_bound"],
            100 * rep["cells_trusted"], 100 * rep["val_frac_trusted"]))
    sur.info = rep
    return sur, rep
