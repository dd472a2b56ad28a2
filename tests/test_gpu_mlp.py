"""Surrogate network kernels (gm_mlp.hip, gm_surrogate.hip) against the FP64
oracle (oracle/mlp.py) and the gate / fallback logic.

Tolerances: the GPU feeds activations to the matrix cores as a two-term BF16
split (~2^-17 relative per element) and accumulates in FP32, against FP64 in
the oracle on the same BF16 weights: forward outputs within 2e-4 of max |y|,
gradients within 2e-3 in relative 2-norm per layer, loss within 1e-4.
"""
import numpy as np
import pytest

from oracle import mlp as O

pytestmark = pytest.mark.gpu


def _net(dims, seed):
    from geosmie_amd.surrogate import init_params
    p = init_params(dims, seed)
    rng = np.random.default_rng(seed + 7)
    # nonzero biases so their path is exercised
    lay, _ = O.layout(dims)
    for w_off, b_off, n_in, n_out in lay:
        p[b_off:b_off + n_out] = rng.uniform(-0.3, 0.3, n_out)
    return p


@pytest.mark.parametrize("terms", [1, 2])
@pytest.mark.parametrize("dims,n", [((6, 256, 256, 256, 256, 35), 300), ((3, 70, 5), 100),
                                    ((6, 64, 1), 1), ((1, 33, 2), 65)])
def test_forward_vs_oracle(gpu, dims, n, terms):
    from geosmie_amd import _lib
    from geosmie_amd.surrogate import MLP
    p = _net(dims, n)
    x = np.random.default_rng(n).standard_normal((n, dims[0])).astype(np.float32)
    mlp = MLP(dims, params=p, weight_terms=terms)
    y = mlp.forward(_lib.to_dev(x, np.float32)).cpu().numpy()
    ref = O.forward(p, dims, x, terms=terms)
    assert np.max(np.abs(y - ref)) <= 2e-4 * np.abs(ref).max()


@pytest.mark.parametrize("terms", [1, 2])
@pytest.mark.parametrize("dims,n", [((6, 256, 256, 256, 256, 35), 257), ((3, 70, 5), 100),
                                    ((6, 40, 40, 1), 5000)])
def test_grad_vs_oracle(gpu, dims, n, terms):
    from geosmie_amd import _lib
    from geosmie_amd.surrogate import MLP
    rng = np.random.default_rng(n)
    p = _net(dims, n + 1)
    x = rng.standard_normal((n, dims[0])).astype(np.float32)
    t = rng.standard_normal((n, dims[-1])).astype(np.float32)
    cw = rng.uniform(0.5, 2.0, dims[-1]).astype(np.float32)
    mlp = MLP(dims, params=p, weight_terms=terms)
    f = lambda a: _lib.to_dev(a, np.float32)
    loss = float(mlp.grad(f(x), f(t), f(cw)).item())
    g = mlp.g.cpu().numpy().astype(np.float64)
    lref, gref = O.loss_grad(p, dims, x, t, cw, terms=terms)
    assert abs(loss - lref) <= 1e-4 * abs(lref)
    lay, _ = O.layout(dims)
    for w_off, b_off, n_in, n_out in lay:
        for a, b in ((w_off, b_off), (b_off, b_off + n_out)):
            num = np.linalg.norm(g[a:b] - gref[a:b])
            assert num <= 2e-3 * np.linalg.norm(gref[a:b]) + 1e-12, (a, b)


def test_adam_vs_oracle(gpu):
    from geosmie_amd import _lib
    from geosmie_amd.surrogate import MLP
    dims = (3, 70, 5)
    rng = np.random.default_rng(5)
    mlp = MLP(dims, seed=3)
    p0 = mlp.params_host().astype(np.float64)
    m = np.zeros(p0.size)
    v = np.zeros(p0.size)
    p = p0.copy()
    for step in range(1, 4):
        gr = rng.standard_normal(p0.size).astype(np.float32)
        mlp.g.copy_(_lib.to_dev(gr, np.float32))
        mlp.adam(1e-2)
        p, m, v = O.adam(p, gr.astype(np.float64), m, v, 1e-2, 0.9, 0.999, 1e-8, step)
    np.testing.assert_allclose(mlp.params_host(), p, rtol=1e-5, atol=1e-6)


def test_training_reduces_loss(gpu):
    """A teacher network's outputs are learnt: the loss falls 20x in 400 steps."""
    from geosmie_amd import _lib
    from geosmie_amd.surrogate import MLP, init_params
    dims = (6, 64, 64, 8)
    rng = np.random.default_rng(11)
    x = rng.uniform(-1, 1, (4096, 6)).astype(np.float32)
    teacher = init_params((6, 16, 8), 99)
    t = O.forward(teacher, (6, 16, 8), x).astype(np.float32)
    t = (t - t.mean(0)) / t.std(0)
    f = lambda a: _lib.to_dev(a, np.float32)
    xd, td, cw = f(x), f(t), f(np.ones(8, np.float32))
    mlp = MLP(dims, seed=1)
    first = float(mlp.grad(xd, td, cw).item())
    for it in range(400):
        mlp.grad(xd, td, cw)
        mlp.adam(3e-3)
    last = float(mlp.grad(xd, td, cw).item())
    assert last < first / 20, (first, last)


def _toy_surrogate(cell_ok=None):
    from geosmie_amd.surrogate import DIMS, GATE_CELLS, Surrogate, init_params
    rng = np.random.default_rng(2)
    n_cells = int(np.prod(GATE_CELLS))
    if cell_ok is None:
        cell_ok = (rng.uniform(size=n_cells) < 0.5).astype(np.uint8)
    lo = np.array([np.log(0.01), 1.4, 1.3, np.log(1e-8), np.log(0.2), 0.0])
    hi = np.array([np.log(5.0), 2.6, 1.8, np.log(1.0), np.log(40.0), 1 - 2.16 ** -3])
    return Surrogate(init_params(DIMS, 4), np.zeros(6), np.ones(6), np.zeros(35),
                     np.ones(35) * 0.1, lo, hi, cell_ok,
                     r_dry=np.logspace(np.log10(0.01), np.log10(2.0), 120))


def _host_flags(sur, q):
    from geosmie_amd.surrogate import gate_coords
    gc = gate_coords(q)
    lo, hi = sur.box_lo.astype(np.float64), sur.box_hi.astype(np.float64)
    inside = np.all((gc >= lo) & (gc <= hi), axis=1)
    ci = np.zeros(len(q), np.int64)
    for j, nc in enumerate(sur.cells):
        c = np.floor((gc[:, j] - lo[j]) / (hi[j] - lo[j]) * nc)
        ci = ci * nc + np.clip(np.nan_to_num(c), 0, nc - 1).astype(np.int64)
    return np.where(~inside, 1, np.where(sur.cell_ok[ci] == 0, 2, 0))


def _queries(n, seed, spread=1.0):
    from geosmie_amd.surrogate import make_queries
    r = np.random.default_rng(seed)
    c = lambda lo, hi: (lo + hi) / 2 + spread * (r.uniform(lo, hi, n) - (lo + hi) / 2)
    with np.errstate(invalid="ignore"):    # spread > 1 puts some growth factors out of range
        gf = (1 - c(0.0, 1 - 2.16 ** -3)) ** (-1 / 3)
    return make_queries(np.exp(c(np.log(0.01), np.log(5.0))), c(1.4, 2.6), c(1.3, 1.8),
                        np.exp(c(np.log(1e-8), 0.0)), np.exp(c(np.log(0.2), np.log(40.0))), gf)


def test_gate_flags_match_host(gpu):
    sur = _toy_surrogate()
    q = np.concatenate([_queries(3000, 1), _queries(500, 2, spread=1.3)])
    q[5, 1] = np.nan
    _, flags = sur.predict(q)
    flags = flags.cpu().numpy()
    want = _host_flags(sur, q.astype(np.float64))
    # float32 rounding may move a query that sits on a cell boundary
    assert np.mean(flags == want) > 0.999
    assert flags[5] == 1
    assert set(np.unique(flags)) == {0, 1, 2}


def test_query_fallback_uses_plain_mie(gpu):
    t = __import__("torch")
    sur = _toy_surrogate()
    q = _queries(200, 3)
    q[:, 4] = np.maximum(q[:, 4], 1.0)     # keep the sizes small for a quick test
    rows_net, flags = sur.predict(q)
    rows, nf = sur.query(q)
    f = flags.cpu().numpy()
    assert nf == int(np.sum(f != 0)) and 0 < nf < len(q)
    plain = sur.solve_plain(q[f != 0])
    assert t.equal(rows[t.from_numpy(f != 0).to(rows.device)], plain)
    assert t.equal(rows[t.from_numpy(f == 0).to(rows.device)],
                   rows_net[t.from_numpy(f == 0).to(rows.device)])


def test_targets_post_roundtrip(gpu):
    from geosmie_amd import _lib
    from geosmie_amd._lib import check, ptr, stream
    rng = np.random.default_rng(8)
    n = 1000
    rows = np.empty((n, 35), np.float32)
    rows[:, 0] = np.exp(rng.uniform(-8, 1, n))
    rows[:, 1] = 1 - np.exp(rng.uniform(-9, -0.1, n))
    rows[:, 2] = np.exp(rng.uniform(-12, -0.05, n))
    rows[:, 3:] = rng.uniform(-1, 1, (n, 32)) * (2 * np.arange(32) + 1)
    mean = rng.uniform(-1, 1, 35).astype(np.float32)
    std = rng.uniform(0.5, 2, 35).astype(np.float32)
    f = lambda a: _lib.to_dev(a, np.float32)
    d_rows, d_m, d_s = f(rows), f(mean), f(std)
    t = _lib.torch()
    y = t.empty_like(d_rows)
    back = t.empty_like(d_rows)
    L, h, s = _lib.lib(), _lib.ctx(), stream()
    from geosmie_amd.surrogate import M_WATER, _targets_host
    q = _queries(n, 12)
    d_q = f(q)
    for qq in (None, d_q):     # transform 2 (no baseline) and 3 (Rayleigh baseline)
        check(L.gm_surrogate_targets(h, n, 35, ptr(qq), M_WATER.real, ptr(d_rows), ptr(d_m),
                                     ptr(d_s), ptr(y), s))
        check(L.gm_surrogate_post(h, n, 35, ptr(qq), M_WATER.real, ptr(y), ptr(d_m), ptr(d_s),
                                  ptr(back), s))
        b = back.cpu().numpy()
        np.testing.assert_allclose(b[:, :3], rows[:, :3], rtol=1e-5)
        np.testing.assert_allclose(b[:, 3:], rows[:, 3:], rtol=1e-5, atol=1e-5)
        # the device targets are the host formula's (FP32 vs FP64)
        want = (_targets_host(rows, q if qq is not None else None) - mean) / std
        np.testing.assert_allclose(y.cpu().numpy(), want, rtol=2e-5, atol=2e-4)


def test_rayleigh_baseline_is_the_small_particle_limit(gpu):
    """Bulk Qext of plain GPU Mie over a narrow lognormal of small particles
    (x_eff <= 0.05) equals the baseline B to O(x^2) plus the grid's truncation."""
    from geosmie_amd.surrogate import make_queries, rayleigh_log_b
    sur = _toy_surrogate()
    n = 64
    r = np.random.default_rng(13)
    q = make_queries(np.exp(r.uniform(np.log(0.02), np.log(0.04), n)), r.uniform(1.4, 1.5, n),
                     r.uniform(1.3, 1.8, n), np.exp(r.uniform(np.log(1e-6), 0.0, n)),
                     r.uniform(20.0, 40.0, n), np.ones(n))
    plain = sur.solve_plain(q).cpu().numpy()
    b = np.exp(rayleigh_log_b(q))
    ratio = plain[:, 0] / (1.0 / (1.0 / b - 0.5))      # Q_R from B = Q_R / (1 + Q_R / 2)
    assert np.all(np.abs(ratio - 1) < 0.03), ratio


def test_gate_on_a_finer_cell_grid(gpu):
    """Replacing the cell grid replaces the device cell table with it (a table
    sized for the old grid would be indexed out of bounds)."""
    sur = _toy_surrogate()
    fine = (16, 6, 6, 12, 20, 6)
    r = np.random.default_rng(9)
    sur.set_cells(fine, (r.uniform(size=int(np.prod(fine))) < 0.7).astype(np.uint8))
    assert sur._cell_ok.numel() == int(np.prod(fine))
    q = _queries(4000, 5)
    _, flags = sur.predict(q)
    want = _host_flags(sur, q.astype(np.float64))
    assert np.mean(flags.cpu().numpy() == want) > 0.999
