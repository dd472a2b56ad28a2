"""What the trained surrogate gains over plain GPU Mie (the method without it).

The committed model (geosmie_amd/data/surrogate_lut.npz, trained on the GPU
sweep's own tables by geosmie_amd.train_surrogate) answers random queries in
the sweep's ranges.  Against plain GPU Mie on the same queries:
  * every query the gate routes to plain Mie gets exactly the plain answer;
  * queries the network serves agree with plain Mie within the bound on Qext,
    SSA and g (the gate is empirical: a small stated fraction may exceed it);
  * the network path costs a small fraction of plain Mie per query.
"""
import os
import time

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
MODEL = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "geosmie_amd",
                     "data", "surrogate_lut.npz")


def _uniform_queries(n, seed):
    from geosmie_amd.lut_sweep import full_grid
    from geosmie_amd.surrogate import make_queries
    g = full_grid()
    r = np.random.default_rng(seed)
    gf = (1 - r.uniform(0, 1 - g["gf"].max() ** -3.0, n)) ** (-1.0 / 3.0)
    return make_queries(np.exp(r.uniform(np.log(0.01), np.log(5.0), n)), r.uniform(1.4, 2.6, n),
                        r.uniform(1.3, 1.8, n), np.exp(r.uniform(np.log(1e-8), 0.0, n)),
                        np.exp(r.uniform(np.log(0.2), np.log(40.0), n)), gf)


def _trusted_queries(sur, n, seed):
    """Queries drawn uniformly inside cells the gate trusts."""
    from geosmie_amd.surrogate import make_queries
    r = np.random.default_rng(seed)
    ok = np.flatnonzero(sur.cell_ok)
    cells = r.choice(ok, size=n)
    lo, hi = sur.box_lo.astype(np.float64), sur.box_hi.astype(np.float64)
    idx = np.array(np.unravel_index(cells, sur.cells)).T
    u = (idx + r.uniform(0.02, 0.98, idx.shape)) / np.array(sur.cells)
    gc = lo + u * (hi - lo)     # ln r_mode, sigma, m_r, ln m_i, ln lam, f_w
    gf = (1 - gc[:, 5]) ** (-1.0 / 3.0)
    return make_queries(np.exp(gc[:, 0]), gc[:, 1], gc[:, 2], np.exp(gc[:, 3]), np.exp(gc[:, 4]),
                        gf)


@pytest.fixture(scope="module")
def sur():
    from geosmie_amd.surrogate import Surrogate
    assert os.path.exists(MODEL), "trained surrogate model missing"
    return Surrogate.load(MODEL)


def test_fallback_is_plain_mie(gpu, sur):
    t = __import__("torch")
    q = _uniform_queries(600, 1)
    rows, nf = sur.query(q)
    _, flags = sur.predict(q)
    f = flags.cpu().numpy() != 0
    assert nf == int(f.sum())
    if nf:
        plain = sur.solve_plain(q[f])
        assert t.equal(rows[t.from_numpy(f).to(rows.device)], plain)


def test_served_within_bound_and_faster(gpu, sur):
    from geosmie_amd.surrogate import rel_errors
    t = __import__("torch")
    if not sur.cell_ok.any():
        pytest.skip("the committed model's gate trusts no cell")
    q = _trusted_queries(sur, 2000, 2)
    rows, flags = sur.predict(q)
    assert np.all(flags.cpu().numpy() == 0)
    t.cuda.synchronize()
    t0 = time.perf_counter()
    for _ in range(5):
        sur.predict(q)
    t.cuda.synchronize()
    t_net = (time.perf_counter() - t0) / 5
    t0 = time.perf_counter()
    plain = sur.solve_plain(q)
    t.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    err = rel_errors(rows.cpu().numpy(), plain.cpu().numpy()).max(axis=1)
    frac = float(np.mean(err <= sur.bound))
    print("served queries within %.0e: %.4f (median err %.2e, max %.2e); network %.3g s, "
          "plain GPU Mie %.3g s for %d queries (%.0fx)" % (
              sur.bound, frac, np.median(err), err.max(), t_net, t_plain, len(q),
              t_plain / t_net))
    assert frac >= 0.95
    assert np.median(err) <= sur.bound
    assert t_plain > 20 * t_net


def test_end_to_end_gain_over_plain(gpu, sur):
    """Uniform queries answered by query() (network, gate, plain re-solve of
    the flagged ones) against plain GPU Mie for all of them: the surrogate
    path with its fallback is several times faster, and every answer is
    either the network's (served) or exactly plain Mie's."""
    t = __import__("torch")
    q = _uniform_queries(2048, 3)
    sur.query(q[:128])
    sur.solve_plain(q[:128])
    t.cuda.synchronize()
    t0 = time.perf_counter()
    rows, nf = sur.query(q)
    t.cuda.synchronize()
    t_sur = time.perf_counter() - t0
    t0 = time.perf_counter()
    plain = sur.solve_plain(q)
    t.cuda.synchronize()
    t_plain = time.perf_counter() - t0
    print("uniform queries: %d of %d re-solved by plain Mie; surrogate+fallback %.4f s, plain "
          "GPU Mie %.4f s (%.1fx)" % (nf, len(q), t_sur, t_plain, t_plain / t_sur))
    assert nf < 0.2 * len(q)
    assert t_plain > 4 * t_sur
