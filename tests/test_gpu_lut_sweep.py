"""LUT sweep (bulk Qext, SSA, g, alpha1 moments) on the GPU against the oracle.

The oracle chain per slab: numpy Mie (oracle/mie.py) at the mixed index and
wet sizes, S1/S2 from its pi/tau tables, integratePSD's area weighting, and
the spher_expan.f restatement (oracle/gsf.py, itself pinned to the compiled
reference program) for alpha1.  Outputs are float32: Qext, SSA, g within
2e-6 relative; alpha1_l within 2e-6 of max |alpha1|.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

SLABS = [(0.40, 1.50 + 1e-3j, 1.00), (0.55, 1.45 + 0.1j, 1.50), (1.00, 1.70 + 1e-6j, 2.16),
         (0.30, 1.33 + 1e-8j, 1.20), (2.5, 1.80 + 1.0j, 1.04)]


def _oracle_bulk(r, lam, m_dry, gf, W, ang, ng, n_mom, mw=1.33 + 0j):
    from oracle import gsf as G
    from oracle import mie as M
    m = mw + (m_dry - mw) * (1.0 / gf) ** 3
    x = 2 * np.pi * gf * r / lam
    cost = np.cos(np.radians(ang))
    nm_all = [M.nmax_of(v) for v in x]
    P, T = M.pt_table(cost, max(nm_all))
    qext = np.empty(x.size)
    qsca = np.empty(x.size)
    g = np.empty(x.size)
    p11 = np.empty((x.size, ang.size))
    for i, xi in enumerate(x):
        an, bn, nmax = M.mie_coeffs(m * m, xi)
        qext[i], qsca[i], _, _, g[i], _ = M.mie_props_raw(an, bn, nmax, xi)
        s1, s2 = M.s12_from_pt(an, bn, P, T)
        p11[i] = 0.5 * (np.abs(s1) ** 2 + np.abs(s2) ** 2)
    rw = gf * r
    out = []
    for w in W:
        s_r2 = np.sum(w * rw ** 2)
        qe = np.sum(w * qext * rw ** 2) / s_r2
        qs = np.sum(w * qsca * rw ** 2) / s_r2
        gg = np.sum(w * g * qsca * rw ** 2) / np.sum(w * qsca * rw ** 2)
        pb = w @ p11
        al, _ = G.expand(ang, pb, pb, pb, pb, pb, pb, ng=ng, quantize=False)
        out.append(np.concatenate([[qe, qs / qe, gg], al[0, :n_mom]]))
    return np.array(out)


@pytest.fixture(scope="module")
def sweep():
    from geosmie_amd.lut_sweep import LutSweep
    r = np.logspace(np.log10(0.01), np.log10(5.0), 80)
    return LutSweep(r, rmode=[0.1, 0.5], sigma=[1.6, 2.0])


def test_bulk_vs_oracle(gpu, sweep):
    lam = np.array([s[0] for s in SLABS])
    m = np.array([s[1] for s in SLABS])
    gf = np.array([s[2] for s in SLABS])
    got = sweep.run(lam, m, gf).cpu().numpy().astype(np.float64)
    assert got.shape == (len(SLABS), 4, 35)
    for s in range(len(SLABS)):
        want = _oracle_bulk(sweep.r, lam[s], m[s], gf[s], sweep.w, sweep.ang, sweep.ng,
                            sweep.n_mom)
        np.testing.assert_allclose(got[s, :, :3], want[:, :3], rtol=2e-6, atol=0)
        amax = np.abs(want[:, 3:]).max(axis=1, keepdims=True)
        assert np.all(np.abs(got[s, :, 3:] - want[:, 3:]) <= 2e-6 * amax), s
        assert np.all(got[s, :, 3] == 1.0)


def test_per_slab_weights_match_shared(gpu, sweep):
    t = __import__("torch")
    lam = np.array([0.5, 1.2, 3.0])
    m = np.array([1.5 + 0.01j, 1.6 + 1e-4j, 1.4 + 0.3j])
    gf = np.array([1.0, 1.3, 2.0])
    shared = sweep.run(lam, m, gf)
    per = sweep.run(lam, m, gf, w=np.broadcast_to(sweep.w, (3,) + sweep.w.shape).copy())
    assert t.equal(shared, per)
    one = sweep.run(lam, m, gf, w=sweep.w[None, 2:3].repeat(3, axis=0))
    assert t.equal(one[:, 0], shared[:, 2])


def test_sweep_order_matches_run(gpu, sweep):
    t = __import__("torch")
    mr, mi = np.array([1.35, 1.55]), np.array([1e-6, 0.05])
    lam, gf = np.array([0.35, 0.9]), np.array([1.0, 1.75])
    # tiny batch limits force several batches
    grid = sweep.sweep(mr, mi, lam, gf, max_coeffs=2e4, max_points=400)
    assert tuple(grid.shape) == (2, 2, 2, 2, sweep.n_psd, 35)
    for a in range(2):
        for b in range(2):
            for c in range(2):
                one = sweep.run([lam[c]] * 2, [mr[a] + 1j * mi[b]] * 2, gf)
                assert t.equal(grid[a, b, c], one)


def test_bad_arguments(gpu, sweep):
    with pytest.raises(ValueError):
        sweep.run([0.5], [1.5 - 0.1j], [1.0])
    with pytest.raises(ValueError):
        sweep.run([0.5, 0.6], [1.5], [1.0])
