"""GPU parity of the runoptics chain (rawMie, integratePSD, normalisation, PSD
contraction) against the oracle and the reference's own integratePSD vectors.

Tolerances: integrated P elements and efficiencies 1e-12 relative when fed
identical per-size inputs (pure FP64 reductions, order differs); the full
device chain (Mie + S12 + integration) 1e-9 on Q/g and 1e-9 of max |P| on P.
"""
import os

import numpy as np
import pytest

from oracle import lut as OL, mie as M

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
LV = np.load(os.path.join(HERE, "golden", "lut_vectors.npz"))


def _raw():
    raw = {k: LV["scatvals"][i] for i, k in enumerate(OL.SCATKEYS)}
    for k in OL.SCALARKEYS:
        raw[k] = LV["raw_" + k]
    return raw


def test_integrate_psd_vs_reference_vectors(gpu):
    """Drop-in integratePSD on the reference's own inputs/outputs (two modes)."""
    from geosmie_amd.lut import integratePSD
    raw = _raw()
    out = integratePSD(LV["xx"], [raw, raw], list(LV["psd"]), list(LV["fracs"]), float(LV["lam"]),
                       list(LV["reff0"]), float(LV["rhop0"]), float(LV["rhop"]))
    for k in ["num", "area", "volume", "mass", "rEff", "rMass", "qext", "qsca", "qabs", "qb",
              "g", "csca", "cext", "bsca", "bext", "bbck"] + OL.SCATKEYS:
        ref = LV["out_" + k]
        assert np.allclose(out[k], ref, rtol=1e-12, atol=0), (k, out[k], ref)


def test_integrate_psd_alias_quirk_switch(gpu):
    """With the alias quirk off, csca and cext are area-weighted (the oracle's
    alias=False); every other key is unchanged, and csca, cext move by the
    Qsca weighting the quirk adds."""
    from geosmie_amd.lut import integratePSD
    raw = _raw()
    args = (LV["xx"], [raw, raw], list(LV["psd"]), list(LV["fracs"]), float(LV["lam"]),
            list(LV["reff0"]), float(LV["rhop0"]), float(LV["rhop"]))
    on = integratePSD(*args)
    off = integratePSD(*args, alias_quirk=False)
    ref = OL.integrate_psd(*args, alias=False)
    for k in ["qext", "qsca", "qabs", "qb", "g", "csca", "cext", "bsca", "bext", "rEff"]:
        assert np.allclose(off[k], ref[k], rtol=1e-12, atol=0), (k, off[k], ref[k])
    for k in ["qext", "qsca", "qabs", "qb", "g", "bsca", "bext", "bbck"]:
        assert off[k] == on[k], k
    for k in ["csca", "cext"]:
        assert abs(off[k] - on[k]) > 1e-6 * abs(on[k]), k


def test_rawmie_vs_oracle(gpu):
    from geosmie_amd.lut import MultipleMie, rawMie, scatkeys, scalarkeys
    xx = LV["xx"][::3]
    ang = OL.angle_grid()
    cost = np.cos(np.radians(ang))
    mm = MultipleMie(xx, None, cost)
    mm.preCalculate()
    lam, m = 0.55, 1.50 + 0.01j
    got = rawMie(mm, scatkeys, scalarkeys, lam, m.real, m.imag, None, cost)
    ref = OL.raw_mie(xx, lam, M.size_range(xx, cost, m.real, m.imag))
    scale = np.abs(ref["p11"]).max(axis=1, keepdims=True)   # |P_ij| <= P11
    for k in OL.SCATKEYS:
        assert np.all(np.abs(got[k] - ref[k]) <= 1e-11 * scale), k
    for k in OL.SCALARKEYS:
        assert np.allclose(got[k], ref[k], rtol=1e-9, atol=0), k


def test_size_range_structure(gpu):
    from geosmie_amd.lut import MultipleMie
    xx = np.array([0.3, 2.0, 9.0])
    cost = np.cos(np.radians([0., 45., 90., 180.]))
    mm = MultipleMie(xx, None, cost)
    mm.preCalculate()
    got = mm.calculateS12SizeRange(1.45, 0.002)
    ref = M.size_range(xx, cost, 1.45, 0.002)
    assert len(got["s12"]) == 3 and len(got["s12"][0]) == 4
    for i in range(3):
        for j in range(4):
            s1, s2 = got["s12"][i][j]
            assert abs(s1 - ref["s1"][i, j]) <= 1e-12 * np.abs(ref["s1"][i]).max()
            assert abs(s2 - ref["s2"][i, j]) <= 1e-12 * np.abs(ref["s2"][i]).max()
    assert np.allclose(got["qext"], ref["qext"], rtol=1e-12)


def test_size_range_coated(gpu):
    """A coated size range (core sizes xArr, shell sizes yArr, shell index
    (mr2, mi2)): S1/S2 and efficiencies against the mpmath coated vectors with
    the reference's start (dense core 1.95+0.79i, real shell 1.45)."""
    from geosmie_amd.lut import MultipleMie
    MP = np.load(os.path.join(HERE, "golden", "mp_vectors.npz"))
    sel = np.flatnonzero((MP["c_m2"] == 1.45 + 0j))
    x, y = MP["c_x"][sel], MP["c_y"][sel]
    cost = np.cos(np.radians(MP["theta"]))
    mm = MultipleMie(x, y, cost)
    with pytest.raises(ValueError):
        mm.calculateS12SizeRange(1.95, 0.79)
    got = mm.calculateS12SizeRange(1.95, 0.79, 1.45, 0.0)
    for i, k in enumerate(sel):
        s1 = np.array([v[0] for v in got["s12"][i]])
        s2 = np.array([v[1] for v in got["s12"][i]])
        scale = max(np.abs(MP["c_s1_ref"][k]).max(), np.abs(MP["c_s2_ref"][k]).max())
        assert np.abs(s1 - MP["c_s1_ref"][k]).max() < 1e-9 * scale
        assert np.abs(s2 - MP["c_s2_ref"][k]).max() < 1e-9 * scale
        assert abs(got["qext"][i] - MP["c_props_ref"][k][0]) < 1e-9 * MP["c_props_ref"][k][0]
        assert abs(got["qsca"][i] - MP["c_props_ref"][k][1]) < 1e-9 * MP["c_props_ref"][k][1]


def test_post_integration_vs_oracle(gpu):
    from geosmie_amd.lut import post_integration
    raw = _raw()
    one = OL.integrate_psd(LV["xx"], [raw], [LV["psd"][0]], [1.0], float(LV["lam"]),
                           [LV["reff0"][0]], 1700., 1500.)
    ref = OL.post_integration(one, LV["ang"])
    got = post_integration(one, LV["ang"])
    for k in ["p11", "p12", "p22", "p33", "p34", "p44", "pback"]:
        assert np.allclose(got[k], ref[k], rtol=1e-12, atol=1e-15), k
    assert abs(got["ssa"] - ref["ssa"]) < 1e-15
    assert abs(got["lidar_ratio"] - ref["lidar_ratio"]) < 1e-12 * ref["lidar_ratio"]


def test_lut_entries_end_to_end(gpu):
    """Device chain Mie -> Mueller -> PSD contraction -> scalars -> normalisation."""
    from geosmie_amd.lut import lut_entries
    xx, dr = OL.getXArrCarma(0.05, 60.0, 30)
    lam, m = 0.65, 1.53 + 0.005j
    ang = OL.angle_grid()
    psd = []
    reff0 = []
    for rm, s in [(0.1, 1.6), (0.5, 2.0), (1.2, 1.8)]:
        p, r0 = OL.lognorm_psd(rm, s, xx, dr, lam, 50.0, 0.001)
        psd.append(p)
        reff0.append(r0)
    got = lut_entries(xx, lam, m, np.array(psd), ang, reff0=reff0, rhop0=1700., rhop=1700.)
    raw = OL.raw_mie(xx, lam, M.size_range(xx, np.cos(np.radians(ang)), m.real, m.imag))
    for r in range(3):
        one = OL.integrate_psd(xx, [raw], [psd[r]], [1.0], lam, [reff0[r]], 1700., 1700.)
        ref = OL.post_integration(one, ang)
        for k in ["qext", "qsca", "qabs", "g", "qb", "bext", "bsca"]:
            assert abs(got[k][r] - ref[k]) <= 1e-9 * abs(ref[k]), (r, k)
        for i, k in enumerate(["p11", "p12", "p33", "p34"]):
            scale = np.abs(ref["p11"]).max()
            assert np.abs(got["pmat"][r, i] - ref[k]).max() <= 1e-9 * scale, (r, k)
        assert np.allclose(got["pback"][r], ref["pback"], rtol=1e-8, atol=1e-12)


def test_psd_contract_mfma_and_small(gpu):
    """Batched-PSD MFMA path (n_psd > 8) and split-K VALU path vs numpy."""
    from geosmie_amd.lut import _contract
    rng = np.random.default_rng(11)
    v = rng.normal(size=(1337, 71))
    for n_psd in (1, 3, 8, 9, 40):
        w = rng.uniform(size=(n_psd, 1337))
        got = _contract(w, v)
        ref = w @ v
        assert np.allclose(got, ref, rtol=1e-12, atol=1e-12 * np.abs(ref).max()), n_psd


def test_integrate_psd_rows_matches_per_row(gpu):
    """gm_integrate_psd_rows (all rows in one pass) equals gm_integrate_psd run
    row by row, and the one-pass contraction equals per-row contractions."""
    import ctypes
    from geosmie_amd import _lib
    from geosmie_amd._lib import check, ptr, stream
    from geosmie_amd.lut import getDR, getLogNormPSD
    t = _lib.torch()
    rng = np.random.default_rng(4)
    r = np.logspace(-2, 1, 700)
    lams = np.array([0.35, 0.6, 1.1, 2.2])
    x = np.concatenate([2 * np.pi * r / L for L in lams])
    n = x.size
    W = np.zeros((lams.size, n))
    for i, L in enumerate(lams):
        p = getLogNormPSD(rng.uniform(0.1, 1.0), rng.uniform(1.5, 2.2), 2 * np.pi * r / L, L,
                          1e9, 0.0) * (2 * np.pi / L) * getDR(r)
        W[i, i * r.size:(i + 1) * r.size] = p / p.sum()
    from geosmie_amd.mie import mie_batch
    res = mie_batch(x, m=1.5 + 0.01j)
    props = np.stack([res[k] for k in ("qext", "qsca", "qabs", "qb", "asy", "qratio")])
    rr = x[None, :] * np.repeat(lams, r.size)[None, :] / (2 * np.pi)
    reff0 = (W * rr ** 4).sum(1) / (W * rr ** 3).sum(1)
    d_x, d_props, d_w = _lib.to_dev(x), _lib.to_dev(props), _lib.to_dev(W)
    d_lp = _lib.to_dev(np.repeat(lams, r.size))
    d_r0 = _lib.to_dev(reff0)
    out = t.empty(lams.size * 16, dtype=t.float64, device=_lib.device())
    L, h, s = _lib.lib(), _lib.ctx(), stream()
    check(L.gm_integrate_psd_rows(h, lams.size, n, ptr(d_x), ptr(d_lp), ptr(d_props), ptr(d_w),
                                  ptr(d_r0), 1000.0, 1700.0, None, ptr(out), s))
    rows = out.view(lams.size, 16).cpu().numpy()
    # per-row wet densities (one per rh row in runoptics)
    rho = 1100.0 + 150.0 * np.arange(lams.size)
    d_rho = _lib.to_dev(rho)
    out2 = t.empty_like(out)
    check(L.gm_integrate_psd_rows(h, lams.size, n, ptr(d_x), ptr(d_lp), ptr(d_props), ptr(d_w),
                                  ptr(d_r0), 1000.0, 1700.0, ptr(d_rho), ptr(out2), s))
    rows2 = out2.view(lams.size, 16).cpu().numpy()
    one = _lib.to_dev(np.array([1.0]))
    for i, Lm in enumerate(lams):
        o1 = t.empty(16, dtype=t.float64, device=_lib.device())
        check(L.gm_integrate_psd(h, n, 1, ptr(d_x), ptr(d_props), ptr(d_w[i]), ptr(one),
                                 ptr(d_r0[i:i + 1]), float(Lm), 1000.0, 1700.0, ptr(o1), s))
        np.testing.assert_allclose(rows[i], o1.cpu().numpy(), rtol=1e-12)
        check(L.gm_integrate_psd(h, n, 1, ptr(d_x), ptr(d_props), ptr(d_w[i]), ptr(one),
                                 ptr(d_r0[i:i + 1]), float(Lm), 1000.0, float(rho[i]), ptr(o1),
                                 s))
        np.testing.assert_allclose(rows2[i], o1.cpu().numpy(), rtol=1e-12)


@pytest.mark.parametrize("chunk", [64, 512])
def test_psd_contract_rows_matches_per_row(gpu, chunk):
    """The segmented contraction (every row over its own points, one launch)
    equals per-row gm_psd_contract on the same ranges."""
    import ctypes
    from geosmie_amd import _lib, autotune
    from geosmie_amd._lib import check, ptr, stream
    t = _lib.torch()
    rng = np.random.default_rng(chunk)
    n_pts, n_cols = 3000, 333
    starts = np.array([0, 700, 701, 1500, 2999])
    counts = np.array([700, 1, 799, 1499, 1])
    W = rng.uniform(0, 1, (starts.size, n_pts))
    V = rng.standard_normal((n_pts, n_cols))
    d_w, d_v = _lib.to_dev(W), _lib.to_dev(V)
    out = t.empty(starts.size * n_cols, dtype=t.float64, device=_lib.device())
    L, h, s = _lib.lib(), _lib.ctx(), stream()
    autotune.set_knob(h, "psd_chunk", chunk)
    st = (ctypes.c_int64 * starts.size)(*starts.tolist())
    ct = (ctypes.c_int64 * starts.size)(*counts.tolist())
    check(L.gm_psd_contract_rows(h, starts.size, st, ct, n_pts, n_cols, ptr(d_w), ptr(d_v),
                                 ptr(out), s))
    got = out.view(starts.size, n_cols).cpu().numpy()
    for r, (a, c) in enumerate(zip(starts, counts)):
        want = W[r, a:a + c] @ V[a:a + c]
        np.testing.assert_allclose(got[r], want, rtol=1e-12, atol=1e-12 * np.abs(want).max())
    bad = (ctypes.c_int64 * 1)(2900)
    cnt = (ctypes.c_int64 * 1)(200)
    with pytest.raises(ValueError):
        check(L.gm_psd_contract_rows(h, 1, bad, cnt, n_pts, n_cols, ptr(d_w), ptr(d_v), ptr(out), s))
    autotune.set_knob(h, "psd_chunk", 128)
