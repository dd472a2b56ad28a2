"""Pins of oracle/lut.py (runoptics size grid, PSD, integration, normalisation)."""
import os

import numpy as np
from scipy.integrate import trapezoid

from oracle import lut

HERE = os.path.dirname(os.path.abspath(__file__))
LV = np.load(os.path.join(HERE, "golden", "lut_vectors.npz"))


def _raw_from_vectors():
    raw = {k: LV["scatvals"][i] for i, k in enumerate(lut.SCATKEYS)}
    for k in lut.SCALARKEYS:
        raw[k] = LV["raw_" + k]
    return raw


def test_integrate_psd_matches_reference_vectors():
    """Restatement vs the reference's own integratePSD output (two modes)."""
    raw = _raw_from_vectors()
    out = lut.integrate_psd(LV["xx"], [raw, raw], list(LV["psd"]), list(LV["fracs"]),
                            float(LV["lam"]), list(LV["reff0"]), float(LV["rhop0"]),
                            float(LV["rhop"]))
    for k in ["num", "area", "volume", "mass", "rEff", "rMass", "qext", "qsca", "qabs", "qb",
              "g", "csca", "cext", "bsca", "bext", "bbck"] + lut.SCATKEYS:
        ref = LV["out_" + k]
        assert np.allclose(out[k], ref, rtol=1e-13, atol=0), k


def test_alias_quirk_weights_csca_by_qsca():
    """csca/cext after key 'g' are Qsca-weighted (dointegration.py:1111-1112,
    1154-1157): an independent weighted mean reproduces the vector, the plain
    area-weighted mean does not."""
    raw = _raw_from_vectors()
    w = LV["psd"][0]
    r2 = (LV["xx"] * float(LV["lam"]) / (2 * np.pi)) ** 2
    qsca = raw["qsca"]
    csca_q = np.sum(raw["csca"] * w * r2 * qsca) / np.sum(w * r2 * qsca)
    csca_a = np.sum(raw["csca"] * w * r2) / np.sum(w * r2)
    one = lut.integrate_psd(LV["xx"], [raw], [w], [1.0], float(LV["lam"]), [LV["reff0"][0]],
                            1700., 1500.)
    assert abs(one["csca"] - csca_q) < 1e-13 * csca_q
    assert abs(one["csca"] - csca_a) > 1e-3 * csca_a
    # qext stays purely area weighted
    qext_a = np.sum(raw["qext"] * w * r2) / np.sum(w * r2)
    assert abs(one["qext"] - qext_a) < 1e-13 * qext_a


def test_delta_psd_returns_the_bin():
    raw = _raw_from_vectors()
    j = 40
    w = np.zeros(LV["xx"].size)
    w[j] = 1.0
    r = LV["xx"][j] * float(LV["lam"]) / (2 * np.pi)
    one = lut.integrate_psd(LV["xx"], [raw], [w], [1.0], float(LV["lam"]), [r], 1000., 1000.)
    for k in ["qext", "qsca", "qabs", "g"]:
        assert abs(one[k] - raw[k][j]) < 1e-14 * abs(raw[k][j])
    assert np.allclose(one["p11"], raw["p11"][j], rtol=1e-15)
    assert abs(one["qb"] - raw["qb"][j]) < 1e-13 * raw["qb"][j]     # back-out identity
    assert abs(one["rEff"] - r) < 1e-15 * r
    assert abs(one["bsca"] - 0.75 * raw["qsca"][j] / 1000. / r) < 1e-12 * one["bsca"]


def test_lognormal_closed_form_moments():
    """sum r^k n(r) dr -> r_mode^k exp(k^2 ln^2 s / 2) on a fine wide grid."""
    lam = 1.0
    xx, dr = lut.getXArrCarma(1e-3, 1e3, 400)
    rm, s = 0.3, 1.8
    p, reff_mass = lut.lognorm_psd(rm, s, xx, dr, lam, 1e9, 0.0)
    r = xx * lam / (2 * np.pi)
    for k in (1, 2, 3):
        exact = rm ** k * np.exp(k * k * np.log(s) ** 2 / 2)
        assert abs(np.sum(r ** k * p) - exact) < 2e-3 * exact
    assert abs(reff_mass - rm * np.exp(3.5 * np.log(s) ** 2)) < 2e-3 * reff_mass


def test_carma_grid_is_geometric():
    xx, dr = lut.getXArrCarma(0.01, 100.0, 50)
    rat = xx[1:] / xx[:-1]
    assert np.allclose(rat, rat[0], rtol=1e-12)
    assert abs(rat[0] ** 3 - (1e4 ** 3) ** (1 / 200.)) < 1e-12
    assert xx.size == 200


def test_post_integration_normalisation():
    raw = _raw_from_vectors()
    one = lut.integrate_psd(LV["xx"], [raw], [LV["psd"][0]], [1.0], float(LV["lam"]),
                            [LV["reff0"][0]], 1700., 1500.)
    post = lut.post_integration(one, LV["ang"])
    th = np.radians(LV["ang"])
    assert abs(trapezoid(post["p11"] * np.sin(th), th) - 2.0) < 1e-13
    ratio = post["p12"] / post["p11"]
    assert np.allclose(ratio, one["p12"] / one["p11"], rtol=1e-13)
    assert post["pback"][0] == post["p11"][-1]
    assert abs(post["lidar_ratio"] - 4 * np.pi * one["qext"] / one["qb"]) < 1e-12 * post["lidar_ratio"]


def test_angle_grid():
    a = lut.angle_grid()
    assert a.size == 371 and a[0] == 0.0 and a[-1] == 180.0
    assert np.all(np.diff(a) > 0)


def test_integrate_psd_alias_switch_closed_form():
    """alias=False: csca = sum w (pi r^2)^2 Qsca / sum w pi r^2 (area weighted);
    alias=True keeps the reference's extra Qsca factor on csca, cext."""
    xx = np.logspace(-1, 1, 40)
    lam = 0.5
    rr = xx * lam / (2 * np.pi)
    rng = np.random.default_rng(2)
    q = rng.uniform(0.5, 2.0, (3, xx.size))
    from oracle import lut as OL
    raw = {k: np.ones((xx.size, 3)) for k in OL.SCATKEYS}
    raw.update(qext=q[0] + q[1], qsca=q[1], qabs=q[0], qb=q[2], g=np.full(xx.size, 0.6),
               csca=q[1] * np.pi * rr ** 2, cext=(q[0] + q[1]) * np.pi * rr ** 2)
    w = rng.uniform(0.1, 1.0, xx.size)
    off = OL.integrate_psd(xx, [raw], [w], [1.0], lam, [rr.mean()], 1000., 1000., alias=False)
    on = OL.integrate_psd(xx, [raw], [w], [1.0], lam, [rr.mean()], 1000., 1000.)
    a = w * np.pi * rr ** 2
    assert np.isclose(off["csca"], np.sum(a * raw["csca"]) / a.sum(), rtol=1e-13)
    assert np.isclose(on["csca"], np.sum(a * q[1] * raw["csca"]) / np.sum(a * q[1]), rtol=1e-13)
    assert off["qext"] == on["qext"] and off["g"] == on["g"]
