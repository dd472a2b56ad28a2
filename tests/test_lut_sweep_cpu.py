"""Host-side pieces of the LUT sweep (no GPU): PSD weights, series-length
rule and the sweep grid's size."""
import numpy as np

from oracle import lut as OL
from oracle import mie as OM


def test_lognormal_weights_match_calculate_psd():
    from geosmie_amd.lut_sweep import lognormal_weights
    r = np.logspace(np.log10(0.005), np.log10(50.0), 3000)
    for rmode, sigma in [(0.01, 1.4), (0.3, 2.0), (5.0, 2.6)]:
        # calculatePSD in size-parameter space with lambda = 2 pi (x = r)
        want, _ = OL.lognorm_psd(rmode, sigma, r, OL.getDR(r), 2 * np.pi, 1e30, 0.0)
        np.testing.assert_allclose(lognormal_weights(r, rmode, sigma), want, rtol=1e-13,
                                   atol=1e-300)


def test_nmax_rule_half_even():
    from geosmie_amd.lut_sweep import nmax_rule
    xs = np.concatenate([np.logspace(-3, 3.6, 2000), [0.5, 1.0, 8.0, 27.0, 64.0, 125.0]])
    assert np.array_equal(nmax_rule(xs), [OM.nmax_of(v) for v in xs])


def test_full_grid_size():
    from geosmie_amd.lut_sweep import SU_GF, SU_RH, full_grid
    g = full_grid()
    n_slab = g["mr"].size * g["mi"].size * g["lam"].size * g["gf"].size
    assert SU_GF.size == SU_RH.size == 36
    assert np.all(np.diff(SU_GF) > 0) and SU_GF[0] == 1.0
    assert n_slab * g["rmode"].size * g["sigma"].size == 56_160_000
    assert g["r"].size * g["mr"].size * g["mi"].size * g["lam"].size == 117_000_000


def test_query_encoding_roundtrip():
    from geosmie_amd.surrogate import decode_queries, gate_coords, make_queries
    r = np.random.default_rng(0)
    n = 500
    rm = np.exp(r.uniform(np.log(0.01), np.log(5), n))
    sg = r.uniform(1.4, 2.6, n)
    mr = r.uniform(1.3, 1.8, n)
    mi = np.exp(r.uniform(np.log(1e-8), 0, n))
    lam = np.exp(r.uniform(np.log(0.2), np.log(40), n))
    gf = r.uniform(1.0, 2.16, n)
    q = make_queries(rm, sg, mr, mi, lam, gf)
    a, b, m, l, g = decode_queries(q)
    np.testing.assert_allclose(a, rm, rtol=2e-5)
    np.testing.assert_allclose(b, sg, rtol=1e-6)
    np.testing.assert_allclose(m.real, mr, rtol=1e-6)
    np.testing.assert_allclose(m.imag, mi, rtol=2e-5)
    np.testing.assert_allclose(l, lam, rtol=1e-6)
    np.testing.assert_allclose(g, gf, rtol=1e-5)
    gc = gate_coords(q)
    np.testing.assert_allclose(gc[:, 0], np.log(rm), atol=2e-5)
    # r_eff of the lognormal: exp(2.5 ln^2 sigma) r_mode, grown by gf
    np.testing.assert_allclose(np.exp(q[:, 0]), gf * rm * np.exp(2.5 * np.log(sg) ** 2), rtol=1e-6)
