"""Pins of oracle/gsf.py against the reference's own GSF program and closed forms."""
import os
import subprocess

import numpy as np
import pytest

from oracle import gsf

HERE = os.path.dirname(os.path.abspath(__file__))
GV = np.load(os.path.join(HERE, "golden", "gsf_vectors.npz"))
REF_EXE = os.path.join(os.path.dirname(HERE), "oracle", "_ref", "spher_expan.x")


@pytest.mark.parametrize("i", range(len(GV["names"])))
def test_expand_matches_reference_program(i):
    """Bit-for-bit equal to spher_expan.x's F17.10 output (tests/golden/gsf_vectors.npz)."""
    al, _ = gsf.expand(GV["ang"], *GV["F"][i])
    assert np.array_equal(al, GV["coeffs"][i]), np.abs(al - GV["coeffs"][i]).max()


def test_gauss_vs_leggauss():
    for n in (5, 64, 129, 300):
        z, w = gsf.gauss(n)
        zl, wl = np.polynomial.legendre.leggauss(n)
        assert np.abs(z - zl).max() < 1e-14
        assert np.abs(w - wl).max() < 1e-13


def test_rayleigh_closed_form():
    """Rayleigh: alpha1 = (1, 0, 1/2), alpha2_2 = 3, alpha4_1 = 3/2, beta1_2 = sqrt(6)/2
    -- exact with F given at the Gauss nodes (no interpolation error)."""
    z, w = gsf.gauss(129)
    mu = z
    p11 = 0.75 * (1 + mu ** 2)
    al = gsf.spher_expan(z, w, p11, p11, 1.5 * mu, 1.5 * mu, -0.75 * (1 - mu ** 2), 0 * mu, 129)
    al = al / al[0, 0]
    assert np.allclose(al[0, :3], [1.0, 0.0, 0.5], atol=1e-14)
    assert abs(al[1, 2] - 3.0) < 1e-13
    assert abs(al[3, 1] - 1.5) < 1e-13
    assert abs(abs(al[4, 2]) - np.sqrt(6) / 2) < 1e-13
    assert np.abs(al[0, 3:]).max() < 1e-13


def test_henyey_greenstein_moments():
    """alpha1_l = (2l+1) g^l for the HG phase function (at the nodes)."""
    g = 0.6
    z, w = gsf.gauss(129)
    hg = (1 - g * g) / (1 + g * g - 2 * g * z) ** 1.5
    al = gsf.spher_expan(z, w, hg, hg, hg, hg, 0 * z, 0 * z, 129)
    al = al / al[0, 0]
    l = np.arange(60)
    assert np.allclose(al[0, :60], (2 * l + 1) * g ** l, atol=1e-10)


def test_matr_round_trip():
    """Coefficients -> MATR at the nodes reproduces F11 and F44: with N Gauss
    nodes and l < N the discrete Legendre (P^00) transform is invertible
    (Christoffel-Darboux).  The P^22/P^02 rows are not, so only F11, F44."""
    z, w = gsf.gauss(129)
    i = list(GV["names"]).index("mie_0.5_2")
    F = GV["F"][i]
    ang = np.radians(GV["ang"])
    Fn = np.array([[gsf.linterpol(ang, f, np.arccos(x)) for x in z] for f in F])
    al = gsf.spher_expan(z, w, *Fn, l1max=129)
    back = gsf.matr(al, np.arccos(z[::8]))
    for row in (0, 3):
        assert np.allclose(back[row], Fn[row, ::8], rtol=1e-9, atol=1e-11 * np.abs(Fn[row]).max())


def test_hbleg_vs_legval():
    x = np.linspace(-1, 1, 37)
    c = np.random.default_rng(3).normal(size=20)
    assert np.allclose(gsf.hbleg(x, c), np.polynomial.legendre.legval(x, c), atol=1e-12)


@pytest.mark.skipif(not os.path.exists(REF_EXE), reason="oracle/_ref not built")
def test_oracle_vs_reference_program_live(tmp_path):
    """Fresh random smooth matrix through the compiled reference program."""
    rng = np.random.default_rng(7)
    ang = GV["ang"]
    mu = np.cos(np.radians(ang))
    c = rng.uniform(0.1, 1.0, 6)
    F = np.stack([1 + c[k] * mu + 0.3 * c[k] * mu ** 2 for k in range(6)])
    fn = tmp_path / "m.txt"
    np.savetxt(fn, np.vstack([ang, F]).T)
    subprocess.run([REF_EXE, str(fn)], stdout=subprocess.DEVNULL, check=True)
    ref = np.loadtxt(str(fn) + ".expan_coeff", skiprows=1, unpack=True)[1:]
    al, _ = gsf.expand(ang, *F)
    assert np.array_equal(al, ref)


def test_fit_check_vs_reference_program():
    """MATR of the unnormalised coefficients at the input angles against the
    reference program's '.expan_matr' (printed E15.5 / F11.5), and the fit
    error against the value it prints (MAXABS over both check grids)."""
    for i, name in enumerate(GV["names"]):
        m, err = gsf.fit(GV["ang"], *GV["F"][i])
        ref = GV["matr"][i]
        np.testing.assert_allclose(ref[0], GV["ang"], atol=0.006)
        # F11..F44 printed E15.5 (mantissa 0.ddddd: half a unit is <= 5e-5 relative;
        # at a zero crossing the printed value is rounding noise of the fit),
        # F12 and F34 F11.5 (5e-6 absolute)
        for k in range(4):
            assert np.all(np.abs(m[k] - ref[k + 1]) <= 5.01e-5 * np.abs(ref[k + 1]) +
                          1e-12 * np.abs(ref[k + 1]).max()), (name, k)
        for k in (4, 5):
            assert np.all(np.abs(m[k] - ref[k + 1]) <= 5.01e-6), (name, k)
        assert abs(err - GV["fiterr"][i]) <= 1e-9 * GV["fiterr"][i] + 1e-15, (name, err)
