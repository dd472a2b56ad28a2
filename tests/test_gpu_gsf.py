"""GPU parity of the GSF expansion against the reference program's output and
the oracle; Legendre evaluation against numpy.

The reference writes coefficients rounded to 10 decimals (F17.10,
spher_expan.f:104); the device rounds with rint(v 1e10) 1e-10, which can land
one unit of the last decimal away on ties, so the bound is 1.01e-10 absolute.
"""
import os

import numpy as np
import pytest

from oracle import gsf as OG

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GV = np.load(os.path.join(HERE, "golden", "gsf_vectors.npz"))


def test_spher_expan_vs_reference_program(gpu):
    from geosmie_amd.gsf import spher_expan
    for i, name in enumerate(GV["names"]):
        got = spher_expan(GV["ang"], *GV["F"][i])
        ref = GV["coeffs"][i]
        assert np.abs(got - ref).max() <= 1.01e-10, (name, np.abs(got - ref).max())


def test_pmom_batch_layout(gpu):
    """All golden matrices in one batch; pmom p order 11, 12, 33, 34, 22, 44
    (convertncdf.py:381-394)."""
    from geosmie_amd.gsf import pmom_batch
    F = GV["F"]   # program order F11, F22, F33, F44, F12, F34
    table = np.stack([F[:, 0], F[:, 4], F[:, 2], F[:, 5], F[:, 1], F[:, 3]], axis=1)
    got = pmom_batch(GV["ang"], table)
    for i in range(len(F)):
        ref = np.zeros((6, 129))
        for ii, npol in enumerate(OG.PMOM_ORDER):
            ref[npol] = GV["coeffs"][i][ii]
        assert np.abs(got[i] - ref).max() <= 1.01e-10


def test_unquantized_300_moments_vs_oracle(gpu):
    """NG = 300 nodes/moments (the large-x workload's moment count), no rounding."""
    from geosmie_amd.gsf import pmom_batch
    i = list(GV["names"]).index("mie_0.5_2")
    F = GV["F"][i]
    table = np.stack([F[0], F[4], F[2], F[5], F[1], F[3]])
    got = pmom_batch(GV["ang"], table, ng=300, quantize=False)[0]
    al, _ = OG.expand(GV["ang"], *F, ng=300, quantize=False)
    ref = np.zeros_like(got)
    for ii, npol in enumerate(OG.PMOM_ORDER):
        ref[npol] = al[ii]
    assert np.abs(got - ref).max() <= 1e-11 * np.abs(ref).max()


def test_hbleg(gpu):
    from geosmie_amd.gsf import hbleg
    x = np.cos(np.radians(np.linspace(0, 180, 181)))
    c = np.random.default_rng(5).normal(size=(3, 64))
    got = hbleg(x, c)
    for s in range(3):
        assert np.allclose(got[s], OG.hbleg(x, c[s]), rtol=1e-12, atol=1e-12)


def test_fit_check_vs_reference_program(gpu):
    """The fitted matrix back from the device coefficients (MATR on the GPU)
    against the reference program's '.expan_matr' (E15.5 / F11.5 text), and the
    fit error against the value the program prints."""
    from geosmie_amd.gsf import expan_matr
    for i, name in enumerate(GV["names"]):
        m, err = expan_matr(GV["ang"], *GV["F"][i])
        ref = GV["matr"][i]
        for k in range(4):
            assert np.all(np.abs(m[k] - ref[k + 1]) <= 5.01e-5 * np.abs(ref[k + 1]) +
                          1e-12 * np.abs(ref[k + 1]).max()), \
                (name, k)
        for k in (4, 5):
            assert np.all(np.abs(m[k] - ref[k + 1]) <= 5.01e-6), (name, k)
        assert abs(err - GV["fiterr"][i]) <= 1e-9 * GV["fiterr"][i], (name, err, GV["fiterr"][i])


def test_matr_vs_oracle_300_moments(gpu):
    """MATR of 300 normalised moments on the device against the oracle's MATR."""
    from geosmie_amd.gsf import matr_batch, pmom_batch
    i = list(GV["names"]).index("mie_0.5_2")
    F = GV["F"][i]
    table = np.stack([F[0], F[4], F[2], F[5], F[1], F[3]])
    c = pmom_batch(GV["ang"], table, ng=300, quantize=False)
    got = matr_batch(c, GV["ang"])[0]
    prog = np.stack([c[0, 0], c[0, 4], c[0, 2], c[0, 5], c[0, 1], c[0, 3]])
    ref = OG.matr(prog, np.radians(GV["ang"]))          # F11, F22, F33, F44, F12, F34
    want = np.stack([ref[0], ref[4], ref[2], ref[5], ref[1], ref[3]])
    assert np.abs(got - want).max() <= 1e-12 * np.abs(want).max()
