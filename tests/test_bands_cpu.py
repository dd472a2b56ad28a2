"""Band averaging on the CPU: the oracle and the band tables against vectors
written by the reference's own doAverage/getBands, and the host-side sample
wavenumbers the GPU kernel consumes."""
import os

import numpy as np
import pytest

from geosmie_amd import bands as B
from oracle import bands as OB

V = np.load(os.path.join(os.path.dirname(__file__), "golden", "band_vectors.npz"))
MODES = ["GEOS5", "RRTMG", "RRTMGP", "PURDUE"]


@pytest.mark.parametrize("mode", MODES)
def test_band_tables(mode):
    lo, up, mean, use_wn, nb = B.getBands(mode)
    np.testing.assert_array_equal(np.array([np.asarray(lo, float), np.asarray(up, float)]),
                                  V["bands_%s" % mode])
    np.testing.assert_array_equal(np.asarray(mean, float), V["mean_%s" % mode])
    assert [int(use_wn), nb] == list(V["meta_%s" % mode])


@pytest.mark.parametrize("mode", MODES)
def test_oracle_vs_reference(mode):
    lam, y = V["lam"], V["y"]
    lo, up, _, use_wn, _ = B.getBands(mode)
    got = np.array([[OB.do_average(lam, y[s], lo[b], up[b], use_wn) for b in range(len(lo))]
                    for s in range(y.shape[0])])
    np.testing.assert_array_equal(got, V["avg_%s" % mode])


def test_samples_and_emulated_kernel():
    """The host samples plus numpy.interp on the sorted grid and the pairwise
    sum (the kernel's arithmetic, emulated) reproduce the reference means."""
    lam, y = V["lam"], V["y"]
    for mode in MODES:
        lo, up, _, use_wn, _ = B.getBands(mode)
        xp, order, xs = B.band_samples(lam, lo, up, use_wn)
        assert np.all(np.diff(xp) > 0)
        for s in range(3):
            for b in range(len(lo)):
                v = np.interp(xs[b], xp, y[s][order])
                r = list(v[:8])
                for i in range(8, 96, 8):
                    for k in range(8):
                        r[k] += v[i + k]
                res = ((r[0] + r[1]) + (r[2] + r[3])) + ((r[4] + r[5]) + (r[6] + r[7]))
                for i in range(96, 100):
                    res += v[i]
                assert res / 100 == V["avg_%s" % mode][s, b], (mode, s, b)
