"""numpy restatement of runbands' band averaging.

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).

Follows src/geosmie/bandaverage.py:18-50 (doAverage): band edges to
wavenumbers for wavelength band sets, 100 samples from numpy.linspace across
the band, clipped to the table's wavenumbers, scipy interp1d (linear) of the
spectrum at the samples, sum / 100.  Pinned bit for bit against vectors the
reference's own doAverage wrote (tests/golden/make_band_vectors.py,
tests/test_bands_cpu.py).
"""
import numpy as np
from scipy.interpolate import interp1d


def do_average(lam, y, bandl, bandr, use_wavenumber, n_sub=100):
    if not use_wavenumber:
        bandl = (100. * bandl) ** (-1)
        bandr = (100. * bandr) ** (-1)
    wnum = (100. * lam) ** (-1)
    beg, end = (bandl, bandr) if bandl < bandr else (bandr, bandl)
    xs = np.clip(np.linspace(beg, end, n_sub), np.min(wnum), np.max(wnum))
    v = interp1d(wnum, y)(xs)
    return np.sum(v) / len(v)
