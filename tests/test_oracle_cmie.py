"""Pins of the C restatement (oracle/cmie.c) against the scipy-based oracle and
the arbitrary-precision golden vectors."""
import os

import numpy as np
import pytest

from oracle import cmie, mie as M

HERE = os.path.dirname(os.path.abspath(__file__))
MP = np.load(os.path.join(HERE, "golden", "mp_vectors.npz"))


def test_homog_vs_mp_reference_start():
    r = cmie.batch(MP["h_x"], MP["h_m"])
    ref = MP["h_props_ref"]
    with np.errstate(divide="ignore", invalid="ignore"):
        rel = np.abs(r.T - ref) / np.abs(ref)
    assert rel[:, [0, 1]].max() < 1e-11
    assert np.abs(r[4] - ref[:, 4]).max() < 1e-11
    assert np.nanmax(rel[:, [3, 5]]) < 1e-8


def test_coated_vs_mp_reference_start():
    r = cmie.batch(MP["c_x"], MP["c_m1"], MP["c_y"], MP["c_m2"])
    ref = MP["c_props_ref"]
    rel = np.abs(r.T - ref) / np.abs(ref)
    real_shell = np.abs(MP["c_m2"].imag) < 0.05
    assert rel[real_shell][:, [0, 1]].max() < 1e-10
    # absorbing shell (m2 = 1.5+0.5i): the shell combination of mie_coeffs.py:230-244
    # divides quantities of size e^{+-Im(v)}; any FP64 evaluation loses digits
    assert rel[~real_shell][:, [0, 1]].max() < 1e-6


@pytest.mark.parametrize("m", [1.53 + 0.005j, 1.33 + 1e-8j, 1.8 + 0.6j])
def test_homog_vs_scipy_oracle(m):
    x = np.logspace(-2, np.log10(300), 60)
    r = cmie.batch(x, m)
    for i in range(x.size):
        an, bn, nm = M.single_mie_coeff(m * m, 1.0, float(x[i]))
        q = np.array(M.mie_props_raw(an, bn, nm, float(x[i])))
        assert np.all(np.abs(r[[0, 1], i] - q[[0, 1]]) <= 1e-11 * np.abs(q[[0, 1]]))


def test_threads_deterministic():
    x = np.logspace(-2, 2, 500)
    a = cmie.batch(x, 1.5 + 0.01j, threads=1)
    b = cmie.batch(x, 1.5 + 0.01j, threads=4)
    assert np.array_equal(a, b)
