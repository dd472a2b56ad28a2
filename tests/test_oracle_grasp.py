"""Pins of the kernel-mode restatement (oracle/grasp.py): against vectors the
reference's own bilinear_interpolation, get_interpolated and integrateShapes
wrote (tests/golden/make_grasp_vectors.py), and against what the mathematics
fixes (bilinear functions are reproduced exactly; outside the grid the value
is that of the clipped index; the kernel grid widths keep r/dr constant)."""
import os

import numpy as np

from oracle import grasp as G

HERE = os.path.dirname(os.path.abspath(__file__))
GV = np.load(os.path.join(HERE, "golden", "grasp_vectors.npz"))


def test_integrate_shapes_vs_reference_vectors():
    avg = G.integrate_shapes({"qext": GV["sca"], "scama": GV["scama"], "mr": GV["mr"]},
                             GV["fracs0"])
    np.testing.assert_array_equal(avg["qext"], GV["avg_qext"])
    np.testing.assert_array_equal(avg["scama"], GV["avg_scama"])
    np.testing.assert_array_equal(avg["mr"], GV["avg_mr"])


def test_get_interpolated_vs_reference_vectors():
    for i, (a, b) in enumerate(zip(GV["q_mr"], GV["q_mi"])):
        s = G.get_interpolated(GV["avg_qext"], a, b, GV["mr"], GV["mi"])
        p = G.get_interpolated(GV["avg_scama"][:, :, :, 2, :], a, b, GV["mr"], GV["mi"])
        np.testing.assert_array_equal(s, GV["interp_qext"][i])
        np.testing.assert_array_equal(p, GV["interp_p"][i])


def test_corner_weights_reproduce_get_interpolated():
    """The (indices, weights) form the device kernel is given gives the same
    values when combined in the reference's order."""
    d = GV["avg_qext"]
    for a, b in zip(GV["q_mr"], GV["q_mi"]):
        i, j, w1, w2, w3, w4 = G.corner_weights(a, b, GV["mr"], GV["mi"])
        got = w1 * d[i, j] + w2 * d[i + 1, j] + w4 * d[i, j + 1] + w3 * d[i + 1, j + 1]
        np.testing.assert_array_equal(got, G.get_interpolated(d, a, b, GV["mr"], GV["mi"]))


def test_bilinear_functions_exact_and_clamped():
    mr = np.linspace(1.3, 1.8, 6)
    mi = np.array([0.0, 1e-3, 1e-2, 0.1, 0.5])
    f = lambda a, b: 0.7 + 2.0 * a - 3.0 * b + 1.5 * a * b
    table = f(mr[:, None], mi[None, :])[:, :, None] * np.ones(4)
    for a, b in [(1.41, 0.004), (1.3, 0.0), (1.8, 0.5), (1.55, 0.3)]:
        got = G.get_interpolated(table, a, b, mr, mi)
        assert np.allclose(got, f(a, b), rtol=1e-14, atol=1e-14)
    # outside the grid: the value at the clipped refractive index
    for a, b in [(1.2, 0.004), (1.9, 0.7), (1.5, -1.0)]:
        got = G.get_interpolated(table, a, b, mr, mi)
        assert np.allclose(got, f(np.clip(a, 1.3, 1.8), np.clip(b, 0.0, 0.5)), rtol=1e-14)


def test_kernel_grid_widths_constant_ratio():
    x = 0.05 * 1.2 ** np.arange(30)
    dr = G.kernel_grid_widths(x)
    ratio = dr / x
    assert np.allclose(ratio, ratio[0], rtol=1e-13)
    # the width of a bin of volume ratio rmrat between volume-equivalent edges
    rmrat = 1.2 ** 3
    r_lo = x * (2. / (1. + rmrat)) ** (1. / 3.)
    assert np.allclose(dr, r_lo * (rmrat ** (1. / 3.) - 1.) * (4. * np.pi / 3.) ** (1. / 3.) *
                       (3. / 4. / np.pi) ** (1. / 3.), rtol=1e-12)
