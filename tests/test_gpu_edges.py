"""Edge cases of the Mie path on the GPU, through the C ABI (the Python drop-in).

Batches of 0, 1 and ragged sizes around a 64-lane wavefront; no index
contrast (m = 1); the Rayleigh limit against its closed form; very strongly
absorbing cores; x = 2e4 (n_max 20111, past every workload) against the
oracle, one point per lane and 64 lanes per point; S1/S2 there.
Tolerances as tests/test_gpu_mie.py: Qext, Qsca, g 1e-9 relative against the
oracle; coefficients 1e-9 of each point's largest.
"""
import numpy as np
import pytest

from oracle import mie as M

pytestmark = pytest.mark.gpu


def _check_vs_oracle(r, x, m, i, rtol=1e-9):
    an, bn, nmax = M.single_mie_coeff(m * m, 1.0, float(x))
    assert r["nmax"][i] == nmax
    q = M.mie_props_raw(an, bn, nmax, float(x))
    for j, k in enumerate(("qext", "qsca")):
        assert abs(r[k][i] - q[j]) <= rtol * abs(q[j]), (x, m, k, r[k][i], q[j])
    assert abs(r["asy"][i] - q[4]) <= rtol, (x, m)
    if "an" in r:
        scale = max(np.abs(an).max(), np.abs(bn).max())
        assert np.abs(r["an"][i] - an).max() <= rtol * scale, (x, m)
        assert np.abs(r["bn"][i] - bn).max() <= rtol * scale, (x, m)


def test_empty_batch(gpu):
    from geosmie_amd.mie import mie_batch
    r = mie_batch(np.zeros(0), m=1.5 + 0.01j, coeffs=True)
    for k in ("qext", "qsca", "qabs", "qb", "asy", "qratio", "nmax"):
        assert r[k].shape == (0,)
    assert r["an"] == [] and r["bn"] == []
    c = mie_batch(np.zeros(0), m=1.95 + 0.79j, y=np.zeros(0), m2=1.45 + 0j)
    assert c["qext"].shape == (0,)


@pytest.mark.parametrize("n", [1, 63, 64, 65, 129])
def test_ragged_batch_sizes(gpu, n):
    """A last wavefront with 1, 63, 0, 1 or 1 of 64 lanes busy: every point
    matches the oracle (each lane's point, none lost or duplicated)."""
    from geosmie_amd.mie import mie_batch
    x = np.exp(np.linspace(np.log(0.05), np.log(300.0), n))[::-1].copy()
    m = 1.5 + 0.02j
    r = mie_batch(x, m=m, coeffs=True)
    for i in range(n):
        _check_vs_oracle(r, x[i], m, i)


def test_no_index_contrast(gpu):
    """m = 1: psi_{n-1} = (D_n + n/x) psi_n makes every a_n, b_n numerator
    vanish, so nothing scatters or absorbs, to rounding: the numerators are
    differences of O(1) terms, and the reference algorithm (the oracle) itself
    leaves residuals growing with the row count, 2e-12 at x = 1000; each
    point's residual stays within 4x the oracle's (+1e-14)."""
    from geosmie_amd.mie import mie_batch
    x = np.logspace(-3, 3, 50)
    r = mie_batch(x, m=1.0 + 0j, coeffs=True)
    assert np.all(np.abs(r["qext"]) < 1e-12)
    assert np.all(np.abs(r["qsca"]) < 1e-20)
    for i, (a, b) in enumerate(zip(r["an"], r["bn"])):
        oa, ob, _ = M.single_mie_coeff(1.0 + 0j, 1.0, float(x[i]))
        ref = max(np.abs(oa).max(), np.abs(ob).max())
        assert max(np.abs(a).max(), np.abs(b).max()) <= 4 * ref + 1e-14, x[i]


def test_rayleigh_limit_closed_form(gpu):
    """x -> 0: Qsca = (8/3) x^4 |K|^2 and Qext = 4 x Im K + (8/3) x^4 Re K^2
    with K = (m^2 - 1)/(m^2 + 2), to relative O(x^2); g -> 0."""
    from geosmie_amd.mie import mie_batch
    x = np.array([1e-5, 1e-4, 3e-4])
    for m in (1.33 + 0j, 1.5 + 0.01j, 2.0 + 1.0j):
        K = (m * m - 1) / (m * m + 2)
        r = mie_batch(x, m=m)
        qsca = 8.0 / 3.0 * x ** 4 * abs(K) ** 2
        qext = 4 * x * K.imag + 8.0 / 3.0 * x ** 4 * (K * K).real
        assert np.all(np.abs(r["qsca"] - qsca) <= 1e-5 * qsca), m
        assert np.all(np.abs(r["qext"] - qext) <= 1e-5 * qext), m
        assert np.all(np.abs(r["asy"]) < 1e-5), m


def test_strongly_absorbing(gpu):
    from geosmie_amd.mie import mie_batch
    x = np.logspace(-1, 2.3, 40)
    for m in (1.5 + 5j, 2.0 + 10j, 3.0 + 4j):
        r = mie_batch(x, m=m, coeffs=True)
        for i in range(x.size):
            _check_vs_oracle(r, x[i], m, i)


@pytest.mark.parametrize("lanes", [1, 64])
def test_very_large_size_parameter(gpu, lanes):
    """x = 2e4 (n_max = 20111) with one point per lane and with the rows
    split over 64 lanes; S1/S2 on 37 angles within 1e-11 of max |S|."""
    from geosmie_amd import _lib, autotune
    from geosmie_amd.mie import mie_batch
    from geosmie_amd.s12 import s12_batch
    h = _lib.ctx()
    x = np.array([20000.0, 12345.6, 5000.0])
    m = 1.45 + 1e-4j
    try:
        autotune.set_knob(h, "mie_lanes", lanes)
        r = mie_batch(x, m=m, coeffs=True)
    finally:
        autotune.set_knob(h, "mie_lanes", 0)
    for i in range(x.size):
        _check_vs_oracle(r, x[i], m, i)
    cost = np.cos(np.radians(np.linspace(0.0, 180.0, 37)))
    s1, s2 = s12_batch(r["an"], r["bn"], cost=cost)
    P, T = M.pt_table(cost, int(r["nmax"].max()))
    for i in range(x.size):
        o1, o2 = M.s12_from_pt(r["an"][i], r["bn"][i], P, T)
        scale = max(np.abs(o1).max(), np.abs(o2).max())
        assert np.abs(s1[i] - o1).max() <= 1e-11 * scale, i
        assert np.abs(s2[i] - o2).max() <= 1e-11 * scale, i


@pytest.mark.parametrize("ckpt", [4, 0])
def test_coated_large_shells(gpu, ckpt):
    """Coated spheres out to y = 1000 (5x the coated workload's largest
    shell), real and weakly absorbing shells, strongly and weakly absorbing
    cores, with the checkpointed kernel (4) and the one that keeps every row
    (0), against the C restatement's cancellation-free form (its literal form
    agrees with it to 5e-11 on these points): Qext, Qsca 1e-9 relative, g
    1e-9 absolute."""
    from oracle import cmie
    from geosmie_amd import _lib, autotune
    from geosmie_amd.mie import mie_batch
    h = _lib.ctx()
    r = np.random.default_rng(8)
    y = np.exp(np.linspace(np.log(200.0), np.log(1000.0), 60))
    x = y * r.uniform(0.3, 0.9, y.size)
    m1 = np.where(r.uniform(size=y.size) < 0.5, 1.95 + 0.79j, 1.5 + 0.01j)
    m2 = np.where(r.uniform(size=y.size) < 0.5, 1.33 + 0j, 1.45 + 1e-4j)
    try:
        autotune.set_knob(h, "mie_ckpt", ckpt)
        got = mie_batch(x, m=m1, y=y, m2=m2)
    finally:
        autotune.set_knob(h, "mie_ckpt", 4)
    c = cmie.batch(x, m1, y, m2, threads=8, stable=True)
    for j, k in enumerate(("qext", "qsca")):
        rel = np.abs(got[k] - c[j]) / np.abs(c[j])
        assert rel.max() < 1e-9, (k, rel.max(), y[rel.argmax()])
    assert np.abs(got["asy"] - c[4]).max() < 1e-9
