"""Pins of the C restatement (oracle/cmie.c) against the scipy-based oracle and
the arbitrary-precision golden vectors."""
import os
import subprocess

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


MPL = np.load(os.path.join(HERE, "golden", "mp_largex_vectors.npz"))


def _coated_sets():
    """Every coated mpmath vector with the reference's start: the dense-core
    set (m1 = 1.95+0.79i) and the dense-shell set (|m2| >= |m1|), whose
    start lies close enough to |m2 x| or |m2 y| to need the start correction."""
    return (np.concatenate([MP["c_x"], MPL["cd_x"]]), np.concatenate([MP["c_y"], MPL["cd_y"]]),
            np.concatenate([MP["c_m1"], MPL["cd_m1"]]), np.concatenate([MP["c_m2"], MPL["cd_m2"]]),
            np.concatenate([MP["c_props_ref"], MPL["cd_props"]]))


def _start_correction_on(x, y, m1, m2):
    """Whether the stable form's dD terms are nonzero (cmie.c converged_start)."""
    nmax = np.rint(2 + y + 4 * y ** (1.0 / 3.0))
    nmx = np.rint(np.maximum(nmax, np.maximum(np.abs(m1 * y), np.abs(m2 * y))) + 16)
    av, aw = np.abs(m2 * x), np.abs(m2 * y)
    return (nmx - av < 8 * np.cbrt(av)) | (nmx - aw < 8 * np.cbrt(aw))


def test_coated_stable_form_vs_mp_every_vector():
    """cmie.batch(stable=True), the checker of the GPU's coated kernels, against
    every coated mpmath vector (reference start).  Real and weakly absorbing
    shells: Qext, Qsca 1e-11, g 1e-11 absolute; the absorbing-shell points
    (|Im m2| y > 1) use the literal form in both and keep its 1e-6."""
    x, y, m1, m2, ref = _coated_sets()
    on = _start_correction_on(x, y, m1, m2)
    assert on.sum() >= 90          # the dD branch is exercised, not only the converged case
    r = cmie.batch(x, m1, y, m2, stable=True)
    rel = np.abs(r[:2].T - ref[:, :2]) / np.abs(ref[:, :2])
    weak = np.abs(m2.imag) * y <= 1.0
    assert rel[weak].max() < 1e-11, rel[weak].max()
    assert np.abs(r[4] - ref[:, 4])[weak].max() < 1e-11
    assert rel[~weak].max() < 1e-6
    # the start correction matters: the converged coefficients differ from the
    # reference's by far more than the tolerance on these points
    conv = MPL["cd_props_conv"][:, :2]
    assert (np.abs(conv - MPL["cd_props"][:, :2]) / np.abs(MPL["cd_props"][:, :2])).max() > 1e-4


def test_coated_literal_form_vs_mp_dense_shell():
    """The literal restatement of mie_coeffs.py:230-249 on the dense-shell set:
    within 1e-9 (its FP64 conditioning on these points is ~2.5e-10)."""
    r = cmie.batch(MPL["cd_x"], MPL["cd_m1"], MPL["cd_y"], MPL["cd_m2"], stable=False)
    ref = MPL["cd_props"]
    rel = np.abs(r[:2].T - ref[:, :2]) / np.abs(ref[:, :2])
    assert rel.max() < 1e-9, rel.max()


def test_chain_homog_vs_numpy_oracle():
    """The host chain baseline (coefficients, S1/S2, Mueller elements,
    weighted row sums) against the numpy/scipy oracle composed the same way,
    on two rows of a small homogeneous size grid at 37 angles."""
    rng = np.random.default_rng(7)
    lam = np.array([0.4, 1.1])
    rad = np.logspace(-2, 1, 40)
    x = np.concatenate([2 * np.pi * rad / L for L in lam])
    r = np.concatenate([rad, rad])
    row = np.repeat([0, 1], rad.size)
    w = rng.uniform(0.1, 1.0, x.size)
    m = 1.53 + 0.005j
    cost = np.cos(np.radians(np.linspace(0, 180, 37)))
    pm, sc = cmie.chain_homog(x, m, r, row, w, 2, cost, threads=3)
    from oracle import lut as OL
    for k in range(2):
        sel = row == k
        o = M.size_range(x[sel], cost, m.real, m.imag)
        sv = OL.scat_vals(o["s1"], o["s2"])
        want = np.stack([w[sel] @ sv[0], w[sel] @ sv[1], w[sel] @ sv[3], w[sel] @ sv[4]])
        scale = np.abs(want[0]).max()
        assert np.abs(pm[k] - want).max() < 1e-12 * scale
        area = w[sel] * np.pi * r[sel] ** 2
        want_sc = [area.sum(), area @ o["qext"], area @ o["qsca"], area @ (o["qsca"] * o["asy"])]
        np.testing.assert_allclose(sc[k], want_sc, rtol=1e-12)
    a, _ = cmie.chain_homog(x, m, r, row, w, 2, cost, threads=1)
    b, _ = cmie.chain_homog(x, m, r, row, w, 2, cost, threads=5)
    np.testing.assert_allclose(a, b, rtol=1e-13)


def test_sanitizer_build_runs_clean():
    """The C restatement built with AddressSanitizer and UndefinedBehaviorSanitizer
    (make -C oracle sanitize) runs its driver (every coated and homogeneous path,
    the chain, several OpenMP threads) without a report and with finite output."""
    import shutil
    if shutil.which("gcc") is None:
        pytest.skip("gcc not available")
    oracle_dir = os.path.join(os.path.dirname(HERE), "oracle")
    subprocess.run(["make", "-s", "-C", oracle_dir, "sanitize"], check=True)
    # verify_asan_link_order=0: the run may start with other libraries preloaded
    env = dict(os.environ,
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=1:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(oracle_dir, "_san", "cmie_check")], env=env,
                       stdout=subprocess.PIPE, stderr=subprocess.PIPE, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "ok" in r.stdout
