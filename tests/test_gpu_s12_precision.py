"""Reduced-precision S1/S2 contractions (FP32 MFMA, split-BF16 MFMA).

The coefficients stay FP64; only the angular contraction
S1 = sum_n a_n pi'_n + b_n tau'_n, S2 = sum_n a_n tau'_n + b_n pi'_n
(mie_props.py:133-149) runs at reduced precision.

1. Inputs that both formats hold exactly (small integers, multiples of 1/8)
   must give the FP64 result bit for bit: this pins the MFMA lane layouts, the
   K/angle padding and the hi/lo split independently of rounding.
2. On real Mie coefficients the error against the FP64 path (itself pinned to
   the oracle at 1e-11) is measured and bounded per quantity: P11 relative
   error; S1/S2 error relative to max |S| on the angle grid.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _exact_case(rng, n_sets, nmax_cap, n_ang):
    nm = rng.integers(1, nmax_cap + 1, n_sets)
    nm[0] = nmax_cap
    an = [(rng.integers(-16, 17, k) + 1j * rng.integers(-16, 17, k)) / 8.0 for k in nm]
    bn = [(rng.integers(-16, 17, k) + 1j * rng.integers(-16, 17, k)) / 8.0 for k in nm]
    pi_t = rng.integers(-8, 9, (nmax_cap, n_ang)).astype(np.float64)
    tau_t = rng.integers(-8, 9, (nmax_cap, n_ang)).astype(np.float64)
    s1 = np.array([a @ pi_t[:len(a)] + b @ tau_t[:len(a)] for a, b in zip(an, bn)])
    s2 = np.array([a @ tau_t[:len(a)] + b @ pi_t[:len(a)] for a, b in zip(an, bn)])
    return an, bn, pi_t, tau_t, s1, s2


@pytest.mark.parametrize("precision", ["fp64", "fp32", "bf16x3"])
@pytest.mark.parametrize("shape", [(37, 300, 300), (8, 33, 64), (5, 7, 1)])
def test_exact_inputs_bitwise(gpu, precision, shape):
    from geosmie_amd import _lib
    from geosmie_amd.s12 import s12_batch
    rng = np.random.default_rng(sum(shape))
    an, bn, pi_t, tau_t, s1, s2 = _exact_case(rng, *shape)
    g1, g2 = s12_batch(an, bn, pi_tab=_lib.to_dev(pi_t), tau_tab=_lib.to_dev(tau_t),
                       precision=precision)
    np.testing.assert_array_equal(g1, s1)
    np.testing.assert_array_equal(g2, s2)


def _real_case():
    from geosmie_amd.mie import mie_batch
    x = np.logspace(-1, np.log10(800.0), 61)
    m = 1.5 + 0.01j
    r = mie_batch(x, m=m, coeffs=True)
    return r["an"], r["bn"], np.cos(np.radians(np.linspace(0.0, 180.0, 181)))


# FP32 rounds each operand at 2^-24 and accumulates in FP32; split-BF16 keeps
# ~17 bits per operand.  Measured on this grid (x <= 800, 181 angles):
#   fp32    P11 rel median 1.2e-7, max 3.6e-6; S err / max|S| 5.3e-7
#   bf16x3  P11 rel median 5.3e-6, max 2.2e-4; S err / max|S| 1.7e-5
# (deterministic inputs; the bounds leave ~4x).  Where |S| is small next to
# its largest terms (side and back scattering at large x) the relative error
# grows with that cancellation.
BOUNDS = {"fp32": {"p11_median": 5e-7, "p11_max": 2e-5, "s_max": 3e-6},
          "bf16x3": {"p11_median": 2e-5, "p11_max": 1e-3, "s_max": 7e-5}}


@pytest.mark.parametrize("precision", ["fp32", "bf16x3"])
def test_error_vs_fp64(gpu, precision):
    from geosmie_amd.s12 import s12_batch
    an, bn, cost = _real_case()
    r1, r2 = s12_batch(an, bn, cost=cost, precision="fp64")
    g1, g2 = s12_batch(an, bn, cost=cost, precision=precision)
    p11r = 0.5 * (np.abs(r1) ** 2 + np.abs(r2) ** 2)
    p11g = 0.5 * (np.abs(g1) ** 2 + np.abs(g2) ** 2)
    rel = np.abs(p11g - p11r) / p11r
    smax = np.maximum(np.abs(r1).max(axis=1), np.abs(r2).max(axis=1))[:, None]
    serr = max(np.max(np.abs(g1 - r1) / smax), np.max(np.abs(g2 - r2) / smax))
    print("%s: P11 rel median %.2e max %.2e; S err / max|S| %.2e"
          % (precision, np.median(rel), rel.max(), serr))
    b = BOUNDS[precision]
    assert np.median(rel) < b["p11_median"]
    assert rel.max() < b["p11_max"]
    assert serr < b["s_max"]


def test_unknown_precision_rejected(gpu):
    from geosmie_amd import _lib
    from geosmie_amd.s12 import s12_batch
    an, bn, pi_t, tau_t, _, _ = _exact_case(np.random.default_rng(1), 3, 5, 4)
    with pytest.raises(ValueError, match="unknown precision"):
        s12_batch(an, bn, pi_tab=_lib.to_dev(pi_t), tau_tab=_lib.to_dev(tau_t), precision=7)
