"""gm_s12_contract_rows (S1/S2 -> Mueller elements -> weighted row sums in one
pass) against the unfused pair gm_s12_mueller + gm_psd_contract_rows on the
same coefficients, and against a numpy contraction of the Mueller planes.
Rows cover ragged ranges (single points, the last point, groups of 1, 2, 4
and 5 rows over the same range, so every accumulator width and the split of a
long group run), with an angle count that is not a multiple of 64."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _batch(n, seed):
    from geosmie_amd.lut import _DevMie
    from geosmie_amd.s12 import angle_tables_dev
    rng = np.random.default_rng(seed)
    x = np.sort(10 ** rng.uniform(-2, 2.2, n))
    dm = _DevMie(x, complex(1.53, 0.008))
    ang = np.linspace(0, 180, 181)
    cap = int(np.rint(2 + x.max() + 4 * x.max() ** (1. / 3.)))
    pi_t, tau_t = angle_tables_dev(np.cos(np.radians(ang)), cap)
    return dm, ang, cap, pi_t, tau_t, rng


def _fused(dm, ang, cap, pi_t, tau_t, starts, counts, W, precision=0):
    from geosmie_amd import _lib
    from geosmie_amd._lib import check, ptr, stream
    t = _lib.torch()
    k = len(starts)
    out = t.empty(k * 4 * ang.size, dtype=t.float64, device=_lib.device())
    st = (ctypes.c_int64 * k)(*[int(a) for a in starts])
    ct = (ctypes.c_int64 * k)(*[int(c) for c in counts])
    d_w = _lib.to_dev(W)
    cl = None
    if precision:
        n_coef = dm.d_an.numel() // 2
        d_cl = t.empty(2 * n_coef, dtype=t.float64, device=_lib.device())   # 16 B per pair
        check(_lib.lib().gm_s12_split_coef(_lib.ctx(), n_coef, ptr(dm.d_an), ptr(dm.d_bn),
                                           precision, ptr(d_cl), stream()), "gm_s12_split_coef")
        cl = ptr(d_cl)
    check(_lib.lib().gm_s12_contract_rows(_lib.ctx(), dm.n, ptr(dm.d_off), ptr(dm.d_nmax),
                                          None if precision else ptr(dm.d_an),
                                          None if precision else ptr(dm.d_bn), ang.size,
                                          ptr(pi_t), ptr(tau_t), cap, k, st, ct, ptr(d_w),
                                          ptr(out), precision, cl, stream()),
          "gm_s12_contract_rows")
    return out.view(k, 4, ang.size).cpu().numpy()


@pytest.mark.parametrize("precision", [0, 1, 2])
@pytest.mark.parametrize("layout", ["mixed", "single", "pairs"])
def test_fused_rows_match_unfused(gpu, layout, precision):
    """At FP32 and split-BF16 the fused pass multiplies the same operands in
    the same order as gm_s12_mueller at that precision, so only the order of
    the FP64 row sums differs from the unfused pair."""
    from geosmie_amd import _lib
    from geosmie_amd._lib import check, ptr, stream
    t = _lib.torch()
    n = 3000
    dm, ang, cap, pi_t, tau_t, rng = _batch(n, 7)
    if layout == "mixed":
        ranges = [(0, 700), (700, 1), (700, 1), (701, 799), (701, 799), (701, 799), (701, 799),
                  (1500, 1499), (1500, 1499), (1500, 1499), (1500, 1499), (1500, 1499),
                  (2999, 1)]
    elif layout == "single":
        ranges = [(0, 3000), (5, 64), (69, 65), (1000, 1), (2999, 1), (100, 2899)]
    else:
        ranges = [(0, 1500), (0, 1500), (1500, 1500), (1500, 1500), (17, 9), (17, 9)]
    starts = np.array([a for a, _ in ranges])
    counts = np.array([c for _, c in ranges])
    W = rng.uniform(0, 1, (len(ranges), n))
    from geosmie_amd import autotune
    gots = []
    for tiles in (1, 2, 4):    # 8-point tiles per work unit
        print("fused rows", layout, "precision", precision, "tiles", tiles, flush=True)
        autotune.set_knob(_lib.ctx(), "s12_row_tiles", tiles)
        try:
            gots.append(_fused(dm, ang, cap, pi_t, tau_t, starts, counts, W, precision))
        finally:
            autotune.set_knob(_lib.ctx(), "s12_row_tiles", 0)
    mu = dm.mueller(pi_t, tau_t, precision)
    ref = t.empty(len(ranges) * 4 * ang.size, dtype=t.float64, device=_lib.device())
    st = (ctypes.c_int64 * len(ranges))(*starts.tolist())
    ct = (ctypes.c_int64 * len(ranges))(*counts.tolist())
    d_w = _lib.to_dev(W)
    check(_lib.lib().gm_psd_contract_rows(_lib.ctx(), len(ranges), st, ct, n, 4 * ang.size,
                                          ptr(d_w), ptr(mu), ptr(ref), stream()))
    ref = ref.view(len(ranges), 4, ang.size).cpu().numpy()
    M = mu.view(n, 4 * ang.size).cpu().numpy()
    for got in gots:
        for r, (a, c) in enumerate(ranges):
            scale = np.abs(ref[r]).max()
            assert np.abs(got[r] - ref[r]).max() <= 1e-12 * scale, (layout, r)
            want = (W[r, a:a + c] @ M[a:a + c]).reshape(4, ang.size)
            assert np.abs(got[r] - want).max() <= 1e-12 * scale, (layout, r)


def test_shared_table_tiles_bitwise(gpu):
    """Split-BF16 fused rows, one 8-point tile per work unit, with the
    angle-table tiles shared by four work units through LDS against each
    wavefront fetching its own: the same
    operands in the same order, so the results are identical bit for bit.
    The ranges give a work-unit count that is not a multiple of four (idle
    wavefronts in the last block), blocks whose units have different series
    lengths, and a partial last angle group (181 angles)."""
    from geosmie_amd import _lib, autotune
    n = 3000
    dm, ang, cap, pi_t, tau_t, rng = _batch(n, 11)
    ranges = [(0, 700), (700, 1), (701, 799), (701, 799), (1500, 1499), (2999, 1), (5, 37)]
    starts = np.array([a for a, _ in ranges])
    counts = np.array([c for _, c in ranges])
    W = rng.uniform(0, 1, (len(ranges), n))
    h = _lib.ctx()
    got = {}
    try:
        autotune.set_knob(h, "s12_row_tiles", 1)
        for share in (1, 0):
            autotune.set_knob(h, "s12_share", share)
            print("split-BF16 rows, one tile per unit, share", share, flush=True)
            got[share] = _fused(dm, ang, cap, pi_t, tau_t, starts, counts, W, 2)
    finally:
        autotune.set_knob(h, "s12_share", 1)
        autotune.set_knob(h, "s12_row_tiles", 0)
    assert np.isfinite(got[1]).all()
    assert np.array_equal(got[1], got[0])


def test_fused_rows_rejects_bad_ranges(gpu):
    dm, ang, cap, pi_t, tau_t, rng = _batch(200, 3)
    W = np.ones((1, 200))
    with pytest.raises(ValueError):
        _fused(dm, ang, cap, pi_t, tau_t, [150], [51], W)
    with pytest.raises(ValueError):
        _fused(dm, ang, cap, pi_t, tau_t, [0], [0], W)
    W65 = np.ones((65, 200))
    with pytest.raises(ValueError):
        _fused(dm, ang, cap, pi_t, tau_t, [0] * 65, [10] * 65, W65)
    # reduced precision without the converted coefficients
    from geosmie_amd import _lib
    t = _lib.torch()
    out = t.empty(4 * ang.size, dtype=t.float64, device=_lib.device())
    st, ct = (ctypes.c_int64 * 1)(0), (ctypes.c_int64 * 1)(200)
    d_w = _lib.to_dev(W)
    with pytest.raises(ValueError, match="gm_s12_split_coef"):
        _lib.check(_lib.lib().gm_s12_contract_rows(
            _lib.ctx(), dm.n, _lib.ptr(dm.d_off), _lib.ptr(dm.d_nmax), None, None, ang.size,
            _lib.ptr(pi_t), _lib.ptr(tau_t), cap, 1, st, ct, _lib.ptr(d_w), _lib.ptr(out), 2,
            None, _lib.stream()), "gm_s12_contract_rows")


@pytest.mark.parametrize("precision", [0, 1, 2])
def test_chain_fused_equals_unfused(gpu, precision):
    """The runoptics chain's step with and without the fused pass, at each
    precision: the row integrals, normalised P, pback, scalars and GSF
    moments agree."""
    from geosmie_amd.lut import getLogNormPSD
    from geosmie_amd.pipeline import RunopticsChain
    t = __import__("torch")
    r = np.logspace(np.log10(0.02), np.log10(5.0), 500)
    lam = np.array([0.35, 0.6, 1.0, 2.2])
    x = (2 * np.pi * r[None, :] / lam[:, None]).ravel()
    w = np.zeros((4, x.size))
    for i, L in enumerate(lam):
        p = getLogNormPSD(0.4, 1.9, 2 * np.pi * r / L, L, 1e9, 0.0) * np.gradient(r)
        w[i, i * 500:(i + 1) * 500] = p / p.sum()
    ang = np.linspace(0.0, 180.0, 201)
    res = []
    for fused in (True, False):
        ch = RunopticsChain(x, 1.5 + 0.02j, ang, w, lam, [(i * 500, 500) for i in range(4)],
                            gsf_ng=64, fused=fused, precision=precision)
        ch.step()
        t.cuda.synchronize()
        res.append({k: getattr(ch, k).cpu().numpy() for k in ("d_pm", "d_back", "d_sc", "d_pmom")})
    assert res[0]["d_pm"].size == 4 * 4 * ang.size
    for k in res[0]:
        a, b = res[0][k], res[1][k]
        assert np.abs(a - b).max() <= 1e-12 * np.abs(b).max(), k
