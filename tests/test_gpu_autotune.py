"""Launch-shape knobs change timing only: results are bitwise identical for
every persistent-grid multiple and S1/S2 block width, and the autotuner
leaves a valid setting on the context."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def test_knobs_do_not_change_results(gpu):
    from geosmie_amd import _lib, autotune
    from geosmie_amd.mie import mie_batch
    from geosmie_amd.s12 import s12_batch
    h = _lib.ctx()
    x = np.logspace(-2, 2.5, 700)
    ref = None
    for mult, waves, pts, lay in [(1, 4, 8, 0), (2, 4, 8, 0), (4, 4, 8, 0), (1, 1, 8, 0),
                                  (1, 2, 16, 0), (2, 4, 16, 0), (1, 1, 16, 0), (1, 4, 8, 1),
                                  (1, 2, 8, 1), (1, 4, 16, 1), (1, 1, 16, 1), (1, 4, 8, 2),
                                  (2, 1, 16, 2)]:
        autotune.set_knob(h, "mie_grid_mult", mult)
        autotune.set_knob(h, "s12_waves", waves)
        autotune.set_knob(h, "s12_points", pts)
        autotune.set_knob(h, "s12_layout", lay)
        r = mie_batch(x, m=1.5 + 0.02j, coeffs=True)
        s1, s2 = s12_batch(r["an"][::37], r["bn"][::37],
                           cost=np.cos(np.radians(np.linspace(0, 180, 300))))
        got = (np.stack([r[k] for k in ("qext", "qsca", "asy")]), s1, s2)
        if ref is None:
            ref = got
        else:
            for a, b in zip(ref, got):
                assert np.array_equal(a, b)
    autotune.set_knob(h, "mie_grid_mult", 1)
    autotune.set_knob(h, "s12_waves", 4)
    autotune.set_knob(h, "s12_points", 8)
    autotune.set_knob(h, "s12_layout", 0)


def test_mie_register_variants_agree(gpu):
    """The register-budget variants of the homogeneous and coated recurrence
    kernels (0 = unconstrained, 3 and 4 waves per SIMD) run the same source;
    under register pressure the compiler may fuse a different set of
    multiply-adds, so they agree to rounding level, not bit for bit: within
    the parity bound against the oracle (1e-9 relative on the efficiencies,
    where near-resonant points amplify ulp differences; 1e-9 of each point's
    largest coefficient)."""
    from geosmie_amd import _lib, autotune
    from geosmie_amd.mie import mie_batch
    h = _lib.ctx()
    r = np.random.default_rng(3)
    y = np.exp(r.uniform(np.log(0.05), np.log(200.0), 3000))
    x = y * r.uniform(0.3, 0.9, y.size)
    xs = np.logspace(-2, 3, 2000)
    ref = None
    for waves in (0, 3, 4):
        autotune.set_knob(h, "mie_waves", waves)
        c = mie_batch(x, m=1.95 + 0.79j, y=y, m2=1.45 + 0j, coeffs=True)
        hm = mie_batch(xs, m=1.5 + 0.01j, coeffs=True)
        got = [c[k] for k in ("qext", "qsca", "asy", "qb", "an", "bn")] + \
              [hm[k] for k in ("qext", "qsca", "asy", "an", "bn")]
        if ref is None:
            ref = got
        else:
            for a, b in zip(ref, got):
                if isinstance(a, (list, tuple)):       # per-point coefficient arrays
                    for ap, bp in zip(a, b):
                        assert ap.shape == bp.shape
                        assert np.all(np.abs(ap - bp) <= 1e-9 * np.abs(ap).max()), waves
                else:
                    assert np.all(np.abs(a - b) <= 1e-9 * np.abs(a)), waves
    autotune.set_knob(h, "mie_waves", 0)


def test_tune_picks_a_candidate(gpu):
    from geosmie_amd import _lib, autotune
    from geosmie_amd.pipeline import RunopticsChain
    r = np.logspace(-2, 0.5, 400)
    x = 2 * np.pi * r / 0.55
    w = np.exp(-np.log(r / 0.3) ** 2)[None, :]
    ch = RunopticsChain(x, 1.5 + 0.01j, np.linspace(0, 180, 181), w / w.sum(), [0.55],
                        [(0, x.size)])
    res = autotune.tune(ch.ctx, ch.step, reps=2)
    for name, r_ in res.items():
        assert r_["best"] in autotune.KNOBS[name][1]
        assert autotune.get(ch.ctx, name) == r_["best"]
    with pytest.raises(ValueError):
        autotune.set_knob(ch.ctx, "s12_waves", 3)


def test_coated_checkpoint_variants_agree(gpu):
    """The checkpointed coated kernel (chain states kept every 4, 5 or 8 rows,
    each block of rows regenerated in LDS) against the kernel that keeps every
    row in scratch: the regenerated rows repeat pass 1's operations from the
    same states, so the results agree to the last bits (asserted at 1e-12, and
    the bitwise outcome printed), including points that need the start
    correction (|m2| >= |m1|) and stored coefficients."""
    from geosmie_amd import _lib, autotune
    from geosmie_amd.mie import mie_batch
    h = _lib.ctx()
    r = np.random.default_rng(5)
    y = np.exp(r.uniform(np.log(0.05), np.log(209.0), 4000))
    x = y * r.uniform(0.3, 0.95, y.size)
    m1 = np.where(r.uniform(size=y.size) < 0.5, 1.95 + 0.79j, 1.33 + 1e-4j)
    m2 = np.where(m1.real > 1.5, 1.45 + 0j, 1.5 + 1e-3j)
    res = {}
    try:
        for ck in (0, 4, 5, 8):
            autotune.set_knob(h, "mie_ckpt", ck)
            res[ck] = mie_batch(x, m=m1, y=y, m2=m2, coeffs=True)
    finally:
        autotune.set_knob(h, "mie_ckpt", 4)
    for ck in (4, 5, 8):
        same = True
        for k in ("qext", "qsca", "asy", "qb"):
            a, b = res[0][k], res[ck][k]
            same &= bool(np.array_equal(a, b))
            assert np.all(np.abs(a - b) <= 1e-12 * np.abs(a)), (ck, k)
        for a, b in zip(res[0]["an"], res[ck]["an"]):
            same &= bool(np.array_equal(a, b))
            assert np.all(np.abs(a - b) <= 1e-12 * np.abs(a).max()), ck
        print("checkpoint every %d rows: bitwise identical to the full-row kernel: %s" % (ck, same))


def test_homog_checkpoint_agrees(gpu):
    """The checkpointed lane-per-point homogeneous kernel (D_n and psi-ratio
    states kept every 4 rows, blocks regenerated in LDS) against the kernel
    that keeps every row in scratch: the regenerated rows repeat pass 1's
    operations from the same states and the upward step is the same source,
    so the results agree to the last bits (asserted at 1e-12, the bitwise
    outcome printed); sizes from the Rayleigh limit to x = 3400 (n_max
    3455), weakly to strongly absorbing indices, coefficients stored."""
    from geosmie_amd import _lib, autotune
    from geosmie_amd.mie import mie_batch
    h = _lib.ctx()
    r = np.random.default_rng(21)
    x = np.concatenate([np.logspace(-3, np.log10(3400.0), 2500), [0.5, 1.0, 4.0, 5.0, 8.0]])
    m = (r.uniform(1.01, 1.9, x.size) + 1j * 10 ** r.uniform(-8, 0.3, x.size))
    res = {}
    try:
        autotune.set_knob(h, "mie_lanes", 1)
        for ck in (0, 4):
            autotune.set_knob(h, "mie_ckpt", ck)
            res[ck] = mie_batch(x, m=m, coeffs=True)
    finally:
        autotune.set_knob(h, "mie_ckpt", 4)
        autotune.set_knob(h, "mie_lanes", 0)
    same = True
    for k in ("qext", "qsca", "asy", "qb"):
        a, b = res[0][k], res[4][k]
        same &= bool(np.array_equal(a, b))
        assert np.all(np.abs(a - b) <= 1e-12 * np.abs(a)), k
    for key in ("an", "bn"):
        for a, b in zip(res[0][key], res[4][key]):
            assert a.shape == b.shape
            same &= bool(np.array_equal(a, b))
            assert np.all(np.abs(a - b) <= 1e-12 * np.abs(a).max()), key
    print("homogeneous, checkpoint every 4 rows: bitwise identical to the full-row kernel: %s" % same)


def test_homog_lane_variants_agree(gpu):
    """The homogeneous kernel with one point per lane and with 16, 32 or 64
    lanes per point (segments joined by parallel scans) agree to rounding
    level: efficiencies within 1e-10 relative (g 1e-10 absolute), stored
    coefficients within 1e-10 of each point's largest, on sizes from the
    Rayleigh limit to x = 2094 and indices from weakly to strongly absorbing."""
    from geosmie_amd import _lib, autotune
    from geosmie_amd.mie import mie_batch
    h = _lib.ctx()
    r = np.random.default_rng(11)
    x = np.concatenate([np.logspace(-3, np.log10(2094.0), 900), [0.5, 1.0, 1.7, 3.0]])
    m = (r.uniform(1.01, 1.9, x.size) + 1j * 10 ** r.uniform(-8, 0.3, x.size))
    res = {}
    try:
        for lanes in (1, 16, 32, 64):
            autotune.set_knob(h, "mie_lanes", lanes)
            res[lanes] = mie_batch(x, m=m, coeffs=True)
    finally:
        autotune.set_knob(h, "mie_lanes", 0)
    ref = res[1]
    for lanes in (16, 32, 64):
        got = res[lanes]
        for k in ("qext", "qsca", "qb"):
            rel = np.abs(got[k] - ref[k]) / np.abs(ref[k])
            assert rel.max() < 1e-10, (lanes, k, rel.max(), x[rel.argmax()], m[rel.argmax()])
        assert np.abs(got["asy"] - ref["asy"]).max() < 1e-10, lanes
        # each point's coefficients against its largest |a_n|, |b_n| (in the
        # Rayleigh limit b_1 ~ x^5 is formed by cancellation in both kernels)
        for a, b, ga, gb in zip(ref["an"], ref["bn"], got["an"], got["bn"]):
            scale = max(np.abs(a).max(), np.abs(b).max())
            assert np.abs(a - ga).max() <= 1e-10 * scale, lanes
            assert np.abs(b - gb).max() <= 1e-10 * scale, lanes


def test_tune_cache_reuses_choices(gpu, tmp_path):
    """A second tune() with the same workload name reads the cached choices
    (no timing) and leaves them on a fresh context."""
    from geosmie_amd import _lib, autotune
    from geosmie_amd.mie import mie_batch
    x = np.logspace(-1, 2, 300)
    path = str(tmp_path / "tune.json")
    h = _lib.ctx()
    first = autotune.tune(h, lambda: mie_batch(x, m=1.5 + 0.01j), knobs=("mie_lanes",), reps=1,
                          cache_key="homog300", cache=path)
    assert not first["mie_lanes"].get("cached")
    h2 = _lib.private_ctx()
    again = autotune.tune(h2, lambda: None, knobs=("mie_lanes",), cache_key="homog300",
                          cache=path)
    assert again["mie_lanes"]["cached"] and again["mie_lanes"]["best"] == first["mie_lanes"]["best"]
    assert autotune.get(h2, "mie_lanes") == first["mie_lanes"]["best"]
    autotune.set_knob(h, "mie_lanes", 0)
