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
    for mult, waves, pts in [(1, 4, 8), (2, 4, 8), (4, 4, 8), (1, 1, 8), (1, 2, 16),
                             (2, 4, 16), (1, 1, 16)]:
        autotune.set_knob(h, "mie_grid_mult", mult)
        autotune.set_knob(h, "s12_waves", waves)
        autotune.set_knob(h, "s12_points", pts)
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
