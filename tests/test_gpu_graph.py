"""The runoptics chain captured as a HIP graph replays to the eager result.

Replays are interleaved with unrelated eager work on the default context and
checked bitwise each time against one eager pass."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _chain(precision=0, gsf_ng=40):
    from geosmie_amd.lut import getLogNormPSD
    from geosmie_amd.pipeline import RunopticsChain
    r = np.logspace(np.log10(0.02), np.log10(3.0), 300)
    lam = np.array([0.4, 0.7, 1.1])
    x = (2 * np.pi * r[None, :] / lam[:, None]).ravel()
    w = np.zeros((3, x.size))
    for i, L in enumerate(lam):
        p = getLogNormPSD(0.3, 1.8, 2 * np.pi * r / L, L, 1e9, 0.0) * np.gradient(r)
        w[i, i * 300:(i + 1) * 300] = p / p.sum()
    ang = np.linspace(0.0, 180.0, 91)
    return RunopticsChain(x, 1.45 + 0.01j, ang, w, lam, [(i * 300, 300) for i in range(3)],
                          precision=precision, gsf_ng=gsf_ng)


def _same(a, b):
    for k in a:
        if a[k] is None:
            assert b[k] is None
        else:
            assert np.array_equal(a[k], b[k]), k


@pytest.mark.parametrize("precision", [0, 2])
def test_graph_replay_matches_eager(gpu, precision):
    from geosmie_amd.mie import mie_batch as mb
    t = __import__("torch")
    ch = _chain(precision)
    ch.step()
    t.cuda.synchronize()
    eager = ch.results()
    assert np.isfinite(eager["pmat"]).all() and np.isfinite(eager["pmom"]).all()
    ch.capture()
    for k in range(3):
        # unrelated eager work between replays, and a clobbered output buffer
        mb(np.linspace(0.1, 50.0, 333), m=1.6 + 0.2j)
        ch.d_pm.fill_(-1.0)
        ch.replay()
        t.cuda.synchronize()
        _same(eager, ch.results())
