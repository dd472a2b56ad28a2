"""The runoptics chain and a coated Mie batch captured as HIP graphs replay to
the eager result.

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


def test_coated_batch_graph_replays(gpu):
    """A coated gm_mie_batch (plan kernels, then the checkpointed kernel,
    whose work-queue counter the plan's scan kernel resets) captured as a HIP
    graph, as bench.py times it: three replays, each after unrelated coated
    work on the default context and with the output clobbered, reproduce the
    eager step bit for bit."""
    t = __import__("torch")
    from geosmie_amd import _lib
    from geosmie_amd._lib import check, ptr
    from geosmie_amd.mie import mie_batch as mb
    r = np.random.default_rng(6)
    y = np.exp(r.uniform(np.log(0.05), np.log(150.0), 20000))
    x = y * r.uniform(0.3, 0.9, y.size)
    n = y.size
    d_x, d_y = _lib.to_dev(x), _lib.to_dev(y)
    d_eps = _lib.cplx_dev(np.full(n, (1.95 + 0.79j) ** 2))
    d_eps2 = _lib.cplx_dev(np.where(r.uniform(size=n) < 0.5, 1.45 + 0j, 1.33 + 1e-4j) ** 2)
    d_props = t.empty(6 * n, dtype=t.float64, device="cuda")
    L = _lib.lib()
    h = _lib.private_ctx()

    def step():
        check(L.gm_mie_batch(h, n, ptr(d_x), ptr(d_y), ptr(d_eps), None, ptr(d_eps2),
                             float(y.max()), None, None, None, ptr(d_props), None,
                             _lib.GM_MIE_SORT, _lib.stream()), "gm_mie_batch")
    side = t.cuda.Stream()
    side.wait_stream(t.cuda.current_stream())
    with t.cuda.stream(side):
        step()
    t.cuda.current_stream().wait_stream(side)
    t.cuda.synchronize()
    eager = d_props.cpu().numpy().copy()
    assert np.isfinite(eager).all()
    graph = t.cuda.CUDAGraph()
    with t.cuda.graph(graph):
        step()
    for _ in range(3):
        mb(np.linspace(0.1, 50.0, 333), m=1.6 + 0.2j, y=np.linspace(0.2, 60.0, 333), m2=1.4 + 0j)
        d_props.fill_(-1.0)
        graph.replay()
        t.cuda.synchronize()
        assert np.array_equal(d_props.cpu().numpy(), eager)
