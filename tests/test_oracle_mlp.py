"""The surrogate-network oracle (oracle/mlp.py) against torch autograd and
finite differences (CPU)."""
import numpy as np
import pytest

from oracle import mlp as O


def _case(dims, n, seed):
    rng = np.random.default_rng(seed)
    _, total = O.layout(dims)
    p = (rng.standard_normal(total) * 0.4).astype(np.float32)
    x = rng.standard_normal((n, dims[0]))
    t = rng.standard_normal((n, dims[-1]))
    cw = rng.uniform(0.5, 2.0, dims[-1])
    return p, x, t, cw


def test_bf16_round():
    v = np.array([1.0, 1.00390625, 1.005859375, 3.1415927, -2.0e-3, 65504.0], dtype=np.float32)
    r = O.bf16_round(v)
    # representable values stay; ties go to even; |rel err| <= half an ulp = 2^-8
    assert r[0] == 1.0 and r[1] == 1.0 and r[2] == 1.0078125
    assert np.all(np.abs(r - v) <= np.abs(v) * 2.0 ** -8)
    u = r.astype(np.float32).view(np.uint32)
    assert np.all(u & 0xFFFF == 0)


@pytest.mark.parametrize("dims", [(6, 16, 16, 5), (6, 256, 256, 256, 256, 35)])
def test_grad_vs_torch_autograd(dims):
    torch = pytest.importorskip("torch")
    p, x, t, cw = _case(dims, 24, 1)
    loss, g = O.loss_grad(p, dims, x, t, cw)
    layers = O.unpack(p, dims)
    Ws = [torch.tensor(W, dtype=torch.float64, requires_grad=True) for W, _ in layers]
    bs = [torch.tensor(b, dtype=torch.float64, requires_grad=True) for _, b in layers]
    a = torch.tensor(x)
    for i, (W, b) in enumerate(zip(Ws, bs)):
        z = a @ W.T + b
        a = torch.nn.functional.silu(z) if i < len(Ws) - 1 else z
    lt = (torch.tensor(cw) * (a - torch.tensor(t)) ** 2).sum() / x.shape[0]
    lt.backward()
    lay, _ = O.layout(dims)
    ref = np.zeros_like(g)
    for (w_off, b_off, n_in, n_out), W, b in zip(lay, Ws, bs):
        ref[w_off:b_off] = W.grad.numpy().ravel()
        ref[b_off:b_off + n_out] = b.grad.numpy()
    assert abs(loss - lt.item()) <= 1e-12 * abs(loss)
    np.testing.assert_allclose(g, ref, rtol=1e-10, atol=1e-13 * np.abs(ref).max())


def test_grad_vs_finite_differences():
    dims = (3, 7, 4)
    p, x, t, cw = _case(dims, 5, 2)
    # differentiate with respect to the rounded weights: perturb in float64
    layers = O.unpack(p, dims)
    lay, total = O.layout(dims)

    def loss_of(flat):
        a = x
        for li, (w_off, b_off, n_in, n_out) in enumerate(lay):
            W = flat[w_off:b_off].reshape(n_out, n_in)
            z = a @ W.T + flat[b_off:b_off + n_out]
            a = O.silu(z) if li < len(lay) - 1 else z
        return float(np.sum(cw * (a - t) ** 2) / x.shape[0])

    base = np.zeros(total)
    for (w_off, b_off, n_in, n_out), (W, b) in zip(lay, layers):
        base[w_off:b_off] = W.ravel()
        base[b_off:b_off + n_out] = b
    _, g = O.loss_grad(p, dims, x, t, cw)
    h = 1e-6
    fd = np.array([(loss_of(base + h * e) - loss_of(base - h * e)) / (2 * h)
                   for e in np.eye(total)])
    np.testing.assert_allclose(g, fd, rtol=1e-6, atol=1e-8)


def test_adam_matches_definition():
    rng = np.random.default_rng(3)
    p, g = rng.standard_normal(10), rng.standard_normal(10)
    m, v = np.zeros(10), np.zeros(10)
    p1, m1, v1 = O.adam(p, g, m, v, 1e-3, 0.9, 0.999, 1e-8, 1)
    # first step of bias-corrected Adam moves every weight by lr * sign(g) (up to eps)
    np.testing.assert_allclose(p - p1, 1e-3 * np.sign(g), rtol=1e-6)


def test_rayleigh_baseline_closed_forms():
    """The surrogate's Qext baseline: one sphere (sigma_g = 1) at small x is
    the Rayleigh extinction of the Mie series (oracle Mie to O(x^2)); B
    saturates at 2 for large particles; a dropped factor of the moments
    (exp(6 s^2)) changes B for broad distributions."""
    from geosmie_amd.surrogate import make_queries, rayleigh_log_b
    from oracle import mie as OM
    for m in (1.5 + 0.01j, 1.33 + 1e-6j, 1.8 + 0.5j):
        for x in (1e-3, 3e-3):
            lam = 10.0
            q = make_queries(np.array([x * lam / (2 * np.pi)]), np.array([1.0 + 1e-12]),
                             np.array([m.real]), np.array([m.imag]), np.array([lam]),
                             np.array([1.0]))
            an, bn, nmax = OM.single_mie_coeff(m * m, 1.0, x)
            qext = OM.mie_props_raw(an, bn, nmax, x)[0]
            b = np.exp(rayleigh_log_b(q))[0]
            q_r = 1.0 / (1.0 / b - 0.5)          # B = Q_R / (1 + Q_R / 2)
            # queries are float32: ~1e-5 from the rounded radius and index
            assert abs(q_r / qext - 1) < 10 * x * x + 5e-5, (m, x, q_r, qext)
    big = make_queries(np.array([5.0]), np.array([2.0]), np.array([1.5]), np.array([1e-3]),
                       np.array([0.2]), np.array([1.0]))
    assert abs(np.exp(rayleigh_log_b(big))[0] - 2) < 1e-6
    # broad distribution: Q_sca scales with exp(6 ln^2 sigma) at fixed r_eff
    q1 = make_queries(np.array([0.01]), np.array([1.0 + 1e-12]), np.array([1.5]),
                      np.array([1e-9]), np.array([3.0]), np.array([1.0]))
    q2 = make_queries(np.array([0.01 * np.exp(-2.5 * np.log(2.0) ** 2)]), np.array([2.0]),
                      np.array([1.5]), np.array([1e-9]), np.array([3.0]), np.array([1.0]))
    d = rayleigh_log_b(q2)[0] - rayleigh_log_b(q1)[0]
    assert abs(d - 6 * np.log(2.0) ** 2) < 1e-3
