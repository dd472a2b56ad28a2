"""FP64 numpy restatement of the surrogate network (forward, loss, gradients, Adam).

TEST INFRASTRUCTURE ONLY (see ``oracle/__init__.py``).  The surrogate has no
counterpart in the reference; this module defines the arithmetic the GPU
kernels (geosmie_amd/csrc/gm_mlp.hip) must reproduce:

  * parameters: one FP32 vector, per layer W [out][in] row-major then b [out];
  * the model's weights are bf16(W) (round to nearest even of the FP32 value),
    or, in the two-term setting, bf16(W) + bf16(W - bf16(W)); biases stay FP32;
  * z_l = a_{l-1} bf16(W_l)^T + b_l,  a_l = SiLU(z_l) = z_l / (1 + exp(-z_l))
    for every layer but the last, y = z_L;
  * loss = sum_j cw[j] mean_b (y[b][j] - t[b][j])^2;
  * gradients with respect to the FP32 master weights equal those with respect
    to bf16(W) (the rounding is treated as identity in the backward pass);
  * Adam with bias correction.

Pinned by tests/test_oracle_mlp.py: torch autograd in float64 on the same
rounded weights (to 1e-12) and central finite differences on a small net.
"""
import numpy as np


def bf16_round(a):
    """Round FP32 values to BF16 (nearest even); returned as float64."""
    u = np.asarray(a, dtype=np.float32).view(np.uint32).astype(np.uint64)
    u = ((u + 0x7FFF + ((u >> 16) & 1)) >> 16) << 16
    return u.astype(np.uint32).view(np.float32).astype(np.float64)


def layout(dims):
    """[(w_offset, b_offset, n_in, n_out)] per layer and the total length."""
    out, off = [], 0
    for i in range(len(dims) - 1):
        n_in, n_out = int(dims[i]), int(dims[i + 1])
        out.append((off, off + n_in * n_out, n_in, n_out))
        off += n_in * n_out + n_out
    return out, off


def weights(w, terms=1):
    """The weights the network computes with: bf16(W), or with terms=2 the
    pair bf16(W) + bf16(W - bf16(W))."""
    hi = bf16_round(w)
    if terms == 1:
        return hi
    return hi + bf16_round((np.asarray(w, np.float32).astype(np.float64) - hi).astype(np.float32))


def unpack(params, dims, terms=1):
    lay, total = layout(dims)
    p = np.asarray(params, dtype=np.float32)
    assert p.size == total
    return [(weights(p[w:b].reshape(o, i), terms), p[b:b + o].astype(np.float64))
            for w, b, i, o in lay]


def silu(z):
    return z / (1.0 + np.exp(-z))


def dsilu(z):
    s = 1.0 / (1.0 + np.exp(-z))
    return s * (1.0 + z * (1.0 - s))


def forward(params, dims, x, keep=False, terms=1):
    layers = unpack(params, dims, terms)
    a = np.asarray(x, dtype=np.float64)
    zs, acts = [], [a]
    for li, (W, b) in enumerate(layers):
        z = a @ W.T + b
        zs.append(z)
        a = silu(z) if li < len(layers) - 1 else z
        acts.append(a)
    return (a, zs, acts, layers) if keep else a


def loss_grad(params, dims, x, t, cw, terms=1):
    """(loss, flat gradient) in FP64."""
    y, zs, acts, layers = forward(params, dims, x, keep=True, terms=terms)
    n = y.shape[0]
    cw = np.asarray(cw, dtype=np.float64)
    d = y - np.asarray(t, dtype=np.float64)
    loss = float(np.sum(cw * d * d) / n)
    dz = 2.0 * cw * d / n
    lay, total = layout(dims)
    g = np.zeros(total)
    for li in range(len(layers) - 1, -1, -1):
        w_off, b_off, n_in, n_out = lay[li]
        g[w_off:b_off] = (dz.T @ acts[li]).ravel()
        g[b_off:b_off + n_out] = dz.sum(0)
        if li > 0:
            dz = (dz @ layers[li][0]) * dsilu(zs[li - 1])
    return loss, g


def adam(params, grad, m, v, lr, b1, b2, eps, step):
    """One bias-corrected Adam step in FP64; returns (params, m, v)."""
    m = b1 * m + (1 - b1) * grad
    v = b2 * v + (1 - b2) * grad * grad
    c1 = 1.0 / (1.0 - b1 ** step)
    c2 = 1.0 / (1.0 - b2 ** step)
    return params - lr * (m * c1) / (np.sqrt(v * c2) + eps), m, v
