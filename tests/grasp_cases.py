"""Synthetic GRASP-style kernel tables for the kernel-mode tests: the
variables convertkernels.py writes (mr, mi stored negative, x on a geometric
grid, qext, qabs, qsca, qb, g [shape, m_r, m_i, x], scama [shape, m_r, m_i, x,
6, 181] with elements p11, p22, p33, p44, p12, p34), filled with smooth made-up
values (no kernel files ship with the reference)."""
import numpy as np


def kernel_variables(n_shape=2, n_x=36):
    mr = np.linspace(1.30, 1.70, 5)
    mi = -np.array([0.0, 1e-4, 1e-3, 1e-2, 0.1])
    x = 0.02 * 1.25 ** np.arange(n_x)
    th = np.radians(np.linspace(0., 180., 181))
    R, I, X = np.meshgrid(mr, -mi, x, indexing="ij")
    out = {k: [] for k in ("qext", "qabs", "qsca", "qb", "g", "scama")}
    for s in range(n_shape):
        f = 1.0 + 0.15 * s
        qext = 2.0 * f * X ** 4 / (1.0 + X ** 4) * (1.0 + 0.3 * (R - 1.3)) + 0.01 * I
        qabs = qext * (I / (I + 0.05)) * (0.5 + 0.1 * s)
        qsca = qext - qabs
        g = 0.7 * X ** 2 / (1.0 + X ** 2) + 0.02 * s
        qb = 0.1 * qsca * (1.0 + 0.2 * np.sin(X)) / f
        out["qext"].append(qext)
        out["qabs"].append(qabs)
        out["qsca"].append(qsca)
        out["qb"].append(qb)
        out["g"].append(g)
        gg = g[..., None]
        ct = np.cos(th)
        p11 = (1 - gg ** 2) / (1 + gg ** 2 - 2 * gg * ct) ** 1.5 + 0.05 * f
        p22 = p11 * (0.9 - 0.05 * s + 0.05 * ct)
        p33 = p11 * ct * 0.8
        p44 = p11 * ct * (0.75 + 0.02 * s)
        p12 = -p11 * 0.3 * np.sin(th) ** 2 / (1 + ct ** 2)
        p34 = p11 * 0.1 * np.sin(th) * (R[..., None] - 1.2)
        out["scama"].append(np.stack([p11, p22, p33, p44, p12, p34], axis=3))
    v = {k: np.array(a) for k, a in out.items()}
    v.update(mr=mr, mi=mi, x=x)
    return v


FRACS = np.array([0.6, 1.4])


def kernel_params(base):
    """A particle description in kernel mode built on a runoptics test case."""
    p = dict(base)
    p["mode"] = "kernel"
    return p
