"""Write tests/golden/grasp_vectors.npz from the reference's own kernel-mode
helpers.

bilinear_interpolation, get_interpolated and integrateShapes
(src/geosmie/dointegration.py:28-91) -- GRASP kernel mode's refractive-index
interpolation and shape averaging -- have no test in the reference, and the
kernel files they read are not shipped with it.  The module cannot be
imported here (netCDF4 and numba at import), so this script parses the
source file, compiles just those three functions, and runs them on synthetic
kernel arrays: a shape-resolved table [shape][m_r][m_i][x] of one scalar key
and [shape][m_r][m_i][x][element][angle] of phase-matrix kernels, averaged
over shapes with unnormalised fractions, then interpolated at refractive
indices inside the grid, on grid nodes, and outside it on every side (where
the reference clamps the bracketing indices and the index itself).  Inputs
and outputs are stored as data.
Run:  python tests/golden/make_grasp_vectors.py   (needs /root/reference)
"""
import ast
import bisect
import os
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/geosmie/dointegration.py"
WANT_FUNCS = ("bilinear_interpolation", "get_interpolated", "integrateShapes")


def load_reference_functions():
    tree = ast.parse(open(SRC).read(), SRC)
    body = [n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name in WANT_FUNCS]
    ns = {"np": np, "bisect": bisect}
    exec(compile(ast.Module(body=body, type_ignores=[]), SRC, "exec"), ns)
    return ns


def main():
    ns = load_reference_functions()
    rng = np.random.default_rng(20261203)
    n_shape, n_mr, n_mi, n_x, n_el, n_ang = 3, 5, 6, 6, 6, 7
    mr = np.linspace(1.30, 1.70, n_mr)
    mi = np.concatenate([[0.0], np.logspace(-4, -1, n_mi - 1)])     # ascending, as bisect needs
    sca = rng.uniform(0.1, 3.0, (n_shape, n_mr, n_mi, n_x))
    scama = rng.uniform(-1.0, 2.0, (n_shape, n_mr, n_mi, n_x, n_el, n_ang))
    fracs0 = np.array([0.2, 1.3, 0.5])
    data = types.SimpleNamespace(variables={"qext": sca, "scama": scama, "mr": mr, "mi": -mi})
    avg = ns["integrateShapes"](data, fracs0)
    # query points: interior, on nodes, on the last node, and beyond each edge
    q_mr = np.array([1.42, 1.30, 1.70, 1.55, 1.25, 1.90, 1.61, 1.333, 1.70, 1.49])
    q_mi = np.array([3e-3, 0.0, 0.1, 1e-4, 5e-2, 2e-3, 0.3, 2e-5, 1e-5, 0.01])
    out_s, out_p = [], []
    for a, b in zip(q_mr, q_mi):
        # kernel mode passes one-element arrays (getHumidRefractiveIndex's lists)
        out_s.append(ns["get_interpolated"](avg["qext"], np.array([a]), np.array([b]), mr, mi))
        out_p.append(ns["get_interpolated"](avg["scama"][:, :, :, 2, :], np.array([a]),
                                            np.array([b]), mr, mi))
    np.savez_compressed(os.path.join(HERE, "grasp_vectors.npz"), mr=mr, mi=mi, sca=sca,
                        scama=scama, fracs0=fracs0, avg_qext=avg["qext"], avg_scama=avg["scama"],
                        avg_mr=avg["mr"], q_mr=q_mr, q_mi=q_mi, interp_qext=np.array(out_s),
                        interp_p=np.array(out_p))
    print("wrote", os.path.join(HERE, "grasp_vectors.npz"))


if __name__ == "__main__":
    main()
