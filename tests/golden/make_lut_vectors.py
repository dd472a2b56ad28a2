"""Write tests/golden/lut_vectors.npz from the reference's own integratePSD.

integratePSD and calculateScatVals (src/geosmie/dointegration.py:1044-1050,
1064-1209) have no test in the reference and their quirks (the aliased
thisweight that Qsca-weights g, csca and cext; the qb back-out; 'dumix'
normalisation) are pinned only by the source.  The module itself cannot be
imported here (it needs netCDF4 and numba at import), so this script parses
the source file, compiles just those two functions plus the key lists they
read (scatkeys, scalarkeys), and runs them on synthetic inputs: per-size
values from the numpy Mie oracle on a CARMA-style size grid, two lognormal
modes with unequal fractions.  Inputs and outputs are stored as data.
Run:  python tests/golden/make_lut_vectors.py   (needs /root/reference)
"""
import ast
import os
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
from oracle import lut, mie as M  # noqa: E402

SRC = "/root/reference/src/geosmie/dointegration.py"
WANT_FUNCS = ("integratePSD", "calculateScatVals")
WANT_NAMES = ("scatkeys", "scalarkeys")


def load_reference_functions():
    tree = ast.parse(open(SRC).read(), SRC)
    body = []
    for node in tree.body:
        if isinstance(node, ast.FunctionDef) and node.name in WANT_FUNCS:
            body.append(node)
        elif isinstance(node, ast.Assign) and any(
                isinstance(t, ast.Name) and t.id in WANT_NAMES for t in node.targets):
            body.append(node)
    ns = {"np": np}
    exec(compile(ast.Module(body=body, type_ignores=[]), SRC, "exec"), ns)
    return ns


def main():
    ns = load_reference_functions()
    lam = 0.55
    ang = lut.angle_grid()
    cost = np.cos(np.radians(ang))
    xx, dr = lut.getXArrCarma(0.05, 40.0, 40)
    m = 1.50 + 0.01j
    res = M.size_range(xx, cost, m.real, m.imag)
    # rawMie's dict, keys in scatkeys + scalarkeys order (dointegration.py:1245)
    n_x, n_ang = res["s1"].shape
    sv = np.empty((6, n_x, n_ang))
    ns["calculateScatVals"](None, ns["scatkeys"], sv, n_x, res["s1"], res["s2"])
    rr = xx * lam / (2 * np.pi)
    raw = {}
    for i, k in enumerate(ns["scatkeys"]):
        raw[k] = sv[i]
    vals = {"qext": np.array(res["qext"]), "qsca": np.array(res["qsca"]),
            "qabs": np.array(res["qabs"]), "qb": np.array(res["qb"]),
            "g": np.array(res["asy"]),
            "csca": np.array(res["qsca"]) * np.pi * rr ** 2,
            "cext": np.array(res["qext"]) * np.pi * rr ** 2}
    for k in ns["scalarkeys"]:
        raw[k] = vals[k]
    psd = []
    reff0 = []
    for (rm, sg) in [(0.08, 1.7), (0.6, 2.1)]:
        p, re = lut.lognorm_psd(rm, sg, xx, dr, lam, 50.0, 0.001)
        psd.append(p)
        reff0.append(re)
    fracs = [0.7, 0.3]
    out = ns["integratePSD"](xx, [raw, raw], psd, fracs, lam, reff0, 1700.0, 1500.0)
    save = {"xx": xx, "dr": dr, "lam": lam, "ang": ang, "m": m, "psd": np.array(psd),
            "reff0": np.array(reff0), "fracs": np.array(fracs), "rhop0": 1700.0, "rhop": 1500.0,
            "scatvals": sv}
    for k in ns["scalarkeys"]:
        save["raw_" + k] = raw[k]
    for k, v in out.items():
        save["out_" + k] = np.asarray(v)
    np.savez_compressed(os.path.join(HERE, "lut_vectors.npz"), **save)
    print("keys:", sorted(out.keys()))


if __name__ == "__main__":
    main()
