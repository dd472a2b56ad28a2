"""Write tests/golden/band_vectors.npz from the reference's own band averaging.

bandaverage.doAverage and getBands (src/geosmie/bandaverage.py:18-50,
73-124) have no test in the reference; the module cannot be imported here
(it needs netCDF4 at import), so this script compiles just those two
functions from the source file and runs them on synthetic spectra over a
61-wavelength 0.25-40 um grid for every band set (GEOS5, RRTMG, RRTMGP,
PURDUE).  Inputs and outputs are stored as data.
Run:  python tests/golden/make_band_vectors.py   (needs /root/reference)
"""
import ast
import os

import numpy as np
from scipy.interpolate import interp1d

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/geosmie/bandaverage.py"


def main():
    tree = ast.parse(open(SRC).read(), SRC)
    body = [n for n in tree.body if isinstance(n, ast.FunctionDef)
            and n.name in ("doAverage", "getBands")]
    ns = {"np": np, "interp1d": interp1d}
    exec(compile(ast.Module(body=body, type_ignores=[]), SRC, "exec"), ns)
    lam = np.concatenate([np.linspace(0.25, 1.0, 16), np.linspace(1.25, 10.0, 30),
                          np.linspace(11.0, 40.0, 15)]) * 1e-6
    r = np.random.default_rng(5)
    n_series = 12
    y = np.exp(r.uniform(-3, 3, (n_series, lam.size))) * \
        np.sin(np.linspace(0, 7, lam.size))[None, :] ** 2 + r.uniform(0, 1, (n_series, 1))
    out = {"lam": lam, "y": y}
    for mode in ("GEOS5", "RRTMG", "RRTMGP", "PURDUE"):
        lo, up, mean, use_wn, nb = ns["getBands"](mode)
        out["bands_%s" % mode] = np.array([np.asarray(lo, np.float64), np.asarray(up, np.float64)])
        out["mean_%s" % mode] = np.asarray(mean, np.float64)
        out["meta_%s" % mode] = np.array([int(use_wn), nb])
        out["avg_%s" % mode] = np.array([[ns["doAverage"](lam, y[s], lo[b], up[b], use_wn, None)
                                          for b in range(len(lo))] for s in range(n_series)])
    np.savez_compressed(os.path.join(HERE, "band_vectors.npz"), **out)


if __name__ == "__main__":
    main()
