"""Synthetic particle descriptions for the runoptics table tests (the
reference's particle-file structure, with refractive-index and water tables
made up here so no reference data file is needed)."""
import copy

import numpy as np

LAM = np.array([0.30e-6, 0.55e-6, 1.0e-6, 2.0e-6])


def water():
    lam = np.linspace(0.2e-6, 3.0e-6, 15)
    return np.array([lam, 1.34 - 0.01 * lam / 1e-6, 1e-9 + 1e-4 * (lam / 1e-6) ** 3])


def _mlist(mr, k):
    # tables store the imaginary part negative (OPAC sign)
    return np.array([LAM, np.full(LAM.size, mr) + 0.02 * np.arange(LAM.size),
                     -np.array(k, dtype=np.float64)])


def case(name):
    rh = [0.0, 0.5, 0.9, 0.99]
    if name == "lognorm_simple":
        return {"rhop0": 1800.0, "rh": rh,
                "rhDep": {"type": "simple", "params": {"gf": [1.0, 1.2, 1.6, 2.4]}},
                "psd": {"type": "lognorm", "params": {
                    "r0": [[0.0212e-6]], "rmax0": [[0.3e-6]], "rmin0": [[0.005e-6]],
                    "sigma": [[2.2]], "numperdec": [20], "fracs": [[1.0]]}},
                "mList": [_mlist(1.53, [0.006, 0.005, 0.02, 0.05])]}
    if name == "lognorm_su":
        c = case("lognorm_simple")
        c["rhDep"] = {"type": "su", "params": {"temp": 230.0}}
        c["rhop0"] = 1700.0
        return c
    if name == "ss_maxrh":
        return {"rhop0": 2200.0, "rh": [0.0, 0.5, 0.95, 0.99], "maxrh": 0.95,
                "rhDep": {"type": "ss", "params": {"c1": 0.7674, "c2": 3.079, "c3": 2.573e-11,
                                                   "c4": -1.424}},
                "psd": {"type": "ss", "params": {
                    "rMinMaj": [0.03e-6, 0.1e-6], "rMaxMaj": [0.1e-6, 0.5e-6],
                    "fracs": [[1.0], [1.0]], "numperdec": [12, 12]}},
                "mList": [_mlist(1.50, [1e-8, 1e-8, 1e-5, 2e-3])]}
    if name == "bimodal_trivial":
        return {"rhop0": [2500.0, 2650.0], "rh": rh,
                "rhDep": {"type": "trivial", "params": {"gf": [1.0]}},
                "psd": {"type": "lognorm", "params": {
                    "r0": [[0.05e-6, 0.2e-6], [0.4e-6, 0.8e-6]],
                    "rmax0": [[0.3e-6, 0.6e-6], [1.0e-6, 1.5e-6]],
                    "rmin0": [[0.01e-6, 0.02e-6], [0.05e-6, 0.1e-6]],
                    "sigma": [[1.8, 2.0], [1.6, 1.7]], "numperdec": [15, 10],
                    "fracs": [[0.3, 0.7], [0.6, 0.4]]}},
                "mList": [_mlist(1.53, [0.003, 0.002, 0.004, 0.01]),
                          _mlist(1.45, [0.0, 1e-6, 1e-4, 1e-3])]}
    if name == "du_trivial":
        # dust: r^-4 sub-bins on the reference's non-monotonic du grid, which
        # for the first bin passes through x ~ 1e-30 (log10 x changes sign)
        return {"rhop0": [2500.0, 2650.0], "rh": rh,
                "rhDep": {"type": "trivial", "params": {"gf": [1.0]}},
                "psd": {"type": "du", "params": {
                    "rMinMaj": [[0.1e-6, 0.3e-6], [1.0e-6]],
                    "rMaxMaj": [[0.3e-6, 1.0e-6], [2.0e-6]],
                    "fracs": [[0.3, 0.7], [1.0]]}},
                "mList": [_mlist(1.53, [0.004, 0.003, 0.002, 0.004])]}
    raise KeyError(name)


NAMES = ["lognorm_simple", "lognorm_su", "ss_maxrh", "bimodal_trivial", "du_trivial"]


def params(name):
    return copy.deepcopy(case(name))
