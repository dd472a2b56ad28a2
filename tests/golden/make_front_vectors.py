"""Write tests/golden/front_vectors.npz from the reference's own front-end code.

The readers, humidity rules and PSD builders runoptics runs before the Mie
batch (src/geosmie/particleparams.py:18-127, src/geosmie/carma_utils.py:
136-314, src/geosmie/dointegration.py:93-153, 425-447, 493-622) have no test
in the reference.  Their modules cannot be imported here (dointegration needs
netCDF4 and numba at import, carma_utils matplotlib), so this script parses
the source files, compiles just those functions and runs them: on the
reference's particle files (with numperdec lowered to keep the vectors
small), its refractive-index and water tables, and a synthetic 'su' humidity
rule.  Inputs and outputs are stored as data.
Run:  python tests/golden/make_front_vectors.py   (needs /root/reference)
"""
import ast
import json
import os
import types

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = "/root/reference/src/geosmie"
CFG = "/root/reference/src/config/geosparticles"


def extract(path, names, ns):
    tree = ast.parse(open(path).read(), path)
    body = [n for n in tree.body if isinstance(n, ast.FunctionDef) and n.name in names]
    exec(compile(ast.Module(body=body, type_ignores=[]), path, "exec"), ns)
    return ns


def main():
    carma = extract(os.path.join(SRC, "carma_utils.py"), ("wtpct", "dens", "grow_v75"),
                    {"np": np})
    pp = extract(os.path.join(SRC, "particleparams.py"),
                 ("getM", "getMSep", "getWaterM", "humidityGrowth", "getLogNormPSD"),
                 {"np": np, "carma_utils": types.SimpleNamespace(**carma)})
    di = extract(os.path.join(SRC, "dointegration.py"),
                 ("getDR", "getXArrCarma", "initializeXarr", "getHumidRefractiveIndex",
                  "calculatePSD"), {"np": np, "pp": types.SimpleNamespace(**pp)})
    out = {}
    # CARMA sulfate growth
    rhs = np.array([0.0, 1e-7, 0.01, 0.049, 0.05, 0.3, 0.6, 0.85, 0.86, 0.95, 0.99, 1.0])
    temps = np.array([190., 220., 250.])
    rds = np.array([1e-9, 5e-8, 1e-6])
    out["carma_rh"], out["carma_temp"], out["carma_rd"] = rhs, temps, rds
    out["wtpct"] = np.array([[carma["wtpct"](r, temp=t) for r in rhs] for t in temps])
    out["dens"] = np.array([[carma["dens"](r, temp=t) for r in rhs] for t in temps])
    out["grow_v75"] = np.array([[[carma["grow_v75"](r, d, temp=t) for r in rhs[1:]]
                                 for d in rds] for t in temps])
    # readers
    cwd = os.getcwd()
    os.chdir(SRC)
    try:
        out["water"] = pp["getWaterM"]()
        out["gads_soot00"] = pp["getM"]("data/soot00")
        out["wsv_brc"] = pp["getMSep"]("data/ri-brc.wsv", None)
    finally:
        os.chdir(cwd)
    water = out["water"]
    cases = []
    files = ["oc.json", "su.json", "bc.json", "ss.json", "ni.json", "du-mie.json"]
    for fn in files + ["su.json#su"]:
        name, _, variant = fn.partition("#")
        params = json.load(open(os.path.join(CFG, name)))
        if variant == "su":
            params["rhDep"] = {"type": "su", "params": {"temp": 230.0}}
        if "numperdec" in params["psd"]["params"]:
            params["psd"]["params"]["numperdec"] = [40] * len(params["psd"]["params"]["numperdec"])
        ri = params["ri"]
        os.chdir(SRC)
        try:
            if ri["format"] == "gads":
                mlist = [pp["getM"](p) for p in ri["path"]]
            else:
                mlist = [pp["getMSep"](p, None if ri["format"] == "wsv" else ",")
                         for p in ri["path"]]
        finally:
            os.chdir(cwd)
        lams = mlist[0][0]
        rh = list(params["rh"])
        if "maxrh" in params:
            rh = list(np.minimum(np.array(rh), params["maxrh"]))
        n_bin = len(params["psd"]["params"]["fracs"])
        for radind in range(min(n_bin, 2)):
            xx, dr = di["initializeXarr"](params, radind, lams[0], lams[-1])
            for li in (0, len(lams) // 2, len(lams) - 1):
                lam = lams[li]
                nref0 = [complex(np.interp(lam, m[0], m[1]), -np.interp(lam, m[0], m[2]))
                         for m in mlist]
                nw = complex(np.interp(lam, water[0], water[1]), np.interp(lam, water[0],
                                                                           water[2]))
                for rhi in (0, 12, len(rh) - 1):
                    if params["rhDep"]["type"] == "trivial" and rhi > 0:
                        continue
                    mr, mi, gf, rrat = di["getHumidRefractiveIndex"](params, radind, rhi, rh,
                                                                     nref0, nw)
                    psd, ref, rlo, rup = di["calculatePSD"](params, radind, rh[rhi], rh, xx, dr,
                                                           rrat, lam)
                    cases.append(dict(file=fn, radind=radind, lam=lam, rhi=rhi,
                                      nref0=[[c.real, c.imag] for c in nref0],
                                      nw=[nw.real, nw.imag], mr=list(mr), mi=list(mi),
                                      gf=float(gf), rrat=float(rrat), ref=list(ref),
                                      rlo=float(rlo), rup=float(rup)))
                    key = "psd_%d" % (len(cases) - 1)
                    out[key] = np.array(psd)
            out["xx_%s_%d" % (fn, radind)] = xx
            if dr is not None:      # the du grid has no widths
                out["dr_%s_%d" % (fn, radind)] = dr
        cases_params = {fn: params}
        out["params_%s" % fn] = np.array(json.dumps(cases_params[fn]))
    out["cases"] = np.array(json.dumps(cases))
    np.savez_compressed(os.path.join(HERE, "front_vectors.npz"), **out)
    print("cases", len(cases))


if __name__ == "__main__":
    main()
