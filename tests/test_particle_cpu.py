"""The runoptics front end (geosmie_amd.particle) against vectors written by
the reference's own functions (tests/golden/make_front_vectors.py): CARMA
sulfate growth, the refractive-index and water readers, humid refractive
index and the lognormal / sea-salt / dust PSDs of the reference's particle
files."""
import json
import os

import numpy as np
import pytest

from geosmie_amd import particle as P

V = np.load(os.path.join(os.path.dirname(__file__), "golden", "front_vectors.npz"))
REF = "/root/reference/src/geosmie"


def test_carma_growth():
    rhs, temps, rds = V["carma_rh"], V["carma_temp"], V["carma_rd"]
    w = np.array([[P.wtpct(r, temp=t) for r in rhs] for t in temps])
    d = np.array([[P.dens(r, temp=t) for r in rhs] for t in temps])
    g = np.array([[[P.grow_v75(r, rd, temp=t) for r in rhs[1:]] for rd in rds] for t in temps])
    np.testing.assert_array_equal(w, V["wtpct"])
    np.testing.assert_array_equal(d, V["dens"])
    np.testing.assert_allclose(g, V["grow_v75"], rtol=1e-15, atol=0)
    # growth is > 1 for humid air and grows with rh
    assert np.all(np.diff(g, axis=2) >= 0) and np.all(g > 1)


@pytest.mark.skipif(not os.path.isdir(REF), reason="reference data tables not present")
def test_readers():
    np.testing.assert_array_equal(P.getWaterM(data_dir=REF), V["water"])
    np.testing.assert_array_equal(P.getM(os.path.join(REF, "data/soot00")), V["gads_soot00"])
    np.testing.assert_array_equal(P.getMSep(os.path.join(REF, "data/ri-brc.wsv"), None),
                                  V["wsv_brc"])


def test_humid_index_and_psd():
    cases = json.loads(str(V["cases"]))
    assert len(cases) > 40
    kinds = set()
    for ci, c in enumerate(cases):
        params = json.loads(str(V["params_%s" % c["file"]]))
        kinds.add((params["psd"]["type"], params["rhDep"]["type"]))
        rh = list(params["rh"])
        if "maxrh" in params:
            rh = list(np.minimum(np.array(rh), params["maxrh"]))
        key = "%s_%d" % (c["file"], c["radind"])
        xx, dr = P.initializeXarr(params, c["radind"], *_lam_range(params, c))
        np.testing.assert_array_equal(xx, V["xx_" + key])
        if "dr_" + key in V.files:
            np.testing.assert_array_equal(dr, V["dr_" + key])
        else:
            assert dr is None
        nref0 = [complex(a, b) for a, b in c["nref0"]]
        nw = complex(*c["nw"])
        mr, mi, gf, rrat = P.getHumidRefractiveIndex(params, c["radind"], c["rhi"], rh, nref0, nw)
        assert (mr, mi, gf, rrat) == (c["mr"], c["mi"], c["gf"], c["rrat"]), ci
        psd, ref, rlo, rup = P.calculatePSD(params, c["radind"], rh[c["rhi"]], rh, xx, dr, rrat,
                                            c["lam"])
        np.testing.assert_array_equal(np.array(psd), V["psd_%d" % ci])
        assert (list(ref), rlo, rup) == (c["ref"], c["rlo"], c["rup"]), ci
    assert kinds >= {("lognorm", "simple"), ("ss", "ss"), ("lognorm", "su"), ("du", "trivial")}


def _lam_range(params, c):
    return params["_lam"] if "_lam" in params else LAMS[c["file"]]


LAMS = {}


@pytest.fixture(autouse=True, scope="module")
def _lams():
    # the wavelength range of each particle file's table (first and last rows)
    for k in V.files:
        if k.startswith("params_"):
            fn = k[len("params_"):]
            cs = [c for c in json.loads(str(V["cases"])) if c["file"] == fn]
            lam = sorted({c["lam"] for c in cs})
            LAMS[fn] = (lam[0], lam[-1])
