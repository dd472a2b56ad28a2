"""The optics table writer (runoptics layout, dointegration.py:205-422,
1006-1031; pmom convertncdf.py:70-79) on the CPU: dimensions, orders,
metadata and the placement of every written entry, in both layouts."""
import numpy as np
import pytest

from geosmie_amd.table import NL_SCALAR_KEYS, P_ORDER, SCALAR_KEYS, SCAT_KEYS, OpticsTable, read


def _ret(seed, n_ang):
    r = np.random.default_rng(seed)
    ret = {k: r.uniform(size=n_ang) for k in SCAT_KEYS}
    ret.update({k: float(r.uniform()) for k in SCALAR_KEYS + NL_SCALAR_KEYS})
    ret["pback"] = r.uniform(size=6)
    return ret


@pytest.mark.parametrize("fmt", ["NETCDF4", "NETCDF3_64BIT_OFFSET"])
@pytest.mark.parametrize("classic", [False, True])
def test_layout_and_round_trip(tmp_path, classic, fmt):
    rh = np.array([0.0, 0.5, 0.9])
    lam = np.array([0.44e-6, 0.55e-6])
    ang = np.linspace(0, 180, 7)
    t = OpticsTable(4, rh, lam, ang, classic=classic, n_mom=5)
    placed = {}
    for b in range(4):
        for li in range(2):
            for ri in range(3):
                ret = _ret(100 * b + 10 * li + ri, ang.size)
                pm = np.arange(24.0).reshape(6, 4) + b + li + ri
                t.set_entry(b, li, ri, ret, pmom=pm)
                if li == 0:
                    t.set_bin_rh(b, ri, ret)
                placed[b, li, ri] = (ret, pm)
    path = str(tmp_path / "optics_TEST.nomom.nc")
    t.write(path, format=fmt)
    v = read(path)
    r, lm, p = ("radius", "lambda", "nPol") if classic else ("bin", "wavelength", "p")
    assert list(v[r][0]) == [1, 2, 3, 4] and list(v[p][0]) == P_ORDER
    assert v[lm][2] == "m" and np.allclose(v[lm][0], lam)
    assert v["p11"][1] == ((r, "rh", lm, "ang") if classic else (r, lm, "rh", "ang"))
    assert v["qext"][1] == ((r, "rh", lm) if classic else (r, lm, "rh"))
    assert v["pback"][1] == ((p, r, "rh", lm) if classic else (r, lm, "rh", p))
    assert v["rEff"][1] == (r, "rh")
    assert v["pmom"][1] == (("nPol", "nMom", r, "rh", lm) if classic else (r, lm, "rh", p, "m"))
    assert v["bbck"][2] == "m2 (kg dry mass)-1 sr-1"
    for (b, li, ri), (ret, pm) in placed.items():
        idx = (b, ri, li) if classic else (b, li, ri)
        assert np.array_equal(v["p34"][0][idx], ret["p34"])
        assert v["g"][0][idx] == ret["g"]
        pb = v["pback"][0][:, b, ri, li] if classic else v["pback"][0][b, li, ri]
        assert np.array_equal(pb, ret["pback"])
        pmw = v["pmom"][0][:, :, b, ri, li] if classic else v["pmom"][0][b, li, ri]
        assert np.array_equal(pmw[:, :4], pm) and np.all(pmw[:, 4] == 0)
        if li == 0:
            assert v["rhop"][0][b, ri] == ret["rhop"]


@pytest.mark.parametrize("classic", [False, True])
def test_set_entries_equals_set_entry(classic):
    """The all-rh writer of the table loop puts every value where the per-entry
    writer does, in both layouts, pmom zero-padded."""
    rng = np.random.default_rng(5)
    rh = [0.0, 0.3, 0.6, 0.9, 0.99]
    lam = np.array([0.4e-6, 0.8e-6])
    ang = np.linspace(0, 180, 7)
    a = OpticsTable(2, rh, lam, ang, classic=classic, n_mom=9)
    b = OpticsTable(2, rh, lam, ang, classic=classic, n_mom=9)
    rhis = [4, 1, 2]
    ret = {k: rng.standard_normal((3, ang.size)) for k in ("p11", "p12", "p22", "p33", "p34",
                                                            "p44")}
    for k in ("qsca", "qabs", "qext", "g", "ssa", "qb", "bsca", "bext", "csca", "cext", "bbck",
              "lidar_ratio", "refreal", "refimag"):
        ret[k] = rng.standard_normal(3)
    ret["pback"] = rng.standard_normal((3, 6))
    pm = rng.standard_normal((3, 6, 7))
    nl = {k: rng.standard_normal(3) for k in ("mass", "volume", "area", "rEff", "rMass", "rUp",
                                              "rLow", "rhop", "growth_factor")}
    a.set_entries(1, 1, rhis, ret, pmom=pm)
    a.set_bins_rh(1, rhis, nl)
    for e, rhi in enumerate(rhis):
        b.set_entry(1, 1, rhi, {k: v[e] for k, v in ret.items()}, pmom=pm[e])
        b.set_bin_rh(1, rhi, {k: v[e] for k, v in nl.items()})
    for k in a.data:
        np.testing.assert_array_equal(a.data[k], b.data[k], err_msg=k)
