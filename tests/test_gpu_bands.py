"""Band averaging on the GPU (gm_band_average through geosmie_amd.bands):
bit-identical to the reference's doAverage vectors and to the oracle, per
spectrum and for whole tables in both layouts, and the runbands file."""
import os

import numpy as np
import pytest

from tests.runoptics_cases import params, water

pytestmark = pytest.mark.gpu
V = np.load(os.path.join(os.path.dirname(__file__), "golden", "band_vectors.npz"))
MODES = ["GEOS5", "RRTMG", "RRTMGP", "PURDUE"]


@pytest.mark.parametrize("mode", MODES)
def test_band_means_bit_exact(gpu, mode):
    from geosmie_amd import bands as B
    lo, up, _, use_wn, _ = B.getBands(mode)
    got = B.band_means(V["lam"], V["y"], lo, up, use_wn)
    np.testing.assert_array_equal(got, V["avg_%s" % mode])
    assert B.doAverage(V["lam"], V["y"][3], lo[4], up[4], use_wn) == V["avg_%s" % mode][3, 4]


def _reference_loop(tab, mode, classic, noIR):
    """runbands' loops (bandaverage.py:159-215) over the oracle's doAverage."""
    from geosmie_amd.bands import VARS, getBands
    from oracle.bands import do_average
    lo, up, mean, use_wn, nb = getBands(mode)
    d = tab.data
    nbin, nrh = tab.n_bin, tab.n_rh
    out = {v: np.zeros([nbin, nrh, len(lo)] if classic else [nbin, len(lo), nrh]) for v in VARS}
    for irh in range(nrh):
        for il in range(len(lo)):
            for ib in range(nbin):
                for v in VARS:
                    y = d[v][ib, irh, :] if classic else d[v][ib, :, irh]
                    r = do_average(tab.lam, y, lo[il], up[il], use_wn)
                    if classic:
                        out[v][ib, irh, il] = r
                    else:
                        out[v][ib, il, irh] = r
    if noIR and mode == "GEOS5":
        ind = np.where(lo > 3e-6)
        for v, val in (("qsca", 0), ("bsca", 0), ("qext", 1e-32), ("bext", 1e-32)):
            if classic:
                out[v][:, :, ind] = val
            else:
                out[v][:, ind, :] = val
    if mode == "GEOS5":
        bw1, bw2 = up[0] - lo[0], up[2] - lo[2]
        for v, val in out.items():
            if classic:
                val[:, :, 2] = (bw1 * val[:, :, 0] + bw2 * val[:, :, 2]) / (bw1 + bw2)
                out[v] = val[:, :, 1:]
            else:
                val[:, 2, :] = (bw1 * val[:, 0, :] + bw2 * val[:, 2, :]) / (bw1 + bw2)
                out[v] = val[:, 1:, :]
    return out


@pytest.mark.parametrize("classic", [False, True])
@pytest.mark.parametrize("mode,noIR", [("GEOS5", True), ("GEOS5", False), ("RRTMG", False),
                                       ("PURDUE", False)])
def test_average_table(gpu, classic, mode, noIR):
    from geosmie_amd.bands import VARS, average_table
    from geosmie_amd.runoptics import build_table
    tab = build_table(params("ss_maxrh"), water=water(), classic=classic)
    got, mean, _, _, nb = average_table({v: tab.data[v] for v in VARS}, tab.lam, mode, classic,
                                        noIR)
    ref = _reference_loop(tab, mode, classic, noIR)
    for v in VARS:
        np.testing.assert_array_equal(got[v], ref[v], err_msg=v)
        assert got[v].shape[2 if classic else 1] == nb
    assert len(mean) == nb


def test_runbands_file(gpu, tmp_path):
    from geosmie_amd import bands as B
    from geosmie_amd.runoptics import build_table
    from geosmie_amd.table import read
    tab = build_table(params("lognorm_simple"), water=water())
    src = str(tmp_path / "optics_T.nomom.nc4")
    tab.write(src)
    out = B.fun(src, "T", str(tmp_path / "bands.nc4"), "RRTMG")
    v = read(out)
    got, mean, _, _, nb = B.average_table({k: tab.data[k] for k in B.VARS}, tab.lam, "RRTMG")
    np.testing.assert_array_equal(v["qext"][0], got["qext"])
    assert v["qext"][1] == ("bin", "wavelength", "rh")
    np.testing.assert_array_equal(v["wavelength"][0], np.asarray(mean))
    # the reference's bandUp holds the lower edges (bandaverage.py:266-268)
    np.testing.assert_array_equal(v["bandUp"][0], v["bandLow"][0])
    assert v["bandUp"][2] == "micrometers"
    assert b"".join(v["qname"][0][0]).rstrip(b"\0") == b"T001"
    out3 = B.fun(src, "T", str(tmp_path / "bands3.nc"), "RRTMG", format="NETCDF3_64BIT_OFFSET",
                 band_up_quirk=False)
    v3 = read(out3)
    assert np.all(v3["bandUp"][0] > v3["bandLow"][0])
    np.testing.assert_array_equal(v3["bandLow"][0], v["bandLow"][0])
    for k in B.VARS + ["rEff", "rMass", "rLow", "rUp", "bin", "wavelength", "rh", "qname"]:
        np.testing.assert_array_equal(v3[k][0], v[k][0], err_msg=k)
        assert v3[k][1] == v[k][1], k
