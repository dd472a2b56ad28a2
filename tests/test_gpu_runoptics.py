"""runoptics tables on the GPU (geosmie_amd.runoptics) against the oracle's
CPU restatement of the same loop (oracle/runoptics.py): every variable of
every (bin, wavelength, rh) entry, both layouts, and the file written by
fun() from a particle file on disk."""
import json
import os

import numpy as np
import pytest

from tests.runoptics_cases import LAM, NAMES, params, water

pytestmark = pytest.mark.gpu
SCAL = ["qext", "qsca", "qabs", "g", "ssa", "qb", "bsca", "bext", "csca", "cext", "bbck",
        "lidar_ratio", "refreal", "refimag"]
NL = ["mass", "volume", "area", "rEff", "rMass", "rUp", "rLow", "rhop", "growth_factor"]
P = ["p11", "p12", "p22", "p33", "p34", "p44"]


@pytest.fixture(scope="module")
def oracle_tables():
    from oracle import runoptics as ORO
    return {n: ORO.table(params(n), water()) for n in NAMES}


@pytest.mark.parametrize("name", NAMES)
def test_table_vs_oracle(gpu, oracle_tables, name):
    from geosmie_amd.runoptics import build_table
    print("runoptics case", name, flush=True)
    tab = build_table(params(name), water=water())
    ent, nl = oracle_tables[name]
    d = tab.data
    for (b, li, rhi), e in ent.items():
        scale = np.abs(e["p11"]).max()
        for k in P:
            err = np.abs(d[k][b, li, rhi] - e[k]).max()
            assert err <= 1e-9 * scale, (name, b, li, rhi, k, err)
        np.testing.assert_allclose(d["pback"][b, li, rhi], e["pback"], rtol=1e-9,
                                   atol=1e-9 * scale)
        for k in SCAL:
            # qabs = qext - qsca cancels for weak absorbers: its error is relative to qext
            ref = abs(e["qext"]) if k == "qabs" else abs(e[k])
            assert abs(d[k][b, li, rhi] - e[k]) <= 1e-9 * ref + 1e-300, (name, k)
    for (b, rhi), v in nl.items():
        for k in NL:
            assert abs(d[k][b, rhi] - v[k]) <= 1e-9 * abs(v[k]), (name, b, rhi, k)


def test_classic_layout_and_file(gpu, tmp_path):
    from geosmie_amd.runoptics import build_table
    from geosmie_amd.table import read
    new = build_table(params("bimodal_trivial"))
    old = build_table(params("bimodal_trivial"), classic=True)
    np.testing.assert_array_equal(old.data["p34"], new.data["p34"].transpose(0, 2, 1, 3))
    np.testing.assert_array_equal(old.data["qext"], new.data["qext"].transpose(0, 2, 1))
    np.testing.assert_array_equal(old.data["pback"], new.data["pback"].transpose(3, 0, 2, 1))
    path = str(tmp_path / "t.nc4")
    new.write(path)
    v = read(path)
    np.testing.assert_array_equal(v["g"][0], new.data["g"])
    assert v["p11"][1] == ("bin", "wavelength", "rh", "ang")


def test_fun_from_particle_file(gpu, tmp_path):
    """runoptics' entry point: particle JSON + wsv refractive-index table +
    water table on disk -> optics_<name>.nomom.nc4 equal to build_table."""
    from geosmie_amd.runoptics import build_table, fun
    from geosmie_amd.table import read
    p = params("lognorm_simple")
    m = p.pop("mList")[0]
    os.makedirs(tmp_path / "data")
    np.savetxt(tmp_path / "data" / "ri-test.wsv", np.array([m[0] * 1e6, m[1], m[2]]).T,
               header="lambda[um] m_real m_imaginary")
    w = water()
    with open(tmp_path / "data" / "refrac.water.txt", "w") as f:
        f.write("\n" * 13)
        for row in np.array([w[0] * 1e6, w[1], w[2]]).T:
            f.write("  %.6f  %.6f  %.6e\n" % tuple(row))
    p["ri"] = {"format": "wsv", "path": ["data/ri-test.wsv"]}
    with open(tmp_path / "tst.json", "w") as f:
        json.dump(p, f)
    out = fun(str(tmp_path / "tst.json"), oppfx=str(tmp_path), data_dir=str(tmp_path))
    assert os.path.basename(out) == "optics_tst.nomom.nc4"
    v = read(out)
    from geosmie_amd import particle
    pj = particle.getPPJSON(str(tmp_path / "tst.json"), str(tmp_path))
    ref = build_table(pj, water=particle.getWaterM(data_dir=str(tmp_path)))
    for k in ("p11", "qext", "bbck", "rhop", "pback"):
        np.testing.assert_array_equal(v[k][0], ref.data[k])
    np.testing.assert_allclose(v["wavelength"][0], LAM, rtol=1e-12)


@pytest.mark.parametrize("classic", [False, True])
def test_table_with_gsf_moments(gpu, tmp_path, classic):
    """n_mom adds pmom (the rungsf step): every entry's moments are the GSF
    expansion of that entry's normalised phase matrix (oracle/gsf.py, which
    restates spher_expan.f), in the table's p order, in both layouts."""
    from geosmie_amd.runoptics import build_table
    from geosmie_amd.table import read
    from oracle import gsf as OG
    tab = build_table(params("ss_maxrh"), water=water(), classic=classic, n_mom=129)
    d = tab.data
    for (b, li, rhi) in [(0, 0, 0), (1, 3, 2), (0, 2, 3)]:
        idx = (b, rhi, li) if classic else (b, li, rhi)
        ref = OG.pmom_from_table_slab(tab.ang, d["p11"][idx], d["p12"][idx], d["p33"][idx],
                                      d["p34"][idx], d["p22"][idx], d["p44"][idx])
        got = d["pmom"][:, :, b, rhi, li] if classic else d["pmom"][b, li, rhi]
        assert np.abs(got - ref).max() <= 1.01e-10, (b, li, rhi)
    path = str(tmp_path / "optics_T.nc4")
    tab.write(path)
    v = read(path)
    assert v["pmom"][1] == (("nPol", "nMom", "radius", "rh", "lambda") if classic
                            else ("bin", "wavelength", "rh", "p", "m"))
    np.testing.assert_array_equal(v["pmom"][0], d["pmom"])


def test_lognorm_rows_match_calculate_psd(gpu):
    """gm_lognorm_rows (the table loop's device PSDs) against the front end's
    calculatePSD for every rh and mode of a bimodal lognormal bin with CARMA
    growth (grown mode radius and upper cut-off), written at block offsets of
    a wider weight matrix whose other entries stay untouched."""
    from geosmie_amd import _lib, particle
    from geosmie_amd._lib import check, ptr, stream
    t = _lib.torch()
    p = params("bimodal_trivial")
    p["rhDep"] = {"type": "su", "params": {"temp": 230.0}}
    rh = list(p["rh"])
    pp = p["psd"]["params"]
    lam = 0.55e-6
    kx = 2 * np.pi / lam
    xx, dr = particle.initializeXarr(p, 0, LAM[0], LAM[-1])
    spec, want = [], []
    for rhi, r in enumerate(rh):
        psd, _, _, _ = particle.calculatePSD(p, 0, r, rh, xx, dr, 1.0, lam, want_ref=False)
        for k, r0 in enumerate(pp["r0"][0]):
            spec.append((particle.humidityGrowth(p["rhDep"], r0, r, rh) * kx,
                         np.log(pp["sigma"][0][k]), pp["rmin0"][0][k] * kx,
                         particle.humidityGrowth(p["rhDep"], pp["rmax0"][0][k], r, rh) * kx))
            want.append(psd[k])
    nr, nx = len(spec), xx.size
    ldw = nx * 3 + 5
    col0 = np.array([(i % 3) * nx + (5 if i % 2 else 0) for i in range(nr)], np.int64)
    d_w = t.full((nr, ldw), -7.0, dtype=t.float64, device=_lib.device())
    # keep every input tensor referenced until the kernel has run
    d_x, d_dr = _lib.to_dev(xx), _lib.to_dev(dr)
    d_par, d_col = _lib.to_dev(np.array(spec)), _lib.to_dev(col0, dtype=np.int64)
    check(_lib.lib().gm_lognorm_rows(_lib.ctx(), nr, nx, ptr(d_x), ptr(d_dr), ptr(d_par),
                                     ptr(d_col), ptr(d_w), ldw, stream()), "gm_lognorm_rows")
    W = d_w.cpu().numpy()
    for i in range(nr):
        got = W[i, col0[i]:col0[i] + nx]
        np.testing.assert_allclose(got, want[i], rtol=1e-13, atol=1e-16 * want[i].max())
        rest = np.delete(W[i], np.arange(col0[i], col0[i] + nx))
        assert np.all(rest == -7.0)


def test_table_reduced_precision_path(gpu):
    """build_table at FP32 S1/S2 (Mueller planes + row contraction instead of
    the fused FP64 pass): P within FP32 rounding of the FP64 table, the
    efficiencies (FP64 series) unchanged."""
    from geosmie_amd.runoptics import build_table
    p64 = build_table(params("lognorm_simple"), water=water()).data
    p32 = build_table(params("lognorm_simple"), water=water(), precision=1).data
    for k in ("p11", "p12", "p33", "p34"):
        scale = np.abs(p64["p11"]).max()
        assert np.abs(p32[k] - p64[k]).max() <= 1e-5 * scale, k
    for k in ("qext", "qsca", "g", "bext"):
        np.testing.assert_allclose(p32[k], p64[k], rtol=1e-12)
