"""GRASP kernel mode on the GPU (geosmie_amd.grasp) against the reference's own
helpers' vectors (tests/golden/grasp_vectors.npz) and the oracle's restatement
of the kernel branch of fun (oracle/grasp.py): shape averaging, bilinear
interpolation in refractive index (bit for bit), six-element normalisation,
and whole kernel-mode tables (every variable of every entry), read from
NetCDF-3 and .npz kernel files too."""
import os

import numpy as np
import pytest

from oracle import grasp as OG
from tests.grasp_cases import FRACS, kernel_params, kernel_variables
from tests.runoptics_cases import params, water

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))
GV = np.load(os.path.join(HERE, "golden", "grasp_vectors.npz"))


def test_shape_average_vs_reference_vectors(gpu):
    import torch
    from geosmie_amd import _lib
    w = GV["fracs0"] / np.sum(GV["fracs0"])
    for src, want in (("sca", "avg_qext"), ("scama", "avg_scama")):
        d_in = _lib.to_dev(GV[src])
        out = torch.empty(GV[want].size, dtype=torch.float64, device="cuda")
        _lib.check(_lib.lib().gm_shape_average(_lib.ctx(), GV[src].shape[0], out.numel(),
                                               _lib.ptr(_lib.to_dev(w)), _lib.ptr(d_in),
                                               _lib.ptr(out), _lib.stream()), "gm_shape_average")
        got = out.cpu().numpy().reshape(GV[want].shape)
        # sums of 3 terms of either sign: within 2 ulp of the largest term
        np.testing.assert_allclose(got, GV[want], rtol=0,
                                   atol=4.5e-16 * np.abs(GV[src]).max())


def test_ri_bilinear_bitwise_vs_reference_vectors(gpu):
    """get_interpolated at interior points, grid nodes and beyond every edge:
    the device blend equals the reference's numpy result bit for bit."""
    from geosmie_amd.grasp import KernelTables
    var = {"mr": GV["mr"], "mi": -GV["mi"], "x": np.arange(GV["avg_qext"].shape[2], dtype=float)}
    # one shape with weight 1 reproduces the averaged tables exactly
    for k in ("qext", "qsca", "qabs", "qb", "g"):
        var[k] = GV["avg_qext"][None]
    var["scama"] = GV["avg_scama"][None]
    kt = KernelTables(var, [1.0])
    props, scama = kt.interpolate(GV["q_mr"], GV["q_mi"])
    nq, nx = GV["q_mr"].size, kt.n_x
    got = props.cpu().numpy().reshape(6, nq, nx)[0]
    np.testing.assert_array_equal(got, GV["interp_qext"])
    sp = scama.cpu().numpy().reshape(nq, nx, 6, kt.n_ang)[:, :, 2, :]
    np.testing.assert_array_equal(sp, GV["interp_p"])


def test_normalize6_vs_oracle(gpu):
    import torch
    from geosmie_amd import _lib
    from oracle import lut as OL
    ang = OL.angle_grid()
    rng = np.random.default_rng(4)
    p = rng.uniform(0.1, 2.0, (3, 6, ang.size))
    d_p = _lib.to_dev(p)
    d_b = torch.empty(18, dtype=torch.float64, device="cuda")
    _lib.check(_lib.lib().gm_pmatrix_normalize6(_lib.ctx(), 3, ang.size,
                                                _lib.ptr(_lib.to_dev(np.radians(ang))),
                                                _lib.ptr(d_p), _lib.ptr(d_b), _lib.stream()),
               "gm_pmatrix_normalize6")
    got = d_p.cpu().numpy().reshape(p.shape)
    back = d_b.cpu().numpy().reshape(3, 6)
    th = np.radians(ang)
    for e in range(3):
        p11n = 2. * p[e, 0] / np.trapezoid(p[e, 0] * np.sin(th), th)
        np.testing.assert_allclose(got[e, 0], p11n, rtol=1e-13)
        for el in range(1, 6):
            np.testing.assert_allclose(got[e, el], p[e, el] * p11n / p[e, 0], rtol=1e-13)
        np.testing.assert_array_equal(back[e], got[e, [0, 1, 3, 4, 2, 5], -1])


KCASES = ["lognorm_simple", "du_trivial"]


@pytest.fixture(scope="module")
def kernel_oracle():
    v = kernel_variables()
    return v, {n: OG.table(kernel_params(params(n)), v, FRACS, water()) for n in KCASES}


SCAL = ["qext", "qsca", "qabs", "g", "ssa", "qb", "bsca", "bext", "csca", "cext", "bbck",
        "lidar_ratio", "refreal", "refimag"]
NL = ["mass", "volume", "area", "rEff", "rMass", "rUp", "rLow", "rhop", "growth_factor"]
P = ["p11", "p12", "p22", "p33", "p34", "p44"]


def _compare(tab, ent, nl, name):
    d = tab.data
    for (b, li, rhi), e in ent.items():
        scale = np.abs(e["p11"]).max()
        for k in P:
            err = np.abs(d[k][b, li, rhi] - e[k]).max()
            assert err <= 1e-11 * scale, (name, b, li, rhi, k, err)
        np.testing.assert_allclose(d["pback"][b, li, rhi], e["pback"], rtol=1e-11,
                                   atol=1e-11 * scale)
        for k in SCAL:
            ref = abs(e["qext"]) if k == "qabs" else abs(e[k])
            assert abs(d[k][b, li, rhi] - e[k]) <= 1e-12 * ref + 1e-300, (name, k, b, li, rhi)
    for (b, rhi), v in nl.items():
        for k in NL:
            assert abs(d[k][b, rhi] - v[k]) <= 1e-12 * abs(v[k]), (name, b, rhi, k)


@pytest.mark.parametrize("name", KCASES)
def test_kernel_table_vs_oracle(gpu, kernel_oracle, name):
    from geosmie_amd.grasp import KernelTables
    from geosmie_amd.runoptics import build_table
    v, oracle = kernel_oracle
    print("kernel-mode case", name, flush=True)
    tab = build_table(kernel_params(params(name)), water=water(),
                      kernels=KernelTables(v, FRACS))
    ent, nl = oracle[name]
    _compare(tab, ent, nl, name)
    # p22, p44 are the kernels' own elements, not copies of p11, p33
    assert not np.allclose(tab.data["p22"], tab.data["p11"])


def test_kernel_files_npz_and_netcdf(gpu, kernel_oracle, tmp_path):
    """KernelTables.read on .npz, NetCDF-3 and NetCDF-4 kernel files with a
    shape distribution file, through params['kernel_params'] as fun() sees
    them."""
    from geosmie_amd import nc4
    from scipy.io import netcdf_file
    from geosmie_amd.runoptics import build_table
    v, oracle = kernel_oracle
    dist = tmp_path / "shapes.txt"
    np.savetxt(dist, np.stack([FRACS, np.arange(FRACS.size)], axis=1))
    npz = tmp_path / "k.npz"
    np.savez(npz, **v)
    nc = tmp_path / "k.nc"
    with netcdf_file(str(nc), "w") as f:
        dims = {"shape": v["qext"].shape[0], "mr": v["mr"].size, "mi": v["mi"].size,
                "x": v["x"].size, "elem": 6, "ang": v["scama"].shape[-1]}
        for k, n in dims.items():
            f.createDimension(k, n)
        for k in ("mr", "mi", "x"):
            f.createVariable(k, "d", (k,))[:] = v[k]
        for k in ("qext", "qabs", "qsca", "qb", "g"):
            f.createVariable(k, "d", ("shape", "mr", "mi", "x"))[:] = v[k]
        f.createVariable("scama", "d", ("shape", "mr", "mi", "x", "elem", "ang"))[:] = v["scama"]
    nc4f = tmp_path / "k.nc4"
    with nc4.Writer(str(nc4f)) as w:
        for k, n in dims.items():
            w.dim(k, n)
        for k in ("mr", "mi", "x"):
            w.var(k, "f8", (k,))
        for k in ("qext", "qabs", "qsca", "qb", "g"):
            w.var(k, "f8", ("shape", "mr", "mi", "x"))
        w.var("scama", "f8", ("shape", "mr", "mi", "x", "elem", "ang"))
        for k in ("mr", "mi", "x", "qext", "qabs", "qsca", "qb", "g", "scama"):
            w.put(k, v[k])
    ent, nl = oracle["lognorm_simple"]
    for path in (npz, nc, nc4f):
        p = kernel_params(params("lognorm_simple"))
        p["kernel_params"] = {"path": str(path), "shape_dist": str(dist)}
        tab = build_table(p, water=water())
        _compare(tab, ent, nl, str(path))


def test_kernel_mode_errors(gpu):
    from geosmie_amd.runoptics import build_table
    p = kernel_params(params("bimodal_trivial"))       # two refractive-index components
    with pytest.raises(ValueError):
        build_table(p, water=water())
