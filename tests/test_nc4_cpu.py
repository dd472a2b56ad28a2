"""NetCDF-4 table files (libgeosmie_nc4.so) on the CPU.

The reference writes its tables with netCDF4-python in the NetCDF-4 format
with compression='zlib' (dointegration.py:205-354, 418; convertncdf.py:70-79;
bandaverage.py:222-290).  netCDF4-python and the netCDF C library are not in
this image, so the files are checked against netCDF's HDF5 conventions with
an independent HDF5 reader: h5py in the image's conda interpreter, run as a
separate process, reports every dataset's type, shape, filters, dimension
scales and attributes.  Our own reader is checked on files h5py writes.
"""
import json
import os
import subprocess

import numpy as np
import pytest

from geosmie_amd import nc4
from geosmie_amd.table import META, P_ORDER, SCAT_KEYS, OpticsTable, read

CONDA_PY = "/opt/conda/bin/python3.9"


def _h5py_ok():
    if not os.path.exists(CONDA_PY):
        return False
    r = subprocess.run([CONDA_PY, "-c", "import h5py"], capture_output=True)
    return r.returncode == 0


needs_h5py = pytest.mark.skipif(not _h5py_ok(), reason="no h5py interpreter in this image")

DUMP = r"""
import hashlib, json, sys
import h5py, numpy as np
f = h5py.File(sys.argv[1], "r")
out = {"order": [], "attrs": {k: (v.decode() if isinstance(v, bytes) else str(v))
                              for k, v in f.attrs.items()}}
def txt(v):
    if isinstance(v, h5py.Empty):
        return ""
    if isinstance(v, bytes):
        return v.decode()
    if isinstance(v, np.ndarray):
        return v.tolist()
    return v.item() if hasattr(v, "item") else v
f.id.links.iterate(lambda n: out["order"].append(n.decode()) or None,
                   idx_type=h5py.h5.INDEX_CRT_ORDER)
for k in f:
    d = f[k]
    e = {"dtype": d.dtype.str, "shape": list(d.shape), "compression": d.compression,
         "level": d.compression_opts, "shuffle": bool(d.shuffle), "chunks": d.chunks,
         "fill": txt(d.fillvalue) if d.dtype.kind != "S" else None}
    if "DIMENSION_LIST" in d.attrs:
        e["dims"] = [[f[r].name for r in d.attrs["DIMENSION_LIST"][i]] for i in range(d.ndim)]
    e["attrs"] = {a: txt(d.attrs[a]) for a in d.attrs
                  if a not in ("DIMENSION_LIST", "REFERENCE_LIST")}
    if "REFERENCE_LIST" in d.attrs:
        e["refs"] = sorted([f[r[0]].name, int(r[1])] for r in d.attrs["REFERENCE_LIST"])
    e["sha"] = hashlib.sha256(np.ascontiguousarray(d[()]).tobytes()).hexdigest() \
        if d.size and "CLASS" not in d.attrs or "units" in d.attrs else None
    out[k] = e
json.dump(out, sys.stdout)
"""

MAKE = r"""
import sys
import h5py, numpy as np
with h5py.File(sys.argv[1], "w") as f:
    lam = f.create_dataset("lambda", data=np.array([4.4e-7, 5.5e-7, 8.8e-7]))
    lam.make_scale("lambda")
    lam.attrs["units"] = np.bytes_("m")
    nb = f.create_dataset("radius", data=np.array([1, 2], dtype="<i8"))
    nb.make_scale("radius")
    ph = f.create_dataset("nchar", shape=(4,), dtype=">f4")
    ph.make_scale("This is a netCDF dimension but not a netCDF variable.         4")
    q = f.create_dataset("qext", data=np.arange(6.0).reshape(2, 3), compression="gzip",
                         compression_opts=2)
    q.dims[0].attach_scale(nb)
    q.dims[1].attach_scale(lam)
    q.attrs["long_name"] = np.bytes_("extinction efficiency")
    c = f.create_dataset("qname", data=np.array([[b"a", b"b", b"", b""], [b"c", b"", b"", b""]],
                                                dtype="S1"))
    c.dims[0].attach_scale(nb)
    c.dims[1].attach_scale(ph)
"""


def _table(classic, n_mom):
    rng = np.random.default_rng(3)
    t = OpticsTable(3, [0.0, 0.5, 0.95], np.array([0.44e-6, 0.55e-6]), np.linspace(0, 180, 5),
                    classic=classic, n_mom=n_mom)
    for k in t.data:
        t.data[k][...] = rng.standard_normal(t.data[k].shape)
    return t


def _dump(path):
    r = subprocess.run([CONDA_PY, "-c", DUMP, path], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return json.loads(r.stdout)


@needs_h5py
@pytest.mark.parametrize("classic,n_mom", [(False, 0), (False, 7), (True, 7)])
def test_table_file_follows_netcdf4_conventions(tmp_path, classic, n_mom):
    """Every variable of the table as netCDF4-python writes it: a dataset per
    variable in definition order, coordinate variables as their dimension's
    scale, a phony scale for m/nMom, DIMENSION_LIST in dimension order,
    _Netcdf4Dimid in definition order, f8 data and i8 indices compressed with
    shuffle + deflate 4, netCDF default fill values, units/long_name."""
    t = _table(classic, n_mom)
    path = str(tmp_path / "optics_T.nomom.nc4")
    t.write(path)
    assert nc4.is_hdf5(path)
    d = _dump(path)
    r, lm, p = t.radius, t.lamname, t.pname
    dims = ["rh", lm, r, p, "ang"] + ([t.mname] if n_mom else [])
    coords = ["rh", lm, r, p, "ang"]
    # definition order: coordinates, then the data variables, then the phony dimension
    assert d["order"] == coords + list(t.dims) + ([t.mname] if n_mom else [])
    for i, name in enumerate(dims):
        e = d[name]
        assert e["attrs"]["CLASS"] == "DIMENSION_SCALE"
        assert e["attrs"]["_Netcdf4Dimid"] == i
        assert e["shape"] == [t.size[name]]
    for name in coords:
        e = d[name]
        assert e["attrs"]["NAME"] == name
        assert e["compression"] == "gzip" and e["level"] == 4 and e["shuffle"]
    assert d[r]["dtype"] == "<i8" and d[p]["dtype"] == "<i8"
    assert d[r]["fill"] == -9223372036854775806
    assert d["rh"]["fill"] == 9.969209968386869e36
    if n_mom:
        e = d[t.mname]
        assert e["dtype"] == ">f4" and e["attrs"]["NAME"] == \
            "This is a netCDF dimension but not a netCDF variable.%10d" % n_mom
    for k, kd in t.dims.items():
        e = d[k]
        assert e["dtype"] == "<f8" and e["shape"] == list(t.data[k].shape)
        assert e["compression"] == "gzip" and e["level"] == 4 and e["shuffle"]
        assert e["dims"] == [["/" + x] for x in kd]
        units, lname = META[k]
        assert e["attrs"]["long_name"] == lname
        assert e["attrs"].get("units") == units
        import hashlib
        assert e["sha"] == hashlib.sha256(np.ascontiguousarray(t.data[k]).tobytes()).hexdigest()
    # each scale lists every (variable, axis) that uses it
    for name in dims:
        want = sorted(["/" + k, i] for k, kd in t.dims.items() for i, x in enumerate(kd)
                      if x == name)
        assert d[name].get("refs", []) == want
    assert d[r]["attrs"]["long_name"] == "radius bin index (1-indexed)"
    assert d["ang"]["attrs"]["units"] == "degrees"


def test_read_back_bit_exact(tmp_path):
    t = _table(False, 5)
    path = str(tmp_path / "t.nc4")
    t.write(path)
    v = nc4.read(path, with_filters=True)
    assert list(v) == ["rh", "wavelength", "bin", "p", "ang"] + list(t.dims)
    assert v["bin"][0].dtype == np.int64 and list(v["p"][0]) == P_ORDER
    for k in t.dims:
        a, dims, at, filt = v[k]
        assert np.array_equal(a, t.data[k]) and dims == t.dims[k] and filt == (1, 4)
    assert v["lidar_ratio"][2]["units"] == ""


@needs_h5py
def test_reads_files_from_another_hdf5_writer(tmp_path):
    """A NetCDF-4-style file written by h5py (no creation-order index, a
    phony dimension scale, a char variable): variables listed by name, their
    dimension names from the attached scales, data and text attributes."""
    path = str(tmp_path / "k.nc4")
    r = subprocess.run([CONDA_PY, "-c", MAKE, path], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    v = nc4.read(path)
    assert list(v) == ["lambda", "qext", "qname", "radius"]
    assert v["qext"][1] == ("radius", "lambda")
    assert np.array_equal(v["qext"][0], np.arange(6.0).reshape(2, 3))
    assert v["qext"][2] == {"long_name": "extinction efficiency"}
    assert v["lambda"][2] == {"units": "m"} and v["lambda"][1] == ("lambda",)
    assert v["radius"][0].dtype == np.int64
    assert v["qname"][1] == ("radius", "nchar")
    assert v["qname"][0].dtype == np.dtype("S1") and v["qname"][0][0, 1] == b"b"


def test_writer_errors(tmp_path):
    w = nc4.Writer(str(tmp_path / "e.nc4"))
    w.dim("x", 3)
    with pytest.raises(nc4.NC4Error, match="defined twice"):
        w.dim("x", 4)
    w.var("a", "f8", ("x",))
    with pytest.raises(nc4.NC4Error, match="data of shape"):
        w.put("a", np.zeros(4))
    w.put("a", [1.0, 2.0, 3.0])
    with pytest.raises(nc4.NC4Error, match="define mode"):
        w.dim("y", 2)
    w.close()
    with pytest.raises(nc4.NC4Error, match="not an HDF5"):
        nc4.read(str(tmp_path / "missing.nc4"))
    with pytest.raises(ValueError, match="format"):
        _table(False, 0).write(str(tmp_path / "f.nc"), format="NETCDF5")


def test_empty_dimension_and_chunking(tmp_path):
    """A zero-length dimension and a variable larger than one 4 MiB chunk."""
    path = str(tmp_path / "z.nc4")
    big = np.random.default_rng(0).standard_normal((9, 300, 400))
    with nc4.Writer(path) as w:
        w.dim("e", 0)
        w.dim("a", 9)
        w.dim("b", 300)
        w.dim("c", 400)
        w.var("empty", "f8", ("e", "a"))
        w.var("big", "f8", ("a", "b", "c"))
        w.put("empty", np.zeros((0, 9)))
        w.put("big", big)
    v = nc4.read(path)
    assert v["empty"][0].shape == (0, 9)
    assert np.array_equal(v["big"][0], big)
