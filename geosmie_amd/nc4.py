"""NetCDF-4 files through libgeosmie_nc4.so (HDF5 underneath).

The reference writes its tables with netCDF4-python
(src/geosmie/dointegration.py:205-354, src/gsf/convertncdf.py:70-79,
src/geosmie/bandaverage.py:222-290); ``Writer`` takes the same steps -
dimensions, variables with compression='zlib' (shuffle and deflate level 4,
netCDF4-python's defaults), text attributes, then the data - and ``read``
returns every variable of a NetCDF-4 file with its dimension names and text
attributes.  The library is built by geosmie_amd.build next to this module.
"""
import ctypes
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB = os.path.join(HERE, "libgeosmie_nc4.so")

BYTE, CHAR, SHORT, INT, FLOAT, DOUBLE, INT64 = 1, 2, 3, 4, 5, 6, 10
GLOBAL = -1
MAX_DIMS = 8
_DTYPE = {BYTE: np.int8, CHAR: np.dtype("S1"), SHORT: np.int16, INT: np.int32,
          FLOAT: np.float32, DOUBLE: np.float64, INT64: np.int64}
_CODE = {"f8": DOUBLE, "d": DOUBLE, "f4": FLOAT, "f": FLOAT, "i8": INT64, "i4": INT, "i": INT,
         "i2": SHORT, "i1": BYTE, "b": BYTE, "c": CHAR, "S1": CHAR}
HDF5_MAGIC = b"\x89HDF\r\n\x1a\n"

_lib = None


class NC4Error(RuntimeError):
    pass


def lib():
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB):
        raise NC4Error("%s is not built (python -m geosmie_amd.build; it needs the HDF5 C library "
                       "and headers, found under GM_HDF5_DIR, /opt/conda, /usr/local or /usr)"
                       % LIB)
    L = ctypes.CDLL(LIB)
    vp, ip, cp = ctypes.c_void_p, ctypes.POINTER(ctypes.c_int), ctypes.c_char_p
    sig = {
        "gmnc_last_error": (cp, []),
        "gmnc_create": (ctypes.c_int, [cp, ctypes.POINTER(vp)]),
        "gmnc_def_dim": (ctypes.c_int, [vp, cp, ctypes.c_int64, ip]),
        "gmnc_def_var": (ctypes.c_int, [vp, cp, ctypes.c_int, ctypes.c_int, ip, ctypes.c_int,
                                        ctypes.c_int, ip]),
        "gmnc_put_att_text": (ctypes.c_int, [vp, ctypes.c_int, cp, cp]),
        "gmnc_enddef": (ctypes.c_int, [vp]),
        "gmnc_put_var": (ctypes.c_int, [vp, ctypes.c_int, vp]),
        "gmnc_open": (ctypes.c_int, [cp, ctypes.POINTER(vp)]),
        "gmnc_inq_nvars": (ctypes.c_int, [vp, ip]),
        "gmnc_inq_var": (ctypes.c_int, [vp, ctypes.c_int, cp, ctypes.c_size_t, ip, ip,
                                        ctypes.POINTER(ctypes.c_int64)]),
        "gmnc_inq_dimname": (ctypes.c_int, [vp, ctypes.c_int, ctypes.c_int, cp, ctypes.c_size_t]),
        "gmnc_inq_filters": (ctypes.c_int, [vp, ctypes.c_int, ip, ip]),
        "gmnc_get_att_text": (ctypes.c_int, [vp, ctypes.c_int, cp, cp, ctypes.c_size_t,
                                             ctypes.POINTER(ctypes.c_size_t)]),
        "gmnc_get_var": (ctypes.c_int, [vp, ctypes.c_int, vp]),
        "gmnc_close": (ctypes.c_int, [vp]),
    }
    for name, (res, args) in sig.items():
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    _lib = L
    return L


def _check(rc, what):
    if rc < 0:
        raise NC4Error("%s: %s" % (what, lib().gmnc_last_error().decode(errors="replace")))
    return rc


def is_hdf5(path):
    with open(path, "rb") as f:
        return f.read(8) == HDF5_MAGIC


class Writer(object):
    """netCDF4.Dataset(path, 'w') for the variables the tables hold.

    Define everything first (dim, var with attrs), then put the data; the
    first put creates the datasets."""

    def __init__(self, path):
        L = lib()
        self._f = ctypes.c_void_p()
        _check(L.gmnc_create(os.fsencode(path), ctypes.byref(self._f)), "create " + str(path))
        self._dims, self._vars = {}, {}

    def dim(self, name, size):
        d = ctypes.c_int()
        _check(lib().gmnc_def_dim(self._f, name.encode(), int(size), ctypes.byref(d)),
               "dimension " + name)
        self._dims[name] = (d.value, int(size))

    def var(self, name, dtype, dims, zlib=True, complevel=4, shuffle=True, attrs=None):
        code = _CODE[dtype] if isinstance(dtype, str) else int(dtype)
        ids = (ctypes.c_int * max(1, len(dims)))(*[self._dims[d][0] for d in dims])
        v = ctypes.c_int()
        _check(lib().gmnc_def_var(self._f, name.encode(), code, len(dims), ids,
                                  int(complevel) if zlib else 0, int(bool(shuffle)),
                                  ctypes.byref(v)), "variable " + name)
        self._vars[name] = (v.value, code, tuple(self._dims[d][1] for d in dims))
        for k, text in (attrs or {}).items():
            self.attr(name, k, text)

    def attr(self, var, name, text):
        vid = GLOBAL if var is None else self._vars[var][0]
        _check(lib().gmnc_put_att_text(self._f, vid, name.encode(), str(text).encode()),
               "attribute %s of %s" % (name, var))

    def put(self, name, data):
        vid, code, shape = self._vars[name]
        a = np.ascontiguousarray(data, dtype=_DTYPE[code])
        if a.shape != shape:
            if a.size != int(np.prod(shape)):
                raise NC4Error("variable %s: data of shape %s for %s" % (name, a.shape, shape))
            a = a.reshape(shape)
        _check(lib().gmnc_put_var(self._f, vid, a.ctypes.data_as(ctypes.c_void_p)),
               "data of " + name)

    def close(self):
        if self._f:
            f, self._f = self._f, None
            _check(lib().gmnc_close(f), "close")

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()


def _text(L, f, vid, name):
    n = ctypes.c_size_t()
    buf = ctypes.create_string_buffer(4096)
    rc = L.gmnc_get_att_text(f, vid, name.encode(), buf, 4096, ctypes.byref(n))
    _check(rc, "attribute " + name)
    if rc > 0:
        return None
    if n.value >= 4096:
        buf = ctypes.create_string_buffer(n.value + 1)
        _check(L.gmnc_get_att_text(f, vid, name.encode(), buf, n.value + 1, ctypes.byref(n)),
               "attribute " + name)
    return buf.value.decode(errors="replace")


def read(path, attrs=("units", "long_name"), with_filters=False):
    """{name: (array, dimension names, {attribute: text})} of a NetCDF-4 file,
    in the file's variable order; with_filters adds (shuffle, deflate level)."""
    L = lib()
    f = ctypes.c_void_p()
    _check(L.gmnc_open(os.fsencode(path), ctypes.byref(f)), "open " + str(path))
    out = {}
    try:
        n = ctypes.c_int()
        _check(L.gmnc_inq_nvars(f, ctypes.byref(n)), "variables")
        for vid in range(n.value):
            name = ctypes.create_string_buffer(256)
            typ, nd = ctypes.c_int(), ctypes.c_int()
            shape = (ctypes.c_int64 * MAX_DIMS)()
            _check(L.gmnc_inq_var(f, vid, name, 256, ctypes.byref(typ), ctypes.byref(nd), shape),
                   "variable %d" % vid)
            dims = []
            for i in range(nd.value):
                dn = ctypes.create_string_buffer(256)
                _check(L.gmnc_inq_dimname(f, vid, i, dn, 256), "dimension")
                dims.append(dn.value.decode() or "dim_%d" % i)
            a = np.empty(tuple(shape[i] for i in range(nd.value)), dtype=_DTYPE[typ.value])
            _check(L.gmnc_get_var(f, vid, a.ctypes.data_as(ctypes.c_void_p)),
                   "data of " + name.value.decode())
            at = {}
            for k in attrs:
                t = _text(L, f, vid, k)
                if t is not None:
                    at[k] = t
            entry = (a, tuple(dims), at)
            if with_filters:
                sh, dl = ctypes.c_int(), ctypes.c_int()
                _check(L.gmnc_inq_filters(f, vid, ctypes.byref(sh), ctypes.byref(dl)), "filters")
                entry = entry + ((sh.value, dl.value),)
            out[name.value.decode()] = entry
    finally:
        L.gmnc_close(f)
    return out
