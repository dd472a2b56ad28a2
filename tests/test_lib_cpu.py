"""The C-ABI library loads and exports every entry point include/geosmie.h declares
(no compute calls: this runs without a GPU)."""
import ctypes
import os
import re

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "geosmie.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(gm_[a-z0-9_]+)\s*\(", src)))


def test_header_declares_entry_points():
    names = declared_functions()
    for must in ["gm_ctx_create", "gm_mie_batch", "gm_mie_s12", "gm_angle_tables"]:
        assert must in names


def test_library_exports_every_declared_symbol():
    from geosmie_amd import _lib
    L = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in declared_functions() if not hasattr(L, n)]
    assert not missing, missing
    # and the ctypes layer binds each of them
    unbound = [n for n in declared_functions() if n not in _lib.SIGNATURES]
    assert not unbound, unbound


def test_nc4_library_exports_every_declared_symbol():
    """libgeosmie_nc4.so (NetCDF-4 tables) against include/geosmie_nc4.h."""
    from geosmie_amd import nc4
    src = open(os.path.join(ROOT, "include", "geosmie_nc4.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    names = sorted(set(re.findall(r"\b(gmnc_[a-z0-9_]+)\s*\(", src)))
    assert "gmnc_create" in names and "gmnc_get_var" in names
    L = nc4.lib()
    assert not [n for n in names if not hasattr(L, n)]


def test_one_hip_runtime_per_process():
    """Loading the library first (before any torch import) still leaves one HIP
    runtime in the process: torch's, which the library's libamdhip64 binds to.
    With two, the second runtime finds no GPU."""
    import subprocess
    import sys
    code = ("import re, sys; sys.path.insert(0, %r)\n"
            "from geosmie_amd import _lib\n"
            "_lib.lib()\n"
            "import torch\n"
            "maps = open('/proc/self/maps').read()\n"
            "print(len(set(re.findall(r'/\\S*libamdhip64\\S*', maps))))\n" % ROOT)
    r = subprocess.run([sys.executable, "-c", code], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr
    assert r.stdout.strip().splitlines()[-1] == "1"


def test_host_only_entry_points():
    from geosmie_amd import _lib
    L = _lib.lib()
    assert L.gm_version().startswith(b"geosmie_amd")
    # n_max rule with half-to-even rounding (mie_coeffs.py:99)
    from oracle.mie import nmax_of
    for s in [0.01, 0.5, 1.0, 2.5, 10.0, 123.456, 2100.0]:
        assert L.gm_nmax(s) == nmax_of(s)


def test_no_gpu_raises_loudly():
    import torch
    if torch.cuda.is_available():
        return
    import pytest
    from geosmie_amd import _lib
    from geosmie_amd.mie import mie_batch
    with pytest.raises(_lib.GMError):
        mie_batch([1.0], m=1.5)


def test_surrogate_prep_rejects_cell_table_size_mismatch():
    """The gate's cell table must cover every cell index the kernel can form:
    a table sized for another cell grid is refused before anything launches."""
    import ctypes
    import pytest
    from geosmie_amd import _lib
    L = _lib.lib()
    fine = (ctypes.c_int * 6)(16, 6, 6, 12, 20, 6)       # 829440 cells
    rc = L.gm_surrogate_prep(None, 10, None, None, None, None, None, fine, None, 172800, 1.33,
                             None, None, None)
    assert rc < 0 and b"cell_ok has 172800 entries, the cell grid 829440" in L.gm_last_error()
    rc = L.gm_surrogate_prep(None, 10, None, None, None, None, None, fine, None, 829440, 1.33,
                             None, None, None)
    assert rc < 0 and b"bad arguments" in L.gm_last_error()


def test_missing_declared_symbol_raises_at_load(monkeypatch):
    """A library that lacks an entry point the binding declares is refused when
    it is loaded, not when the missing function is first called."""
    import pytest
    from geosmie_amd import _lib
    sig = dict(_lib.SIGNATURES)
    sig["gm_not_an_export"] = (ctypes.c_int, [])
    monkeypatch.setattr(_lib, "SIGNATURES", sig)
    monkeypatch.setattr(_lib, "_lib", None)
    with pytest.raises(_lib.GMError, match="gm_not_an_export"):
        _lib.lib()


def test_autotune_cache_file(tmp_path):
    """The tuning cache: entries merge per (device, workload) key, survive a
    reload, and a damaged file reads as empty."""
    from geosmie_amd import autotune
    path = str(tmp_path / "sub" / "tune.json")
    assert autotune.load_cache(path) == {}
    autotune.store_cache(path, "gfx950/256|coated", {"mie_waves": 0})
    autotune.store_cache(path, "gfx950/256|coated", {"mie_grid_mult": 2})
    autotune.store_cache(path, "gfx950/256|homog", {"mie_lanes": 32})
    d = autotune.load_cache(path)
    assert d["gfx950/256|coated"] == {"mie_waves": 0, "mie_grid_mult": 2}
    assert d["gfx950/256|homog"] == {"mie_lanes": 32}
    with open(path, "w") as f:
        f.write("{not json")
    assert autotune.load_cache(path) == {}
