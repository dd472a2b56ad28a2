"""Build libgeosmie_hip.so in-tree with hipcc for gfx950.

Usage: python -m geosmie_amd.build   (or geosmie_amd.build.build())
"""
import os
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgeosmie_hip.so")
INCLUDE = os.path.join(os.path.dirname(HERE), "include")

SOURCES = ["gm_api.hip", "gm_util.hip", "gm_mie.hip", "gm_angles.hip", "gm_s12.hip",
           "gm_integrate.hip", "gm_gsf.hip", "gm_lut.hip", "gm_mlp.hip",
           "gm_surrogate.hip", "gm_band.hip", "gm_grasp.hip"]

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
CC = os.environ.get("CC", "gcc")
NC4_LIB = os.path.join(HERE, "libgeosmie_nc4.so")

# sources whose results must match numpy bit for bit: no fused multiply-add
NO_CONTRACT = {"gm_band.hip", "gm_grasp.hip"}

# -packed-fp32-ops: no packed FP32 instructions from the compiler (some of their
# forms have given wrong results on these GPUs without an error).
# -amdgpu-mfma-vgpr-form: matrix-instruction accumulators in ordinary VGPRs
# (gfx950 reads and writes them there): without it the compiler keeps a
# K loop's accumulators in VGPRs and copies them to and from AGPRs around
# every step's MFMAs (128 extra instructions per step of the FP64 S1/S2
# kernel), and AGPR allocations cost occupancy.
FLAGS = ["-O3", "-std=c++17", "--offload-arch=gfx950", "-fPIC", "-shared",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
         "-mllvm", "-amdgpu-mfma-vgpr-form=1",
         "-ffp-contract=fast", "-Wno-unused-result", "-I" + INCLUDE, "-I" + CSRC]


def sources():
    return [os.path.join(CSRC, s) for s in SOURCES if os.path.exists(os.path.join(CSRC, s))]


def build(verbose=False):
    """Compile every HIP source into one shared library (always recompiles)."""
    objs = []
    tmp = os.path.join(HERE, "_objs")
    os.makedirs(tmp, exist_ok=True)
    procs = []
    for src in sources():
        obj = os.path.join(tmp, os.path.basename(src) + ".o")
        flags = [f for f in FLAGS if f != "-shared"]
        if os.path.basename(src) in NO_CONTRACT:
            flags = [("-ffp-contract=off" if f.startswith("-ffp-contract") else f) for f in flags]
        cmd = [HIPCC] + flags + ["-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd))
        procs.append((subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT), src))
        objs.append(obj)
    for p, src in procs:
        out, _ = p.communicate()
        if p.returncode != 0:
            sys.stderr.write(out.decode(errors="replace"))
            raise RuntimeError("hipcc failed on %s" % src)
    tmp_lib = LIB + ".tmp"
    cmd = [HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC", "-o", tmp_lib] + objs
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("link failed")
    os.replace(tmp_lib, LIB)
    build_nc4(verbose)
    return LIB


def hdf5_dir():
    """Prefix holding the HDF5 C library and headers (include/hdf5.h,
    lib/libhdf5.so, lib/libhdf5_hl.so), or None."""
    cands = [os.environ.get("GM_HDF5_DIR", ""), "/opt/conda", "/usr/local", "/usr"]
    for c in cands:
        if c and os.path.exists(os.path.join(c, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(c, "lib", "libhdf5.so")) and \
                os.path.exists(os.path.join(c, "lib", "libhdf5_hl.so")):
            return c
    return None


def build_nc4(verbose=False):
    """The NetCDF-4 table writer/reader (plain C on HDF5); None when no HDF5
    is found (tables can then be written only as NetCDF-3)."""
    h5 = hdf5_dir()
    if h5 is None:
        sys.stderr.write("geosmie_amd.build: no HDF5 C library found (set GM_HDF5_DIR); "
                         "libgeosmie_nc4.so not built\n")
        return None
    tmp_lib = NC4_LIB + ".tmp"
    cmd = [CC, "-O2", "-std=c11", "-D_GNU_SOURCE", "-Wall", "-fPIC", "-shared",
           "-I" + os.path.join(h5, "include"), "-I" + INCLUDE,
           os.path.join(CSRC, "gm_nc4.c"), "-L" + os.path.join(h5, "lib"),
           "-Wl,-rpath," + os.path.join(h5, "lib"), "-lhdf5_hl", "-lhdf5", "-o", tmp_lib]
    if verbose:
        print(" ".join(cmd))
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)
    if r.returncode != 0:
        sys.stderr.write(r.stdout.decode(errors="replace"))
        raise RuntimeError("gcc failed on gm_nc4.c")
    os.replace(tmp_lib, NC4_LIB)
    return NC4_LIB


if __name__ == "__main__":
    print(build(verbose="-v" in sys.argv))
