"""Build recipe for the se_amd HIP library (gfx950).

Compiles every se_amd/csrc/*.hip translation unit with hipcc in parallel and
links them into se_amd/lib/libse_amd.so, in-tree, so the shared object travels
with the source tree.  Run as a module or through __graft_entry__.build().
"""
import concurrent.futures as cf
import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
LIBDIR = os.path.join(HERE, "lib")
OBJDIR = os.path.join(HERE, "_build")
ARCH = os.environ.get("SE_OFFLOAD_ARCH", "gfx950")

HIPFLAGS = [
    "--offload-arch=" + ARCH, "-O3", "-std=c++17", "-fPIC",
    # packed FP32 instructions stay off (only ordinary scalar / float2 code)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # matrix-core accumulators in VGPRs: with the AGPR form the compiler copies
    # every GEMM accumulator VGPR -> AGPR -> VGPR around each K chunk
    "-mllvm", "-amdgpu-mfma-vgpr-form",
    "-I" + os.path.join(ROOT, "include"), "-I" + CSRC,
    "-Wno-unused-command-line-argument",
]


def _hipcc():
    for c in ("/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def _compile(src):
    obj = os.path.join(OBJDIR, os.path.basename(src) + ".o")
    cmd = [_hipcc()] + HIPFLAGS + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("compile failed: %s\n%s" % (" ".join(cmd), r.stderr))
    return obj


def build(verbose=True):
    os.makedirs(LIBDIR, exist_ok=True)
    os.makedirs(OBJDIR, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")))
    with cf.ThreadPoolExecutor(max_workers=max(1, jobs)) as ex:
        objs = list(ex.map(_compile, srcs))
    out = os.path.join(LIBDIR, "libse_amd.so")
    tmp = out + ".tmp"
    cmd = [_hipcc(), "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("link failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    build_io(verbose)
    return out


def build_io(verbose=True):
    """libse_amd_io.so: field files in the reference's HDF5 layout (host code,
    g++ against HDF5 1.10 at $SE_HDF5_PREFIX, default /opt/conda)."""
    prefix = os.environ.get("SE_HDF5_PREFIX", "/opt/conda")
    inc, lib = os.path.join(prefix, "include"), os.path.join(prefix, "lib")
    if not os.path.exists(os.path.join(inc, "hdf5.h")) or not os.path.exists(os.path.join(lib, "libhdf5.so")):
        if verbose:
            print("field I/O library skipped: no HDF5 under", prefix)
        return None
    out = os.path.join(LIBDIR, "libse_amd_io.so")
    tmp = out + ".tmp"
    cmd = ["g++", "-O2", "-std=c++17", "-fPIC", "-shared", "-I" + os.path.join(ROOT, "include"), "-I" + inc,
           os.path.join(CSRC, "field_io.cpp"), "-o", tmp, os.path.join(lib, "libhdf5.so"), "-Wl,-rpath," + lib]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError("field I/O build failed: %s\n%s" % (" ".join(cmd), r.stderr))
    os.replace(tmp, out)
    if verbose:
        print("built", out)
    return out


if __name__ == "__main__":
    build()
    sys.exit(0)
