"""Build the gfx950 HIP library ``libvlasov1d2v.so`` in-tree with hipcc.

Usage: ``python -m vlasov1d2v_amd.build`` (or ``__graft_entry__.build()``).
The shared library is written next to this file so that it travels with the
source tree; it is never installed into site-packages.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libvlasov1d2v.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)
SOURCES = ["vlasov.hip"]
ARCH = "gfx950"


def hipcc():
    cand = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not os.path.exists(cand):
        raise RuntimeError("hipcc not found; the HIP library cannot be built")
    return cand


def compile_cmd(out=LIB_PATH, extra=()):
    return [hipcc(), f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
            # packed FP32 instructions are kept out of the device code
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument",
            # the f^n staging asm sets M0 itself (no compiler-generated code in
            # these kernels keeps a value in M0 across it)
            "-Wno-inline-asm",
            "-I", INCLUDE, "-I", CSRC,
            *extra,
            "-o", out] + [os.path.join(CSRC, s) for s in SOURCES]


IO_LIB_NAME = "libvlasov1d2v_io.so"
IO_LIB_PATH = os.path.join(HERE, IO_LIB_NAME)
IO_SOURCES = ["vl_io.c"]


def hdf5_root():
    """Prefix of the HDF5 1.10+ C library (``include/hdf5.h``, ``lib/libhdf5.so``):
    ``$HDF5_ROOT``, else the image's /opt/conda, else /usr; None if absent."""
    for cand in (os.environ.get("HDF5_ROOT"), "/opt/conda", "/usr"):
        if cand and os.path.exists(os.path.join(cand, "include", "hdf5.h")) and \
                os.path.exists(os.path.join(cand, "lib", "libhdf5.so")):
            return cand
    return None


def io_compile_cmd(root, out=IO_LIB_PATH):
    lib = os.path.join(root, "lib")
    return ["gcc", "-O2", "-std=gnu99", "-fPIC", "-shared", "-Wall", "-I", INCLUDE,
            "-I", os.path.join(root, "include"), "-o", out] + \
        [os.path.join(CSRC, s) for s in IO_SOURCES] + \
        ["-L", lib, "-Wl,-rpath," + lib, "-lhdf5", "-lpthread"]


def build_io(verbose=True):
    """Build the run-file library (HDF5) next to the HIP library.  Returns
    its path, or None when no HDF5 C library is installed (the stepping
    library does not need it)."""
    root = hdf5_root()
    if root is None:
        if verbose:
            print("HDF5 C library not found: libvlasov1d2v_io.so not built", flush=True)
        return None
    tmp = IO_LIB_PATH + ".tmp"
    cmd = io_compile_cmd(root, tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"gcc failed with exit code {r.returncode}")
    os.replace(tmp, IO_LIB_PATH)
    return IO_LIB_PATH


def build(verbose=True):
    tmp = LIB_PATH + ".tmp"
    cmd = compile_cmd(tmp)
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        sys.stderr.write(r.stdout + r.stderr)
        raise RuntimeError(f"hipcc failed with exit code {r.returncode}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


if __name__ == "__main__":
    build()
    build_io()
