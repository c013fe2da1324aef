"""Builds libmrx_hip.so in-tree with hipcc for gfx950.

``python -m meraxes_amd.build`` compiles every HIP source of the package.
Packed-FP32 instructions are disabled (-packed-fp32-ops), and nothing is
built with fast-math: the window and threshold arithmetic must round exactly
as the reference's C does.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OUT = os.path.join(HERE, "libmrx_hip.so")
IO_OUT = os.path.join(HERE, "libmrx_io.so")
# the HDF5 C library (serial, with the high-level H5LT routines) of this image
HDF5_ROOT = os.environ.get("MRX_HDF5_ROOT", "/opt/conda")
ARCH = os.environ.get("MRX_OFFLOAD_ARCH", "gfx950")


def hipcc():
    for c in (shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found")


def flags():
    return [
        f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-shared",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        "-Wno-unused-command-line-argument",
    ]


def build(verbose=False):
    src = os.path.join(CSRC, "mrx_reion.hip")
    tmp = OUT + ".tmp"
    # host-side spin-temperature tables: plain C, no fused multiply-adds
    objs = []
    for name in ("ts_tables", "recomb_tables"):  # host tables of the spin temperature and the recombinations
        obj = os.path.join(CSRC, name + ".o")
        ccmd = ["gcc", "-O2", "-std=gnu11", "-fPIC", "-pthread", "-ffp-contract=off", "-Wall", "-c",
                os.path.join(CSRC, name + ".c"), "-o", obj]
        if verbose:
            print(" ".join(ccmd), flush=True)
        r = subprocess.run(ccmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"gcc failed:\n{r.stderr}")
        objs.append(obj)
    # the HIP translation unit and the C object are linked together through
    # the linker so hipcc does not treat the object as a HIP source
    grid = os.path.join(CSRC, "grid.hip")  # galaxy gridding (rocPRIM sort): its own translation unit
    util = os.path.join(CSRC, "util.hip")  # device utilities (the streaming copy)
    cmd = ([hipcc()] + flags() + ["-o", tmp, src, grid, util] + [a for o in objs for a in ("-Xlinker", o)] +
           ["-lpthread"])
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed:\n{err}")
    if verbose and err:
        print(err, file=sys.stderr)
    os.replace(tmp, OUT)
    build_io(verbose)
    return OUT


def build_io(verbose=False):
    """libmrx_io.so: the grid files (include/mrx_io.h) over the HDF5 C library.

    A library of its own so that libmrx_hip.so does not depend on HDF5; it
    links libmrx_hip.so (the context) from its own directory.
    """
    tmp = IO_OUT + ".tmp"
    cmd = ([hipcc()] + flags() + ["-idirafter", os.path.join(HDF5_ROOT, "include"), "-o", tmp,
                                  os.path.join(CSRC, "io.hip"), "-L" + HERE, "-lmrx_hip",
                                  "-L" + os.path.join(HDF5_ROOT, "lib"), "-lhdf5_hl", "-lhdf5",
                                  # the system's own libraries first: the HDF5 directory also holds
                                  # an older libstdc++ and libgcc_s
                                  "-Wl,-rpath,$ORIGIN", "-Wl,-rpath,/usr/lib/x86_64-linux-gnu",
                                  "-Wl,-rpath," + os.path.join(HDF5_ROOT, "lib")])
    if verbose:
        print(" ".join(cmd), flush=True)
    r = subprocess.run(cmd, capture_output=True, text=True)
    err = "\n".join(l for l in r.stderr.splitlines() if "packed-fp32-ops" not in l and l.strip())
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed (libmrx_io.so):\n{err}")
    if verbose and err:
        print(err, file=sys.stderr)
    os.replace(tmp, IO_OUT)
    return IO_OUT


if __name__ == "__main__":
    print(build(verbose=True))
