"""The Meraxes drop-in (dropin/meraxes_hip_dropin.c) as a Meraxes GPU build would
take it: compiled with -DUSE_HIP against the run_globals fields it reads
(tests/dropin_stub/meraxes.h declares exactly those, with the reference's names
and types), linked against libmrx_hip.so, defining the symbols Meraxes calls
(init_CUDA and _find_HII_bubbles_gpu with the signatures of src/meraxes_gpu.h:53-54,
find_HII_bubbles.c:503-513) and reporting a library failure the reference's way
(mlog_error, then ABORT).  CPU only: no call reaches a GPU."""
import os
import shutil
import subprocess

import pytest

from meraxes_amd import _lib

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SHIM = os.path.join(ROOT, "dropin", "meraxes_hip_dropin.c")
STUB = os.path.join(ROOT, "tests", "dropin_stub")
DEFINES = ("init_CUDA", "_find_HII_bubbles_gpu", "_ComputeTs_gpu", "ComputeBrightnessTemperatureBox_gpu",
           "construct_baryon_grids_gpu")


def symbols(path, dynamic=False):
    out = subprocess.run(["nm", "-D" if dynamic else "-g", path], capture_output=True, text=True, check=True).stdout
    defined, undefined = set(), set()
    for line in out.splitlines():
        parts = line.split()
        if len(parts) >= 2 and parts[-2] == "U":
            undefined.add(parts[-1])
        elif len(parts) == 3 and parts[1] in "TDBRW":
            defined.add(parts[2])
    return defined, undefined


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a C compiler")
def test_dropin_compiles_links_and_reports_errors(tmp_path):
    _lib.load()  # the library is built
    obj = str(tmp_path / "dropin.o")
    flags = ["-std=c99", "-O2", "-Wall", "-Wextra", "-Werror", "-DUSE_HIP", "-I", STUB, "-I",
             os.path.join(ROOT, "include")]
    subprocess.run(["gcc"] + flags + ["-c", SHIM, "-o", obj], check=True)
    defined, undefined = symbols(obj)
    assert set(DEFINES) <= defined
    lib_defined, _ = symbols(_lib.LIB_PATH, dynamic=True)
    wanted = {u for u in undefined if u.startswith("mrx_")}
    assert wanted and wanted <= lib_defined, wanted - lib_defined
    exe = str(tmp_path / "dropin_main")
    subprocess.run(["gcc"] + flags + [os.path.join(STUB, "main.c"), obj, "-L", _lib.HERE, "-lmrx_hip",
                                      "-Wl,-rpath," + _lib.HERE, "-o", exe], check=True)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    # no context (init_CUDA never ran): the library refuses, the shim aborts as Meraxes does
    assert r.returncode == 3, (r.returncode, r.stdout, r.stderr)
    assert "MI355X path, find_HII_bubbles" in r.stderr and "no context" in r.stderr
    assert "ABORT(1)" in r.stderr
