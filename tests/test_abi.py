"""The C-ABI library loads and exports every entry point include/mrx_reion.h declares.

CPU-only: no call here touches a GPU (only the host-side ladder and sizeof).
"""
import ctypes as C
import os
import re

import numpy as np

from meraxes_amd import _lib, abi
from meraxes_amd import params as P
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    src = open(os.path.join(ROOT, "include", "mrx_reion.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    return sorted(set(re.findall(r"\b(mrx_[a-z0-9_]+)\s*\(", src)))


def test_library_exports_every_declared_symbol():
    L = _lib.load()
    names = declared_functions()
    assert len(names) >= 12
    for n in names:
        assert hasattr(L, n), n
    assert set(names) == set(_lib.EXPORTS)


def declared_arity():
    """{function: number of parameters} from include/mrx_reion.h."""
    src = open(os.path.join(ROOT, "include", "mrx_reion.h")).read()
    src = re.sub(r"/\*.*?\*/", "", src, flags=re.S)
    out = {}
    for m in re.finditer(r"\b(mrx_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", src, flags=re.S):
        args = m.group(2).strip()
        out[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    return out


def test_bindings_match_declared_arity():
    L = _lib.load()
    ar = declared_arity()
    assert set(ar) == set(_lib.EXPORTS)
    for name, n in ar.items():
        at = getattr(L, name).argtypes
        assert at is not None, name
        assert len(at) == n, (name, len(at), n)


def test_struct_mirrors_match_library():
    L = _lib.load()
    assert L.mrx_abi_sizeof(0) == C.sizeof(abi.HiiParams)
    assert L.mrx_abi_sizeof(1) == C.sizeof(abi.HiiGlobals)
    assert L.mrx_abi_sizeof(2) == C.sizeof(abi.TbParams)
    assert L.mrx_abi_sizeof(3) == C.sizeof(abi.TsParams)
    assert L.mrx_abi_sizeof(4) == C.sizeof(abi.TsGlobals)
    assert L.mrx_abi_sizeof(5) == C.sizeof(abi.HiiBatchBox)
    assert L.mrx_abi_sizeof(6) == C.sizeof(abi.GridParams)
    assert L.mrx_abi_sizeof(7) == C.sizeof(abi.GridStats)
    assert L.mrx_abi_sizeof(9) == C.sizeof(abi.TsTables)
    assert L.mrx_abi_sizeof(8) == C.sizeof(abi.TbVelParams)


def test_radius_ladder_matches_oracle_and_closed_form():
    from meraxes_amd import reion

    # (N, L, expected S) of the measured configurations
    for N, L, S in [(64, 100.0, 42), (256, 200.0, 44), (512, 400.0, 44), (128, 100.0, 44), (192, 100.0, 48)]:
        p = P.hii_params(N, L, 8.0)
        lib_R = reion.radius_ladder(p)
        orc_R = O.radius_ladder(p)
        py_R = np.array(P.radius_ladder(p))
        assert len(lib_R) == S, (N, L, len(lib_R))
        assert np.array_equal(lib_R, orc_R) and np.array_equal(lib_R, py_R)
        clf = 1.0 if L / N < 1.0 else 0.620350491
        assert lib_R[-1] == clf * L / N
        assert lib_R[0] == min(50.0, 0.620350491 * L)


def test_missing_library_fails_loudly(monkeypatch):
    monkeypatch.setattr(_lib, "_lib", None)
    monkeypatch.setattr(_lib, "LIB_PATH", "/nonexistent/libmrx_hip.so")
    try:
        _lib.load()
    except OSError as e:
        assert "not built" in str(e)
    else:
        raise AssertionError("load() must raise without the library")
