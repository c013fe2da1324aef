"""Grid files, host side: names, the C ABI of libmrx_io.so, the reader, the oracle.

CPU-only: nothing here opens a context.  The file name follows
gen_grids_fname (reionization.c:1640-1646); libmrx_io.so exports what
include/mrx_io.h declares with matching struct sizes; the ctypes reader
(tests/h5read.py) returns what the HDF5 library wrote; the oracle's de-padding
is the index map of misc_tools.c:200-220 and its unit conversion turns one
solar mass per year, expressed in internal units, back into 1.
"""
import ctypes as C
import os
import re

import numpy as np
import pytest

import h5read
from meraxes_amd import _lib, abi, reion
from meraxes_amd import params as P
from oracle import grid_files as GF

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header(name):
    src = open(os.path.join(ROOT, "include", name)).read()
    return re.sub(r"/\*.*?\*/", "", src, flags=re.S)


def test_io_library_exports_every_declared_symbol_with_its_arity():
    L = _lib.load_io()
    src = header("mrx_io.h")
    ar = {}
    for m in re.finditer(r"\b(mrx_[a-z0-9_]+)\s*\(([^;{]*?)\)\s*;", src, flags=re.S):
        args = m.group(2).strip()
        ar[m.group(1)] = 0 if args in ("", "void") else args.count(",") + 1
    assert set(ar) == set(_lib.IO_EXPORTS)
    for name, n in ar.items():
        fn = getattr(L, name)
        assert fn.argtypes is not None and len(fn.argtypes) == n, name
    assert L.mrx_io_abi_sizeof(0) == C.sizeof(abi.SaveParams)
    assert L.mrx_io_abi_sizeof(1) == C.sizeof(abi.OutputGrids)
    assert L.mrx_io_abi_sizeof(2) == 0


def test_grids_fname_follows_the_reference_format():
    assert reion.gen_grids_fname("/scratch/run", "meraxes", 57) == "/scratch/run/meraxes_grids_57.hdf5"
    assert reion.gen_grids_fname(None, "meraxes", 0, relative=True) == "meraxes_grids_0.hdf5"
    L = _lib.load_io()
    small = C.create_string_buffer(8)
    assert L.mrx_grids_fname(small, 8, b"/x", b"meraxes", 1, 0) == abi.MRX_EINVAL  # does not fit
    buf = C.create_string_buffer(64)
    n = L.mrx_grids_fname(buf, 64, b"out", b"g", 123, 0)
    assert n == len("out/g_grids_123.hdf5") and buf.value == b"out/g_grids_123.hdf5"


def test_reader_returns_what_the_library_wrote(tmp_path):
    rng = np.random.default_rng(3)
    a = rng.normal(size=(3, 4, 5)).astype(np.float32)
    b = np.arange(7, dtype=np.float32)
    path = str(tmp_path / "plain.hdf5")
    h5read.write_plain_file(path, {"a": a, "b": b})
    f = h5read.read_file(path)
    assert sorted(f) == ["a", "b"]
    assert f["a"].shape == (3, 4, 5) and np.array_equal(f["a"].data, a)
    assert f["b"].shape == (7,) and np.array_equal(f["b"].data, b)
    for d in f.values():
        assert (d.type_class, d.type_size, d.order) == (h5read.H5T_FLOAT, 4, h5read.H5T_ORDER_LE)
        assert d.layout == h5read.H5D_CONTIGUOUS and d.chunk is None and d.attrs == {}


def test_oracle_unpad_is_the_padded_index_map():
    N = 6
    padded = np.arange(N * N * (N + 2), dtype=np.float32)
    real = GF.unpad(padded, N)
    i, j, k = np.meshgrid(np.arange(N), np.arange(N), np.arange(N), indexing="ij")
    assert np.array_equal(real, (k + (N + 2) * (j + N * i)).astype(np.float32))


def test_oracle_unit_conversion_inverts_internal_units():
    p = P.save_params(8)
    N = 8
    one = GF.SOLAR_MASS / GF.SEC_PER_YEAR / (p.unit_mass_g / p.unit_time_s)  # 1 M_sun/yr in internal units
    g = np.zeros((N, N, N + 2), np.float32)
    g[...] = np.float32(one)
    out = GF.msun_per_yr(g, N, p.unit_mass_g, p.unit_time_s)
    np.testing.assert_allclose(out, 1.0, rtol=1e-6)
    # padding never reaches the file
    g[:, :, N:] = np.nan
    assert np.isfinite(GF.msun_per_yr(g, N, p.unit_mass_g, p.unit_time_s)).all()


@pytest.mark.parametrize("uvb,spin,tb,ps", [(1, 1, 1, 1), (0, 0, 0, 0), (1, 0, 1, 0), (0, 1, 0, 1)])
def test_oracle_dataset_names_per_flag(uvb, spin, tb, ps):
    """The names save_reion_output_grids writes for each flag (reionization.c:1830-2040)."""
    N = 4
    p = P.save_params(N, uvb_flag=uvb, include_spin_temp=spin, compute_21cm_bright_temp=tb, compute_ps=ps,
                      ps_length=3)
    real = np.zeros((N, N, N), np.float32)
    grids = {n: real for n in abi.OUTPUT_GRID_NAMES}
    grids["x_e_box_prev"] = np.zeros((N, N, N + 2), np.float32)
    grids.update(PS_k=np.ones(3, np.float32), PS_data=np.ones(3, np.float32), PS_error=np.ones(3, np.float32))
    ds, at = GF.output_file(p, grids, {n: 0.0 for n in abi.OUTPUT_AVERAGES})
    expect = ["xH", "z_at_ionization", "r_bubble"]
    expect += ["J_21", "J_21_at_ionization", "Mvir_crit"] if uvb else []
    expect += ["TS_box", "Tk_box", "x_e_box"] if spin else []
    expect += ["delta_T"] if tb else []
    expect += ["k_bins", "PS_data", "PS_error"] if ps else []
    assert list(ds) == expect
    n_attr = 2 + (1 if uvb else 0) + (7 if spin else 0) + (1 if tb else 0)
    assert sum(len(v) for v in at.values()) == n_attr
