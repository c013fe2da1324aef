"""A small HDF5 reader over the HDF5 C library (ctypes), for the grid-file tests.

h5py is not installed; the HDF5 1.10 C library of this image is.  Reading
the files back through the library itself checks what any HDF5 reader
(h5py in the reference's output/show_xH_slice.py) sees: dataset names,
shapes, element types, chunking, values and attributes.
"""
import ctypes as C
import os

import numpy as np

HDF5_LIB = os.environ.get("MRX_HDF5_ROOT", "/opt/conda") + "/lib"

hid_t = C.c_int64
hsize_t = C.c_uint64
H5P_DEFAULT = 0
H5S_ALL = 0
H5E_DEFAULT = 0
H5F_ACC_RDONLY = 0
H5F_ACC_TRUNC = 2
H5_INDEX_NAME = 0
H5_ITER_INC = 0
H5T_FLOAT = 1
H5T_ORDER_LE = 0
H5D_CONTIGUOUS = 1
H5D_CHUNKED = 2


class H5GInfo(C.Structure):
    _fields_ = [("storage_type", C.c_int), ("nlinks", hsize_t), ("max_corder", C.c_int64), ("mounted", C.c_bool)]


_h5 = None


def lib():
    global _h5
    if _h5 is not None:
        return _h5
    L = C.CDLL(os.path.join(HDF5_LIB, "libhdf5.so.103"), mode=C.RTLD_GLOBAL)
    hl = C.CDLL(os.path.join(HDF5_LIB, "libhdf5_hl.so.100"))
    assert L.H5open() >= 0
    L.H5Eset_auto2.argtypes = [hid_t, C.c_void_p, C.c_void_p]
    L.H5Eset_auto2(H5E_DEFAULT, None, None)  # failures are return codes here
    for name, res, args in [
        ("H5Fopen", hid_t, [C.c_char_p, C.c_uint, hid_t]),
        ("H5Fcreate", hid_t, [C.c_char_p, C.c_uint, hid_t, hid_t]),
        ("H5Fclose", C.c_int, [hid_t]),
        ("H5Gget_info", C.c_int, [hid_t, C.POINTER(H5GInfo)]),
        ("H5Lget_name_by_idx", C.c_ssize_t, [hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p, C.c_size_t,
                                             hid_t]),
        ("H5Dopen2", hid_t, [hid_t, C.c_char_p, hid_t]),
        ("H5Dcreate2", hid_t, [hid_t, C.c_char_p, hid_t, hid_t, hid_t, hid_t, hid_t]),
        ("H5Dwrite", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        ("H5Dclose", C.c_int, [hid_t]),
        ("H5Dget_space", hid_t, [hid_t]),
        ("H5Dget_type", hid_t, [hid_t]),
        ("H5Dget_create_plist", hid_t, [hid_t]),
        ("H5Dread", C.c_int, [hid_t, hid_t, hid_t, hid_t, hid_t, C.c_void_p]),
        ("H5Screate_simple", hid_t, [C.c_int, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
        ("H5Sget_simple_extent_ndims", C.c_int, [hid_t]),
        ("H5Sget_simple_extent_dims", C.c_int, [hid_t, C.POINTER(hsize_t), C.POINTER(hsize_t)]),
        ("H5Sclose", C.c_int, [hid_t]),
        ("H5Tget_class", C.c_int, [hid_t]),
        ("H5Tget_size", C.c_size_t, [hid_t]),
        ("H5Tget_order", C.c_int, [hid_t]),
        ("H5Tclose", C.c_int, [hid_t]),
        ("H5Pget_layout", C.c_int, [hid_t]),
        ("H5Pget_chunk", C.c_int, [hid_t, C.c_int, C.POINTER(hsize_t)]),
        ("H5Pclose", C.c_int, [hid_t]),
        ("H5Aget_name_by_idx", C.c_ssize_t, [hid_t, C.c_char_p, C.c_int, C.c_int, hsize_t, C.c_char_p, C.c_size_t,
                                             hid_t]),
        ("H5Aopen_by_name", hid_t, [hid_t, C.c_char_p, C.c_char_p, hid_t, hid_t]),
        ("H5Aget_space", hid_t, [hid_t]),
        ("H5Aget_type", hid_t, [hid_t]),
        ("H5Aread", C.c_int, [hid_t, hid_t, C.c_void_p]),
        ("H5Aclose", C.c_int, [hid_t]),
    ]:
        fn = getattr(L, name)
        fn.restype, fn.argtypes = res, args
    L.native_float = hid_t.in_dll(L, "H5T_NATIVE_FLOAT_g").value
    L.native_double = hid_t.in_dll(L, "H5T_NATIVE_DOUBLE_g").value
    L.hl = hl
    _h5 = L
    return L


def _dims(L, space):
    nd = L.H5Sget_simple_extent_ndims(space)
    d = (hsize_t * max(nd, 1))()
    assert L.H5Sget_simple_extent_dims(space, d, None) == nd
    return tuple(int(x) for x in d[:nd])


class Dataset:
    """What a reader sees of one dataset: shape, element type, layout, values, attributes."""

    def __init__(self, shape, type_class, type_size, order, layout, chunk, data, attrs):
        self.shape, self.type_class, self.type_size, self.order = shape, type_class, type_size, order
        self.layout, self.chunk, self.data, self.attrs = layout, chunk, data, attrs


def read_file(path):
    """{dataset name: Dataset} of the file's root group (float datasets read as float32)."""
    L = lib()
    f = L.H5Fopen(os.fsencode(path), H5F_ACC_RDONLY, H5P_DEFAULT)
    if f < 0:
        raise OSError(f"cannot open {path}")
    out = {}
    try:
        info = H5GInfo()
        assert L.H5Gget_info(f, C.byref(info)) >= 0
        buf = C.create_string_buffer(256)
        for i in range(info.nlinks):
            assert L.H5Lget_name_by_idx(f, b".", H5_INDEX_NAME, H5_ITER_INC, i, buf, 256, H5P_DEFAULT) > 0
            name = buf.value.decode()
            out[name] = _read_dataset(L, f, name)
    finally:
        L.H5Fclose(f)
    return out


def _read_dataset(L, f, name):
    d = L.H5Dopen2(f, name.encode(), H5P_DEFAULT)
    assert d >= 0, name
    try:
        sp = L.H5Dget_space(d)
        shape = _dims(L, sp)
        L.H5Sclose(sp)
        t = L.H5Dget_type(d)
        tclass, tsize, order = L.H5Tget_class(t), int(L.H5Tget_size(t)), L.H5Tget_order(t)
        L.H5Tclose(t)
        pl = L.H5Dget_create_plist(d)
        layout = L.H5Pget_layout(pl)
        chunk = None
        if layout == H5D_CHUNKED:
            c = (hsize_t * 8)()
            nd = L.H5Pget_chunk(pl, 8, c)
            chunk = tuple(int(x) for x in c[:nd])
        L.H5Pclose(pl)
        data = np.empty(shape, np.float32)
        assert L.H5Dread(d, L.native_float, H5S_ALL, H5S_ALL, H5P_DEFAULT, data.ctypes.data_as(C.c_void_p)) >= 0
        attrs = _read_attrs(L, f, name)
    finally:
        L.H5Dclose(d)
    return Dataset(shape, tclass, tsize, order, layout, chunk, data, attrs)


def _read_attrs(L, f, obj):
    """{attribute name: (shape, type class, type size, values as float64)} of one object."""
    out = {}
    buf = C.create_string_buffer(256)
    i = 0
    while L.H5Aget_name_by_idx(f, obj.encode(), H5_INDEX_NAME, H5_ITER_INC, i, buf, 256, H5P_DEFAULT) > 0:
        an = buf.value.decode()
        a = L.H5Aopen_by_name(f, obj.encode(), an.encode(), H5P_DEFAULT, H5P_DEFAULT)
        sp = L.H5Aget_space(a)
        shape = _dims(L, sp)
        L.H5Sclose(sp)
        t = L.H5Aget_type(a)
        tclass, tsize = L.H5Tget_class(t), int(L.H5Tget_size(t))
        L.H5Tclose(t)
        v = np.empty(int(np.prod(shape)) if shape else 1, np.float64)
        assert L.H5Aread(a, L.native_double, v.ctypes.data_as(C.c_void_p)) >= 0
        L.H5Aclose(a)
        out[an] = (shape, tclass, tsize, v)
        i += 1
    return out


def write_plain_file(path, datasets):
    """Writes float32 arrays as contiguous datasets with the C library directly (checks the reader)."""
    L = lib()
    f = L.H5Fcreate(os.fsencode(path), H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT)
    assert f >= 0
    try:
        for name, a in datasets.items():
            a = np.ascontiguousarray(a, np.float32)
            dims = (hsize_t * a.ndim)(*a.shape)
            sp = L.H5Screate_simple(a.ndim, dims, None)
            d = L.H5Dcreate2(f, name.encode(), L.native_float, sp, H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT)
            assert d >= 0
            assert L.H5Dwrite(d, L.native_float, H5S_ALL, H5S_ALL, H5P_DEFAULT, a.ctypes.data_as(C.c_void_p)) >= 0
            L.H5Dclose(d)
            L.H5Sclose(sp)
    finally:
        L.H5Fclose(f)
