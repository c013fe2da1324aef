"""Loader of the HIP library libmrx_hip.so (the C ABI of include/mrx_reion.h).

There is no CPU fallback: if the library is missing or cannot load, every
entry point raises.  Build it with ``python -m meraxes_amd.build``.
"""
import ctypes as C
import os

from .abi import (ERRORS, GridParams, GridStats, HiiBatchBox, HiiGlobals, HiiParams, OutputGrids, SaveParams,
                  TbParams, TbVelParams, TsGlobals, TsParams, TsTables)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libmrx_hip.so")
IO_LIB_PATH = os.path.join(HERE, "libmrx_io.so")

# every symbol include/mrx_reion.h declares
EXPORTS = (
    "mrx_init", "mrx_free", "mrx_last_error", "mrx_synchronize", "mrx_set_stream", "mrx_radius_ladder",
    "mrx_find_hii_bubbles", "mrx_filter", "mrx_fft_r2c_3d", "mrx_fft_c2r_3d", "mrx_filtered_field",
    "mrx_brightness_temp", "mrx_abi_sizeof", "mrx_profile", "mrx_profile_read", "mrx_compute_ts",
    "mrx_hii_batch_create", "mrx_hii_batch_run", "mrx_hii_batch_describe", "mrx_hii_batch_free",
    "mrx_power_spectrum", "mrx_power_spectrum_adjoint", "mrx_snapshot_adjoint", "mrx_construct_baryon_grids",
    "mrx_brightness_temp_vel", "mrx_find_hii_bubbles_rec", "mrx_brightness_temp_noalloc", "mrx_ts_snapshot_tables",
    "mrx_ts_last_shells", "mrx_stream_copy", "mrx_device_count",
)

# every symbol include/mrx_io.h declares (libmrx_io.so)
IO_EXPORTS = ("mrx_grids_fname", "mrx_save_reion_input_grids", "mrx_save_reion_output_grids", "mrx_io_abi_sizeof")

_lib = None
_io = None


class MrxError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


def load():
    """The loaded library; raises OSError if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise OSError(f"{LIB_PATH} is not built (python -m meraxes_amd.build)")
    L = C.CDLL(LIB_PATH)
    vp, i, d, f = C.c_void_p, C.c_int, C.c_double, C.c_float
    ctxp = C.c_void_p
    L.mrx_init.argtypes = [i, C.POINTER(ctxp)]
    L.mrx_device_count.argtypes = []
    L.mrx_free.argtypes = [ctxp]
    L.mrx_last_error.argtypes = [ctxp]
    L.mrx_last_error.restype = C.c_char_p
    L.mrx_synchronize.argtypes = [ctxp]
    L.mrx_set_stream.argtypes = [ctxp, vp]
    L.mrx_radius_ladder.argtypes = [C.POINTER(HiiParams), vp, i]
    L.mrx_find_hii_bubbles.argtypes = [ctxp, C.POINTER(HiiParams)] + [vp] * 9 + [C.POINTER(HiiGlobals)]
    L.mrx_filter.argtypes = [ctxp, vp, i, i, i, f, i, d, i]
    L.mrx_fft_r2c_3d.argtypes = [ctxp, i, vp, vp, i]
    L.mrx_fft_c2r_3d.argtypes = [ctxp, i, vp, vp, i]
    L.mrx_filtered_field.argtypes = [ctxp, i, d, i, vp, d, i, i, vp, i]
    L.mrx_brightness_temp.argtypes = [ctxp, C.POINTER(TbParams), vp, vp, vp, vp, C.POINTER(C.c_double)]
    L.mrx_brightness_temp_noalloc.argtypes = [ctxp, C.POINTER(TbParams), vp, vp, vp, vp, C.POINTER(C.c_double)]
    L.mrx_compute_ts.argtypes = [ctxp, C.POINTER(TsParams)] + [vp] * 8 + [C.POINTER(TsGlobals)]
    L.mrx_ts_snapshot_tables.argtypes = [C.POINTER(TsParams), d, d, C.POINTER(TsTables)]
    L.mrx_ts_last_shells.argtypes = [ctxp, i, i, vp, i]
    L.mrx_stream_copy.argtypes = [ctxp, vp, vp, C.c_size_t, i, C.POINTER(C.c_double)]
    L.mrx_profile.argtypes = [ctxp, i]
    L.mrx_profile_read.argtypes = [ctxp, i, C.c_char_p, i, vp, vp]
    L.mrx_hii_batch_create.argtypes = [ctxp, C.POINTER(HiiBatchBox), i, i, C.POINTER(vp)]
    L.mrx_hii_batch_run.argtypes = [vp, C.POINTER(HiiGlobals)]
    L.mrx_hii_batch_describe.argtypes = [vp, C.c_char_p, i]
    L.mrx_hii_batch_free.argtypes = [vp]
    L.mrx_power_spectrum.argtypes = [ctxp, i, d, f, vp, i, vp, vp, vp, vp]
    L.mrx_power_spectrum_adjoint.argtypes = [ctxp, i, d, f, vp, vp, vp]
    L.mrx_snapshot_adjoint.argtypes = [ctxp, vp, C.POINTER(HiiParams), C.POINTER(TbParams)] + [vp] * 16
    L.mrx_construct_baryon_grids.argtypes = ([ctxp, C.POINTER(GridParams), C.c_long] + [vp] * 10 +
                                             [C.POINTER(GridStats)])
    L.mrx_find_hii_bubbles_rec.argtypes = [ctxp, C.POINTER(HiiParams)] + [vp] * 11 + [C.POINTER(HiiGlobals)]
    L.mrx_brightness_temp_vel.argtypes = [ctxp, C.POINTER(TbVelParams)] + [vp] * 5 + [C.POINTER(C.c_double)]
    L.mrx_abi_sizeof.argtypes = [i]
    L.mrx_abi_sizeof.restype = C.c_size_t
    _lib = L
    return L


def load_io():
    """The grid-file library libmrx_io.so (over HDF5); raises OSError if it is not built."""
    global _io
    if _io is not None:
        return _io
    load()  # the context library it links
    if not os.path.exists(IO_LIB_PATH):
        raise OSError(f"{IO_LIB_PATH} is not built (python -m meraxes_amd.build)")
    L = C.CDLL(IO_LIB_PATH)
    vp, i = C.c_void_p, C.c_int
    L.mrx_grids_fname.argtypes = [C.c_char_p, C.c_size_t, C.c_char_p, C.c_char_p, i, i]
    L.mrx_save_reion_input_grids.argtypes = [vp, C.POINTER(SaveParams), C.c_char_p] + [vp] * 4
    L.mrx_save_reion_output_grids.argtypes = [vp, C.POINTER(SaveParams), C.c_char_p, C.POINTER(OutputGrids)]
    L.mrx_io_abi_sizeof.argtypes = [i]
    L.mrx_io_abi_sizeof.restype = C.c_size_t
    _io = L
    return L


class Context:
    """One mrx_ctx bound to a device (replaces init_CUDA, meraxes_gpu.cu:127-194)."""

    def __init__(self, device=0):
        self.lib = load()
        self.ptr = C.c_void_p()
        rc = self.lib.mrx_init(device, C.byref(self.ptr))
        if rc != 0:
            raise MrxError(rc, f"mrx_init(device={device}) failed")

    def check(self, rc, what):
        if rc != 0:
            msg = self.lib.mrx_last_error(self.ptr)
            raise MrxError(rc, f"{what}: {msg.decode() if msg else ''}")
        return rc

    def close(self):
        if self.ptr:
            self.lib.mrx_free(self.ptr)
            self.ptr = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()
