"""ctypes binding of libvsfeel_bm3d.so (the C ABI of include/vsfeel_bm3d.h).

The library is the gfx950 build in this directory.  Loading fails loudly when
it is missing: there is no CPU fallback for the filter.
"""

from __future__ import annotations

import ctypes
import os

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libvsfeel_bm3d.so")

OK = 0
EINVAL = 1
EDEVICE = 2
ERANGE = 3

# every function the header declares, for the export check
EXPORTS = (
    "vsfeel_bm3d_create",
    "vsfeel_bm3d_free",
    "vsfeel_bm3d_get_info",
    "vsfeel_bm3d_frame_window",
    "vsfeel_bm3d_get_frame",
    "vsfeel_bm3d_process_device",
    "vsfeel_bm3d_debug_groups",
    "vsfeel_bm3d_debug_estimate",
    "vsfeel_bm3d_set_timing",
    "vsfeel_bm3d_kernel_times",
    "vsfeel_bm3d_forward_saved",
    "vsfeel_bm3d_backward",
    "vsfeel_bm3d_set_graphs",
    "vsfeel_bm3d_process_device_static",
    "vsfeel_bm3d_fail_next_request",
)

KERNEL_NAMES = ("fill_zero", "bm3d_match", "bm3d_estimate", "bm3d_aggregate", "bm3d_tile_reduce",
                "bm3d_weight_box", "bm3d_gather", "bm3d_planes", "bm3d_match_temporal")


class Format(ctypes.Structure):
    _fields_ = [
        ("color_family", ctypes.c_int32),
        ("sample_type", ctypes.c_int32),
        ("bits_per_sample", ctypes.c_int32),
        ("sub_sampling_w", ctypes.c_int32),
        ("sub_sampling_h", ctypes.c_int32),
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("num_frames", ctypes.c_int32),
    ]


_D = ctypes.POINTER(ctypes.c_double)
_I64 = ctypes.POINTER(ctypes.c_int64)


class Args(ctypes.Structure):
    _fields_ = [
        ("clip", Format),
        ("has_ref", ctypes.c_int32),
        ("ref", Format),
        ("sigma", _D), ("num_sigma", ctypes.c_int32),
        ("block_step", _I64), ("num_block_step", ctypes.c_int32),
        ("bm_range", _I64), ("num_bm_range", ctypes.c_int32),
        ("radius", _I64), ("num_radius", ctypes.c_int32),
        ("ps_num", _I64), ("num_ps_num", ctypes.c_int32),
        ("ps_range", _I64), ("num_ps_range", ctypes.c_int32),
        ("num_streams", _I64), ("num_num_streams", ctypes.c_int32),
        ("extractor_exp", _I64), ("num_extractor_exp", ctypes.c_int32),
        ("fp_mode", ctypes.c_int32),
        ("device", ctypes.c_int32),
        ("precision", ctypes.c_int32),
    ]


class Info(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32), ("height", ctypes.c_int32), ("num_frames", ctypes.c_int32),
        ("chroma", ctypes.c_int32), ("final_pass", ctypes.c_int32),
        ("sigma", ctypes.c_float), ("extractor", ctypes.c_float),
        ("block_step", ctypes.c_int32), ("bm_range", ctypes.c_int32), ("radius", ctypes.c_int32),
        ("ps_num", ctypes.c_int32), ("ps_range", ctypes.c_int32), ("num_streams", ctypes.c_int32),
        ("temporal_width", ctypes.c_int32),
        ("ref_blocks_x", ctypes.c_int32), ("ref_blocks_y", ctypes.c_int32),
        ("fp_mode", ctypes.c_int32), ("device", ctypes.c_int32), ("pitch", ctypes.c_int32),
        ("precision", ctypes.c_int32),
    ]


class DeviceClip(ctypes.Structure):
    _fields_ = [
        ("data", ctypes.c_void_p),
        ("frame_stride", ctypes.c_int64),
        ("pitch", ctypes.c_int32),
        ("ref", ctypes.c_void_p),
    ]


class Saved(ctypes.Structure):
    _fields_ = [
        ("groups", ctypes.c_void_p),
        ("masks", ctypes.c_void_p),
        ("weights", ctypes.c_void_p),
        ("denom", ctypes.c_void_p),
        ("denom_pitch", ctypes.c_int32),
    ]


# vsfeel_bm3d_args.precision: the arithmetic of the block-matching distances
PRECISION = {"fp32": 0, "fp16": 1, "bf16": 2}


def precision_code(p) -> int:
    """0 / 1 / 2 from an int or one of "fp32", "fp16", "bf16"."""
    if isinstance(p, str):
        if p not in PRECISION:
            raise ValueError(f"precision must be one of {sorted(PRECISION)}")
        return PRECISION[p]
    return int(p)


_lib = None


def load() -> ctypes.CDLL:
    """The loaded library; raises when the gfx950 build is missing."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"vsfeel: {LIB_PATH} is not built; run __graft_entry__.build() "
            "(make -C vapoursynth-feel_amd/csrc)")
    L = ctypes.CDLL(LIB_PATH)
    vp = ctypes.c_void_p
    cp = ctypes.c_char_p
    sz = ctypes.c_size_t
    i32 = ctypes.c_int32
    fpp = ctypes.POINTER(ctypes.c_void_p)
    pdp = ctypes.POINTER(ctypes.c_ssize_t)
    L.vsfeel_bm3d_create.argtypes = [ctypes.POINTER(Args), ctypes.POINTER(vp), cp, sz]
    L.vsfeel_bm3d_create.restype = i32
    L.vsfeel_bm3d_free.argtypes = [vp]
    L.vsfeel_bm3d_free.restype = None
    L.vsfeel_bm3d_get_info.argtypes = [vp, ctypes.POINTER(Info)]
    L.vsfeel_bm3d_get_info.restype = i32
    L.vsfeel_bm3d_frame_window.argtypes = [vp, i32, ctypes.POINTER(i32), ctypes.POINTER(i32)]
    L.vsfeel_bm3d_frame_window.restype = i32
    L.vsfeel_bm3d_get_frame.argtypes = [vp, i32, fpp, pdp, fpp, pdp, vp, ctypes.c_ssize_t, cp, sz]
    L.vsfeel_bm3d_get_frame.restype = i32
    L.vsfeel_bm3d_process_device.argtypes = [vp, ctypes.POINTER(DeviceClip), i32, i32, vp, i32,
                                             ctypes.c_int64, vp, cp, sz]
    L.vsfeel_bm3d_process_device.restype = i32
    L.vsfeel_bm3d_process_device_static.argtypes = [ctypes.POINTER(Args), ctypes.POINTER(DeviceClip), i32, i32,
                                                    vp, i32, ctypes.c_int64, vp, cp, sz]
    L.vsfeel_bm3d_process_device_static.restype = i32
    L.vsfeel_bm3d_debug_groups.argtypes = [vp, ctypes.POINTER(DeviceClip), i32, vp, vp, cp, sz]
    L.vsfeel_bm3d_debug_groups.restype = i32
    L.vsfeel_bm3d_debug_estimate.argtypes = [vp, ctypes.POINTER(DeviceClip), i32, vp, i32, cp, sz]
    L.vsfeel_bm3d_debug_estimate.restype = i32
    L.vsfeel_bm3d_set_timing.argtypes = [vp, i32]
    L.vsfeel_bm3d_set_timing.restype = i32
    L.vsfeel_bm3d_set_graphs.argtypes = [vp, i32]
    L.vsfeel_bm3d_set_graphs.restype = i32
    L.vsfeel_bm3d_fail_next_request.argtypes = [vp, i32]
    L.vsfeel_bm3d_fail_next_request.restype = i32
    L.vsfeel_bm3d_kernel_times.argtypes = [vp, ctypes.POINTER(ctypes.c_double),
                                           ctypes.POINTER(ctypes.c_int64)]
    L.vsfeel_bm3d_kernel_times.restype = i32
    L.vsfeel_bm3d_forward_saved.argtypes = [vp, ctypes.POINTER(DeviceClip), i32, i32, vp, i32,
                                            ctypes.c_int64, ctypes.POINTER(Saved), vp, cp, sz]
    L.vsfeel_bm3d_forward_saved.restype = i32
    L.vsfeel_bm3d_backward.argtypes = [vp, vp, i32, ctypes.c_int64, i32, ctypes.POINTER(Saved), vp, i32,
                                       ctypes.c_int64, vp, cp, sz]
    L.vsfeel_bm3d_backward.restype = i32
    _lib = L
    return L


def err_buf():
    return ctypes.create_string_buffer(512)
