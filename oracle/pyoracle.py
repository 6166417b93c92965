"""ctypes bindings of the BM3Dv2 CPU oracle -- test infrastructure only.

Loaded by tests/, the smoke check and the benchmark's CPU-baseline leg as
the checker.  The product package never imports this module.
"""

from __future__ import annotations

import ctypes
import os
import struct
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libbm3d_oracle.so")

# sigma scaling of the reference's BM3DCreate (src/bm3d.cpp:1105-1110)
SIGMA_FACTOR_BASIC = struct.unpack("<f", struct.pack("<I", 0x3F021BB6))[0]
SIGMA_FACTOR_FINAL = struct.unpack("<f", struct.pack("<I", 0x3E40C0C1))[0]


class Params(ctypes.Structure):
    _fields_ = [
        ("width", ctypes.c_int32),
        ("height", ctypes.c_int32),
        ("pitch", ctypes.c_int32),
        ("block_step", ctypes.c_int32),
        ("bm_range", ctypes.c_int32),
        ("radius", ctypes.c_int32),
        ("ps_num", ctypes.c_int32),
        ("ps_range", ctypes.c_int32),
        ("sigma", ctypes.c_float),
        ("extractor", ctypes.c_float),
        ("final_pass", ctypes.c_int32),
        ("fp_mode", ctypes.c_int32),
        ("nthreads", ctypes.c_int32),
    ]


def build() -> str:
    subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            build()
        L = ctypes.CDLL(LIB_PATH)
        fp = ctypes.POINTER(ctypes.c_float)
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int32)
        pp = ctypes.POINTER(Params)
        fpp = ctypes.POINTER(fp)
        L.bm3d_oracle_ref_grid.argtypes = [pp, ip, ip]
        L.bm3d_oracle_ref_grid.restype = ctypes.c_int32
        L.bm3d_oracle_agg_z.argtypes = [ctypes.c_int32] * 4
        L.bm3d_oracle_agg_z.restype = ctypes.c_int32
        L.bm3d_oracle_f32_estimate.argtypes = [pp, fpp, fpp, fp, ip]
        L.bm3d_oracle_f32_estimate.restype = ctypes.c_int32
        L.bm3d_oracle_f64_estimate.argtypes = [pp, fpp, fpp, dp, ip]
        L.bm3d_oracle_f64_estimate.restype = ctypes.c_int32
        L.bm3d_oracle_f32_clip.argtypes = [pp, ctypes.c_int32, fp, fp, ctypes.c_int32, ctypes.c_int32, fp]
        L.bm3d_oracle_f32_clip.restype = ctypes.c_int32
        L.bm3d_oracle_f64_clip.argtypes = [pp, ctypes.c_int32, fp, fp, ctypes.c_int32, ctypes.c_int32, dp]
        L.bm3d_oracle_f64_clip.restype = ctypes.c_int32
        L.bm3d_oracle_f32_dct8.argtypes = [fp, ctypes.c_int32, ctypes.c_int32]
        L.bm3d_oracle_f64_dct8.argtypes = [dp, ctypes.c_int32, ctypes.c_int32]
        L.bm3d_oracle_f32_group.argtypes = [fp, fp, ctypes.c_float, ctypes.c_int32, ip]
        L.bm3d_oracle_f32_group.restype = ctypes.c_float
        L.bm3d_oracle_f64_group.argtypes = [dp, dp, ctypes.c_double, ctypes.c_int32, ip]
        L.bm3d_oracle_f64_group.restype = ctypes.c_double
        L.bm3d_oracle_f32_window_dist.argtypes = [fp] + [ctypes.c_int32] * 10 + [fp]
        L.bm3d_oracle_f32_window_dist.restype = ctypes.c_int32
        _lib = L
    return _lib


def _fptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_float))


def _dptr(a: np.ndarray):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def make_params(width, height, *, sigma=3.0, block_step=8, bm_range=9, radius=0, ps_num=2,
                ps_range=4, extractor_exp=0, final=False, fp_mode=1, nthreads=1,
                pitch=None) -> Params:
    """Parameters as the reference's BM3DCreate derives them from the user's
    arguments (sigma scaling, extractor = 2^exp or 0)."""
    factor = SIGMA_FACTOR_FINAL if final else SIGMA_FACTOR_BASIC
    s = float(np.float32(np.float32(sigma) * np.float32(factor)))
    e = float(np.ldexp(np.float32(1.0), extractor_exp)) if extractor_exp != 0 else 0.0
    return Params(width, height, width if pitch is None else pitch, block_step, bm_range,
                  radius, ps_num, ps_range, s, e, int(bool(final)), fp_mode, nthreads)


def ref_grid(p: Params):
    nx = ctypes.c_int32()
    ny = ctypes.c_int32()
    n = lib().bm3d_oracle_ref_grid(ctypes.byref(p), ctypes.byref(nx), ctypes.byref(ny))
    return n, nx.value, ny.value


def agg_z(i, n, nframes, radius):
    return lib().bm3d_oracle_agg_z(i, n, nframes, radius)


def estimate(p: Params, search: list[np.ndarray], inputs: list[np.ndarray] | None = None,
             f64: bool = False, want_groups: bool = True):
    """One estimation dispatch.  Returns (res [tw,2,H,pitch], groups [ny,nx,8,3])."""
    tw = 2 * p.radius + 1
    assert len(search) == tw
    search = [np.ascontiguousarray(s, np.float32) for s in search]
    FP = ctypes.POINTER(ctypes.c_float)
    sarr = (FP * tw)(*[_fptr(s) for s in search])
    if inputs is not None:
        inputs = [np.ascontiguousarray(s, np.float32) for s in inputs]
        iarr = (FP * tw)(*[_fptr(s) for s in inputs])
    else:
        iarr = None
    n, nx, ny = ref_grid(p)
    groups = np.zeros((ny, nx, 8, 3), np.int32) if want_groups else None
    gptr = groups.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)) if want_groups else None
    if f64:
        res = np.zeros((tw, 2, p.height, p.pitch), np.float64)
        rc = lib().bm3d_oracle_f64_estimate(ctypes.byref(p), sarr, iarr, _dptr(res), gptr)
    else:
        res = np.zeros((tw, 2, p.height, p.pitch), np.float32)
        rc = lib().bm3d_oracle_f32_estimate(ctypes.byref(p), sarr, iarr, _fptr(res), gptr)
    assert rc == 0
    return res, groups


def clip(p: Params, frames: np.ndarray, ref: np.ndarray | None = None, first: int = 0,
         count: int | None = None, f64: bool = False) -> np.ndarray:
    """Whole-clip output frames [first, first+count) of a (N,H,pitch) clip."""
    frames = np.ascontiguousarray(frames, np.float32)
    nf = frames.shape[0]
    count = nf - first if count is None else count
    refp = _fptr(np.ascontiguousarray(ref, np.float32)) if ref is not None else None
    if ref is not None:
        ref = np.ascontiguousarray(ref, np.float32)
        refp = _fptr(ref)
    if f64:
        out = np.zeros((count, p.height, p.pitch), np.float64)
        rc = lib().bm3d_oracle_f64_clip(ctypes.byref(p), nf, _fptr(frames), refp, first, count, _dptr(out))
    else:
        out = np.zeros((count, p.height, p.pitch), np.float32)
        rc = lib().bm3d_oracle_f32_clip(ctypes.byref(p), nf, _fptr(frames), refp, first, count, _fptr(out))
    assert rc == 0
    return out


def dct8(v, inverse=False, fp_mode=1, f64=False):
    if f64:
        a = np.array(v, np.float64)
        lib().bm3d_oracle_f64_dct8(_dptr(a), int(inverse), fp_mode)
    else:
        a = np.array(v, np.float32)
        lib().bm3d_oracle_f32_dct8(_fptr(a), int(inverse), fp_mode)
    return a


def group(cube: np.ndarray, sigma: float, ref_cube: np.ndarray | None = None, fp_mode=1,
          f64=False):
    """Collaborative filter of one [8,8,8] group; returns (out, weight, kept)."""
    kept = ctypes.c_int32()
    if f64:
        a = np.array(cube, np.float64)
        r = None if ref_cube is None else np.array(ref_cube, np.float64)
        w = lib().bm3d_oracle_f64_group(_dptr(a), None if r is None else _dptr(r), sigma,
                                        fp_mode, ctypes.byref(kept))
    else:
        a = np.array(cube, np.float32)
        r = None if ref_cube is None else np.array(ref_cube, np.float32)
        w = lib().bm3d_oracle_f32_group(_fptr(a), None if r is None else _fptr(r), sigma,
                                        fp_mode, ctypes.byref(kept))
    return a, w, kept.value


def window_dist(plane: np.ndarray, x, y, rng, cx0, cy0, exact=False, fp_mode=1):
    plane = np.ascontiguousarray(plane, np.float32)
    h, w = plane.shape
    left, right = max(cx0 - rng, 0), min(cx0 + rng, w - 8)
    top, bottom = max(cy0 - rng, 0), min(cy0 + rng, h - 8)
    out = np.zeros((bottom - top + 1, right - left + 1), np.float32)
    lib().bm3d_oracle_f32_window_dist(_fptr(plane), w, w, h, x, y, rng, cx0, cy0, int(exact),
                                      fp_mode, _fptr(out))
    return out, (left, top)
