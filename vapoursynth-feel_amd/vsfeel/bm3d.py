"""Host-side mirror of ``core.vsfeel.BM3Dv2`` over the C ABI.

``BM3Dv2(clip, ...)`` takes the reference's arguments with their meaning,
defaults, validation order and error messages (src/bm3d.cpp:1043-1192,
registration string :1561-1571) and returns a clip whose frames are computed
on the GPU on request, like a VapourSynth filter node.  Clips are plain
numpy planes here: ``Clip.gray(array)`` for GrayS and ``Clip.yuv(y, u, v)``
for YUV float clips.  Only the luma plane is denoised; chroma planes are
passed through bit-identically (src/bm3d.cpp:715-726).
"""

from __future__ import annotations

import ctypes
from typing import Sequence

import numpy as np

from . import _lib

GRAY, RGB, YUV = 1, 2, 3
INTEGER, FLOAT = 0, 1

__all__ = ["Error", "Clip", "BM3Dv2", "GRAY", "RGB", "YUV", "INTEGER", "FLOAT"]


class Error(Exception):
    """Raised where the reference raises vapoursynth.Error."""


class Clip:
    """A clip of planar frames held as numpy arrays (num_frames, h, w)."""

    def __init__(self, planes: Sequence[np.ndarray], color_family: int, sample_type: int = FLOAT,
                 bits_per_sample: int = 32, sub_sampling_w: int = 0, sub_sampling_h: int = 0):
        # numpy arrays, or any plane with a shape and per-frame indexing
        self._planes = [p if hasattr(p, "shape") and hasattr(p, "__getitem__") else np.asarray(p) for p in planes]
        self.color_family = color_family
        self.sample_type = sample_type
        self.bits_per_sample = bits_per_sample
        self.sub_sampling_w = sub_sampling_w
        self.sub_sampling_h = sub_sampling_h

    @classmethod
    def gray(cls, luma: np.ndarray) -> "Clip":
        luma = np.asarray(luma)
        if luma.ndim == 2:
            luma = luma[None]
        st = FLOAT if luma.dtype.kind == "f" else INTEGER
        return cls([luma], GRAY, st, luma.dtype.itemsize * 8)

    @classmethod
    def yuv(cls, y: np.ndarray, u: np.ndarray, v: np.ndarray) -> "Clip":
        y, u, v = (np.asarray(p) for p in (y, u, v))
        ssw = int(round(np.log2(y.shape[-1] / u.shape[-1]))) if u.shape[-1] else 0
        ssh = int(round(np.log2(y.shape[-2] / u.shape[-2]))) if u.shape[-2] else 0
        st = FLOAT if y.dtype.kind == "f" else INTEGER
        return cls([y, u, v], YUV, st, y.dtype.itemsize * 8, ssw, ssh)

    @property
    def num_frames(self) -> int:
        return int(self._planes[0].shape[0])

    @property
    def width(self) -> int:
        return int(self._planes[0].shape[-1])

    @property
    def height(self) -> int:
        return int(self._planes[0].shape[-2])

    @property
    def num_planes(self) -> int:
        return len(self._planes)

    def get_frame(self, n: int) -> list[np.ndarray]:
        if not 0 <= n < self.num_frames:
            raise Error(f"frame {n} out of range")
        return [p[n] for p in self._planes]

    def plane_array(self, plane: int = 0) -> np.ndarray:
        """All frames of one plane as a (num_frames, h, w) array."""
        return np.stack([self.get_frame(n)[plane] for n in range(self.num_frames)])

    def _format(self) -> _lib.Format:
        return _lib.Format(self.color_family, self.sample_type, self.bits_per_sample,
                           self.sub_sampling_w, self.sub_sampling_h, self.width, self.height,
                           self.num_frames)


def _float_array(v):
    if v is None:
        return None
    if np.isscalar(v):
        v = [v]
    return np.ascontiguousarray(np.asarray(v, np.float64).ravel())


def _int_array(v):
    if v is None:
        return None
    if np.isscalar(v):
        v = [v]
    return np.ascontiguousarray(np.asarray(v, np.int64).ravel())


class _Bm3dClip(Clip):
    """Output node: frames are computed on request through the C ABI."""

    def __init__(self, src: Clip, ref: Clip | None, handle, keep):
        self._src = src
        self._ref = ref
        self._handle = handle
        self._keep = keep
        super().__init__(src._planes, src.color_family, src.sample_type, src.bits_per_sample,
                         src.sub_sampling_w, src.sub_sampling_h)

    def __del__(self):
        h = getattr(self, "_handle", None)
        # at interpreter exit the module globals may already be gone
        lib = getattr(_lib, "_lib", None) if _lib is not None else None
        if h and lib is not None:
            lib.vsfeel_bm3d_free(h)
        self._handle = None

    def get_frame(self, n: int) -> list[np.ndarray]:
        L = _lib.load()
        if not 0 <= n < self.num_frames:
            raise Error(f"frame {n} out of range")
        first = ctypes.c_int32()
        last = ctypes.c_int32()
        rc = L.vsfeel_bm3d_frame_window(self._handle, n, ctypes.byref(first), ctypes.byref(last))
        if rc != _lib.OK:
            raise Error("BM3D: frame window")
        frames = range(first.value, last.value + 1)
        luma = [np.ascontiguousarray(self._src._planes[0][f], np.float32) for f in frames]
        VP = ctypes.c_void_p
        planes = (VP * len(luma))(*[a.ctypes.data for a in luma])
        strides = (ctypes.c_ssize_t * len(luma))(*[a.strides[0] for a in luma])
        if self._ref is not None:
            rl = [np.ascontiguousarray(self._ref._planes[0][f], np.float32) for f in frames]
            rplanes = (VP * len(rl))(*[a.ctypes.data for a in rl])
            rstrides = (ctypes.c_ssize_t * len(rl))(*[a.strides[0] for a in rl])
        else:
            rl, rplanes, rstrides = None, None, None
        dst = np.empty((self.height, self.width), np.float32)
        err = _lib.err_buf()
        rc = L.vsfeel_bm3d_get_frame(self._handle, n, planes, strides, rplanes, rstrides,
                                     dst.ctypes.data, dst.strides[0], err, len(err))
        if rc != _lib.OK:
            raise Error(err.value.decode())
        out = [dst]
        for p in self._src._planes[1:]:
            out.append(np.array(p[n], copy=True))
        return out


def BM3Dv2(clip: Clip, ref: Clip | None = None, sigma=None, block_step=None, bm_range=None,
           radius=None, ps_num=None, ps_range=None, num_streams=None, extractor_exp=None, *,
           fp_mode: int = 1, device: int = 0, precision="fp32") -> Clip:
    """core.vsfeel.BM3Dv2 (README.md:44-77, src/bm3d.cpp:1558-1574).

    sigma (float[], default 3.0), block_step (int[], [1, 8], default 8),
    bm_range (int[], > 0, default 9), radius (int, [0, 4], default 0),
    ps_num (int[], [1, 8], default 2), ps_range (int[], > 0, default 4),
    num_streams (int, > 0, default 4), extractor_exp (int, default 0: off).
    ``ref`` switches to the final (Wiener) pass with block matching on ref.
    fp_mode and device are extensions (multiply-add rounding mode, HIP device),
    and so is precision: "fp32" (the reference's distances, default), "fp16" or
    "bf16" (MFMA distances with an FP32-rescored shortlist, radius 0-2).
    """
    L = _lib.load()
    args = _lib.Args()
    args.clip = clip._format()
    keep = []

    def put(name, arr, ctype):
        if arr is None:
            return
        keep.append(arr)
        setattr(args, name, arr.ctypes.data_as(ctypes.POINTER(ctype)))
        setattr(args, "num_" + name, int(arr.size))

    if ref is not None:
        args.has_ref = 1
        args.ref = ref._format()
    put("sigma", _float_array(sigma), ctypes.c_double)
    put("block_step", _int_array(block_step), ctypes.c_int64)
    put("bm_range", _int_array(bm_range), ctypes.c_int64)
    put("radius", _int_array(radius), ctypes.c_int64)
    put("ps_num", _int_array(ps_num), ctypes.c_int64)
    put("ps_range", _int_array(ps_range), ctypes.c_int64)
    put("num_streams", _int_array(num_streams), ctypes.c_int64)
    put("extractor_exp", _int_array(extractor_exp), ctypes.c_int64)
    args.fp_mode = fp_mode
    args.device = device
    args.precision = _lib.precision_code(precision)
    handle = ctypes.c_void_p()
    err = _lib.err_buf()
    rc = L.vsfeel_bm3d_create(ctypes.byref(args), ctypes.byref(handle), err, len(err))
    if rc != _lib.OK:
        raise Error(err.value.decode())
    return _Bm3dClip(clip, ref, handle, keep)


def set_graphs(node: Clip, requests: bool, device: bool = True) -> None:
    """Graph launches of a BM3Dv2 output node: per-request graphs of its frame
    requests (off by default) and graph replay of device batches (on by
    default at radius 0, off at radius > 0)."""
    if not isinstance(node, _Bm3dClip):
        raise TypeError("not a BM3Dv2 node")
    flags = (1 if device else 0) | (2 if requests else 0)
    if _lib.load().vsfeel_bm3d_set_graphs(node._handle, flags) != _lib.OK:
        raise Error("BM3D: set_graphs")


def info(node: Clip) -> dict:
    """Derived parameters of a BM3Dv2 output node (scaled sigma, grid, ...)."""
    if not isinstance(node, _Bm3dClip):
        raise TypeError("not a BM3Dv2 node")
    inf = _lib.Info()
    _lib.load().vsfeel_bm3d_get_info(node._handle, ctypes.byref(inf))
    return {name: getattr(inf, name) for name, _ in _lib.Info._fields_}
