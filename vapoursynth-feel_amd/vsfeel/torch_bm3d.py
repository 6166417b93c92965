"""BM3Dv2 on device-resident torch tensors (PyTorch-ROCm), over the C ABI.

``bm3d(x, ...)`` denoises a (num_frames, height, width) float32 tensor that
lives on a HIP device, enqueueing the whole clip's launch sequence on the
current torch stream.  PyTorch only supplies device memory and the stream;
every kernel is the library's own.
"""

from __future__ import annotations

import ctypes
from collections import OrderedDict

import torch

from . import _lib
from .bm3d import Error, GRAY, FLOAT

__all__ = ["bm3d", "Instance"]


class Instance:
    """One native BM3Dv2 instance for a fixed clip shape and parameter set."""

    def __init__(self, num_frames: int, height: int, width: int, *, final: bool = False,
                 sigma=3.0, block_step=8, bm_range=9, radius=0, ps_num=2, ps_range=4,
                 num_streams=4, extractor_exp=0, fp_mode=1, device=0, precision="fp32"):
        L = _lib.load()
        fmt = _lib.Format(GRAY, FLOAT, 32, 0, 0, width, height, num_frames)
        a = _lib.Args()
        a.clip = fmt
        if final:
            a.has_ref = 1
            a.ref = fmt
        self._keep = []

        def put(name, values, ctype, np_dtype):
            import numpy as np
            arr = np.ascontiguousarray(np.atleast_1d(np.asarray(values, np_dtype)))
            self._keep.append(arr)
            setattr(a, name, arr.ctypes.data_as(ctypes.POINTER(ctype)))
            setattr(a, "num_" + name, int(arr.size))

        import numpy as np
        put("sigma", sigma, ctypes.c_double, np.float64)
        for name, v in (("block_step", block_step), ("bm_range", bm_range), ("radius", radius),
                        ("ps_num", ps_num), ("ps_range", ps_range), ("num_streams", num_streams),
                        ("extractor_exp", extractor_exp)):
            put(name, v, ctypes.c_int64, np.int64)
        a.fp_mode = fp_mode
        a.device = device
        a.precision = _lib.precision_code(precision)
        h = ctypes.c_void_p()
        err = _lib.err_buf()
        rc = L.vsfeel_bm3d_create(ctypes.byref(a), ctypes.byref(h), err, len(err))
        if rc != _lib.OK:
            raise Error(err.value.decode())
        self.handle = h
        self.num_frames, self.height, self.width = num_frames, height, width
        self.final = final
        self.device = device
        inf = _lib.Info()
        L.vsfeel_bm3d_get_info(h, ctypes.byref(inf))
        self.info = {name: getattr(inf, name) for name, _ in _lib.Info._fields_}

    def __del__(self):
        h = getattr(self, "handle", None)
        # at interpreter exit the module globals may already be gone
        lib = getattr(_lib, "_lib", None) if _lib is not None else None
        if h and lib is not None:
            lib.vsfeel_bm3d_free(h)
        self.handle = None

    def _clip(self, x: torch.Tensor, ref: torch.Tensor | None) -> _lib.DeviceClip:
        _check(x, self)
        c = _lib.DeviceClip()
        c.data = x.data_ptr()
        c.frame_stride = x.stride(0)
        c.pitch = x.stride(1)
        if self.final:
            if ref is None:
                raise Error("BM3D: the final pass needs the ref clip")
            _check(ref, self)
            if ref.stride() != x.stride():
                raise Error("BM3D: ref must have the layout of the clip")
            c.ref = ref.data_ptr()
        return c

    def process(self, x: torch.Tensor, ref: torch.Tensor | None = None, first: int = 0,
                count: int | None = None, out: torch.Tensor | None = None) -> torch.Tensor:
        count = self.num_frames - first if count is None else count
        c = self._clip(x, ref)
        if out is None:
            out = torch.empty((count, self.height, self.width), dtype=torch.float32, device=x.device)
        if out.shape[0] < count or out.shape[1:] != (self.height, self.width) or out.stride(2) != 1:
            raise Error("BM3D: bad output tensor")
        stream = torch.cuda.current_stream(x.device).cuda_stream
        err = _lib.err_buf()
        rc = _lib.load().vsfeel_bm3d_process_device(
            self.handle, ctypes.byref(c), first, count, out.data_ptr(), out.stride(1),
            out.stride(0), ctypes.c_void_p(stream), err, len(err))
        if rc != _lib.OK:
            raise Error(err.value.decode())
        return out

    # ------------------------------------------------------- differentiable form
    def forward_saved(self, x: torch.Tensor):
        """Forward pass that also returns the state its backward needs:
        (out, (groups int32 [N, nref, 8], masks int64 [N, nref, 8],
        weights float32 [N, nref], denom float32 [N, H, W]))."""
        n = self.num_frames
        nref = self.info["ref_blocks_x"] * self.info["ref_blocks_y"]
        dev = x.device
        groups = torch.empty((n, nref, 8), dtype=torch.int32, device=dev)
        masks = torch.empty((n, nref, 8), dtype=torch.int64, device=dev)
        weights = torch.empty((n, nref), dtype=torch.float32, device=dev)
        denom = torch.empty((n, self.height, self.width), dtype=torch.float32, device=dev)
        out = torch.empty((n, self.height, self.width), dtype=torch.float32, device=dev)
        c = self._clip(x, None)
        sv = self._saved(groups, masks, weights, denom)
        stream = torch.cuda.current_stream(dev).cuda_stream
        err = _lib.err_buf()
        rc = _lib.load().vsfeel_bm3d_forward_saved(
            self.handle, ctypes.byref(c), 0, n, out.data_ptr(), out.stride(1), out.stride(0),
            ctypes.byref(sv), ctypes.c_void_p(stream), err, len(err))
        if rc != _lib.OK:
            raise Error(err.value.decode())
        return out, (groups, masks, weights, denom)

    def backward(self, grad_out: torch.Tensor, saved) -> torch.Tensor:
        """d(loss)/dx = A (grad_out / denom) with the saved groups, masks, weights."""
        groups, masks, weights, denom = saved
        g = grad_out.contiguous()
        _check(g, self)
        gin = torch.empty_like(g)
        sv = self._saved(groups, masks, weights, denom)
        stream = torch.cuda.current_stream(g.device).cuda_stream
        err = _lib.err_buf()
        rc = _lib.load().vsfeel_bm3d_backward(
            self.handle, g.data_ptr(), g.stride(1), g.stride(0), self.num_frames, ctypes.byref(sv),
            gin.data_ptr(), gin.stride(1), gin.stride(0), ctypes.c_void_p(stream), err, len(err))
        if rc != _lib.OK:
            raise Error(err.value.decode())
        return gin

    @staticmethod
    def _saved(groups, masks, weights, denom) -> _lib.Saved:
        for t in (groups, masks, weights, denom):
            if not t.is_contiguous():
                raise Error("BM3D: saved state must be contiguous")
        sv = _lib.Saved()
        sv.groups = groups.data_ptr()
        sv.masks = masks.data_ptr()
        sv.weights = weights.data_ptr()
        sv.denom = denom.data_ptr()
        sv.denom_pitch = denom.stride(1)
        return sv

    def set_timing(self, enable: bool) -> None:
        """Record HIP events around every kernel launch (resets the totals)."""
        _lib.load().vsfeel_bm3d_set_timing(self.handle, int(bool(enable)))

    def set_graphs(self, enable: bool) -> None:
        """Replay process() from a captured HIP graph while its arguments repeat (default on at
        radius 0; off at radius > 0, where direct launches overlap the two lanes better)."""
        _lib.load().vsfeel_bm3d_set_graphs(self.handle, int(bool(enable)))

    def kernel_times(self) -> dict:
        """{kernel: (total ms, launches)} of the launches recorded so far."""
        n = len(_lib.KERNEL_NAMES)
        ms = (ctypes.c_double * n)()
        cnt = (ctypes.c_int64 * n)()
        rc = _lib.load().vsfeel_bm3d_kernel_times(self.handle, ms, cnt)
        if rc != _lib.OK:
            raise Error("BM3D: kernel timing failed")
        return {name: (ms[i], cnt[i]) for i, name in enumerate(_lib.KERNEL_NAMES)}

    def groups(self, x: torch.Tensor, m: int, ref: torch.Tensor | None = None):
        """(ny, nx, 8, 3) int32 group table and (ny, nx, 8) distances of frame m."""
        import numpy as np
        c = self._clip(x, ref)
        torch.cuda.current_stream(x.device).synchronize()
        ny, nx = self.info["ref_blocks_y"], self.info["ref_blocks_x"]
        g = np.zeros((ny, nx, 8, 3), np.int32)
        d = np.zeros((ny, nx, 8), np.float32)
        err = _lib.err_buf()
        rc = _lib.load().vsfeel_bm3d_debug_groups(self.handle, ctypes.byref(c), m, g.ctypes.data,
                                                  d.ctypes.data, err, len(err))
        if rc != _lib.OK:
            raise Error(err.value.decode())
        return g, d

    def estimate(self, x: torch.Tensor, m: int, ref: torch.Tensor | None = None) -> torch.Tensor:
        """Estimate stack of frame m: (2r+1, 2, height, width) float32 on the device."""
        c = self._clip(x, ref)
        torch.cuda.current_stream(x.device).synchronize()
        tw = self.info["temporal_width"]
        pitch = (self.width + 3) // 4 * 4
        res = torch.empty((tw, 2, self.height, pitch), dtype=torch.float32, device=x.device)
        err = _lib.err_buf()
        rc = _lib.load().vsfeel_bm3d_debug_estimate(self.handle, ctypes.byref(c), m, res.data_ptr(),
                                                    pitch, err, len(err))
        if rc != _lib.OK:
            raise Error(err.value.decode())
        return res[..., :self.width]


def _check(x: torch.Tensor, inst: Instance):
    if not isinstance(x, torch.Tensor) or x.device.type != "cuda":
        raise Error("BM3D: expected a float32 tensor on a HIP device")
    if x.dtype != torch.float32:
        raise Error("BM3D: only constant format 32 bit float input supported")
    if tuple(x.shape) != (inst.num_frames, inst.height, inst.width) or x.stride(2) != 1:
        raise Error("BM3D: tensor shape/layout does not match the instance")


_cache: "OrderedDict[tuple, Instance]" = OrderedDict()


def _instance(x: torch.Tensor, final: bool, kw: dict) -> Instance:
    key = (tuple(x.shape), x.device.index, final, tuple(sorted(kw.items())))
    inst = _cache.get(key)
    if inst is None:
        inst = Instance(x.shape[0], x.shape[1], x.shape[2], final=final,
                        device=x.device.index or 0, **kw)
        _cache[key] = inst
        while len(_cache) > 4:
            _cache.popitem(last=False)
    else:
        _cache.move_to_end(key)
    return inst


def bm3d(x: torch.Tensor, ref: torch.Tensor | None = None, *, sigma=3.0, block_step=8, bm_range=9,
         radius=0, ps_num=2, ps_range=4, extractor_exp=0, fp_mode=1, precision="fp32") -> torch.Tensor:
    """Denoise every frame of x (N, H, W); ref selects the Wiener final pass."""
    if x.dim() == 2:
        x = x.unsqueeze(0)
        if ref is not None and ref.dim() == 2:
            ref = ref.unsqueeze(0)
    kw = dict(sigma=sigma, block_step=block_step, bm_range=bm_range, radius=radius,
              ps_num=ps_num, ps_range=ps_range, extractor_exp=extractor_exp, fp_mode=fp_mode,
              precision=precision)
    inst = _instance(x, ref is not None, kw)
    return inst.process(x, ref)
