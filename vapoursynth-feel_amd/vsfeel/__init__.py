"""vsfeel -- MI355X (gfx950) BM3Dv2, a drop-in for core.vsfeel.BM3Dv2.

The filter runs in libvsfeel_bm3d.so (HIP kernels + C ABI, declared in
include/vsfeel_bm3d.h).  This package mirrors the reference's plugin
interface for Python callers:

* ``vsfeel.BM3Dv2(clip, ...)`` on numpy-plane clips (``vsfeel.Clip``), with
  the reference's arguments, defaults and "BM3D: ..." errors;
* ``vsfeel.torch_bm3d.bm3d(x, ...)`` on device-resident torch tensors.
"""

from .bm3d import BM3Dv2, Clip, Error, GRAY, RGB, YUV, INTEGER, FLOAT, info  # noqa: F401

__all__ = ["BM3Dv2", "Clip", "Error", "GRAY", "RGB", "YUV", "INTEGER", "FLOAT", "info"]
