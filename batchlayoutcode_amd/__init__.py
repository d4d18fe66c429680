"""MI355X-native BatchLayout: force-directed minibatch graph layout on HIP/CDNA4.

The compute path lives in ``lib/libbatchlayout.so`` (HIP kernels for gfx950 behind the
C ABI declared in ``include/batchlayout.h``); :mod:`batchlayoutcode_amd.api` binds it
with ctypes.  :mod:`batchlayoutcode_amd.graphs` generates the synthetic workloads.
"""
__all__ = ["graphs"]
