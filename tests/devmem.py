"""Device buffers for GPU tests, through the HIP runtime that
libslaph_amd.so itself links (torch bundles a second HIP runtime, which
cannot initialise next to it in one process).  Test plumbing only."""
import ctypes

import numpy as np

_hip = None


def _lib():
    global _hip
    if _hip is None:
        import slaph  # noqa: F401  (loads libslaph_amd.so and with it its HIP runtime)
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    return _hip


class DeviceArray:
    """A device copy of a numpy array (hipMalloc + hipMemcpy)."""

    def __init__(self, host: np.ndarray):
        self.host = np.ascontiguousarray(host)
        p = ctypes.c_void_p()
        rc = _lib().hipMalloc(ctypes.byref(p), max(1, self.host.nbytes))
        if rc != 0:
            raise RuntimeError("hipMalloc failed (%d)" % rc)
        self.ptr = p.value
        self._copy(self.ptr, self.host.ctypes.data, 1)  # hipMemcpyHostToDevice

    def _copy(self, dst, src, kind):
        if self.host.nbytes and _lib().hipMemcpy(dst, src, self.host.nbytes, kind) != 0:
            raise RuntimeError("hipMemcpy failed")

    def get(self) -> np.ndarray:
        out = np.empty_like(self.host)
        self._copy(out.ctypes.data, self.ptr, 2)  # hipMemcpyDeviceToHost
        return out

    def free(self):
        if self.ptr:
            _lib().hipFree(self.ptr)
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass
