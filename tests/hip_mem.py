"""Device memory for tests that act as a caller owning its buffers: hipMalloc / hipMemcpy /
hipMemset from the HIP runtime libpfh.so is linked with (the same loaded library, found by its
soname), so that a pointer handed to libpfh belongs to the runtime that launches its kernels."""
import ctypes

import numpy as np

_H2D, _D2H = 1, 2
_hip = None


def _lib():
    global _hip
    if _hip is None:
        import pfh

        pfh.load()  # libpfh first: its libamdhip64 is the one looked up below
        _hip = ctypes.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        _hip.hipFree.argtypes = [ctypes.c_void_p]
        _hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        _hip.hipMemset.argtypes = [ctypes.c_void_p, ctypes.c_int, ctypes.c_size_t]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset"):
            getattr(_hip, f).restype = ctypes.c_int
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP status {rc}")


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        _ok(_lib().hipMalloc(ctypes.byref(p), self.nbytes), "hipMalloc")
        self.ptr = int(p.value)

    def fill(self, byte):
        _ok(_lib().hipMemset(self.ptr, int(byte), self.nbytes), "hipMemset")

    def upload(self, host):
        a = np.ascontiguousarray(host).view(np.uint8).reshape(-1)
        assert a.size == self.nbytes
        _ok(_lib().hipMemcpy(self.ptr, a.ctypes.data, self.nbytes, _H2D), "hipMemcpy(H2D)")

    def download(self):
        out = np.empty(self.nbytes, np.uint8)
        _ok(_lib().hipMemcpy(out.ctypes.data, self.ptr, self.nbytes, _D2H), "hipMemcpy(D2H)")
        return out

    def free(self):
        if self.ptr:
            _ok(_lib().hipFree(self.ptr), "hipFree")
            self.ptr = 0
