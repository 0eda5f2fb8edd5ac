"""Device buffers for GPU tests that call device-pointer entry points.

Test plumbing: hipMalloc / hipMemcpy / hipFree through ctypes on the HIP
runtime libes_hip.so itself links (libamdhip64.so.7), so the pointers belong
to the runtime the library uses.  (torch ships its own HIP runtime; two
runtimes in one process do not share a device.)
"""
import ctypes as C

import numpy as np

_H2D, _D2H = 1, 2


class Hip:
    def __init__(self):
        import elasticscattering_amd as es

        es.load_library()  # loads libamdhip64.so.7 first; dlopen below returns that same runtime
        self.lib = C.CDLL("libamdhip64.so.7")
        self.lib.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.lib.hipFree.argtypes = [C.c_void_p]
        self.lib.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.lib.hipDeviceSynchronize.argtypes = []
        self.held = []

    def _ok(self, rc, what):
        if rc != 0:
            raise RuntimeError(f"{what} failed: hipError {rc}")

    def alloc(self, nbytes):
        p = C.c_void_p()
        self._ok(self.lib.hipMalloc(C.byref(p), max(int(nbytes), 8)), "hipMalloc")
        self.held.append(p)
        return p

    def upload(self, a):
        a = np.ascontiguousarray(a)
        p = self.alloc(a.nbytes)
        if a.nbytes:
            self._ok(self.lib.hipMemcpy(p, a.ctypes.data, a.nbytes, _H2D), "hipMemcpy H2D")
        return p

    def fill(self, nbytes, value, dtype):
        return self.upload(np.full(max(int(nbytes) // np.dtype(dtype).itemsize, 1), value, dtype=dtype))

    def download(self, p, count, dtype):
        out = np.empty(int(count), dtype=dtype)
        if out.nbytes:
            self._ok(self.lib.hipMemcpy(out.ctypes.data, p, out.nbytes, _D2H), "hipMemcpy D2H")
        return out

    def sync(self):
        self._ok(self.lib.hipDeviceSynchronize(), "hipDeviceSynchronize")

    def close(self):
        for p in self.held:
            self.lib.hipFree(p)
        self.held = []
