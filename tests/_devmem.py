"""Device buffers for GPU tests of the C ABI's device-memory calls, from the HIP runtime the
library itself links (the test allocates; the entry points under test get plain pointers)."""
import ctypes as C

import numpy as np

from crender_amd import lib

H2D, D2H = 1, 2


class DeviceArrays:
    """Device buffers from the HIP runtime the library itself uses (the test allocates; the
    entry point under test does not)."""

    def __init__(self):
        lib()   # libcrt.so first: the runtime it links is the one looked up here by soname
        self.hip = C.CDLL("libamdhip64.so.7")
        self.hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        self.hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.hip.hipFree.argtypes = [C.c_void_p]
        self.ptrs = []

    def put(self, a):
        a = np.ascontiguousarray(a)
        p = C.c_void_p()
        assert self.hip.hipMalloc(C.byref(p), max(a.nbytes, 4)) == 0
        assert self.hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        self.ptrs.append(p)
        return p.value

    def get(self, ptr, n, dtype):
        out = np.empty(n, dtype)
        assert self.hip.hipMemcpy(out.ctypes.data, C.c_void_p(ptr), out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.ptrs:
            self.hip.hipFree(p)
        self.ptrs = []
