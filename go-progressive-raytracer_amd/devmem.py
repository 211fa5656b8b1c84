"""Device buffers, streams and events through the HIP runtime that libprt.so itself links
(ctypes). Used by the benchmark to keep inputs resident in HBM and to time kernels with HIP
events recorded on the same stream the library launches on."""
import ctypes as C

import numpy as np

import prtgpu as P

_hip = None


def hip():
    global _hip
    if _hip is None:
        P.lib()  # resolves libamdhip64.so.7 for libprt.so first
        _hip = C.CDLL("libamdhip64.so.7")
        vp = C.c_void_p
        for name, args in {
            "hipMalloc": [C.POINTER(vp), C.c_size_t],
            "hipFree": [vp],
            "hipMemcpy": [vp, vp, C.c_size_t, C.c_int],
            "hipMemset": [vp, C.c_int, C.c_size_t],
            "hipStreamCreate": [C.POINTER(vp)],
            "hipStreamDestroy": [vp],
            "hipStreamSynchronize": [vp],
            "hipDeviceSynchronize": [],
            "hipEventCreate": [C.POINTER(vp)],
            "hipEventDestroy": [vp],
            "hipEventRecord": [vp, vp],
            "hipEventSynchronize": [vp],
            "hipEventElapsedTime": [C.POINTER(C.c_float), vp, vp],
            "hipSetDevice": [C.c_int],
        }.items():
            f = getattr(_hip, name)
            f.restype = C.c_int
            f.argtypes = args
    return _hip


def _ok(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed with HIP error {rc}")


class DeviceBuffer:
    def __init__(self, nbytes):
        self.nbytes = nbytes
        p = C.c_void_p()
        _ok(hip().hipMalloc(C.byref(p), max(nbytes, 16)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        _ok(hip().hipMemcpy(C.c_void_p(b.ptr), a.ctypes.data_as(C.c_void_p), a.nbytes, 1), "hipMemcpy H2D")
        return b

    def to_array(self, dtype, count):
        out = np.empty(count, dtype)
        _ok(hip().hipMemcpy(out.ctypes.data_as(C.c_void_p), C.c_void_p(self.ptr), out.nbytes, 2), "hipMemcpy D2H")
        return out

    def free(self):
        if self.ptr:
            hip().hipFree(C.c_void_p(self.ptr))
            self.ptr = None

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    def __init__(self):
        s = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        _ok(hip().hipStreamSynchronize(C.c_void_p(self.handle)), "hipStreamSynchronize")


class Event:
    def __init__(self):
        e = C.c_void_p()
        _ok(hip().hipEventCreate(C.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream):
        _ok(hip().hipEventRecord(C.c_void_p(self.handle), C.c_void_p(stream.handle)), "hipEventRecord")

    def elapsed_ms(self, end):
        ms = C.c_float()
        _ok(hip().hipEventSynchronize(C.c_void_p(end.handle)), "hipEventSynchronize")
        _ok(hip().hipEventElapsedTime(C.byref(ms), C.c_void_p(self.handle), C.c_void_p(end.handle)),
            "hipEventElapsedTime")
        return ms.value


def copy_d2d(dst, src, nbytes):
    _ok(hip().hipMemcpy(C.c_void_p(dst.ptr), C.c_void_p(src.ptr), nbytes, 3), "hipMemcpy D2D")


def synchronize():
    _ok(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")
