"""Minimal ctypes access to the HIP runtime that libcuda-crypt.so links
(/opt/rocm/lib/libamdhip64.so.7): device buffers, streams and events for the
device-resident entry points, the benchmark and the tests.

Using the library's own runtime keeps a single HIP runtime in the process
(PyTorch's wheel bundles another copy under the same soname, and the two do
not coexist).
"""
import ctypes
import os

import numpy as np

_CANDIDATES = ("/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so.7", "libamdhip64.so")


def _load():
    for c in _CANDIDATES:
        try:
            return ctypes.CDLL(c, mode=ctypes.RTLD_GLOBAL)
        except OSError:
            continue
    raise OSError("libamdhip64 not found")


_hip = _load()
_vp = ctypes.c_void_p
_hip.hipMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t]
_hip.hipFree.argtypes = [_vp]
_hip.hipMemcpy.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int]
_hip.hipMemcpyAsync.argtypes = [_vp, _vp, ctypes.c_size_t, ctypes.c_int, _vp]
_hip.hipMemsetAsync.argtypes = [_vp, ctypes.c_int, ctypes.c_size_t, _vp]
_hip.hipStreamCreate.argtypes = [ctypes.POINTER(_vp)]
_hip.hipStreamDestroy.argtypes = [_vp]
_hip.hipStreamSynchronize.argtypes = [_vp]
_hip.hipEventCreate.argtypes = [ctypes.POINTER(_vp)]
_hip.hipEventDestroy.argtypes = [_vp]
_hip.hipEventRecord.argtypes = [_vp, _vp]
_hip.hipEventSynchronize.argtypes = [_vp]
_hip.hipEventElapsedTime.argtypes = [ctypes.POINTER(ctypes.c_float), _vp, _vp]
_hip.hipGetErrorString.restype = ctypes.c_char_p
_hip.hipGetDeviceCount.argtypes = [ctypes.POINTER(ctypes.c_int)]
_hip.hipSetDevice.argtypes = [ctypes.c_int]
_hip.hipDeviceGetAttribute.argtypes = [ctypes.POINTER(ctypes.c_int), ctypes.c_int, ctypes.c_int]
_hip.hipHostMalloc.argtypes = [ctypes.POINTER(_vp), ctypes.c_size_t, ctypes.c_uint]
_hip.hipHostFree.argtypes = [_vp]

H2D, D2H, D2D = 1, 2, 3


class HipError(RuntimeError):
    pass


def check(rc, what=""):
    if rc != 0:
        raise HipError(f"{what}: {_hip.hipGetErrorString(rc).decode()} ({rc})")


def device_count():
    n = ctypes.c_int(0)
    rc = _hip.hipGetDeviceCount(ctypes.byref(n))
    return n.value if rc == 0 else 0


def synchronize():
    check(_hip.hipDeviceSynchronize(), "hipDeviceSynchronize")


class DeviceBuffer:
    """A hipMalloc'd buffer; `ptr` is the device address as an int."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = _vp()
        check(_hip.hipMalloc(ctypes.byref(p), max(1, self.nbytes)), "hipMalloc")
        self.ptr = p.value

    @classmethod
    def from_array(cls, a):
        a = np.ascontiguousarray(a)
        b = cls(a.nbytes)
        if a.nbytes:
            check(_hip.hipMemcpy(b.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        return b

    def upload(self, a, stream=None):
        a = np.ascontiguousarray(a)
        assert a.nbytes <= self.nbytes
        if stream is None:
            check(_hip.hipMemcpy(self.ptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")
        else:
            check(_hip.hipMemcpyAsync(self.ptr, a.ctypes.data, a.nbytes, H2D, stream.handle),
                  "hipMemcpyAsync H2D")

    def download(self, dtype=np.uint8, count=None):
        dtype = np.dtype(dtype)
        count = self.nbytes // dtype.itemsize if count is None else count
        out = np.empty(count, dtype)
        if out.nbytes:
            check(_hip.hipMemcpy(out.ctypes.data, self.ptr, out.nbytes, D2H), "hipMemcpy D2H")
        return out

    def zero(self, stream=None):
        check(_hip.hipMemsetAsync(self.ptr, 0, self.nbytes, stream.handle if stream else None),
              "hipMemsetAsync")

    def free(self):
        if self.ptr:
            _hip.hipFree(self.ptr)
            self.ptr = 0

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class _PinnedBlock:
    """hipHostMalloc'd bytes, freed when the last array viewing them goes."""

    def __init__(self, nbytes):
        p = _vp()
        check(_hip.hipHostMalloc(ctypes.byref(p), max(1, nbytes), 0), "hipHostMalloc")
        self.ptr = p.value

    def __del__(self):
        try:
            if self.ptr:
                _hip.hipHostFree(self.ptr)
                self.ptr = 0
        except Exception:
            pass


def pinned_copy(a):
    """A uint8 copy of `a` in pinned host memory (hipHostMalloc), as a numpy
    array that keeps the allocation alive."""
    a = np.ascontiguousarray(a, dtype=np.uint8).reshape(-1)
    block = _PinnedBlock(a.nbytes)
    buf = (ctypes.c_uint8 * max(1, a.nbytes)).from_address(block.ptr)
    buf._block = block
    out = np.frombuffer(buf, dtype=np.uint8, count=a.nbytes)
    out[:] = a
    return out


class Stream:
    def __init__(self):
        s = _vp()
        check(_hip.hipStreamCreate(ctypes.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def synchronize(self):
        check(_hip.hipStreamSynchronize(self.handle), "hipStreamSynchronize")


class Event:
    def __init__(self):
        e = _vp()
        check(_hip.hipEventCreate(ctypes.byref(e)), "hipEventCreate")
        self.handle = e.value

    def record(self, stream=None):
        """stream: a Stream, a raw hipStream_t handle, or None (null stream)."""
        check(_hip.hipEventRecord(self.handle, _stream_handle(stream)), "hipEventRecord")

    def synchronize(self):
        check(_hip.hipEventSynchronize(self.handle), "hipEventSynchronize")

    def elapsed_ms(self, end):
        ms = ctypes.c_float(0)
        check(_hip.hipEventElapsedTime(ctypes.byref(ms), self.handle, end.handle),
              "hipEventElapsedTime")
        return ms.value


def _stream_handle(stream):
    if stream is None:
        return None
    return stream.handle if hasattr(stream, "handle") else stream


def upload_to(dptr, a):
    """Synchronous copy of a host array to a raw device address."""
    a = np.ascontiguousarray(a)
    if a.nbytes:
        check(_hip.hipMemcpy(dptr, a.ctypes.data, a.nbytes, H2D), "hipMemcpy H2D")


def download_from(dptr, count, dtype=np.uint8):
    """Synchronous copy of `count` elements from a raw device address."""
    out = np.empty(count, np.dtype(dtype))
    if out.nbytes:
        check(_hip.hipMemcpy(out.ctypes.data, dptr, out.nbytes, D2H), "hipMemcpy D2H")
    return out


def copy_d2d(dst, src, nbytes, stream=None):
    check(_hip.hipMemcpyAsync(dst.ptr, src.ptr, nbytes, D2D, stream.handle if stream else None),
          "hipMemcpyAsync D2D")


# hipDeviceAttribute_t values (hip/hip_runtime_api.h, ROCm 7)
HIP_ATTR_CUS = 63   # hipDeviceAttributeMultiprocessorCount


def device_attribute(attr, device=0):
    v = ctypes.c_int(0)
    check(_hip.hipDeviceGetAttribute(ctypes.byref(v), attr, device), "hipDeviceGetAttribute")
    return v.value
