"""Device buffers from the HIP runtime through ctypes, for calls that take
device pointers (the viewer feed). Allocation happens here, in the caller's
code, never inside the engine's entry points."""
import ctypes

_HIP = None
_D2H = 2   # hipMemcpyDeviceToHost


def _hip():
    global _HIP
    if _HIP is None:
        try:
            h = ctypes.CDLL("libamdhip64.so")
        except OSError:
            h = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
        h.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        h.hipMalloc.restype = ctypes.c_int
        h.hipFree.argtypes = [ctypes.c_void_p]
        h.hipFree.restype = ctypes.c_int
        h.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        h.hipMemcpy.restype = ctypes.c_int
        _HIP = h
    return _HIP


class DeviceBuffer:
    """nbytes of device memory on the current HIP device; .ptr is its address."""

    def __init__(self, nbytes):
        self.nbytes = int(nbytes)
        p = ctypes.c_void_p()
        rc = _hip().hipMalloc(ctypes.byref(p), self.nbytes)
        if rc != 0:
            raise RuntimeError("hipMalloc(%d) failed with HIP error %d" % (self.nbytes, rc))
        self.ptr = p.value

    def copy_to(self, out):
        """Copy the first out.nbytes bytes into the contiguous numpy array out."""
        if out.nbytes > self.nbytes or not out.flags["C_CONTIGUOUS"]:
            raise ValueError("destination larger than the buffer or not contiguous")
        rc = _hip().hipMemcpy(ctypes.c_void_p(out.ctypes.data), ctypes.c_void_p(self.ptr), out.nbytes, _D2H)
        if rc != 0:
            raise RuntimeError("hipMemcpy failed with HIP error %d" % rc)

    def close(self):
        if self.ptr:
            _hip().hipFree(ctypes.c_void_p(self.ptr))
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()
