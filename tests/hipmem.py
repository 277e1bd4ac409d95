"""Device buffers and streams for the GPU tests, through the HIP runtime that
libcuda_replace.so itself loads (ctypes; no torch).

torch ships its own libamdhip64.so with the same SONAME as ROCm's, so whichever
is loaded first serves the whole process; a test that loads the library first
and torch later gets a torch on the wrong runtime.  These helpers avoid torch."""
import ctypes as C

_hip = None


def hip():
    global _hip
    if _hip is None:
        from cuda_replace_amd.wrapper import DEFAULT_LIB
        C.CDLL(DEFAULT_LIB)                      # loads (or reuses) the library's HIP runtime
        h = C.CDLL("libamdhip64.so.7")           # the already-loaded runtime, by SONAME
        h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        h.hipFree.argtypes = [C.c_void_p]
        h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
        h.hipDeviceSynchronize.argtypes = []
        h.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        h.hipStreamDestroy.argtypes = [C.c_void_p]
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipMemset", "hipDeviceSynchronize", "hipStreamCreate",
                  "hipStreamDestroy"):
            getattr(h, f).restype = C.c_int
        _hip = h
    return _hip


H2D, D2H = 1, 2


def check(rc, what):
    if rc != 0:
        raise RuntimeError(f"{what} failed rc={rc}")


def synchronize():
    check(hip().hipDeviceSynchronize(), "hipDeviceSynchronize")


class DevBuf:
    """nbytes of device memory (hipMalloc: 256-byte aligned), optionally filled
    with one byte value."""

    def __init__(self, nbytes: int, fill=None):
        self.n = int(nbytes)
        p = C.c_void_p()
        check(hip().hipMalloc(C.byref(p), max(self.n, 1)), "hipMalloc")
        self.ptr = p.value
        if fill is not None:
            check(hip().hipMemset(self.ptr, int(fill), max(self.n, 1)), "hipMemset")

    @classmethod
    def of(cls, data: bytes, extra: int = 0) -> "DevBuf":
        b = cls(len(data) + extra)
        b.upload(data)
        return b

    def upload(self, data: bytes, offset: int = 0):
        if data:
            check(hip().hipMemcpy(self.ptr + offset, C.c_char_p(data), len(data), H2D), "hipMemcpy H2D")

    def download(self, n: int, offset: int = 0) -> bytes:
        out = bytearray(n)
        if n:
            check(hip().hipMemcpy((C.c_char * n).from_buffer(out), self.ptr + offset, n, D2H), "hipMemcpy D2H")
        return bytes(out)

    def free(self):
        if self.ptr:
            hip().hipFree(self.ptr)
            self.ptr = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.free()

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Stream:
    """A HIP stream at the default priority."""

    def __init__(self):
        s = C.c_void_p()
        check(hip().hipStreamCreate(C.byref(s)), "hipStreamCreate")
        self.handle = s.value

    def destroy(self):
        if self.handle:
            hip().hipStreamDestroy(self.handle)
            self.handle = None
