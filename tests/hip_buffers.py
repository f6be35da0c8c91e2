"""Device buffers and streams for GPU tests of the device-pointer entry points,
through the HIP runtime that libtwos_hip.so itself links (its libamdhip64
soname, already loaded with the library), so every pointer and stream
belongs to the library's own runtime."""

import ctypes as C

import numpy as np

from twos import _capi

H2D, D2H, D2D = 1, 2, 3
_hip = None


def hip() -> C.CDLL:
    global _hip
    if _hip is None:
        _capi.load()
        for name in ("libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so.7", "libamdhip64.so"):
            try:
                _hip = C.CDLL(name)
                break
            except OSError:
                continue
        if _hip is None:
            raise OSError("libamdhip64 not found")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipHostMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t, C.c_uint]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipHostFree.argtypes = [C.c_void_p]
        _hip.hipMemcpyAsync.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int, C.c_void_p]
        _hip.hipMemsetAsync.argtypes = [C.c_void_p, C.c_int, C.c_size_t, C.c_void_p]
        _hip.hipStreamCreate.argtypes = [C.POINTER(C.c_void_p)]
        _hip.hipStreamDestroy.argtypes = [C.c_void_p]
        _hip.hipStreamSynchronize.argtypes = [C.c_void_p]
    return _hip


def _ok(rc: int, what: str) -> None:
    if rc != 0:
        raise RuntimeError(f"{what} failed: hipError {rc}")


class Stream:
    def __init__(self):
        self.handle = C.c_void_p()
        _ok(hip().hipStreamCreate(C.byref(self.handle)), "hipStreamCreate")

    @property
    def value(self) -> int:
        return self.handle.value or 0

    def synchronize(self) -> None:
        _ok(hip().hipStreamSynchronize(self.handle), "hipStreamSynchronize")

    def close(self) -> None:
        if self.handle.value:
            hip().hipStreamDestroy(self.handle)
            self.handle = C.c_void_p()


class Buffer:
    """nbytes of device memory with a pinned host mirror."""

    def __init__(self, nbytes: int, fill: int | None = 0):
        self.nbytes = int(nbytes)
        self.dev, self.host = C.c_void_p(), C.c_void_p()
        _ok(hip().hipMalloc(C.byref(self.dev), max(self.nbytes, 1)), "hipMalloc")
        _ok(hip().hipHostMalloc(C.byref(self.host), max(self.nbytes, 1), 0), "hipHostMalloc")
        if fill is not None:
            _ok(hip().hipMemsetAsync(self.dev, fill, self.nbytes, None), "hipMemsetAsync")
            hip().hipStreamSynchronize(None)

    @property
    def ptr(self) -> int:
        return self.dev.value

    def host_view(self, dtype=np.uint8) -> np.ndarray:
        return np.ctypeslib.as_array(C.cast(self.host, C.POINTER(C.c_uint8)), (self.nbytes,)).view(dtype)

    def upload(self, data: np.ndarray, stream: Stream) -> None:
        """Queue host data -> device on the stream (through the pinned mirror)."""
        raw = np.ascontiguousarray(data).view(np.uint8).ravel()
        assert raw.nbytes <= self.nbytes
        self.host_view()[: raw.nbytes] = raw
        _ok(hip().hipMemcpyAsync(self.dev, self.host, raw.nbytes, H2D, stream.handle), "hipMemcpyAsync")

    def download(self, stream: Stream) -> bytes:
        """Queue device -> host on the stream, wait for the stream, return the bytes."""
        _ok(hip().hipMemcpyAsync(self.host, self.dev, self.nbytes, D2H, stream.handle), "hipMemcpyAsync")
        stream.synchronize()
        return self.host_view().tobytes()

    def copy_from(self, other: "Buffer", stream: Stream) -> None:
        n = min(self.nbytes, other.nbytes)
        _ok(hip().hipMemcpyAsync(self.dev, other.dev, n, D2D, stream.handle), "hipMemcpyAsync")

    def close(self) -> None:
        if self.dev.value:
            hip().hipFree(self.dev)
            self.dev = C.c_void_p()
        if self.host.value:
            hip().hipHostFree(self.host)
            self.host = C.c_void_p()
