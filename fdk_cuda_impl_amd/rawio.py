"""Raw float32 files of projections ([view][v][u]) and volumes ([z][y][x]),
little-endian with no header, and the device dump behind them (DumpManager,
src/util/YkIoDump.hpp:9-72: ``<prefix>_a<4-digit index>.raw`` / ``.txt``)."""
from __future__ import annotations

import ctypes
import os

import numpy as np

from . import _capi as C


def dump_name(prefix: str, index: int, ext: str = ".raw") -> str:
    return f"{prefix}_a{index:04d}{ext}"


def read_raw(path: str, shape, dtype=np.float32) -> np.ndarray:
    """Read a raw file as ``shape``; its size must match exactly."""
    count = int(np.prod(shape))
    data = np.fromfile(path, dtype=np.dtype(dtype).newbyteorder("<"))
    if data.size != count:
        raise ValueError(f"{path}: {data.size} values, expected {count} for shape {tuple(shape)}")
    return data.astype(dtype, copy=False).reshape(shape)


def write_raw(path: str, array) -> None:
    np.ascontiguousarray(array, dtype=np.dtype("<f4")).tofile(path)


class Dump:
    """Device buffers (torch CUDA float32 tensors or raw device pointers) to
    raw files under ``output_dir``; a disabled dump writes nothing."""

    def __init__(self, output_dir: str, enabled: bool = True):
        os.makedirs(output_dir, exist_ok=True)
        self.handle = ctypes.c_void_p()
        C.call("yk_dump_create", output_dir.encode(), 1 if enabled else 0, ctypes.byref(self.handle))

    def set_enabled(self, on: bool) -> None:
        C.call("yk_dump_set_enabled", self.handle, 1 if on else 0)

    @staticmethod
    def _src(src, count):
        if hasattr(src, "data_ptr"):
            if not src.is_cuda or not src.is_contiguous() or src.element_size() != 4:
                raise TypeError("dump needs a contiguous CUDA float32 tensor")
            return ctypes.c_void_p(src.data_ptr()), src.numel() if count is None else count
        if count is None:
            raise ValueError("a raw device pointer needs a count")
        return ctypes.c_void_p(int(src)), count

    def device(self, name: str, src, count: int | None = None, stream=None, index: int | None = None) -> None:
        ptr, n = self._src(src, count)
        st = ctypes.c_void_p(stream) if stream else None
        if index is None:
            C.call("yk_dump_device_f32", self.handle, name.encode(), ptr, n, st)
        else:
            C.call("yk_dump_device_f32_a", self.handle, name.encode(), index, ptr, n, st)

    def text(self, name: str, text: str, index: int | None = None) -> None:
        if index is None:
            C.call("yk_dump_text", self.handle, name.encode(), text.encode())
        else:
            C.call("yk_dump_text_a", self.handle, name.encode(), index, text.encode())

    def close(self) -> None:
        if self.handle:
            C.call("yk_dump_destroy", self.handle)
            self.handle = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
