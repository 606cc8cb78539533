"""ctypes loading of the in-tree native libraries (no fallback of any kind)."""
from __future__ import annotations

import ctypes
import pathlib

PKG = pathlib.Path(__file__).resolve().parent


class NativeLibraryMissing(RuntimeError):
    pass


_cache: dict[str, ctypes.CDLL] = {}


def load(name: str) -> ctypes.CDLL:
    """Load gip_amd/lib<name>.so; raise when it has not been built."""
    if name in _cache:
        return _cache[name]
    path = PKG / f"lib{name}.so"
    if not path.exists():
        raise NativeLibraryMissing(
            f"{path} is not built; run `python -c 'import __graft_entry__ as g; g.build()'` first")
    lib = ctypes.CDLL(str(path), mode=ctypes.RTLD_GLOBAL)
    _cache[name] = lib
    return lib


c_int = ctypes.c_int
c_int32 = ctypes.c_int32
c_int64 = ctypes.c_int64
c_double = ctypes.c_double
c_float = ctypes.c_float
c_void_p = ctypes.c_void_p
c_char_p = ctypes.c_char_p
c_size_t = ctypes.c_size_t
POINTER = ctypes.POINTER
byref = ctypes.byref
