"""Python view of the operator side (include/gip_cavity.h): generate hex
lid-driven cavities, read/write/merge cases in the reference's InputFiles
layout, and expose the operators as numpy CSR triples."""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from ._lib import POINTER, byref, c_double, c_int32, load

OPS = ["D", "Ec", "L", "Lb", "Mx", "My", "Mz", "Gx", "Gy", "Gz", "SFx", "SFy", "SFz", "SO", "Mres"]
OP_BITS = {"D": 1 << 0, "Ec": 1 << 1, "L": 1 << 2, "Lb": 1 << 3, "M": 1 << 4, "G": 1 << 5, "SF": 1 << 6,
           "SO": 1 << 7, "Mres": 1 << 8}
OP_ALL = 0x1FF


class _Csr(ctypes.Structure):
    _fields_ = [("nnz", c_int32), ("ncols", c_int32), ("nrows", c_int32), ("val", POINTER(c_double)),
                ("col", POINTER(c_int32)), ("rowptr", POINTER(c_int32))]


class _Desc(ctypes.Structure):
    _fields_ = [("nx", c_int32), ("ny", c_int32), ("nz", c_int32), ("beta", c_double), ("gamma", c_double),
                ("lid_u", c_double), ("lid_profile", POINTER(c_double)), ("pin", c_int32),
                ("ops", ctypes.c_uint32)]


class _Cavity(ctypes.Structure):
    _fields_ = ([("n_inner", c_int32), ("n_ghost", c_int32), ("n_all", c_int32), ("n_faces", c_int32),
                 ("nx", c_int32), ("ny", c_int32), ("nz", c_int32)]
                + [(op, _Csr) for op in OPS]
                + [("u", POINTER(c_double)), ("v", POINTER(c_double)), ("w", POINTER(c_double)),
                   ("dxs", POINTER(c_double)), ("vol", POINTER(c_double)), ("lid_ghost", POINTER(c_int32))])


def _lib():
    lib = load("gip_cavity")
    if not getattr(lib, "_gip_typed", False):
        lib.gip_cavity_generate.argtypes = [POINTER(_Desc), POINTER(POINTER(_Cavity))]
        lib.gip_cavity_write.argtypes = [POINTER(_Cavity), ctypes.c_char_p, ctypes.c_char_p]
        lib.gip_cavity_read.argtypes = [ctypes.c_char_p, ctypes.c_char_p, POINTER(POINTER(_Cavity))]
        lib.gip_cavity_merge.argtypes = [ctypes.c_char_p, ctypes.c_char_p, c_int32, POINTER(POINTER(_Cavity))]
        lib.gip_cavity_split.argtypes = [POINTER(_Cavity), c_int32, ctypes.c_char_p, ctypes.c_char_p]
        lib.gip_cavity_free.argtypes = [POINTER(_Cavity)]
        lib.gip_cavity_free.restype = None
        lib.gip_cavity_last_error.restype = ctypes.c_char_p
        lib.gip_cavity_rhs_normal.argtypes = [c_int32, ctypes.c_uint64, POINTER(c_double)]
        lib._gip_typed = True
    return lib


class CavityError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise CavityError(f"gip_cavity error {rc}: {_lib().gip_cavity_last_error().decode()}")


@dataclass
class CSR:
    rowptr: np.ndarray
    col: np.ndarray
    val: np.ndarray
    nrows: int
    ncols: int

    @property
    def nnz(self):
        return int(self.val.size)

    def to_scipy(self, ncols=None):
        import scipy.sparse as sp
        return sp.csr_matrix((self.val, self.col, self.rowptr), shape=(self.nrows, ncols or self.ncols))


class Cavity:
    """Owns a native gip_cavity; arrays are numpy views (copied on request)."""

    def __init__(self, ptr):
        self._p = ptr
        c = ptr.contents
        self.n_inner, self.n_ghost, self.n_all, self.n_faces = c.n_inner, c.n_ghost, c.n_all, c.n_faces
        self.nx, self.ny, self.nz = c.nx, c.ny, c.nz

    def __del__(self):
        p = getattr(self, "_p", None)
        if p:
            _lib().gip_cavity_free(p)
            self._p = None

    @property
    def ptr(self):
        return self._p

    def op(self, name) -> CSR | None:
        a = getattr(self._p.contents, name)
        if not a.rowptr:
            return None
        rp = np.ctypeslib.as_array(a.rowptr, shape=(a.nrows + 1,))
        nnz = a.nnz
        col = np.ctypeslib.as_array(a.col, shape=(nnz,)) if nnz else np.zeros(0, np.int32)
        val = np.ctypeslib.as_array(a.val, shape=(nnz,)) if nnz else np.zeros(0, np.float64)
        return CSR(rp, col, val, a.nrows, a.ncols)

    def field(self, name) -> np.ndarray:
        p = getattr(self._p.contents, name)
        n = self.n_inner if name == "vol" else self.n_all
        if name == "lid_ghost":
            n = self.nx * self.ny
        return np.ctypeslib.as_array(p, shape=(n,)) if p else None

    def write(self, root, name):
        _check(_lib().gip_cavity_write(self._p, str(root).encode(), name.encode()))

    def write_partitions(self, root, name, nparts):
        """Write as an nparts-partition case (<name>_<nparts>p-<rank> files)."""
        _check(_lib().gip_cavity_split(self._p, nparts, str(root).encode(), name.encode()))


def generate(n, ny=None, nz=None, *, beta=0.0, gamma=1e-2, lid_u=1.0, lid_profile=None, pin=True, ops=OP_ALL):
    d = _Desc()
    d.nx, d.ny, d.nz = n, ny or n, nz or n
    d.beta, d.gamma, d.lid_u, d.pin, d.ops = beta, gamma, lid_u, 1 if pin else 0, ops
    keep = None
    if lid_profile is not None:
        keep = np.ascontiguousarray(lid_profile, dtype=np.float64)
        d.lid_profile = keep.ctypes.data_as(POINTER(c_double))
    out = POINTER(_Cavity)()
    _check(_lib().gip_cavity_generate(byref(d), byref(out)))
    return Cavity(out)


def read(root, name):
    out = POINTER(_Cavity)()
    _check(_lib().gip_cavity_read(str(root).encode(), name.encode(), byref(out)))
    return Cavity(out)


def merge(root, name, nparts):
    out = POINTER(_Cavity)()
    _check(_lib().gip_cavity_merge(str(root).encode(), name.encode(), nparts, byref(out)))
    return Cavity(out)


def rhs_normal(n, seed):
    out = np.empty(n, dtype=np.float64)
    _check(_lib().gip_cavity_rhs_normal(n, seed, out.ctypes.data_as(POINTER(c_double))))
    return out
