"""Python view of the gfx950 kernel library's C ABI (include/gip_hip.h).

Thin ctypes wrappers: a Device (context + stream), DeviceArray buffers
moved to/from numpy, Matrix operators, the vector operations and the fused
PCG.  Every call goes through libgip_hip.so; there is no other code path.
"""
from __future__ import annotations

import ctypes
import weakref
from dataclasses import dataclass

import numpy as np

from ._lib import POINTER, byref, c_double, c_float, c_int, c_int32, c_int64, c_size_t, c_void_p, load

FMT_AUTO, FMT_CSR, FMT_ELL, FMT_DIAG, FMT_STENCIL = 0, 1, 2, 3, 4
FORMAT_NAMES = {FMT_CSR: "csr", FMT_ELL: "sliced-ell", FMT_DIAG: "diag", FMT_STENCIL: "stencil"}
MAT_FP32 = 1 << 8
MAT_ELL_SLICE32 = 1 << 9
MAT_ELL_SLICE128 = 1 << 10
ELL_SLICE_FLAGS = {32: MAT_ELL_SLICE32, 64: 0, 128: MAT_ELL_SLICE128}
MAT_NO_SYM = 1 << 11
SYM_STATES = {0: "none", 1: "in use", -1: "not symmetric", 2: "not made"}
PCG_FP64, PCG_MIXED = 0, 1
OK, EINVAL, EHIP, ENOMEM, EBREAKDOWN, ENOCONV = 0, -1, -2, -3, -4, -5


class PcgOpts(ctypes.Structure):
    _fields_ = [("maxit", c_int32), ("tol", c_double), ("precond", c_int32), ("precision", c_int32),
                ("inner_tol", c_double), ("poll", c_int32), ("coarse", c_int32), ("time_products", c_int32)]


class PcgStats(ctypes.Structure):
    _fields_ = [("iters", c_int32), ("corrections", c_int32), ("relres", c_double), ("status", c_int32),
                ("product_ms", c_double), ("product_samples", c_int32)]


class TuneReport(ctypes.Structure):
    _fields_ = [("chosen_format", c_int32), ("chosen_lanes", c_int32), ("chosen_slice", c_int32),
                ("us", c_float * 8)]


# exported entry points and their argument types (kept in step with include/gip_hip.h)
SIGNATURES = {
    "gip_last_error": ([], ctypes.c_char_p),
    "gip_init": ([c_int, POINTER(c_void_p)], c_int),
    "gip_finalize": ([c_void_p], c_int),
    "gip_sync": ([c_void_p], c_int),
    "gip_use_device": ([c_int], c_int),
    "gip_stream": ([c_void_p], c_void_p),
    "gip_device_count": ([POINTER(c_int)], c_int),
    "gip_malloc": ([c_void_p, c_size_t, POINTER(c_void_p)], c_int),
    "gip_free": ([c_void_p, c_void_p], c_int),
    "gip_memcpy_h2d": ([c_void_p, c_void_p, c_void_p, c_size_t], c_int),
    "gip_memcpy_d2h": ([c_void_p, c_void_p, c_void_p, c_size_t], c_int),
    "gip_memcpy_d2d": ([c_void_p, c_void_p, c_void_p, c_size_t], c_int),
    "gip_event_create": ([c_void_p, POINTER(c_void_p)], c_int),
    "gip_event_record": ([c_void_p, c_void_p], c_int),
    "gip_event_elapsed_ms": ([c_void_p, c_void_p, POINTER(c_float)], c_int),
    "gip_event_destroy": ([c_void_p], c_int),
    "gip_mat_create": ([c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_int32,
                        POINTER(c_void_p)], c_int),
    "gip_mat_destroy": ([c_void_p], c_int),
    "gip_mat_info": ([c_void_p, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32), POINTER(c_int64)], c_int),
    "gip_mat_cols_needed": ([c_void_p, POINTER(c_int32)], c_int),
    "gip_mat_product_bytes": ([c_void_p, c_int32, POINTER(c_int64)], c_int),
    "gip_mat_sym_state": ([c_void_p, POINTER(c_int32), POINTER(c_int64)], c_int),
    "gip_mat_set_lanes": ([c_void_p, c_int32], c_int),
    "gip_mat_autotune": ([c_void_p, c_int32, c_int32, c_int32, c_void_p, c_void_p, c_void_p, c_int32, c_void_p,
                          c_void_p, c_int32, POINTER(c_void_p), c_void_p], c_int),
    "gip_mat_get_values": ([c_void_p, c_void_p], c_int),
    "gip_mat_set_values": ([c_void_p, c_void_p], c_int),
    "gip_mat_diag_inv": ([c_void_p, c_void_p], c_int),
    "gip_dcsrmv": ([c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "gip_dcsrmv_ab": ([c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double], c_int),
    "gip_dcsrmv_vals": ([c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "gip_dcsrmv_vals_ab": ([c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double], c_int),
    "gip_dcsrmv_t": ([c_void_p, c_void_p, c_void_p, c_void_p, c_int32], c_int),
    "gip_scsrmv": ([c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "gip_dfill": ([c_void_p, c_int32, c_void_p, c_double], c_int),
    "gip_daxpy": ([c_void_p, c_int32, c_double, c_void_p, c_void_p], c_int),
    "gip_daxpby": ([c_void_p, c_int32, c_double, c_void_p, c_double, c_void_p], c_int),
    "gip_dcopy": ([c_void_p, c_int32, c_void_p, c_void_p], c_int),
    "gip_dcopy_indexed": ([c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p], c_int),
    "gip_ddot": ([c_void_p, c_int32, c_void_p, c_void_p, POINTER(c_double)], c_int),
    "gip_dnrm_inf": ([c_void_p, c_int32, c_void_p, POINTER(c_double)], c_int),
    "gip_cfl": ([c_void_p, c_int32, c_void_p, c_void_p, c_void_p, c_void_p, c_double, c_double,
                 POINTER(c_double)], c_int),
    "gip_ddot_dev": ([c_void_p, c_int32, c_void_p, c_void_p, c_void_p], c_int),
    "gip_dnrm_inf_dev": ([c_void_p, c_int32, c_void_p, c_void_p], c_int),
    "gip_pcg_create": ([c_void_p, c_void_p, c_int32, POINTER(c_void_p)], c_int),
    "gip_pcg_solve": ([c_void_p, c_void_p, c_void_p, c_void_p, POINTER(PcgOpts), POINTER(PcgStats)], c_int),
    "gip_pcg_destroy": ([c_void_p], c_int),
    "gip_stream_probe": ([c_void_p, c_int32, c_int32, c_int32, c_int32, c_size_t, c_void_p, c_void_p, c_void_p,
                          c_void_p, c_size_t, c_int32, POINTER(c_float)], c_int),
    "gip_barrier_probe": ([c_void_p, c_int32, c_int32, c_int32, POINTER(c_float)], c_int),
    "gip_pcg_solve_noalloc": ([c_void_p, c_void_p, c_int32, c_void_p, c_void_p, c_void_p, POINTER(PcgOpts),
                               POINTER(PcgStats)], c_int),
}


def lib():
    L = load("gip_hip")
    if not getattr(L, "_gip_typed", False):
        for name, (args, res) in SIGNATURES.items():
            f = getattr(L, name)
            f.argtypes = args
            f.restype = res
        L._gip_typed = True
    return L


class HipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"gip_hip error {code}: {msg}")
        self.code = code


def check(rc):
    if rc != 0:
        raise HipError(rc, lib().gip_last_error().decode())
    return rc


class Device:
    def __init__(self, device: int = 0):
        self.ctx = c_void_p()
        self._children = weakref.WeakSet()
        check(lib().gip_init(device, byref(self.ctx)))

    def _adopt(self, obj):
        self._children.add(obj)

    def close(self):
        """Release every live buffer, operator and solver, then the context."""
        if self.ctx:
            for ch in list(self._children):
                ch._release()
            lib().gip_finalize(self.ctx)
            self.ctx = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def sync(self):
        check(lib().gip_sync(self.ctx))

    def array(self, n, dtype=np.float64):
        return DeviceArray(self, n, dtype)

    def to_device(self, host):
        host = np.ascontiguousarray(host)
        a = DeviceArray(self, host.size, host.dtype)
        a.upload(host)
        return a


class DeviceArray:
    def __init__(self, dev: Device, n: int, dtype=np.float64):
        self.dev, self.n, self.dtype = dev, int(n), np.dtype(dtype)
        self.ptr = c_void_p()
        check(lib().gip_malloc(dev.ctx, max(1, self.n) * self.dtype.itemsize, byref(self.ptr)))
        dev._adopt(self)

    @property
    def nbytes(self):
        return self.n * self.dtype.itemsize

    def upload(self, host):
        host = np.ascontiguousarray(host, dtype=self.dtype)
        assert host.size == self.n, (host.size, self.n)
        check(lib().gip_memcpy_h2d(self.dev.ctx, self.ptr, host.ctypes.data, self.nbytes))
        return self

    def download(self):
        out = np.empty(self.n, dtype=self.dtype)
        check(lib().gip_memcpy_d2h(self.dev.ctx, out.ctypes.data, self.ptr, self.nbytes))
        return out

    def free(self):
        if self.ptr and self.dev.ctx:
            lib().gip_free(self.dev.ctx, self.ptr)
        self.ptr = c_void_p()

    _release = free

    def __del__(self):
        try:
            self.free()
        except Exception:
            pass


class Matrix:
    """A device operator built from host CSR arrays (int32 rowptr/col, float64 val)."""

    def __init__(self, dev: Device, rowptr, col, val, nrows=None, ncols=0, fmt=FMT_AUTO, fp32=False, slice=64,
                 sym=True):
        self.dev = dev
        rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        col = np.ascontiguousarray(col, dtype=np.int32)
        val = np.ascontiguousarray(val, dtype=np.float64)
        self.nrows = int(nrows if nrows is not None else rowptr.size - 1)
        self.nnz = int(val.size)
        self.ptr = c_void_p()
        if slice not in ELL_SLICE_FLAGS:
            raise ValueError("slice: rows per ELL slice, 32, 64 or 128")
        flags = fmt | (MAT_FP32 if fp32 else 0) | ELL_SLICE_FLAGS[slice] | (0 if sym else MAT_NO_SYM)
        check(lib().gip_mat_create(dev.ctx, self.nrows, int(ncols), self.nnz, rowptr.ctypes.data, col.ctypes.data,
                                   val.ctypes.data, flags, byref(self.ptr)))
        dev._adopt(self)

    @classmethod
    def from_csr(cls, dev, csr, **kw):
        return cls(dev, csr.rowptr, csr.col, csr.val, nrows=csr.nrows, ncols=csr.ncols, **kw)

    @classmethod
    def autotune(cls, dev, csr, reps=5, fp32=False):
        """Fastest of the candidate layouts (gip_mat_autotune); returns (Matrix, report dict)."""
        rowptr = np.ascontiguousarray(csr.rowptr, dtype=np.int32)
        col = np.ascontiguousarray(csr.col, dtype=np.int32)
        val = np.ascontiguousarray(csr.val, dtype=np.float64)
        need = int(col.max()) + 1 if col.size else 1
        x = dev.to_device(np.ones(max(need, 1)))
        y = dev.array(max(csr.nrows, 1))
        self = cls.__new__(cls)
        self.dev, self.nrows, self.nnz = dev, int(csr.nrows), int(val.size)
        self.ptr = c_void_p()
        rep = TuneReport()
        check(lib().gip_mat_autotune(dev.ctx, self.nrows, int(csr.ncols), self.nnz, rowptr.ctypes.data,
                                     col.ctypes.data, val.ctypes.data, MAT_FP32 if fp32 else 0, x.ptr, y.ptr, reps,
                                     byref(self.ptr), ctypes.cast(byref(rep), c_void_p)))
        dev._adopt(self)
        return self, {"format": rep.chosen_format, "lanes": rep.chosen_lanes, "slice": rep.chosen_slice,
                      "us": list(rep.us)}

    def set_lanes(self, lanes):
        check(lib().gip_mat_set_lanes(self.ptr, lanes))

    def info(self):
        nr, nz, fm, st = c_int32(), c_int32(), c_int32(), c_int64()
        check(lib().gip_mat_info(self.ptr, byref(nr), byref(nz), byref(fm), byref(st)))
        return {"nrows": nr.value, "nnz": nz.value, "format": fm.value, "stored": st.value}

    def product_bytes(self, fp32=False):
        """Bytes one product moves in this operator's layout (gip_mat_product_bytes)."""
        b = c_int64()
        check(lib().gip_mat_product_bytes(self.ptr, 1 if fp32 else 0, byref(b)))
        return b.value

    def sym_state(self):
        """(state, bytes) of the symmetric form (gip_mat_sym_state): state 0
        none, 2 not made yet, 1 in use, -1 values not symmetric; bytes one
        FP64 product moves with it."""
        st, b = c_int32(), c_int64()
        check(lib().gip_mat_sym_state(self.ptr, byref(st), byref(b)))
        return st.value, b.value

    def values(self):
        out = np.empty(self.nnz, np.float64)
        check(lib().gip_mat_get_values(self.ptr, out.ctypes.data))
        return out

    def set_values(self, vals):
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        check(lib().gip_mat_set_values(self.ptr, vals.ctypes.data))

    @property
    def cols_needed(self):
        n = c_int32()
        check(lib().gip_mat_cols_needed(self.ptr, byref(n)))
        return n.value

    def _shapes(self, x, y, dtype=np.float64):
        # host-side bounds check before any launch: the kernels trust these lengths
        if x.dtype != dtype or y.dtype != dtype:
            raise ValueError("vector dtype mismatch")
        if x.n < self.cols_needed:
            raise ValueError(f"input has {x.n} entries, operator reads up to column {self.cols_needed - 1}")
        if y.n < self.nrows:
            raise ValueError(f"output has {y.n} entries, operator has {self.nrows} rows")

    def spmv(self, x: DeviceArray, y: DeviceArray):
        self._shapes(x, y)
        check(lib().gip_dcsrmv(self.dev.ctx, self.ptr, x.ptr, y.ptr))

    def spmv_ab(self, x, y, alpha, beta):
        self._shapes(x, y)
        check(lib().gip_dcsrmv_ab(self.dev.ctx, self.ptr, x.ptr, y.ptr, alpha, beta))

    def spmv_vals(self, x, B: "Matrix"):
        if x.n < self.cols_needed or self.nrows != B.nnz:
            raise ValueError("spmv_vals: input too short or rows(A) != nnz(B)")
        check(lib().gip_dcsrmv_vals(self.dev.ctx, self.ptr, x.ptr, B.ptr))

    def spmv_vals_ab(self, x, B: "Matrix", alpha, beta):
        if x.n < self.cols_needed or self.nrows != B.nnz:
            raise ValueError("spmv_vals_ab: input too short or rows(A) != nnz(B)")
        check(lib().gip_dcsrmv_vals_ab(self.dev.ctx, self.ptr, x.ptr, B.ptr, alpha, beta))

    def spmv_t(self, x, y, ncols_out):
        if x.n < self.nrows or y.n < ncols_out or ncols_out < self.cols_needed:
            raise ValueError("spmv_t: vector lengths do not cover the operator")
        check(lib().gip_dcsrmv_t(self.dev.ctx, self.ptr, x.ptr, y.ptr, ncols_out))

    def spmv32(self, x, y):
        self._shapes(x, y, np.float32)
        check(lib().gip_scsrmv(self.dev.ctx, self.ptr, x.ptr, y.ptr))

    def diag_inv(self, out: DeviceArray):
        check(lib().gip_mat_diag_inv(self.ptr, out.ptr))

    def destroy(self):
        for s in list(getattr(self, "_solvers", [])):
            s.destroy()
        if self.ptr and self.dev.ctx:
            lib().gip_mat_destroy(self.ptr)
        self.ptr = c_void_p()

    _release = destroy

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


def _span(n, *arrays):
    """Host-side bounds check: every vector covers the n entries a kernel touches."""
    for a in arrays:
        if a.n < n or a.dtype != np.float64:
            raise ValueError(f"vector of {a.n} {a.dtype} entries used over {n} FP64 entries")
    return n


def fill(x: DeviceArray, value, n=None):
    n = _span(x.n if n is None else n, x)
    check(lib().gip_dfill(x.dev.ctx, n, x.ptr, value))


def axpy(alpha, x, y, n=None):
    n = _span(y.n if n is None else n, x, y)
    check(lib().gip_daxpy(y.dev.ctx, n, alpha, x.ptr, y.ptr))


def axpby(alpha, x, beta, y, n=None):
    n = _span(y.n if n is None else n, x, y)
    check(lib().gip_daxpby(y.dev.ctx, n, alpha, x.ptr, beta, y.ptr))


def copy(x, y, n=None):
    """y[0:n] = x[0:n] (GIP_VectorMulticore.cpp:167-181)."""
    n = _span(y.n if n is None else n, x, y)
    check(lib().gip_dcopy(y.dev.ctx, n, x.ptr, y.ptr))


def dot(x, y, n=None):
    n = _span(x.n if n is None else n, x, y)
    r = c_double()
    check(lib().gip_ddot(x.dev.ctx, n, x.ptr, y.ptr, byref(r)))
    return r.value


def nrm_inf(x, n=None):
    n = _span(x.n if n is None else n, x)
    r = c_double()
    check(lib().gip_dnrm_inf(x.dev.ctx, n, x.ptr, byref(r)))
    return r.value


def cfl(u, v, w, dxs, n, gamma, rho):
    n = _span(n, u, v, w, dxs)
    r = c_double()
    check(lib().gip_cfl(u.dev.ctx, n, u.ptr, v.ptr, w.ptr, dxs.ptr, gamma, rho, byref(r)))
    return r.value


@dataclass
class PcgResult:
    iters: int
    corrections: int
    relres: float
    status: int
    product_ms: float = 0.0  # time_products: median sampled product launch (ms)
    product_samples: int = 0


class PCG:
    def __init__(self, A: Matrix, n_all: int):
        self.A, self.dev, self.n_all = A, A.dev, int(n_all)
        self.ptr = c_void_p()
        check(lib().gip_pcg_create(self.dev.ctx, A.ptr, n_all, byref(self.ptr)))
        if not hasattr(A, "_solvers"):
            A._solvers = weakref.WeakSet()
        A._solvers.add(self)
        self.dev._adopt(self)

    def solve(self, b, x, x0, *, maxit=100, tol=0.0, precond=True, precision=PCG_FP64, inner_tol=0.1,
              poll=16, coarse=True, raise_on_breakdown=True, time_products=False):
        if b.n < self.A.nrows or x.n < self.n_all or x0.n < self.n_all:
            raise ValueError("PCG vectors shorter than the solver's layout")
        o = PcgOpts(maxit, tol, 1 if precond else 0, precision, inner_tol, poll, 0 if coarse else -1,
                    1 if time_products else 0)
        st = PcgStats()
        rc = lib().gip_pcg_solve(self.ptr, b.ptr, x.ptr, x0.ptr, byref(o), byref(st))
        if rc != 0 and (raise_on_breakdown or rc != EBREAKDOWN):
            check(rc)
        return PcgResult(st.iters, st.corrections, st.relres, st.status, st.product_ms, st.product_samples)

    def destroy(self):
        if self.ptr and self.dev.ctx:
            lib().gip_pcg_destroy(self.ptr)
        self.ptr = c_void_p()

    _release = destroy

    def __del__(self):
        try:
            self.destroy()
        except Exception:
            pass


STREAM_KINDS = {"copy": 0, "read": 1, "write": 2, "mix": 3}


def stream_probe(dev: Device, kind, n, *, vec=1, nontemporal=False, grid=0, reps=10, bufs=None):
    """Mean time (ms) of one launch of a streaming access mix over n doubles
    (gip_stream_probe); `bufs` (a, b, c, part) may be passed to reuse arrays."""
    if bufs is None:
        bufs = (dev.array(n), dev.array(n), dev.array(n), dev.array(1 << 22))
    a, b, c, part = bufs
    ms = c_float()
    check(lib().gip_stream_probe(dev.ctx, STREAM_KINDS[kind], vec, 1 if nontemporal else 0, grid, n, a.ptr, b.ptr,
                                 c.ptr, part.ptr, part.n, reps, byref(ms)))
    return ms.value


def barrier_probe(dev: Device, grid, iters=2000, mode=1):
    """Microseconds per device-wide barrier among `grid` workgroups (gip_barrier_probe)."""
    us = c_float()
    check(lib().gip_barrier_probe(dev.ctx, grid, iters, mode, byref(us)))
    return us.value


def pcg_solve_noalloc(A: Matrix, n_all, b, x, x0, *, maxit=100, tol=0.0, precond=True, poll=16):
    """The FP64 solve of PCG.solve without allocations (gip_pcg_solve_noalloc):
    the workspace is the library's per-device scratch array."""
    if b.n < A.nrows or x.n < n_all or x0.n < n_all:
        raise ValueError("PCG vectors shorter than n_all")
    o = PcgOpts(maxit, tol, 1 if precond else 0, PCG_FP64, 0.1, poll, 0, 0)
    st = PcgStats()
    check(lib().gip_pcg_solve_noalloc(A.dev.ctx, A.ptr, int(n_all), b.ptr, x.ptr, x0.ptr, byref(o), byref(st)))
    return PcgResult(st.iters, st.corrections, st.relres, st.status, st.product_ms, st.product_samples)


class Event:
    def __init__(self, dev: Device):
        self.dev = dev
        self.ptr = c_void_p()
        check(lib().gip_event_create(dev.ctx, byref(self.ptr)))
        dev._adopt(self)

    def _release(self):
        if self.ptr:
            lib().gip_event_destroy(self.ptr)
        self.ptr = c_void_p()

    def record(self):
        check(lib().gip_event_record(self.dev.ctx, self.ptr))

    def elapsed_ms(self, later: "Event"):
        ms = c_float()
        check(lib().gip_event_elapsed_ms(self.ptr, later.ptr, byref(ms)))
        return ms.value

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass
