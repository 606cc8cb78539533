"""ORACLE — TEST INFRASTRUCTURE ONLY.

ctypes binding of the C restatement of the reference CPU path
(oracle/gip_oracle.c, built into oracle/_build/liboracle.so).  Imported only
by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, as the
checker; never by the product.
"""
from __future__ import annotations

import ctypes
import pathlib
import subprocess

import numpy as np

HERE = pathlib.Path(__file__).resolve().parent
BUILD = HERE / "_build"
REF = HERE / "_ref"

P = ctypes.POINTER
d_p = P(ctypes.c_double)
i_p = P(ctypes.c_int)


class OrcCsr(ctypes.Structure):
    _fields_ = [("nnz", ctypes.c_int), ("ncols", ctypes.c_int), ("nrows", ctypes.c_int), ("val", d_p),
                ("col", i_p), ("rowptr", i_p)]


_lib = None


def lib():
    global _lib
    if _lib is None:
        path = BUILD / "liboracle.so"
        if not path.exists():
            raise RuntimeError(f"{path} not built (gip_amd.build.build_oracle)")
        L = ctypes.CDLL(str(path))
        L.orc_spmv.argtypes = [P(OrcCsr), d_p, d_p]
        L.orc_spmv_ab.argtypes = [P(OrcCsr), d_p, d_p, ctypes.c_double, ctypes.c_double]
        L.orc_spmv_vals_ab.argtypes = [P(OrcCsr), d_p, d_p, ctypes.c_double, ctypes.c_double]
        L.orc_spmv_vals_ab.restype = None
        L.orc_diag_setup.argtypes = [P(OrcCsr), P(OrcCsr)]
        L.orc_csr_free.argtypes = [P(OrcCsr)]
        L.orc_csr_free.restype = None
        L.orc_dot.argtypes = [d_p, d_p, ctypes.c_int]
        L.orc_dot.restype = ctypes.c_double
        L.orc_norm_inf.argtypes = [d_p, ctypes.c_int]
        L.orc_norm_inf.restype = ctypes.c_double
        L.orc_cfl.argtypes = [ctypes.c_int, d_p, d_p, d_p, d_p, ctypes.c_double, ctypes.c_double]
        L.orc_cfl.restype = ctypes.c_double
        L.orc_pcg.argtypes = [P(OrcCsr), P(OrcCsr), ctypes.c_int, ctypes.c_int, d_p, d_p, d_p, ctypes.c_int,
                              ctypes.c_double, ctypes.c_int, i_p, d_p]
        L.orc_dns_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p, P(ctypes.c_void_p)]
        L.orc_dns_open_mem.argtypes = [ctypes.c_int, ctypes.c_int, ctypes.c_int, P(OrcCsr), d_p, d_p, d_p, d_p,
                                       P(ctypes.c_void_p)]
        L.orc_set_threads.argtypes = [ctypes.c_int]
        L.orc_set_threads.restype = None
        L.orc_set_blocked_dot.argtypes = [ctypes.c_int]
        L.orc_set_blocked_dot.restype = None
        L.orc_dns_close.argtypes = [ctypes.c_void_p]
        L.orc_dns_close.restype = None
        L.orc_dns_params.argtypes = [ctypes.c_void_p, ctypes.c_double, ctypes.c_double, ctypes.c_int,
                                     ctypes.c_double, ctypes.c_int]
        L.orc_dns_params.restype = None
        L.orc_dns_step.argtypes = [ctypes.c_void_p, d_p, d_p, d_p, i_p]
        L.orc_dns_field.argtypes = [ctypes.c_void_p, ctypes.c_char_p, i_p]
        L.orc_dns_field.restype = d_p
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(d_p)


class Csr:
    """Keeps the numpy arrays alive behind an OrcCsr view."""

    def __init__(self, rowptr, col, val, ncols=0):
        self.rowptr = np.ascontiguousarray(rowptr, dtype=np.int32)
        self.col = np.ascontiguousarray(col, dtype=np.int32)
        self.val = np.ascontiguousarray(val, dtype=np.float64)
        self.s = OrcCsr(self.val.size, ncols, self.rowptr.size - 1, _dp(self.val), self.col.ctypes.data_as(i_p),
                        self.rowptr.ctypes.data_as(i_p))

    @classmethod
    def of(cls, csr):
        return cls(csr.rowptr, csr.col, csr.val, csr.ncols)


def spmv(A: Csr, x, y):
    """GIP_MatrixMulticore.cpp:64-83 (y modified in place)."""
    lib().orc_spmv(ctypes.byref(A.s), _dp(x), _dp(y))


def spmv_ab(A: Csr, x, y, alpha, beta):
    lib().orc_spmv_ab(ctypes.byref(A.s), _dp(x), _dp(y), alpha, beta)


def spmv_vals_ab(A: Csr, x, bval, alpha, beta):
    """GIP_MatrixMulticore.cpp:125-143 (bval, another operator's values, modified in place)."""
    lib().orc_spmv_vals_ab(ctypes.byref(A.s), _dp(x), _dp(bval), alpha, beta)


def dot(a, b):
    return lib().orc_dot(_dp(a), _dp(b), a.size)


def norm_inf(a):
    return lib().orc_norm_inf(_dp(a), a.size)


def cfl(u, v, w, dxs, n, gamma, rho):
    return lib().orc_cfl(n, _dp(u), _dp(v), _dp(w), _dp(dxs), gamma, rho)


def pcg(A: Csr, b, x0, n_inner, maxit, tol=0.0, precond=True):
    """GIP_LinearSolver.h:45-117 on rows [0, n_inner); returns (x, status, iters, relres)."""
    D = OrcCsr()
    lib().orc_diag_setup(ctypes.byref(A.s), ctypes.byref(D))
    x = np.empty_like(x0)
    it = ctypes.c_int()
    rr = ctypes.c_double()
    st = lib().orc_pcg(ctypes.byref(A.s), ctypes.byref(D), n_inner, x0.size, _dp(b), _dp(x), _dp(x0), maxit, tol,
                       1 if precond else 0, ctypes.byref(it), ctypes.byref(rr))
    lib().orc_csr_free(ctypes.byref(D))
    return x, st, it.value, rr.value


def set_threads(n):
    """Threads of the oracle's row loops (products, axpy); results are
    bitwise the same at any count (reductions stay serial)."""
    lib().orc_set_threads(int(n))


def set_blocked_dot(on):
    """Dot products as sums of 4096-term blocks added in block order (not
    the reference's serial order: equal up to rounding); off by default."""
    lib().orc_set_blocked_dot(1 if on else 0)


MEM_OPS = ("D", "Ec", "L", "Lb", "Mx", "My", "Mz", "Gx", "Gy", "Gz", "SFx", "SFy", "SFz", "SO", "Mres")


class Dns:
    """GIP_DNS on one partition (the C restatement), read from
    <root>/InputFiles, or from a generated case in memory (`cavity`: an
    object with n_inner, n_all, n_faces, op(name) -> CSR and field(name);
    its operators are borrowed and must outlive this object)."""

    FIELDS = ("u", "v", "w", "p", "mf")

    def __init__(self, root=None, name=None, gamma=1e-3, rho=1.0, pcg_maxit=100, pcg_tol=0.0, cavity=None):
        self.h = ctypes.c_void_p()
        if cavity is not None:
            self._keep = [cavity] + [Csr.of(cavity.op(o)) for o in MEM_OPS]
            ops = (OrcCsr * len(MEM_OPS))(*[c.s for c in self._keep[1:]])
            f = {k: np.ascontiguousarray(cavity.field(k), dtype=np.float64) for k in ("u", "v", "w", "dxs")}
            rc = lib().orc_dns_open_mem(cavity.n_inner, cavity.n_all, cavity.n_faces, ops, _dp(f["u"]), _dp(f["v"]),
                                        _dp(f["w"]), _dp(f["dxs"]), ctypes.byref(self.h))
            if rc != 0:
                raise RuntimeError("oracle cannot take the case from memory")
        elif lib().orc_dns_open(str(root).encode(), name.encode(), ctypes.byref(self.h)) != 0:
            raise RuntimeError(f"oracle cannot open case {name} under {root}")
        lib().orc_dns_params(self.h, gamma, rho, pcg_maxit, pcg_tol, 1)

    def step(self):
        t, dt, mr = ctypes.c_double(), ctypes.c_double(), ctypes.c_double()
        it = ctypes.c_int()
        st = lib().orc_dns_step(self.h, ctypes.byref(t), ctypes.byref(dt), ctypes.byref(mr), ctypes.byref(it))
        return {"time": t.value, "dt": dt.value, "massr": mr.value, "pcg_iters": it.value, "status": st}

    def field(self, name):
        n = ctypes.c_int()
        p = lib().orc_dns_field(self.h, name.encode(), ctypes.byref(n))
        if not p:
            raise KeyError(name)
        return np.ctypeslib.as_array(p, shape=(n.value,)).copy()

    def close(self):
        if self.h:
            lib().orc_dns_close(self.h)
            self.h = ctypes.c_void_p()

    def __del__(self):
        self.close()


def run_cli(root, name, steps, gamma, maxit=100, tol=0.0, dump=None):
    """Run the oracle CLI; returns per-step dicts (and dumps fields when asked)."""
    cmd = [str(BUILD / "orc_cavity"), str(root), name, str(steps), repr(gamma), str(maxit), repr(tol), "--full"]
    if dump:
        cmd += ["--dump", str(dump)]
    out = subprocess.run(cmd, capture_output=True, text=True, check=True).stdout
    return parse_lines(out)


def parse_lines(text):
    rows = []
    for ln in text.splitlines():
        t = ln.replace("[dt::", " ").replace("]", " ").split()
        if not t or t[0] != "time":
            continue
        row = {"time": float(t[1]), "dt": float(t[2]), "massr": float(t[4]), "step": int(t[6])}
        if "pcg" in t:
            row["pcg_iters"] = int(t[t.index("pcg") + 1])
        rows.append(row)
    return rows


def read_vec(path):
    raw = np.fromfile(path, dtype=np.uint8)
    n = int(np.frombuffer(raw[:4].tobytes(), np.int32)[0])
    return np.frombuffer(raw[4:4 + 8 * n].tobytes(), np.float64).copy()
