"""Python view of the integrator's C ABI (include/gip_dns.h): the templated
GIP_DNS instantiated with the HIP specification, stepped from Python."""
from __future__ import annotations

import ctypes

import numpy as np

from . import cavity as _cav
from ._lib import POINTER, byref, c_double, c_int32, c_void_p, load

MODES = {"reference": 0, "fused": 1, "mixed": 2}


class Config(ctypes.Structure):
    _fields_ = [("gamma", c_double), ("rho", c_double), ("pcg_maxit", c_int32), ("pcg_tol", c_double),
                ("pcg_mode", c_int32), ("pcg_inner_tol", c_double), ("print", c_int32), ("device", c_int32)]


class StepInfo(ctypes.Structure):
    _fields_ = [("time", c_double), ("dt", c_double), ("massr", c_double), ("pcg_relres", c_double),
                ("step", c_int32), ("pcg_iters", c_int32), ("pcg_status", c_int32), ("pcg_corrections", c_int32)]


def lib():
    L = load("gip_dns")
    if not getattr(L, "_gip_typed", False):
        L.gip_dns_open.argtypes = [ctypes.c_char_p, ctypes.c_char_p, POINTER(Config), POINTER(c_void_p)]
        L.gip_dns_from_cavity.argtypes = [c_void_p, POINTER(Config), POINTER(c_void_p)]
        L.gip_dns_step.argtypes = [c_void_p, POINTER(StepInfo)]
        L.gip_dns_phase_ms.argtypes = [c_void_p, POINTER(c_double)]
        L.gip_dns_get_field.argtypes = [c_void_p, ctypes.c_char_p, c_void_p, c_int32]
        L.gip_dns_set_field.argtypes = [c_void_p, ctypes.c_char_p, c_void_p, c_int32]
        L.gip_dns_sizes.argtypes = [c_void_p, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32)]
        L.gip_dns_sync.argtypes = [c_void_p]
        L.gip_dns_close.argtypes = [c_void_p]
        L.gip_dns_last_error.restype = ctypes.c_char_p
        L.gip_dns_open_parts.argtypes = [ctypes.c_char_p, ctypes.c_char_p, c_int32, POINTER(Config), POINTER(c_void_p)]
        L.gip_dns_parts_step.argtypes = [c_void_p, POINTER(StepInfo)]
        L.gip_dns_parts_sizes.argtypes = [c_void_p, c_int32, POINTER(c_int32), POINTER(c_int32), POINTER(c_int32),
                                          POINTER(c_int32)]
        L.gip_dns_parts_get_field.argtypes = [c_void_p, c_int32, ctypes.c_char_p, c_void_p, c_int32]
        L.gip_dns_parts_close.argtypes = [c_void_p]
        L._gip_typed = True
    return L


class DnsError(RuntimeError):
    pass


def _check(rc):
    if rc != 0:
        raise DnsError(f"gip_dns error {rc}: {lib().gip_dns_last_error().decode()}")


class Simulation:
    """GIP_DNS<GIP_MatrixHIP, GIP_VectorHIP, GIP_ArchHIP, GIP_NonLinearHIP> on one device."""

    def __init__(self, *, root=None, name=None, cavity=None, gamma=1e-3, rho=1.0, pcg_maxit=100, pcg_tol=0.0,
                 pcg_mode="reference", pcg_inner_tol=0.1, print_mode=0, device=0):
        cfg = Config(gamma, rho, pcg_maxit, pcg_tol, MODES[pcg_mode], pcg_inner_tol, print_mode, device)
        self.h = c_void_p()
        if cavity is not None:
            self._cav = cavity  # keep the host operators alive while they are uploaded
            _check(lib().gip_dns_from_cavity(ctypes.cast(cavity.ptr, c_void_p), byref(cfg), byref(self.h)))
        else:
            _check(lib().gip_dns_open(str(root).encode(), name.encode(), byref(cfg), byref(self.h)))
        i, a, f = c_int32(), c_int32(), c_int32()
        _check(lib().gip_dns_sizes(self.h, byref(i), byref(a), byref(f)))
        self.n_inner, self.n_all, self.n_faces = i.value, a.value, f.value

    def step(self):
        s = StepInfo()
        _check(lib().gip_dns_step(self.h, byref(s)))
        return {"time": s.time, "dt": s.dt, "massr": s.massr, "step": s.step, "pcg_iters": s.pcg_iters,
                "pcg_status": s.pcg_status, "pcg_relres": s.pcg_relres, "pcg_corrections": s.pcg_corrections}

    PHASES = ("predictor", "poisson", "correction", "mass_fluxes", "cfl")

    def phase_ms(self):
        """Device time of each phase of the last step (ms), keyed by PHASES."""
        out = (c_double * 5)()
        _check(lib().gip_dns_phase_ms(self.h, out))
        return dict(zip(self.PHASES, list(out)))

    def field(self, name):
        n = self.n_faces if name == "mf" else self.n_all
        out = np.empty(n, np.float64)
        _check(lib().gip_dns_get_field(self.h, name.encode(), out.ctypes.data, n))
        return out

    def set_field(self, name, values):
        values = np.ascontiguousarray(values, dtype=np.float64)
        _check(lib().gip_dns_set_field(self.h, name.encode(), values.ctypes.data, values.size))

    def sync(self):
        _check(lib().gip_dns_sync(self.h))

    def close(self):
        if self.h:
            lib().gip_dns_close(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class PartitionedSimulation:
    """An <nparts>-partition case (<name>_<nparts>p-<rank> files) stepped on
    one device, one host thread and stream per partition (gip_dns_open_parts)."""

    def __init__(self, *, root, name, nparts, gamma=1e-3, rho=1.0, pcg_maxit=100, pcg_tol=0.0, print_mode=0,
                 device=0):
        cfg = Config(gamma, rho, pcg_maxit, pcg_tol, MODES["reference"], 0.1, print_mode, device)
        self.h = c_void_p()
        _check(lib().gip_dns_open_parts(str(root).encode(), name.encode(), nparts, byref(cfg), byref(self.h)))
        self.nparts = nparts
        self.sizes = []
        for r in range(nparts):
            i, o, a, f = c_int32(), c_int32(), c_int32(), c_int32()
            _check(lib().gip_dns_parts_sizes(self.h, r, byref(i), byref(o), byref(a), byref(f)))
            self.sizes.append({"inner": i.value, "owned": o.value, "all": a.value, "faces": f.value})

    def step(self):
        s = StepInfo()
        _check(lib().gip_dns_parts_step(self.h, byref(s)))
        return {"time": s.time, "dt": s.dt, "massr": s.massr, "step": s.step, "pcg_iters": s.pcg_iters,
                "pcg_status": s.pcg_status}

    def field(self, rank, name):
        n = self.sizes[rank]["faces" if name == "mf" else "all"]
        out = np.empty(n, np.float64)
        _check(lib().gip_dns_parts_get_field(self.h, rank, name.encode(), out.ctypes.data, n))
        return out

    def merged_field(self, name):
        """A cell field in the order gip_cavity_merge numbers the cells: every
        partition's inner cells, then every partition's boundary ghosts."""
        parts = [self.field(r, name) for r in range(self.nparts)]
        inner = [p[:s["inner"]] for p, s in zip(parts, self.sizes)]
        ghosts = [p[s["inner"]:s["owned"]] for p, s in zip(parts, self.sizes)]
        return np.concatenate(inner + ghosts)

    def close(self):
        if self.h:
            lib().gip_dns_parts_close(self.h)
            self.h = c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def generate_and_open(n, **kw):
    """Convenience: an in-memory n^3 cavity (gip_cavity_generate) opened on the device."""
    gen_kw = {k: kw.pop(k) for k in ("beta", "lid_profile", "pin", "lid_u") if k in kw}
    cav = _cav.generate(n, gamma=kw.get("gamma", 1e-2), **gen_kw)
    return Simulation(cavity=cav, **kw)
