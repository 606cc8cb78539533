"""Python view of the batched integrator (include/gip_batch.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import hip
from ._lib import POINTER, byref, c_double, c_int32, c_void_p


class BatchOpts(ctypes.Structure):
    _fields_ = [("pcg_maxit", c_int32), ("pcg_tol", c_double), ("precision", c_int32), ("inner_tol", c_double),
                ("poll", c_int32), ("use_graph", c_int32), ("autotune", c_int32)]


class BatchInfo(ctypes.Structure):
    _fields_ = [("time", c_double), ("dt", c_double), ("massr", c_double), ("pcg_iters", c_int32),
                ("pcg_status", c_int32), ("n_inner", c_int32), ("n_all", c_int32), ("n_faces", c_int32)]


def _lib():
    L = hip.lib()
    if not getattr(L, "_gip_batch_typed", False):
        L.gip_batch_create.argtypes = [c_void_p, c_int32, POINTER(c_void_p), POINTER(c_double), c_double,
                                       POINTER(BatchOpts), POINTER(c_void_p)]
        L.gip_batch_step.argtypes = [c_void_p]
        L.gip_batch_instance.argtypes = [c_void_p, c_int32, POINTER(BatchInfo)]
        L.gip_batch_get_field.argtypes = [c_void_p, c_int32, ctypes.c_char_p, c_void_p, c_int32]
        L.gip_batch_layouts.argtypes = [c_void_p, POINTER(c_int32), POINTER(c_int32)]
        L.gip_batch_phase_ms.argtypes = [c_void_p, POINTER(ctypes.c_float)]
        L.gip_batch_destroy.argtypes = [c_void_p]
        L._gip_batch_typed = True
    return L


class Batch:
    """Many cavities (gip_amd.cavity.Cavity, every operator built) advanced together."""

    def __init__(self, dev: hip.Device, cavities, gammas, *, rho=1.0, pcg_tol=1e-8, pcg_maxit=5000,
                 precision=hip.PCG_MIXED, inner_tol=0.1, poll=16, use_graph=True, autotune=False):
        self.dev = dev
        self.cavities = list(cavities)
        n = len(self.cavities)
        ptrs = (c_void_p * n)(*[ctypes.cast(c.ptr, c_void_p) for c in self.cavities])
        g = (c_double * n)(*[float(x) for x in gammas])
        o = BatchOpts(pcg_maxit, pcg_tol, precision, inner_tol, poll, 1 if use_graph else 0, 1 if autotune else 0)
        self.h = c_void_p()
        hip.check(_lib().gip_batch_create(dev.ctx, n, ptrs, g, rho, byref(o), byref(self.h)))
        self.n = n
        dev._adopt(self)

    def step(self):
        hip.check(_lib().gip_batch_step(self.h))

    PHASES = ("predictor", "pcg", "post")

    def phase_ms(self):
        """Device time (ms) of the last step's phases, keyed by PHASES: the
        predictor with the Poisson right-hand side, the PCG iterations (with
        the host's reads between poll groups), and correction + mass fluxes
        + CFL."""
        out = (ctypes.c_float * 3)()
        hip.check(_lib().gip_batch_phase_ms(self.h, out))
        return dict(zip(self.PHASES, [float(x) for x in out]))

    def info(self, i):
        s = BatchInfo()
        hip.check(_lib().gip_batch_instance(self.h, i, byref(s)))
        return {k: getattr(s, k) for k, _ in BatchInfo._fields_}

    def field(self, i, name):
        inf = self.info(i)
        n = inf["n_faces"] if name == "mf" else inf["n_all"]
        out = np.empty(n, np.float64)
        hip.check(_lib().gip_batch_get_field(self.h, i, name.encode(), out.ctypes.data, n))
        return out

    def layouts(self):
        f = (c_int32 * 15)()
        ln = (c_int32 * 15)()
        hip.check(_lib().gip_batch_layouts(self.h, f, ln))
        return list(f), list(ln)

    def _release(self):
        if self.h and self.dev.ctx:
            _lib().gip_batch_destroy(self.h)
        self.h = c_void_p()

    close = _release

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass
