"""Python view of the reverse-mode derivative (include/gip_adjoint.h)."""
from __future__ import annotations

import ctypes

import numpy as np

from . import hip
from ._lib import POINTER, byref, c_double, c_int32, c_void_p


class AdjointOpts(ctypes.Structure):
    _fields_ = [("steps", c_int32), ("checkpoint_every", c_int32), ("gamma", c_double), ("rho", c_double),
                ("pcg_tol", c_double), ("pcg_maxit", c_int32)]


def _lib():
    L = hip.lib()
    if not getattr(L, "_gip_adj_typed", False):
        L.gip_adjoint_create.argtypes = [c_void_p, c_void_p, POINTER(AdjointOpts), POINTER(c_void_p)]
        L.gip_adjoint_forward.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p, POINTER(c_double)]
        L.gip_adjoint_backward.argtypes = [c_void_p, c_void_p, c_void_p, c_void_p, c_void_p]
        L.gip_adjoint_dts.argtypes = [c_void_p, c_void_p]
        L.gip_adjoint_destroy.argtypes = [c_void_p]
        L._gip_adj_typed = True
    return L


def _ptr(a):
    return None if a is None else np.ascontiguousarray(a, dtype=np.float64).ctypes.data


class Adjoint:
    """Kinetic-energy objective of a cavity run and its gradient wrt the lid
    profile and the initial velocity (dt frozen after the first forward run)."""

    def __init__(self, dev: hip.Device, cav, *, steps, checkpoint_every, gamma, rho=1.0, pcg_tol=1e-13,
                 pcg_maxit=20000):
        self.dev, self.cav = dev, cav
        self.n_inner, self.nlid = cav.n_inner, cav.nx * cav.ny
        o = AdjointOpts(steps, checkpoint_every, gamma, rho, pcg_tol, pcg_maxit)
        self.h = c_void_p()
        hip.check(_lib().gip_adjoint_create(dev.ctx, ctypes.cast(cav.ptr, c_void_p), byref(o), byref(self.h)))
        self.steps = steps
        dev._adopt(self)

    def forward(self, lid, u0=None, v0=None, w0=None):
        keep = [np.ascontiguousarray(x, dtype=np.float64) if x is not None else None for x in (lid, u0, v0, w0)]
        assert keep[0].size == self.nlid
        for x in keep[1:]:
            assert x is None or x.size == self.n_inner
        J = c_double()
        hip.check(_lib().gip_adjoint_forward(self.h, *[None if x is None else x.ctypes.data for x in keep],
                                             byref(J)))
        return J.value

    def backward(self):
        gl = np.empty(self.nlid)
        gu, gv, gw = (np.empty(self.n_inner) for _ in range(3))
        hip.check(_lib().gip_adjoint_backward(self.h, gl.ctypes.data, gu.ctypes.data, gv.ctypes.data,
                                              gw.ctypes.data))
        return gl, gu, gv, gw

    def dts(self):
        out = np.empty(self.steps)
        hip.check(_lib().gip_adjoint_dts(self.h, out.ctypes.data))
        return out

    def _release(self):
        if self.h and self.dev.ctx:
            _lib().gip_adjoint_destroy(self.h)
        self.h = c_void_p()

    close = _release

    def __del__(self):
        try:
            self._release()
        except Exception:
            pass


def fd_check(adj: Adjoint, lid, u0, v0, w0, ndir=16, eps=1e-6, seed=0, space="both"):
    """Central finite differences of J along random directions against the
    adjoint gradient; returns (J, directional derivatives adjoint, fd, rel
    errors).  `space` picks the directions: "lid" (the lid profile only),
    "velocity" (the initial u, v, w only) or "both" (all together), so that
    each gradient can be checked on its own."""
    if space not in ("lid", "velocity", "both"):
        raise ValueError(space)
    J0 = adj.forward(lid, u0, v0, w0)
    gl, gu, gv, gw = adj.backward()
    rng = np.random.default_rng(seed)
    ad, fd = [], []
    for _ in range(ndir):
        dl = rng.standard_normal(lid.size) if space != "velocity" else np.zeros(lid.size)
        du, dv, dw = (rng.standard_normal(u0.size) if space != "lid" else np.zeros(u0.size) for _ in range(3))
        jp = adj.forward(lid + eps * dl, u0 + eps * du, v0 + eps * dv, w0 + eps * dw)
        jm = adj.forward(lid - eps * dl, u0 - eps * du, v0 - eps * dv, w0 - eps * dw)
        fd.append((jp - jm) / (2 * eps))
        ad.append(float(gl @ dl + gu @ du + gv @ dv + gw @ dw))
    ad, fd = np.array(ad), np.array(fd)
    return J0, ad, fd, np.abs(ad - fd) / np.abs(fd)
