"""gip_amd — MI355X (gfx950) backend of GIP's algebraic fractional-step
Navier–Stokes path.

Native parts (built in-tree by gip_amd.build):
  libgip_hip.so     hand-written gfx950 kernels behind the C ABI of include/gip_hip.h
  libgip_dns.so     the HIP specification classes and the templated integrator,
                    C ABI in include/gip_dns.h
  libgip_cavity.so  operator generator and the reference's file formats,
                    C ABI in include/gip_cavity.h

Python modules are thin ctypes views of those C ABIs: `hip`, `dns`, `cavity`.
"""
from ._lib import NativeLibraryMissing  # noqa: F401

__all__ = ["hip", "dns", "cavity", "NativeLibraryMissing"]
