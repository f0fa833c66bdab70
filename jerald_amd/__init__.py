"""MI355X-native LDL (Lagrangian Deep Learning) hot path of JERALD.

* ``jerald_amd.jaxpops``: drop-in for the reference's jaxpops_bind operators
  (ppaint, preadout, ppaint3D, preadout3D, pfft, pifft, buildplan, ...);
* ``jerald_amd.pmesh.PMesh``: the particle-mesh facade;
* ``jerald_amd.model.JERALDModel``: the sm-kind LDL model with fused kernels.
All compute runs in libldl.so (HIP, gfx950) through a C ABI (include/ldl.h).
"""
from . import _lib  # noqa: F401

__all__ = ["jaxpops", "pmesh", "model"]
