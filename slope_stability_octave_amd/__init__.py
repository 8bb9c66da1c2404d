"""MI355X-native SSR Newton-step hot path (Mohr-Coulomb return mapping,
B'*DS*B tangent assembly, Krylov solves) for the slope-stability code
Beremi/slope_stability_octave.

The compute runs in hand-written HIP kernels for gfx950 behind the C ABI
declared in ``include/ssr_hip.h`` (library ``libssr_hip.so``).  The Python
modules mirror the reference's Octave interfaces (``CONSTITUTIVE``,
``DFGMRES``, ``newton_ind_SSR``, ``SSR_indirect_continuation``) on top of it.
"""
from . import workloads  # noqa: F401
