"""MI355X-native cone-beam projection hot path with aprict's Python API.

Modules mirror aprict's layout: ``cb.flatpanel`` (geometry, ``project``,
``backproject``), ``cb.ramp_filter``, ``ramp_kernels``, ``volume``
(``Volume``, ``FBVolume``), ``iterators`` (SIRT, Landweber), plus
``autograd`` (differentiable projector) and ``phantoms`` (seeded inputs).
All computation runs in the hand-written HIP kernels of
``aprict_amd/lib/libaprict_hip.so`` (built by ``python -m aprict_amd.build``).
"""

from . import _abi, ramp_kernels
from ._abi import AprictError, library_path, lib

__version__ = "0.1.0"

__all__ = ["AprictError", "_abi", "lib", "library_path", "ramp_kernels"]
