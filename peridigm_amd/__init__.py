"""peridigm_amd: MI355X-native LPS + critical-stretch peridynamic force path.

Hand-written HIP kernels for gfx950 behind the C ABI of include/pdgpu.h
(libpdgpu.so, built in-tree); this package is the host-side mirror of
Peridigm's Material / DamageModel interfaces and of its explicit and
quasi-static drivers on top of that ABI.
"""
from ._lib import PdgError, load  # noqa: F401
from .core import Context, count_points, list_to_csr, csr_to_list  # noqa: F401
