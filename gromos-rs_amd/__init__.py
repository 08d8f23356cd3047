"""MI355X-native nonbonded hot path of gromos-rs: charge-group pairlist,
LJ + reaction-field pair loops and perturbed soft-core pairs, behind a C ABI
(include/gromos_amd.h) with HIP kernels for gfx950."""
from . import system  # noqa: F401
