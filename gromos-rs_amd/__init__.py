"""MI355X-native nonbonded hot path of gromos-rs: charge-group pairlist,
LJ + reaction-field pair loops and perturbed soft-core pairs, with the
solute's covalent terms beside them, behind a C ABI
(include/gromos_amd.h) with HIP kernels for gfx950.

Modules:
  system      host description of a topology + coordinates (System)
  generators  seeded synthetic systems (SPC water boxes, heterogeneous solute)
  runtime     ctypes binding of libgromos_amd.so, DeviceSystem
  nonbonded   mirror of the gromos-forces nonbonded functions (device tensors)
  bonded      mirror of the gromos-forces covalent-term functions (device tensors)
  gromos_io   GROMOS topology / configuration / input readers, trajectory writers
  server      continuous-batching job server over a session
  build       hipcc recipe of libgromos_amd.so
"""
from . import system  # noqa: F401
from . import runtime  # noqa: F401
