"""Host mirror of the gromos-forces covalent-term API over the HIP library.

calculate_bonded_forces / calculate_bonded_forces_ntf
(crates/gromos-forces/src/bonded/mod.rs:133-185) with device tensors (torch,
ROCm) for the positions and the forces:

    plan = BondedPlan(system.bonded, n_atoms)        # terms laid out once
    fe = calculate_bonded_forces_ntf(plan, pos)     # pos: (n, 3) f64 on the GPU
    fe.energy, fe.forces, fe.virial, fe.e_bond ...

Quartic bonds only: the reference's force field always asks for them
(forcefield.rs:134, md.rs:1009-1012).  The NTF switches are the plan's
(Bonded.ntf, the FORCE block), or given per call, which builds a plan with
those switches.  There is no CPU path: errors raise GromosAmdError.
"""
import ctypes as C
from dataclasses import dataclass, replace

import numpy as np

from . import runtime as rt


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise rt.GromosAmdError("no GPU visible: the covalent-term kernels run on the device only")
    return torch


class BondedPlan:
    """gromos_amd_bonded_plan: the validated terms of one topology on the device."""

    def __init__(self, bonded, n_atoms):
        keep = rt._Keep()
        self._struct = rt.bonded_struct(bonded, n_atoms, keep)
        self._keep = keep
        self.bonded = bonded
        self.n_atoms = int(n_atoms)
        h = C.c_void_p()
        rt.check(rt.lib().gromos_amd_bonded_plan_create(C.byref(self._struct), C.byref(h)))
        self.h = h

    @property
    def terms(self):
        """Terms kept per kind: bonds, angles, dihedrals, impropers."""
        n = np.zeros(4, np.int32)
        rt.check(rt.lib().gromos_amd_bonded_plan_terms(self.h, rt.ptr(n)))
        return tuple(int(x) for x in n)

    def close(self):
        if getattr(self, "h", None):
            rt.lib().gromos_amd_bonded_plan_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


@dataclass
class ForceEnergy:
    """ForceEnergy (bonded/mod.rs:29-36) with the per-kind energies kept apart."""
    energy: float
    forces: object          # (n, 3) f64 torch tensor on the device
    virial: np.ndarray      # 3 x 3, virial[a][b] = sum r[a] f[b]
    e_bond: float
    e_angle: float
    e_dihedral: float
    e_improper: float


def calculate_bonded_forces_ntf(plan, positions, use_quartic_bonds=True, ntf_bond=None, ntf_angle=None,
                                ntf_dihedral=None, ntf_improper=None):
    """calculate_bonded_forces_ntf (bonded/mod.rs:141-185) on device positions."""
    torch = _torch()
    if not use_quartic_bonds:
        raise rt.GromosAmdError("harmonic bonds are not provided: the force field uses quartic bonds")
    flags = (ntf_bond, ntf_angle, ntf_improper, ntf_dihedral)
    if any(f is not None for f in flags):
        ntf = tuple(int(bool(f)) if f is not None else int(d) for f, d in zip(flags, plan.bonded.ntf))
        if ntf != tuple(plan.bonded.ntf):
            plan = BondedPlan(replace(plan.bonded, ntf=ntf), plan.n_atoms)
    if positions.dtype != torch.float64 or not positions.is_cuda or tuple(positions.shape) != (plan.n_atoms, 3):
        raise rt.GromosAmdError(f"positions must be a ({plan.n_atoms}, 3) float64 tensor on the GPU")
    pos = positions.contiguous()
    forces = torch.empty((plan.n_atoms, 3), dtype=torch.float64, device=pos.device)
    out = torch.empty(13, dtype=torch.float64, device=pos.device)
    stream = C.c_void_p(torch.cuda.current_stream(pos.device).cuda_stream)
    rt.check(rt.lib().gromos_amd_bonded_forces(plan.h, C.c_void_p(pos.data_ptr()), C.c_void_p(forces.data_ptr()),
                                               C.c_void_p(out.data_ptr()), stream))
    o = out.cpu().numpy()
    eb, ea, ed, ei = (float(x) for x in o[:4])
    return ForceEnergy(energy=((eb + ea) + ed) + ei, forces=forces, virial=o[4:].reshape(3, 3), e_bond=eb,
                       e_angle=ea, e_dihedral=ed, e_improper=ei)


def calculate_bonded_forces(plan, positions, use_quartic_bonds=True):
    """calculate_bonded_forces (bonded/mod.rs:133-139): every kind switched on."""
    return calculate_bonded_forces_ntf(plan, positions, use_quartic_bonds, True, True, True, True)
