"""Host mirror of the gromos-forces nonbonded API over the HIP library.

Same names, argument meaning and accumulate semantics as the Rust functions
(crates/gromos-forces/src/nonbonded/innerloops.rs:318-733, excluded.rs:25-206,
perturbed.rs:124-707, params.rs), with device tensors (torch, ROCm) in place of
slices:

    storage = ForceStorage(n_atoms)            # zeros on the GPU
    lj_crf_innerloop(pos, q, iac, pairs, lj, crf, Rectangular(box), FPEPSI, storage)

positions (n, 3) f64, charges (n,) f64, iac (n,) int32 (0-based), pair lists
(P, 2) int32/uint32, lj = LJParamMatrix.  Errors from the library raise
GromosAmdError (the Rust versions panic or return Err); there is no CPU path.
"""
import numpy as np

from . import runtime as rt

CRFParameters = rt.Crf


def crf_parameters(cutoff, epsilon, rf_epsilon, rf_kappa):
    """CRFParameters::new (params.rs:78-97)."""
    return rt.crf_new(cutoff, epsilon, rf_epsilon, rf_kappa)


def _torch():
    import torch
    if not torch.cuda.is_available():
        raise rt.GromosAmdError("no GPU visible: the nonbonded kernels run on the device only")
    return torch


class Vacuum:
    kind = 0
    box = (0.0, 0.0, 0.0)


class Rectangular:
    kind = 1

    def __init__(self, box):
        self.box = tuple(float(b) for b in box)


class Triclinic:
    """Triclinic (math.rs:75-141): box_matrix columns a, b, c, lower triangular."""
    kind = 2

    def __init__(self, box_matrix):
        m = np.asarray(box_matrix, float)
        self.box = tuple(float(v) for v in m.T.reshape(-1))


def boundary_of(system):
    """the boundary object of a system.System"""
    if system.boundary == 2:
        return Triclinic(system.box_matrix)
    return Rectangular(system.box) if system.boundary else Vacuum()


class LJParamMatrix:
    """Flat [ti * n + tj] x (c6, c12, cs6, cs12) table (params.rs:26-58) on the device."""

    def __init__(self, flat, n_types):
        torch = _torch()
        self.n_types = int(n_types)
        self.data = torch.as_tensor(np.asarray(flat, np.float64).reshape(-1), device="cuda")
        assert self.data.numel() == self.n_types * self.n_types * 4

    @staticmethod
    def from_nested(rows):
        """rows[ti][tj] = (c6, c12, cs6, cs12)"""
        a = np.asarray(rows, np.float64)
        return LJParamMatrix(a.reshape(-1), a.shape[0])


class ForceStorage:
    """forces (n, 3), e_lj, e_crf, virial[3][3] (params.rs:99-138), device resident."""

    def __init__(self, n_atoms):
        torch = _torch()
        self.forces = torch.zeros((n_atoms, 3), dtype=torch.float64, device="cuda")
        self.energies = torch.zeros(2, dtype=torch.float64, device="cuda")
        self.virial = torch.zeros((3, 3), dtype=torch.float64, device="cuda")

    @property
    def e_lj(self):
        return float(self.energies[0])

    @property
    def e_crf(self):
        return float(self.energies[1])

    def clear(self):
        self.forces.zero_()
        self.energies.zero_()
        self.virial.zero_()


def _stream():
    return _torch().cuda.current_stream().cuda_stream


def _u32(t):
    torch = _torch()
    t = torch.as_tensor(t, device="cuda")
    if t.dtype not in (torch.int32, torch.uint32):
        t = t.to(torch.int32)
    return t.contiguous()


def _f64(t):
    torch = _torch()
    return torch.as_tensor(t, dtype=torch.float64, device="cuda").contiguous()


def _box(bc):
    return (rt.D * len(bc.box))(*bc.box)


def _innerloop(positions, charges, iac, pairlist, lj_params, crf, periodicity, four_pi_eps_i, storage,
               virial):
    pos, q, ia, pl = _f64(positions), _f64(charges), _u32(iac), _u32(pairlist)
    rt.check(rt.lib().gromos_amd_lj_crf_innerloop(
        rt.ptr(pos), rt.ptr(q), rt.ptr(ia), rt.ptr(pl), pl.numel() // 2, rt.ptr(lj_params.data),
        lj_params.n_types, crf, periodicity.kind, _box(periodicity), four_pi_eps_i, int(virial),
        rt.ptr(storage.forces), rt.ptr(storage.energies), rt.ptr(storage.virial), _stream()))


def lj_crf_innerloop(positions, charges, iac, pairlist, lj_params, crf, periodicity, four_pi_eps_i, storage):
    """innerloops.rs:318-340"""
    _innerloop(positions, charges, iac, pairlist, lj_params, crf, periodicity, four_pi_eps_i, storage, True)


def lj_crf_innerloop_novirial(positions, charges, iac, pairlist, lj_params, crf, periodicity, four_pi_eps_i,
                              storage):
    """innerloops.rs:345-368"""
    _innerloop(positions, charges, iac, pairlist, lj_params, crf, periodicity, four_pi_eps_i, storage, False)


def lj_crf_innerloop_cg_grouped(positions, charges, iac, pairlist, groups, lj_params, crf, periodicity,
                                four_pi_eps_i, storage, virial=True):
    """innerloops.rs:374-428; groups (G, 4): ref_atom_i, ref_atom_j, start, end"""
    pos, q, ia, pl, g = _f64(positions), _f64(charges), _u32(iac), _u32(pairlist), _u32(groups)
    rt.check(rt.lib().gromos_amd_lj_crf_innerloop_cg_grouped(
        rt.ptr(pos), rt.ptr(q), rt.ptr(ia), rt.ptr(pl), rt.ptr(g), g.numel() // 4, rt.ptr(lj_params.data),
        lj_params.n_types, crf, periodicity.kind, _box(periodicity), four_pi_eps_i, int(virial),
        rt.ptr(storage.forces), rt.ptr(storage.energies), rt.ptr(storage.virial), _stream()))


def solvent_innerloop(positions, charges, iac, pairlist, lj_params, crf, periodicity, atoms_per_solvent,
                      four_pi_eps_i, storage, virial=True):
    """innerloops.rs:679-733"""
    pos, q, ia, pl = _f64(positions), _f64(charges), _u32(iac), _u32(pairlist)
    rt.check(rt.lib().gromos_amd_solvent_innerloop(
        rt.ptr(pos), rt.ptr(q), rt.ptr(ia), rt.ptr(pl), pl.numel() // 2, rt.ptr(lj_params.data),
        lj_params.n_types, crf, periodicity.kind, _box(periodicity), int(atoms_per_solvent), four_pi_eps_i,
        int(virial), rt.ptr(storage.forces), rt.ptr(storage.energies), rt.ptr(storage.virial), _stream()))


def rf_excluded_interactions(charges, excl_off, excl, positions, crf, boundary, four_pi_eps_i, storage,
                             num_solute_atoms):
    """excluded.rs:25-112; exclusions as CSR (offsets, indices)"""
    q, pos, off, idx = _f64(charges), _f64(positions), _u32(excl_off), _u32(excl)
    rt.check(rt.lib().gromos_amd_rf_excluded_interactions(
        rt.ptr(q), rt.ptr(off), rt.ptr(idx), q.numel(), int(num_solute_atoms), rt.ptr(pos), crf, boundary.kind,
        _box(boundary), four_pi_eps_i, rt.ptr(storage.forces), rt.ptr(storage.energies), rt.ptr(storage.virial),
        _stream()))


def one_four_interaction_loop(o14_off, o14, positions, charges, iac, lj_params, crf, boundary, four_pi_eps_i,
                              storage, coulomb_scaling=1.0):
    """excluded.rs:131-206; one_four_pairs as CSR"""
    pos, q, ia, off, idx = _f64(positions), _f64(charges), _u32(iac), _u32(o14_off), _u32(o14)
    rt.check(rt.lib().gromos_amd_one_four_interaction_loop(
        rt.ptr(off), rt.ptr(idx), q.numel(), rt.ptr(pos), rt.ptr(q), rt.ptr(ia), rt.ptr(lj_params.data),
        lj_params.n_types, crf, boundary.kind, _box(boundary), four_pi_eps_i, rt.ptr(storage.forces),
        rt.ptr(storage.energies), rt.ptr(storage.virial), float(coulomb_scaling), _stream()))


class PertNBCorrection:
    """forces, delta_e_lj, delta_e_crf, dhdl (perturbed.rs:282-301), device resident."""

    def __init__(self, n_atoms):
        torch = _torch()
        self.forces = torch.zeros((n_atoms, 3), dtype=torch.float64, device="cuda")
        self.out = torch.zeros(3, dtype=torch.float64, device="cuda")

    @property
    def delta_e_lj(self):
        return float(self.out[0])

    @property
    def delta_e_crf(self):
        return float(self.out[1])

    @property
    def dhdl(self):
        return float(self.out[2])


def _pert(pert):
    """PertAtomInfo per atom as device arrays (system.Perturbation, host arrays,
    copied); returns the C struct and the tensors it points into."""
    torch = _torch()
    keep = [torch.as_tensor(np.asarray(pert.flag, np.uint8), device="cuda"),
            _u32(pert.a_iac), _u32(pert.b_iac), _f64(pert.a_q), _f64(pert.b_q), _f64(pert.alpha_lj),
            _f64(pert.alpha_crf)]
    return rt.Pert(*[rt.ptr(t) for t in keep]), keep


def perturbed_pairlist_correction(pos, pairlist, pert, charges, iac, lj, crf, lp, bc, four_pi_eps_i, out):
    """perturbed.rs:309-386; pert: system.Perturbation (host arrays, copied)"""
    torch = _torch()
    ps, keep = _pert(pert)
    x, pl, q, ia = _f64(pos), _u32(pairlist), _f64(charges), _u32(iac)
    rt.check(rt.lib().gromos_amd_perturbed_pairlist_correction(
        rt.ptr(x), rt.ptr(pl), pl.numel() // 2, rt.C.byref(ps), rt.ptr(q), rt.ptr(ia), rt.ptr(lj.data), lj.n_types,
        crf, lp, bc.kind, _box(bc), four_pi_eps_i, rt.ptr(out.forces), rt.ptr(out.out), _stream()))
    torch.cuda.current_stream().synchronize()


def perturbed_self_energy_correction(pert, crf, lp, four_pi_eps_i):
    """perturbed.rs:395-425: returns (delta_e_crf, dhdl)."""
    torch = _torch()
    ps, keep = _pert(pert)
    out = torch.zeros(2, dtype=torch.float64, device="cuda")
    rt.check(rt.lib().gromos_amd_perturbed_self_energy_correction(
        rt.C.byref(ps), len(pert.flag), crf, lp, four_pi_eps_i, rt.ptr(out), _stream()))
    v = out.cpu().numpy()
    return float(v[0]), float(v[1])


def perturbed_excluded_correction(excl_off, excl, pos, pert, charges, crf, lp, bc, n_solute, four_pi_eps_i, out):
    """perturbed.rs:431-511; exclusions as CSR (excl_off, excl) over all atoms."""
    torch = _torch()
    ps, keep = _pert(pert)
    off, idx, x, q = _u32(excl_off), _u32(excl), _f64(pos), _f64(charges)
    rt.check(rt.lib().gromos_amd_perturbed_excluded_correction(
        rt.ptr(off), rt.ptr(idx), x.shape[0], int(n_solute), rt.ptr(x), rt.C.byref(ps), rt.ptr(q), crf, lp, bc.kind,
        _box(bc), four_pi_eps_i, rt.ptr(out.forces), rt.ptr(out.out), _stream()))
    torch.cuda.current_stream().synchronize()


def perturbed_one_four_correction(o14_off, o14, pos, pert, charges, iac, lj, crf, lp, bc, four_pi_eps_i, out):
    """perturbed.rs:517-605; one_four_pairs as CSR (o14_off, o14)."""
    torch = _torch()
    ps, keep = _pert(pert)
    off, idx, x, q, ia = _u32(o14_off), _u32(o14), _f64(pos), _f64(charges), _u32(iac)
    rt.check(rt.lib().gromos_amd_perturbed_one_four_correction(
        rt.ptr(off), rt.ptr(idx), x.shape[0], rt.ptr(x), rt.C.byref(ps), rt.ptr(q), rt.ptr(ia), rt.ptr(lj.data),
        lj.n_types, crf, lp, bc.kind, _box(bc), four_pi_eps_i, rt.ptr(out.forces), rt.ptr(out.out), _stream()))
    torch.cuda.current_stream().synchronize()


def perturbed_atom_pair_correction(atom_pairs, o14_off, o14, pos, charges, iac, lj, crf, lp, bc, four_pi_eps_i,
                                   out):
    """perturbed.rs:624-707; atom_pairs (K, 4) int32: i, j, a_type, b_type (-1 = absent in B)."""
    torch = _torch()
    ap = _u32(np.asarray(atom_pairs, np.int32).reshape(-1, 4))
    off, idx, x, q, ia = _u32(o14_off), _u32(o14), _f64(pos), _f64(charges), _u32(iac)
    rt.check(rt.lib().gromos_amd_perturbed_atom_pair_correction(
        rt.ptr(ap), ap.shape[0], rt.ptr(off), rt.ptr(idx), x.shape[0], rt.ptr(x), rt.ptr(q), rt.ptr(ia),
        rt.ptr(lj.data), lj.n_types, crf, lp, bc.kind, _box(bc), four_pi_eps_i, rt.ptr(out.forces),
        rt.ptr(out.out), _stream()))
    torch.cuda.current_stream().synchronize()


def perturbed_lj_crf_interaction(r, a_c6, a_c12, b_c6, b_c12, a_q, b_q, alpha_lj, alpha_crf, lp, crf,
                                 four_pi_eps_i):
    """perturbed.rs:124-266 on n pairs at once: r (n, 3), the other parameters
    scalars or (n,) arrays; returns (n, 5): force / r, e_lj, e_crf, de_lj, de_crf."""
    torch = _torch()
    rr = np.asarray(r, np.float64).reshape(-1, 3)
    n = rr.shape[0]
    prm = np.stack([np.broadcast_to(np.asarray(v, np.float64), (n,))
                    for v in (a_c6, a_c12, b_c6, b_c12, a_q, b_q, alpha_lj, alpha_crf)], axis=1)
    dr, dp = _f64(rr), _f64(prm)
    out = torch.empty((n, 5), dtype=torch.float64, device="cuda")
    rt.check(rt.lib().gromos_amd_perturbed_lj_crf_interaction(rt.ptr(dr), rt.ptr(dp), n, lp, crf, four_pi_eps_i,
                                                              rt.ptr(out), _stream()))
    return out.cpu().numpy()
