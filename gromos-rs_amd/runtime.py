"""ctypes binding of libgromos_amd.so (include/gromos_amd.h).

The library is the only compute path: nothing here falls back to the CPU.  If
the shared library has not been built, every entry point raises.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libgromos_amd.so")

P = C.c_void_p
D = C.c_double
I32 = C.c_int32
I64 = C.c_int64

FP64 = 0
MIXED = 1

_lib = None


class GromosAmdError(RuntimeError):
    pass


class Crf(C.Structure):
    """CRFParameters (crates/gromos-forces/src/nonbonded/params.rs:60-68)."""
    _fields_ = [("crf_cut", D), ("crf_2cut3i", D), ("crf_cut3i", D), ("cutoff_sq", D)]


class Lambda(C.Structure):
    """PerturbedLambdaParams (crates/gromos-forces/src/nonbonded/perturbed.rs:11-50)."""
    _fields_ = [(k, D) for k in (
        "a_lj_lambda_n", "b_lj_lambda_n", "a_crf_lambda_n", "b_crf_lambda_n",
        "a_lj_lambda_n_1", "b_lj_lambda_n_1", "a_crf_lambda_n_1", "b_crf_lambda_n_1",
        "a_ljs_lambda", "b_ljs_lambda", "a_ljs_lambda2", "b_ljs_lambda2",
        "a_crfs_lambda", "b_crfs_lambda", "a_crfs_lambda2", "b_crfs_lambda2")] + [("lambda_exp", I32)]


class Pert(C.Structure):
    _fields_ = [("flag", P), ("a_iac", P), ("b_iac", P), ("a_q", P), ("b_q", P), ("alpha_lj", P),
                ("alpha_crf", P)]


class Bonded(C.Structure):
    _fields_ = [("n_atoms", I32),
                ("n_bonds", I32), ("n_bond_types", I32), ("bonds", P), ("bond_types", P),
                ("n_angles", I32), ("n_angle_types", I32), ("angles", P), ("angle_types", P),
                ("n_impropers", I32), ("n_improper_types", I32), ("impropers", P), ("improper_types", P),
                ("n_dihedrals", I32), ("n_dihedral_types", I32), ("dihedrals", P), ("dihedral_types", P),
                ("ntf", I32 * 4)]


class Topology(C.Structure):
    _fields_ = [("n_atoms", I32), ("n_solute_atoms", I32), ("n_cg", I32), ("n_solute_cg", I32),
                ("cg_start", P), ("charge", P), ("iac", P), ("mass", P), ("n_types", I32), ("lj", P),
                ("excl_off", P), ("excl", P), ("o14_off", P), ("o14", P), ("atoms_per_solvent", I32),
                ("dist_oh", D), ("dist_hh", D),
                ("pert_flag", P), ("pert_a_iac", P), ("pert_b_iac", P), ("pert_a_q", P), ("pert_b_q", P),
                ("pert_alpha_lj", P), ("pert_alpha_crf", P), ("n_pert_atom_pairs", I32),
                ("pert_atom_pairs", P), ("bonded", C.POINTER(Bonded)),
                ("n_pressure_groups", I32), ("pressure_group_last", P),
                ("n_solvent_constraints", I32), ("solvent_constraint_ij", P), ("solvent_constraint_len", P),
                ("n_solute_constraints", I32), ("solute_constraint_ij", P), ("solute_constraint_len", P)]


class SessionParams(C.Structure):
    _fields_ = [("device", I32), ("hbm_budget", I64), ("precision", I32), ("nsnb", I32), ("admit_interval", I32)]


class SessionInfo(C.Structure):
    _fields_ = [("queued", I32), ("running", I32), ("done", I32), ("resident_jobs", I32),
                ("device_bytes", I64), ("reserved_bytes", I64), ("budget", I64), ("clock", I64),
                ("layouts", I64), ("layout_ms", D), ("pair_count", I64), ("layout_concat_ms", D)]


class Params(C.Structure):
    _fields_ = [("boundary", I32), ("box", D * 3), ("rcutp", D), ("rcutl", D), ("skin", D),
                ("nsnb", I32), ("epsrf", D), ("kappa", D), ("four_pi_eps_i", D), ("dt", D),
                ("use_settle", I32), ("do_virial", I32), ("lam", D), ("nlam", I32),
                ("thermostat", I32), ("t_ref", D), ("tau_t", D), ("dof_t", D), ("box_matrix", D * 9),
                ("nhc", I32), ("nticom", I32), ("nscm", I32),
                ("barostat", I32), ("p_ref", D), ("compressibility", D), ("tau_p", D), ("virial", I32),
                ("solvent_shake", I32), ("shake_tolerance", D), ("solute_lincs", I32), ("solvent_lincs", I32)]


def built():
    return os.path.exists(LIB_PATH)


def lib():
    """Load libgromos_amd.so.  torch is imported first so that the process has a
    single HIP runtime: torch's bundled libamdhip64.so.7 then also serves this
    library (both carry the same SONAME, and a second copy loaded first would
    leave torch without a device)."""
    global _lib
    if _lib is None:
        try:
            import torch  # noqa: F401
        except ImportError:
            pass
        if not built():
            raise GromosAmdError(f"{LIB_PATH} is not built: run __graft_entry__.build() "
                                 "(the HIP library is required; there is no CPU fallback)")
        L = C.CDLL(LIB_PATH)
        _declare(L)
        _lib = L
    return _lib


def _declare(L):
    L.gromos_amd_last_error.restype = C.c_char_p
    L.gromos_amd_version.restype = C.c_char_p
    L.gromos_amd_device_count.restype = C.c_int
    L.gromos_amd_crf_new.argtypes = [D, D, D, D, C.POINTER(Crf)]
    L.gromos_amd_lambda_from.argtypes = [D] * 8 + [I32, C.POINTER(Lambda)]
    base = [P, P, P]
    L.gromos_amd_lj_crf_innerloop.argtypes = base + [P, I64, P, I32, C.POINTER(Crf), I32, P, D, I32, P, P, P, P]
    L.gromos_amd_lj_crf_innerloop_noalloc.argtypes = base + [P, I64, P, I32, C.POINTER(Crf), I32, P, D, I32, P, P, P,
                                                             P]
    L.gromos_amd_lj_crf_innerloop_cg_grouped.argtypes = base + [P, P, I64, P, I32, C.POINTER(Crf), I32, P, D,
                                                                I32, P, P, P, P]
    L.gromos_amd_solvent_innerloop.argtypes = base + [P, I64, P, I32, C.POINTER(Crf), I32, P, I32, D, I32, P, P,
                                                      P, P]
    L.gromos_amd_rf_excluded_interactions.argtypes = [P, P, P, I32, I32, P, C.POINTER(Crf), I32, P, D, P, P, P, P]
    L.gromos_amd_one_four_interaction_loop.argtypes = [P, P, I32, P, P, P, P, I32, C.POINTER(Crf), I32, P, D, P,
                                                       P, P, D, P]
    L.gromos_amd_perturbed_pairlist_correction.argtypes = [P, P, I64, C.POINTER(Pert), P, P, P, I32,
                                                           C.POINTER(Crf), C.POINTER(Lambda), I32, P, D, P, P, P]
    L.gromos_amd_perturbed_self_energy_correction.argtypes = [C.POINTER(Pert), I32, C.POINTER(Crf),
                                                              C.POINTER(Lambda), D, P, P]
    L.gromos_amd_perturbed_excluded_correction.argtypes = [P, P, I32, I32, P, C.POINTER(Pert), P, C.POINTER(Crf),
                                                           C.POINTER(Lambda), I32, P, D, P, P, P]
    L.gromos_amd_perturbed_one_four_correction.argtypes = [P, P, I32, P, C.POINTER(Pert), P, P, P, I32,
                                                           C.POINTER(Crf), C.POINTER(Lambda), I32, P, D, P, P, P]
    L.gromos_amd_perturbed_atom_pair_correction.argtypes = [P, I64, P, P, I32, P, P, P, P, I32, C.POINTER(Crf),
                                                            C.POINTER(Lambda), I32, P, D, P, P, P]
    L.gromos_amd_perturbed_lj_crf_interaction.argtypes = [P, P, I64, C.POINTER(Lambda), C.POINTER(Crf), D, P, P]
    L.gromos_amd_system_create.argtypes = [C.POINTER(Topology), C.POINTER(Params), I32, C.POINTER(P)]
    L.gromos_amd_system_destroy.argtypes = [P]
    L.gromos_amd_system_set_state.argtypes = [P, P, P]
    L.gromos_amd_system_get_state.argtypes = [P, P, P]
    L.gromos_amd_system_pairlist_update.argtypes = [P]
    L.gromos_amd_system_forcefield_apply.argtypes = [P, P, P, P]
    L.gromos_amd_system_md.argtypes = [P, I32, P]
    L.gromos_amd_system_covalent_energies.argtypes = [P, P]
    L.gromos_amd_system_box.argtypes = [P, P, P]
    L.gromos_amd_job_box.argtypes = [P, I64, P, P]
    L.gromos_amd_system_md_covalent.argtypes = [P, I32, P, P]
    L.gromos_amd_bonded_plan_create.argtypes = [C.POINTER(Bonded), C.POINTER(P)]
    L.gromos_amd_bonded_plan_destroy.argtypes = [P]
    L.gromos_amd_bonded_plan_terms.argtypes = [P, P]
    L.gromos_amd_bonded_forces.argtypes = [P, P, P, P, P]
    L.gromos_amd_system_pairlist_get.argtypes = [P, P, P, P, P, P]
    L.gromos_amd_system_pair_count.argtypes = [P]
    L.gromos_amd_system_pair_count.restype = I64
    L.gromos_amd_system_kernel_timing.argtypes = [P, I32, P, P]
    L.gromos_amd_system_device_bytes.argtypes = [P]
    L.gromos_amd_system_device_bytes.restype = I64
    L.gromos_amd_system_vv_lanes.argtypes = [P]
    L.gromos_amd_system_vv_lanes.restype = I32
    L.gromos_amd_system_graph_launches.argtypes = [P]
    L.gromos_amd_system_graph_launches.restype = I64
    L.gromos_amd_system_rebuild_stats.argtypes = [P, P, P]
    L.gromos_amd_system_rebuild_stats.restype = C.c_int
    L.gromos_amd_session_create.argtypes = [C.POINTER(SessionParams), C.POINTER(P)]
    L.gromos_amd_session_destroy.argtypes = [P]
    L.gromos_amd_job_footprint.argtypes = [C.POINTER(Topology), C.POINTER(Params)]
    L.gromos_amd_job_footprint.restype = I64
    L.gromos_amd_job_admit.argtypes = [P, C.POINTER(Topology), C.POINTER(Params), P, P, I64, C.POINTER(I64)]
    L.gromos_amd_session_step.argtypes = [P, I32]
    L.gromos_amd_job_query.argtypes = [P, I64, C.POINTER(I32), C.POINTER(I64)]
    L.gromos_amd_job_fetch.argtypes = [P, I64, P, P]
    L.gromos_amd_job_energies.argtypes = [P, I64, I64, I64, P]
    L.gromos_amd_job_evict.argtypes = [P, I64]
    L.gromos_amd_jobs_kinetic_energy.argtypes = [P, I32, P, P]
    L.gromos_amd_jobs_scale_velocities.argtypes = [P, I32, P, P]
    L.gromos_amd_jobs_swap_configurations.argtypes = [P, I64, I64]
    L.gromos_amd_session_get_info.argtypes = [P, C.POINTER(SessionInfo)]
    L.gromos_amd_session_step_times.argtypes = [P, I32, P, I64, C.POINTER(I64)]
    L.gromos_amd_session_kernel_timing.argtypes = [P, I32, C.POINTER(D), C.POINTER(I64), C.POINTER(I64)]
    L.gromos_amd_system_kernel_pairs.argtypes = [P]
    L.gromos_amd_system_kernel_pairs.restype = I64


def check(rc):
    if rc != 0:
        raise GromosAmdError(f"gromos_amd error {rc}: {lib().gromos_amd_last_error().decode()}")
    return rc


def ptr(a):
    if a is None:
        return None
    if hasattr(a, "data_ptr"):
        return C.c_void_p(a.data_ptr())
    return a.ctypes.data_as(C.c_void_p)


def device_count():
    return lib().gromos_amd_device_count()


def crf_new(cutoff, epsilon=1.0, rf_epsilon=61.0, rf_kappa=0.0):
    c = Crf()
    check(lib().gromos_amd_crf_new(cutoff, epsilon, rf_epsilon, rf_kappa, C.byref(c)))
    return c


def lambda_from(lam, n=1, ljs=None, crfs=None, deriv=1.0):
    lp = Lambda()
    ljs = lam if ljs is None else ljs
    crfs = lam if crfs is None else crfs
    check(lib().gromos_amd_lambda_from(lam, ljs, lam, crfs, deriv, deriv, deriv, deriv, n, C.byref(lp)))
    return lp


COVALENT = ("e_bond", "e_angle", "e_dihedral", "e_improper")


class _Keep:
    """Holds numpy arrays alive while C structs point into them."""

    def __init__(self):
        self.arrays = []

    def __call__(self, a, dtype):
        a = np.ascontiguousarray(a, dtype=dtype)
        self.arrays.append(a)
        return a.ctypes.data_as(C.c_void_p)


def topology_struct(s, keep):
    t = Topology()
    t.n_atoms, t.n_solute_atoms = s.n_atoms, s.n_solute_atoms
    t.n_cg, t.n_solute_cg = len(s.cg_start) - 1, s.n_solute_cg
    t.cg_start = keep(s.cg_start, np.int32)
    t.charge = keep(s.charge, np.float64)
    t.iac = keep(s.iac, np.int32)
    t.mass = keep(s.mass, np.float64)
    t.n_types = s.n_types
    t.lj = keep(s.lj, np.float64)
    t.excl_off, t.excl = keep(s.excl_off, np.int32), keep(s.excl, np.int32)
    t.o14_off, t.o14 = keep(s.o14_off, np.int32), keep(s.o14, np.int32)
    t.atoms_per_solvent = s.atoms_per_solvent
    t.dist_oh, t.dist_hh = s.dist_oh, s.dist_hh
    if s.pert is not None and np.any(s.pert.flag):
        p = s.pert
        t.pert_flag = keep(p.flag, np.uint8)
        t.pert_a_iac, t.pert_b_iac = keep(p.a_iac, np.int32), keep(p.b_iac, np.int32)
        t.pert_a_q, t.pert_b_q = keep(p.a_q, np.float64), keep(p.b_q, np.float64)
        t.pert_alpha_lj, t.pert_alpha_crf = keep(p.alpha_lj, np.float64), keep(p.alpha_crf, np.float64)
        ap = np.asarray(p.atom_pairs, np.int32).reshape(-1)
        t.n_pert_atom_pairs = len(ap) // 4
        t.pert_atom_pairs = keep(ap, np.int32)
    tc = getattr(s, "solvent_constraints", None)
    if (getattr(s, "solvent_shake", False) or getattr(s, "solvent_lincs", 0) > 0) and tc is not None and len(tc):
        tc = np.asarray(tc, float).reshape(-1, 3)
        t.n_solvent_constraints = len(tc)
        t.solvent_constraint_ij = keep(tc[:, :2], np.int32)
        t.solvent_constraint_len = keep(tc[:, 2], np.float64)
    sc = getattr(s, "solute_constraints", None)
    if sc is not None and len(sc):
        sc = np.asarray(sc, float).reshape(-1, 3)
        t.n_solute_constraints = len(sc)
        t.solute_constraint_ij = keep(sc[:, :2], np.int32)
        t.solute_constraint_len = keep(sc[:, 2], np.float64)
    pg = getattr(s, "pressure_groups", None)
    if pg is not None and len(pg):
        t.n_pressure_groups = len(pg)
        t.pressure_group_last = keep(np.asarray(pg), np.int32)
    b = getattr(s, "bonded", None)
    if b is not None and b.n_terms > 0:
        bs = bonded_struct(b, s.n_atoms, keep)
        keep.arrays.append(bs)
        t.bonded = C.pointer(bs)
    return t


def bonded_struct(b, n_atoms, keep):
    """gromos_amd_bonded over a system.Bonded (arrays kept alive by keep)."""
    bs = Bonded()
    bs.n_atoms = int(n_atoms)
    for nm, w, tw in (("bond", 3, 3), ("angle", 4, 3), ("improper", 5, 2), ("dihedral", 5, 3)):
        idx = np.asarray(getattr(b, nm + "s"), np.int32).reshape(-1, w)
        ty = np.asarray(getattr(b, nm + "_types"), np.float64).reshape(-1, tw)
        setattr(bs, f"n_{nm}s", len(idx))
        setattr(bs, f"n_{nm}_types", len(ty))
        setattr(bs, nm + "s", keep(idx, np.int32))
        setattr(bs, nm + "_types", keep(ty, np.float64))
    for k in range(4):
        bs.ntf[k] = int(b.ntf[k])
    return bs


def params_struct(s, use_settle=True, do_virial=False):
    p = Params()
    p.boundary = s.boundary
    for a in range(3):
        p.box[a] = float(s.box[a])
    if s.boundary == 2:
        for k, v in enumerate(s.box_values()):
            p.box_matrix[k] = float(v)
    p.rcutp, p.rcutl, p.skin = s.rcutp, s.rcutl, s.skin
    p.nsnb = s.nsnb
    p.epsrf, p.kappa, p.four_pi_eps_i, p.dt = s.epsrf, s.kappa, s.fpepsi, s.dt
    p.use_settle, p.do_virial = int(use_settle), int(do_virial)
    p.lam = s.pert.lam if s.pert is not None else 0.0
    p.nlam = s.pert.nlam if s.pert is not None else 1
    th = getattr(s, "thermostat", None)
    if th is not None:
        nhc = int(getattr(s, "thermostat_nhc", 0))
        p.thermostat = 2 if nhc > 0 else 1     # Nose-Hoover chain / Berendsen
        p.nhc = nhc
        p.t_ref, p.tau_t, p.dof_t = (float(x) for x in th)
    p.nticom, p.nscm = int(getattr(s, "nticom", 0)), int(getattr(s, "nscm", 0))
    ba = getattr(s, "barostat", None)
    if ba is not None:
        p.barostat = 1
        p.p_ref, p.compressibility, p.tau_p = (float(x) for x in ba[:3])
        p.virial = int(ba[3])
    p.solvent_shake = int(bool(getattr(s, "solvent_shake", False)))
    p.shake_tolerance = float(getattr(s, "shake_tolerance", 1e-4))
    p.solute_lincs = int(getattr(s, "solute_lincs", 0))
    p.solvent_lincs = int(getattr(s, "solvent_lincs", 0))
    return p


class DeviceSystem:
    """One MD job resident in HBM (gromos_amd_system)."""

    def __init__(self, s, precision=FP64, use_settle=True, do_virial=False):
        self.s = s
        keep = _Keep()
        self._topo = topology_struct(s, keep)
        self._keep = keep
        self._prm = params_struct(s, use_settle and s.nsm > 0 and s.atoms_per_solvent == 3, do_virial)
        h = P()
        check(lib().gromos_amd_system_create(C.byref(self._topo), C.byref(self._prm), precision, C.byref(h)))
        self.h = h
        self.n = s.n_atoms
        self.set_state(s.pos, s.vel)

    def close(self):
        if getattr(self, "h", None):
            lib().gromos_amd_system_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_state(self, pos, vel=None):
        x = np.ascontiguousarray(pos, np.float64).reshape(-1)
        v = None if vel is None else np.ascontiguousarray(vel, np.float64).reshape(-1)
        check(lib().gromos_amd_system_set_state(self.h, ptr(x), ptr(v)))

    def get_state(self):
        x, v = np.zeros((self.n, 3)), np.zeros((self.n, 3))
        check(lib().gromos_amd_system_get_state(self.h, ptr(x), ptr(v)))
        return x, v

    def pairlist_update(self):
        check(lib().gromos_amd_system_pairlist_update(self.h))

    def forcefield_apply(self):
        f = np.zeros((self.n, 3))
        e = np.zeros(3)
        vir = np.zeros(9)
        check(lib().gromos_amd_system_forcefield_apply(self.h, ptr(f), ptr(e), ptr(vir)))
        out = dict(forces=f, e_lj=e[0], e_crf=e[1], dhdl=e[2], virial=vir.reshape(3, 3))
        out.update(zip(COVALENT, self.covalent_energies()))
        return out

    def box(self):
        """(box edge lengths now, the last md step's pressure in kJ mol^-1 nm^-3)"""
        b, pr = np.zeros(3), np.zeros(1)
        check(lib().gromos_amd_system_box(self.h, ptr(b), ptr(pr)))
        return b, float(pr[0])

    def covalent_energies(self):
        """e_bond, e_angle, e_dihedral, e_improper of the last force evaluation."""
        e = np.zeros(4)
        check(lib().gromos_amd_system_covalent_energies(self.h, ptr(e)))
        return e

    def md(self, nsteps, covalent=False):
        """(nsteps, 5) e_lj, e_crf, e_pot, e_kin, dhdl; with covalent=True also
        (nsteps, 4) e_bond, e_angle, e_dihedral, e_improper."""
        e = np.zeros((nsteps, 5))
        if not covalent:
            check(lib().gromos_amd_system_md(self.h, int(nsteps), ptr(e)))
            return e
        c = np.zeros((nsteps, 4))
        check(lib().gromos_amd_system_md_covalent(self.h, int(nsteps), ptr(e), ptr(c)))
        return e, c

    def pairlist(self):
        n = np.zeros(4, np.int64)
        check(lib().gromos_amd_system_pairlist_get(self.h, ptr(n), None, None, None, None))
        bufs = [np.zeros((max(int(k), 1), 2), np.uint32) for k in n]
        check(lib().gromos_amd_system_pairlist_get(self.h, ptr(n), *[ptr(b) for b in bufs]))
        names = ("solute_short", "solute_long", "solvent_short", "solvent_long")
        return {k: b[:int(c)] for k, b, c in zip(names, bufs, n)}

    @property
    def pair_count(self):
        return lib().gromos_amd_system_pair_count(self.h)

    @property
    def device_bytes(self):
        return lib().gromos_amd_system_device_bytes(self.h)

    @property
    def vv_lanes(self):
        """lanes per solvent molecule chosen by the autotuner"""
        return lib().gromos_amd_system_vv_lanes(self.h)

    @property
    def graph_launches(self):
        """md steps run as a replayed graph so far"""
        return lib().gromos_amd_system_graph_launches(self.h)

    @property
    def rebuild_stats(self):
        """(pair-list builds, builds that refilled with exact slots)"""
        a, b = I64(), I64()
        check(lib().gromos_amd_system_rebuild_stats(self.h, C.byref(a), C.byref(b)))
        return a.value, b.value

    def kernel_timing(self, enable=-1):
        ms, n = D(), I64()
        check(lib().gromos_amd_system_kernel_timing(self.h, int(enable), C.byref(ms), C.byref(n)))
        return ms.value, n.value

    @property
    def kernel_pairs(self):
        """pair interactions the timed solvent-kernel launches evaluated"""
        return lib().gromos_amd_system_kernel_pairs(self.h)


EBUSY = -5
QUEUED, RUNNING, DONE, FAILED = 0, 1, 2, 3


class Busy(GromosAmdError):
    """The job's footprint does not fit the session's remaining HBM budget now."""


class Session:
    """Continuous batching of independent jobs on one device (gromos_amd_session).

    Jobs are admitted against an HBM budget and join / leave the resident batch
    at clock multiples of ``admit_interval`` (a multiple of the shared pair-list
    cadence ``nsnb``); every job computes exactly what it computes alone."""

    def __init__(self, precision=FP64, nsnb=5, admit_interval=0, hbm_budget=0, device=0):
        sp = SessionParams(device=device, hbm_budget=int(hbm_budget), precision=precision, nsnb=nsnb,
                           admit_interval=admit_interval)
        h = P()
        check(lib().gromos_amd_session_create(C.byref(sp), C.byref(h)))
        self.h = h
        self.n_atoms = {}

    def close(self):
        if getattr(self, "h", None):
            lib().gromos_amd_session_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @staticmethod
    def footprint(s, use_settle=True):
        keep = _Keep()
        t = topology_struct(s, keep)
        p = params_struct(s, use_settle and s.nsm > 0 and s.atoms_per_solvent == 3)
        b = lib().gromos_amd_job_footprint(C.byref(t), C.byref(p))
        if b < 0:
            check(int(b))
        return b

    def admit(self, s, nsteps, use_settle=True, do_virial=False):
        """queue job `s` for `nsteps` md steps; raises Busy when the budget is full"""
        keep = _Keep()
        t = topology_struct(s, keep)
        p = params_struct(s, use_settle and s.nsm > 0 and s.atoms_per_solvent == 3, do_virial)
        x = np.ascontiguousarray(s.pos, np.float64).reshape(-1)
        v = None if s.vel is None else np.ascontiguousarray(s.vel, np.float64).reshape(-1)
        jid = I64()
        rc = lib().gromos_amd_job_admit(self.h, C.byref(t), C.byref(p), ptr(x), ptr(v), int(nsteps), C.byref(jid))
        if rc == EBUSY:
            raise Busy(lib().gromos_amd_last_error().decode())
        check(rc)
        self.n_atoms[jid.value] = s.n_atoms
        return jid.value

    def step(self, nsteps):
        check(lib().gromos_amd_session_step(self.h, int(nsteps)))

    def query(self, job):
        st, done = I32(), I64()
        check(lib().gromos_amd_job_query(self.h, int(job), C.byref(st), C.byref(done)))
        return st.value, done.value

    def fetch(self, job):
        n = self.n_atoms[job]
        x, v = np.zeros((n, 3)), np.zeros((n, 3))
        check(lib().gromos_amd_job_fetch(self.h, int(job), ptr(x), ptr(v)))
        return x, v

    def box(self, job):
        """(the job's box edge lengths now, its last md step's pressure)"""
        b, pr = np.zeros(3), np.zeros(1)
        check(lib().gromos_amd_job_box(self.h, int(job), ptr(b), ptr(pr)))
        return b, float(pr[0])

    def energies(self, job, first=0, count=None):
        if count is None:
            count = self.query(job)[1] - first
        e = np.zeros((max(count, 0), 5))
        if count > 0:
            check(lib().gromos_amd_job_energies(self.h, int(job), int(first), int(count), ptr(e)))
        return e

    def evict(self, job):
        check(lib().gromos_amd_job_evict(self.h, int(job)))
        self.n_atoms.pop(job, None)

    def kinetic_energy(self, jobs):
        ids = np.ascontiguousarray(jobs, np.int64)
        out = np.zeros(len(ids))
        check(lib().gromos_amd_jobs_kinetic_energy(self.h, len(ids), ptr(ids), ptr(out)))
        return out

    def scale_velocities(self, jobs, factors):
        ids = np.ascontiguousarray(jobs, np.int64)
        f = np.ascontiguousarray(factors, np.float64)
        check(lib().gromos_amd_jobs_scale_velocities(self.h, len(ids), ptr(ids), ptr(f)))

    def swap_configurations(self, a, b):
        check(lib().gromos_amd_jobs_swap_configurations(self.h, int(a), int(b)))

    def info(self):
        i = SessionInfo()
        check(lib().gromos_amd_session_get_info(self.h, C.byref(i)))
        return {k: getattr(i, k) for k, _ in SessionInfo._fields_}

    def step_times(self, enable=-1):
        n = I64()
        check(lib().gromos_amd_session_step_times(self.h, -1, None, 0, C.byref(n)))
        ms = np.zeros(n.value)
        check(lib().gromos_amd_session_step_times(self.h, int(enable), ptr(ms), n.value, C.byref(n)))
        return ms

    def kernel_timing(self, enable=-1):
        """(total ms, launches, pair interactions) of the timed solvent-kernel launches"""
        ms, n, pr = D(), I64(), I64()
        check(lib().gromos_amd_session_kernel_timing(self.h, int(enable), C.byref(ms), C.byref(n), C.byref(pr)))
        return ms.value, n.value, pr.value
