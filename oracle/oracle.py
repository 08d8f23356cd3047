"""ctypes front end of the CPU oracle (oracle/gromos_oracle.c).

TEST INFRASTRUCTURE: imported only by tests/, by the smoke check and by the
CPU-baseline leg of bench.py, always as the checker.  Product code never
imports this module.
"""
import ctypes as C
import os
import subprocess

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "_build", "libgromos_oracle.so")

_lib = None


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.run(["make", "-s", "-C", HERE], check=True)
    return LIB_PATH


def lib():
    global _lib
    if _lib is None:
        build()
        _lib = C.CDLL(LIB_PATH)
        _declare(_lib)
    return _lib


P = C.c_void_p
D = C.c_double
I32 = C.c_int32
I64 = C.c_int64


class Pbc(C.Structure):
    _fields_ = [("kind", C.c_int), ("L", D * 3), ("invL", D * 3), ("half", D * 3), ("a", D * 3), ("b", D * 3),
                ("c", D * 3)]


class Crf(C.Structure):
    _fields_ = [("crf_cut", D), ("crf_2cut3i", D), ("crf_cut3i", D), ("cutoff_sq", D)]


class Bonded(C.Structure):
    _fields_ = [("nb", I32), ("bonds", P), ("bond_types", P), ("na", I32), ("angles", P), ("angle_types", P),
                ("ni", I32), ("impropers", P), ("improper_types", P), ("nd", I32), ("dihedrals", P),
                ("dihedral_types", P), ("nbt", I32), ("nat", I32), ("nit", I32), ("ndt", I32),
                ("ntf", I32 * 4)]


class BondedE(C.Structure):
    _fields_ = [("e_bond", D), ("e_angle", D), ("e_dihedral", D), ("e_improper", D)]


class Topo(C.Structure):
    _fields_ = [("n_atoms", I32), ("n_solute_atoms", I32), ("n_cg", I32), ("n_solute_cg", I32),
                ("cg_start", P), ("charge", P), ("iac", P), ("mass", P), ("n_types", I32),
                ("lj", P), ("excl_off", P), ("excl", P), ("o14_off", P), ("o14", P),
                ("atoms_per_solvent", I32), ("bonded", C.POINTER(Bonded))]


class PList(C.Structure):
    _fields_ = [("n", I64), ("cap", I64), ("p", C.POINTER(C.c_uint32))]


class PairList(C.Structure):
    _fields_ = [("solute_short", PList), ("solute_long", PList), ("solvent_short", PList),
                ("solvent_long", PList)]


class Storage(C.Structure):
    _fields_ = [("n", I32), ("f", C.POINTER(D)), ("e_lj", D), ("e_crf", D), ("vir", D * 9)]


class Lambda(C.Structure):
    _fields_ = [(k, D) for k in (
        "a_lj_lambda_n", "b_lj_lambda_n", "a_crf_lambda_n", "b_crf_lambda_n",
        "a_lj_lambda_n_1", "b_lj_lambda_n_1", "a_crf_lambda_n_1", "b_crf_lambda_n_1",
        "a_ljs_lambda", "b_ljs_lambda", "a_ljs_lambda2", "b_ljs_lambda2",
        "a_crfs_lambda", "b_crfs_lambda", "a_crfs_lambda2", "b_crfs_lambda2")] + [("lambda_exp", I32)]


class Pert(C.Structure):
    _fields_ = [("flag", P), ("a_iac", P), ("b_iac", P), ("a_q", P), ("b_q", P),
                ("alpha_lj", P), ("alpha_crf", P), ("n_atom_pairs", I32), ("atom_pairs", P)]


class PertCorr(C.Structure):
    _fields_ = [("f", C.POINTER(D)), ("de_lj", D), ("de_crf", D), ("dhdl", D)]


class FFParams(C.Structure):
    _fields_ = [("rcutp", D), ("rcutl", D), ("skin", D), ("nsnb", I32), ("algorithm", I32),
                ("epsrf", D), ("kappa", D), ("fpepsi", D), ("parallel", I32), ("virial", I32),
                ("lam", D), ("nlam", I32)]


def _declare(L):
    L.or_pbc_init.argtypes = [C.POINTER(Pbc), C.c_int, P]
    L.or_crf_new.argtypes = [D, D, D, D, C.POINTER(Crf)]
    L.or_nearest_image.argtypes = [C.POINTER(Pbc), P, P, P]
    L.or_lj_crf_interaction.argtypes = [P, D, D, D, C.POINTER(Crf), P, P, P]
    L.or_pairlist_standard.argtypes = [C.POINTER(Topo), P, C.POINTER(Pbc), D, D, D, C.POINTER(PairList)]
    L.or_pairlist_cell.argtypes = [C.POINTER(Topo), P, C.POINTER(Pbc), D, D, D, C.POINTER(PairList)]
    L.or_pairlist_free.argtypes = [C.POINTER(PairList)]
    L.or_is_excluded_or_14.argtypes = [C.POINTER(Topo), C.c_int, C.c_int]
    L.or_storage_init.argtypes = [C.POINTER(Storage), C.c_int]
    L.or_storage_free.argtypes = [C.POINTER(Storage)]
    for nm in ("or_lj_crf_innerloop", "or_lj_crf_innerloop_parallel"):
        getattr(L, nm).argtypes = [P, P, P, P, I64, P, I32, C.POINTER(Crf), C.POINTER(Pbc), D, C.c_int,
                                   C.POINTER(Storage)]
    L.or_lj_crf_innerloop_cg_grouped.argtypes = [P, P, P, P, P, I64, P, I32, C.POINTER(Crf),
                                                 C.POINTER(Pbc), D, C.c_int, C.POINTER(Storage)]
    for nm in ("or_solvent_innerloop", "or_solvent_innerloop_parallel"):
        getattr(L, nm).argtypes = [P, P, P, P, I64, P, I32, C.POINTER(Crf), C.POINTER(Pbc), I32, D,
                                   C.c_int, C.POINTER(Storage)]
    L.or_rf_excluded_interactions.argtypes = [C.POINTER(Topo), P, C.POINTER(Crf), C.POINTER(Pbc), D,
                                              C.POINTER(Storage)]
    L.or_one_four_interaction_loop.argtypes = [C.POINTER(Topo), P, C.POINTER(Crf), C.POINTER(Pbc), D, D,
                                               C.POINTER(Storage)]
    L.or_lambda_from.argtypes = [D] * 8 + [C.c_int, C.POINTER(Lambda)]
    L.or_perturbed_lj_crf_interaction.argtypes = [P, D, D, D, D, D, D, D, D, C.POINTER(Lambda),
                                                  C.POINTER(Crf), D, P]
    L.or_perturbed_pairlist_correction.argtypes = [P, P, I64, C.POINTER(Pert), P, P, P, I32, C.POINTER(Crf),
                                                   C.POINTER(Lambda), C.POINTER(Pbc), D, C.c_int,
                                                   C.POINTER(PertCorr)]
    L.or_perturbed_self_energy_correction.argtypes = [C.POINTER(Topo), C.POINTER(Pert), C.POINTER(Crf),
                                                      C.POINTER(Lambda), D, P, P]
    for nm in ("or_perturbed_excluded_correction", "or_perturbed_one_four_correction",
               "or_perturbed_atom_pair_correction"):
        getattr(L, nm).argtypes = [C.POINTER(Topo), P, C.POINTER(Pert), C.POINTER(Crf), C.POINTER(Lambda),
                                   C.POINTER(Pbc), D, C.POINTER(PertCorr)]
    L.or_ff_create.argtypes = [C.POINTER(Topo), C.POINTER(Pbc), C.POINTER(FFParams), C.POINTER(Pert)]
    L.or_ff_create.restype = P
    L.or_ff_destroy.argtypes = [P]
    L.or_ff_initial_pairlist.argtypes = [P, P]
    L.or_ff_apply.argtypes = [P, P, P, P, P, P, P]
    L.or_ff_pairlist.argtypes = [P]
    L.or_ff_last_bonded.argtypes = [P, C.POINTER(BondedE)]
    L.or_bonded_forces.argtypes = [C.POINTER(Bonded), C.c_int, P, P, P, C.POINTER(BondedE)]
    L.or_ff_pairlist.restype = C.POINTER(PairList)
    L.or_ff_used_cg_grouped.argtypes = [P]
    L.or_ff_last_pair_count.argtypes = [P]
    L.or_ff_last_pair_count.restype = I64
    L.or_leapfrog.argtypes = [C.c_int, P, D, P, P, P, P, P]
    L.or_settle.argtypes = [C.c_int, C.c_int, D, D, D, D, D, P, P, P, P, P]
    L.or_ekin.argtypes = [C.c_int, P, P, P]
    L.or_ekin.restype = D
    L.or_md_create.argtypes = [C.POINTER(Topo), C.POINTER(Pbc), C.POINTER(FFParams), C.POINTER(Pert), D,
                               C.c_int, P, P]
    L.or_md_create.restype = P
    L.or_md_destroy.argtypes = [P]
    L.or_md_set_settle_geometry.argtypes = [P, D, D]
    L.or_md_set_berendsen.argtypes = [P, D, D, D]
    L.or_md_set_nosehoover.argtypes = [P, D, D, D, C.c_int]
    L.or_md_set_nosehoover.restype = C.c_int
    L.or_nosehoover_step.argtypes = [P, C.c_int, D, D, D, D, D]
    L.or_nosehoover_step.restype = D
    L.or_md_get_nosehoover.argtypes = [P, P]
    L.or_md_set_com.argtypes = [P, C.c_int, C.c_int]
    L.or_md_set_barostat.argtypes = [P, D, D, D, C.c_int, C.c_int, P]
    L.or_md_set_barostat.restype = C.c_int
    L.or_md_set_shake.argtypes = [P, C.c_int, P, P, C.c_int, P, P, D]
    L.or_shake.argtypes = [C.c_int, P, C.c_int, P, P, P, D, C.c_int, D, P, P, P, P]
    L.or_shake.restype = C.c_int
    L.or_md_set_lincs.argtypes = [P, C.c_int, C.c_int, P, P, C.c_int, C.c_int, P, P]
    L.or_md_set_lincs.restype = C.c_int
    L.or_lincs.argtypes = [C.c_int, P, P, P, C.c_int, C.c_int, P, D, P, P, P]
    L.or_lincs.restype = D
    L.or_md_pressure.argtypes = [P]
    L.or_md_pressure.restype = D
    L.or_md_box.argtypes = [P, P]
    L.or_remove_com.argtypes = [C.c_int, P, P, P, C.POINTER(Pbc), D, C.c_int, C.c_int, C.c_int64]
    L.or_berendsen_scale.argtypes = [C.c_int, P, P, D, D, D, D]
    L.or_berendsen_scale.restype = D
    L.or_md_run.argtypes = [P, C.c_int, P]
    L.or_md_get.argtypes = [P, P, P, P]
    L.or_md_pair_count.argtypes = [P]
    L.or_md_pair_count.restype = I64
    L.or_num_threads.restype = C.c_int
    L.or_set_num_threads.argtypes = [C.c_int]


def ptr(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def c64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def c32(a):
    return np.ascontiguousarray(a, dtype=np.int32)


class OracleSystem:
    """Holds the C structs (and the numpy arrays they point into) for a System."""

    def __init__(self, s):
        self.s = s
        self._keep = dict(
            cg_start=c32(s.cg_start), charge=c64(s.charge), iac=c32(s.iac), mass=c64(s.mass),
            lj=c64(s.lj), excl_off=c32(s.excl_off), excl=c32(s.excl),
            o14_off=c32(s.o14_off), o14=c32(s.o14))
        k = self._keep
        self.topo = Topo(s.n_atoms, s.n_solute_atoms, len(s.cg_start) - 1, s.n_solute_cg,
                         ptr(k["cg_start"]), ptr(k["charge"]), ptr(k["iac"]), ptr(k["mass"]),
                         s.n_types, ptr(k["lj"]), ptr(k["excl_off"]), ptr(k["excl"]),
                         ptr(k["o14_off"]), ptr(k["o14"]), s.atoms_per_solvent, None)
        self.n = s.n_atoms
        self.bonded = _bonded_struct(getattr(s, "bonded", None), k)
        if self.bonded is not None:
            self.topo.bonded = C.pointer(self.bonded)
        self.pbc = make_pbc(s.boundary, s.box_values())
        self.pert = None
        if s.pert is not None:
            p = s.pert
            k.update(pflag=np.ascontiguousarray(p.flag, np.uint8), pa_iac=c32(p.a_iac), pb_iac=c32(p.b_iac),
                     pa_q=c64(p.a_q), pb_q=c64(p.b_q), palj=c64(p.alpha_lj), pacrf=c64(p.alpha_crf),
                     pairs=c32(np.asarray(p.atom_pairs, np.int32).reshape(-1)))
            self.pert = Pert(ptr(k["pflag"]), ptr(k["pa_iac"]), ptr(k["pb_iac"]), ptr(k["pa_q"]),
                             ptr(k["pb_q"]), ptr(k["palj"]), ptr(k["pacrf"]),
                             len(k["pairs"]) // 4, ptr(k["pairs"]))

    def ff_params(self, parallel=False, virial=True):
        s = self.s
        lam = s.pert.lam if s.pert is not None else 0.0
        nlam = s.pert.nlam if s.pert is not None else 1
        return FFParams(s.rcutp, s.rcutl, s.skin, s.nsnb, s.algorithm, s.epsrf, s.kappa, s.fpepsi,
                        int(parallel), int(virial), lam, nlam)


def make_pbc(kind, box):
    p = Pbc()
    b = c64(box)
    lib().or_pbc_init(C.byref(p), int(kind), ptr(b))
    return p


def crf_new(cutoff, eps, epsrf, kappa):
    c = Crf()
    lib().or_crf_new(cutoff, eps, epsrf, kappa, C.byref(c))
    return c


def nearest_image(pbc, ri, rj):
    a, b, r = c64(ri), c64(rj), np.zeros(3)
    lib().or_nearest_image(C.byref(pbc), ptr(a), ptr(b), ptr(r))
    return r


def lj_crf_interaction(r, c6, c12, qp, crf):
    rr = c64(r)
    out = (D(), D(), D())
    lib().or_lj_crf_interaction(ptr(rr), c6, c12, qp, C.byref(crf), *[C.byref(x) for x in out])
    return tuple(x.value for x in out)


def _plist(pl):
    if pl.n == 0:
        return np.zeros((0, 2), np.uint32)
    return np.ctypeslib.as_array(pl.p, shape=(pl.n * 2,)).reshape(-1, 2).copy()


def pairlist(osys, pos=None, algorithm=0, rcutp=None, rcutl=None, skin=None):
    s = osys.s
    x = c64(s.pos if pos is None else pos)
    pl = PairList()
    fn = lib().or_pairlist_cell if algorithm == 2 else lib().or_pairlist_standard
    fn(C.byref(osys.topo), ptr(x), C.byref(osys.pbc), s.rcutp if rcutp is None else rcutp,
       s.rcutl if rcutl is None else rcutl, s.skin if skin is None else skin, C.byref(pl))
    out = dict(solute_short=_plist(pl.solute_short), solute_long=_plist(pl.solute_long),
               solvent_short=_plist(pl.solvent_short), solvent_long=_plist(pl.solvent_long))
    lib().or_pairlist_free(C.byref(pl))
    return out


class Store:
    def __init__(self, n):
        self.st = Storage()
        lib().or_storage_init(C.byref(self.st), n)
        self.n = n

    @property
    def forces(self):
        return np.ctypeslib.as_array(self.st.f, shape=(self.n * 3,)).reshape(-1, 3).copy()

    @property
    def virial(self):
        return np.array(self.st.vir[:]).reshape(3, 3)

    def __del__(self):
        try:
            lib().or_storage_free(C.byref(self.st))
        except Exception:
            pass


def innerloop(kind, osys, pairs, crf, pos=None, virial=True, groups=None, parallel=False):
    s = osys.s
    st = Store(s.n_atoms)
    x = c64(s.pos if pos is None else pos)
    k = osys._keep
    pr = np.ascontiguousarray(pairs, np.uint32).reshape(-1)
    np_ = len(pr) // 2
    if kind == "plain":
        fn = lib().or_lj_crf_innerloop_parallel if parallel else lib().or_lj_crf_innerloop
        fn(ptr(x), ptr(k["charge"]), ptr(k["iac"]), ptr(pr), np_, ptr(k["lj"]), s.n_types, C.byref(crf),
           C.byref(osys.pbc), s.fpepsi, int(virial), C.byref(st.st))
    elif kind == "cg_grouped":
        g = np.ascontiguousarray(groups, np.uint32).reshape(-1)
        lib().or_lj_crf_innerloop_cg_grouped(ptr(x), ptr(k["charge"]), ptr(k["iac"]), ptr(pr), ptr(g),
                                             len(g) // 4, ptr(k["lj"]), s.n_types, C.byref(crf),
                                             C.byref(osys.pbc), s.fpepsi, int(virial), C.byref(st.st))
    elif kind == "solvent":
        fn = lib().or_solvent_innerloop_parallel if parallel else lib().or_solvent_innerloop
        fn(ptr(x), ptr(k["charge"]), ptr(k["iac"]), ptr(pr), np_, ptr(k["lj"]), s.n_types, C.byref(crf),
           C.byref(osys.pbc), s.atoms_per_solvent, s.fpepsi, int(virial), C.byref(st.st))
    else:
        raise ValueError(kind)
    return st


def rf_excluded(osys, crf, pos=None):
    s = osys.s
    st = Store(s.n_atoms)
    x = c64(s.pos if pos is None else pos)
    lib().or_rf_excluded_interactions(C.byref(osys.topo), ptr(x), C.byref(crf), C.byref(osys.pbc), s.fpepsi,
                                      C.byref(st.st))
    return st


def one_four(osys, crf, pos=None, coulomb_scaling=1.0):
    s = osys.s
    st = Store(s.n_atoms)
    x = c64(s.pos if pos is None else pos)
    lib().or_one_four_interaction_loop(C.byref(osys.topo), ptr(x), C.byref(crf), C.byref(osys.pbc), s.fpepsi,
                                       coulomb_scaling, C.byref(st.st))
    return st


def lambda_params(lam, n=1, ljs=None, crfs=None, deriv=1.0):
    lp = Lambda()
    ljs = lam if ljs is None else ljs
    crfs = lam if crfs is None else crfs
    lib().or_lambda_from(lam, ljs, lam, crfs, deriv, deriv, deriv, deriv, n, C.byref(lp))
    return lp


def perturbed_interaction(r, a_c6, a_c12, b_c6, b_c12, a_q, b_q, alj, acrf, lp, crf, fpepsi):
    rr = c64(r)
    out = np.zeros(5)
    lib().or_perturbed_lj_crf_interaction(ptr(rr), a_c6, a_c12, b_c6, b_c12, a_q, b_q, alj, acrf,
                                          C.byref(lp), C.byref(crf), fpepsi, ptr(out))
    return out


def _corr(n):
    f = np.zeros(3 * n)
    pc = PertCorr(f.ctypes.data_as(C.POINTER(D)), 0.0, 0.0, 0.0)
    return f, pc


def _corr_out(f, pc):
    return dict(forces=f.reshape(-1, 3).copy(), de_lj=pc.de_lj, de_crf=pc.de_crf, dhdl=pc.dhdl)


def _crf_of(s):
    return crf_new(s.rcutl, 1.0, s.epsrf, s.kappa)


def pert_pairlist_correction(osys, pairs, lp, pos=None):
    """perturbed_pairlist_correction (perturbed.rs:309-386) over one atom-pair list."""
    s = osys.s
    k = osys._keep
    x = c64(s.pos if pos is None else pos)
    pr = np.ascontiguousarray(pairs, np.uint32).reshape(-1)
    f, pc = _corr(s.n_atoms)
    crf = _crf_of(s)
    lib().or_perturbed_pairlist_correction(ptr(x), ptr(pr), len(pr) // 2, C.byref(osys.pert), ptr(k["charge"]),
                                           ptr(k["iac"]), ptr(k["lj"]), s.n_types, C.byref(crf), C.byref(lp),
                                           C.byref(osys.pbc), s.fpepsi, s.n_atoms, C.byref(pc))
    return _corr_out(f, pc)


def pert_self_correction(osys, lp):
    """perturbed_self_energy_correction (perturbed.rs:395-425): (delta E_crf, dH/dlambda)."""
    s = osys.s
    crf = _crf_of(s)
    de, dh = D(), D()
    lib().or_perturbed_self_energy_correction(C.byref(osys.topo), C.byref(osys.pert), C.byref(crf), C.byref(lp),
                                              s.fpepsi, C.byref(de), C.byref(dh))
    return de.value, dh.value


def _pass(name, osys, lp, pos):
    s = osys.s
    x = c64(s.pos if pos is None else pos)
    f, pc = _corr(s.n_atoms)
    crf = _crf_of(s)
    getattr(lib(), name)(C.byref(osys.topo), ptr(x), C.byref(osys.pert), C.byref(crf), C.byref(lp),
                         C.byref(osys.pbc), s.fpepsi, C.byref(pc))
    return _corr_out(f, pc)


def pert_excluded_correction(osys, lp, pos=None):
    """perturbed_excluded_correction (perturbed.rs:431-511)."""
    return _pass("or_perturbed_excluded_correction", osys, lp, pos)


def pert_one_four_correction(osys, lp, pos=None):
    """perturbed_one_four_correction (perturbed.rs:517-605)."""
    return _pass("or_perturbed_one_four_correction", osys, lp, pos)


def pert_atom_pair_correction(osys, lp, pos=None):
    """perturbed_atom_pair_correction (perturbed.rs:624-707)."""
    return _pass("or_perturbed_atom_pair_correction", osys, lp, pos)


def _bonded_struct(b, keep):
    """or_bonded over a system.Bonded (arrays kept in the dict keep); None
    without terms"""
    if b is None or b.n_terms == 0:
        return None
    for nm in ("bonds", "angles", "impropers", "dihedrals"):
        keep[nm] = c32(getattr(b, nm))
    for nm in ("bond_types", "angle_types", "improper_types", "dihedral_types"):
        keep[nm] = c64(getattr(b, nm))
    ntf = (I32 * 4)(*[int(x) for x in b.ntf])
    return Bonded(len(b.bonds), ptr(keep["bonds"]), ptr(keep["bond_types"]),
                  len(b.angles), ptr(keep["angles"]), ptr(keep["angle_types"]),
                  len(b.impropers), ptr(keep["impropers"]), ptr(keep["improper_types"]),
                  len(b.dihedrals), ptr(keep["dihedrals"]), ptr(keep["dihedral_types"]),
                  len(b.bond_types), len(b.angle_types), len(b.improper_types), len(b.dihedral_types), ntf)


class BondedOnly:
    """Covalent terms without a system (positions x give the atom count)."""

    def __init__(self, x, bonded):
        self._keep = {}
        self.n = len(x)
        self.bonded = _bonded_struct(bonded, self._keep)


def bonded_forces(osys, pos=None):
    """calculate_bonded_forces_ntf (bonded/mod.rs:141-185) alone: forces, virial
    and the four covalent energies (osys: OracleSystem or BondedOnly)."""
    n = osys.n
    x = c64(osys.s.pos if pos is None else pos)
    f, vir, be = np.zeros((n, 3)), np.zeros(9), BondedE()
    if osys.bonded is None:
        return dict(forces=f, virial=vir.reshape(3, 3), e_bond=0.0, e_angle=0.0, e_dihedral=0.0, e_improper=0.0)
    lib().or_bonded_forces(C.byref(osys.bonded), n, ptr(x), ptr(f), ptr(vir), C.byref(be))
    return dict(forces=f, virial=vir.reshape(3, 3), e_bond=be.e_bond, e_angle=be.e_angle,
                e_dihedral=be.e_dihedral, e_improper=be.e_improper)


class ForceField:
    """Forcefield::apply restated: nonbonded, plus the covalent terms when the
    system has them."""

    def __init__(self, osys, parallel=False, virial=True):
        self.osys = osys
        self.prm = osys.ff_params(parallel, virial)
        self.h = lib().or_ff_create(C.byref(osys.topo), C.byref(osys.pbc), C.byref(self.prm),
                                    C.byref(osys.pert) if osys.pert is not None else None)
        self.n = osys.s.n_atoms

    def initial_pairlist(self, pos):
        x = c64(pos)
        lib().or_ff_initial_pairlist(self.h, ptr(x))

    def apply(self, pos):
        x = c64(pos)
        f = np.zeros((self.n, 3))
        vals = [D() for _ in range(3)]
        vir = np.zeros(9)
        lib().or_ff_apply(self.h, ptr(x), ptr(f), C.byref(vals[0]), C.byref(vals[1]), ptr(vir),
                          C.byref(vals[2]))
        be = BondedE()
        lib().or_ff_last_bonded(self.h, C.byref(be))
        return dict(forces=f, e_lj=vals[0].value, e_crf=vals[1].value, virial=vir.reshape(3, 3),
                    dhdl=vals[2].value, pairs=lib().or_ff_last_pair_count(self.h),
                    e_bond=be.e_bond, e_angle=be.e_angle, e_dihedral=be.e_dihedral, e_improper=be.e_improper)

    def pairlist(self):
        pl = lib().or_ff_pairlist(self.h).contents
        return dict(solute_short=_plist(pl.solute_short), solute_long=_plist(pl.solute_long),
                    solvent_short=_plist(pl.solvent_short), solvent_long=_plist(pl.solvent_long))

    @property
    def used_cg_grouped(self):
        return bool(lib().or_ff_used_cg_grouped(self.h))

    def __del__(self):
        try:
            lib().or_ff_destroy(self.h)
        except Exception:
            pass


def single_point(osys, parallel=False, virial=True, pos=None):
    ff = ForceField(osys, parallel, virial)
    p = osys.s.pos if pos is None else pos
    ff.initial_pairlist(p)
    return ff.apply(p)


class MD:
    """Forcefield -> leap-frog -> [SETTLE] -> E_kin sequence on the CPU."""

    def __init__(self, osys, use_settle=True, parallel=False, virial=False):
        s = osys.s
        self.osys = osys
        self.prm = osys.ff_params(parallel, virial)
        self.x0, self.v0 = c64(s.pos), c64(s.vel)
        self.h = lib().or_md_create(C.byref(osys.topo), C.byref(osys.pbc), C.byref(self.prm),
                                    C.byref(osys.pert) if osys.pert is not None else None, s.dt,
                                    int(use_settle), ptr(self.x0), ptr(self.v0))
        lib().or_md_set_settle_geometry(self.h, s.dist_oh, s.dist_hh)
        self.nhc = 0
        if getattr(s, "thermostat", None) is not None:
            t_ref, tau, dof = s.thermostat
            nhc = int(getattr(s, "thermostat_nhc", 0))
            if nhc > 0:
                if lib().or_md_set_nosehoover(self.h, t_ref, tau, dof, nhc) != 0:
                    raise ValueError(f"Nose-Hoover chain length {nhc} out of range")
                self.nhc = nhc
            else:
                lib().or_md_set_berendsen(self.h, t_ref, tau, dof)
        if getattr(s, "nticom", 0) or getattr(s, "nscm", 0):
            lib().or_md_set_com(self.h, int(s.nticom), int(s.nscm))
        sol_c = getattr(s, "solute_constraints", None)
        tmpl = getattr(s, "solvent_constraints", None) if getattr(s, "solvent_shake", False) else None
        lu, lv = int(getattr(s, "solute_lincs", 0)), int(getattr(s, "solvent_lincs", 0))
        if lu > 0 or lv > 0:
            # LINCS takes the solute's constraints (order lu) and / or the solvent
            # template (order lv) in place of SHAKE
            lsc = np.zeros((0, 3)) if (lu == 0 or sol_c is None) else np.asarray(sol_c, float).reshape(-1, 3)
            ltc = np.zeros((0, 3)) if lv == 0 else np.asarray(s.solvent_constraints, float).reshape(-1, 3)
            self.lc = (np.ascontiguousarray(lsc[:, :2], np.int32), np.ascontiguousarray(lsc[:, 2]),
                       np.ascontiguousarray(ltc[:, :2], np.int32), np.ascontiguousarray(ltc[:, 2]))
            if lib().or_md_set_lincs(self.h, lu, len(lsc), ptr(self.lc[0]), ptr(self.lc[1]), lv, len(ltc),
                                     ptr(self.lc[2]), ptr(self.lc[3])) != 0:
                raise ValueError("a LINCS solvent template holds at most 6 constraints")
            if lu > 0:
                sol_c = None
        if (sol_c is not None and len(sol_c)) or tmpl is not None:
            sc = np.zeros((0, 3)) if sol_c is None else np.asarray(sol_c, float).reshape(-1, 3)
            tc = np.zeros((0, 3)) if tmpl is None else np.asarray(tmpl, float).reshape(-1, 3)
            self.sh = (np.ascontiguousarray(sc[:, :2], np.int32), np.ascontiguousarray(sc[:, 2]),
                       np.ascontiguousarray(tc[:, :2], np.int32), np.ascontiguousarray(tc[:, 2]))
            lib().or_md_set_shake(self.h, len(sc), ptr(self.sh[0]), ptr(self.sh[1]), len(tc), ptr(self.sh[2]),
                                  ptr(self.sh[3]), float(s.shake_tolerance))
        if getattr(s, "barostat", None) is not None:
            p_ref, comp, tau_p, vtype = s.barostat
            self.pg = np.ascontiguousarray(s.pressure_group_last(), dtype=np.int32)
            if lib().or_md_set_barostat(self.h, p_ref, comp, tau_p, int(vtype), len(self.pg), ptr(self.pg)) != 0:
                raise ValueError("the barostat needs a periodic box and virial kind 0, 1 or 2")
        self.n = s.n_atoms

    def run(self, nsteps):
        e = np.zeros((nsteps, 5))
        rc = lib().or_md_run(self.h, nsteps, ptr(e))
        if rc != 0:
            raise RuntimeError("SETTLE or SHAKE failed in the oracle")
        return e

    def state(self):
        x, v, f = np.zeros((self.n, 3)), np.zeros((self.n, 3)), np.zeros((self.n, 3))
        lib().or_md_get(self.h, ptr(x), ptr(v), ptr(f))
        return x, v, f

    def pressure(self):
        """the last step's isotropic pressure (kJ mol^-1 nm^-3)"""
        return lib().or_md_pressure(self.h)

    def box(self):
        L = np.zeros(3)
        lib().or_md_box(self.h, ptr(L))
        return L

    def zeta(self):
        """the Nose-Hoover chain's state"""
        z = np.zeros(max(self.nhc, 1))
        lib().or_md_get_nosehoover(self.h, ptr(z))
        return z[:self.nhc]

    @property
    def pair_count(self):
        return lib().or_md_pair_count(self.h)

    def __del__(self):
        try:
            lib().or_md_destroy(self.h)
        except Exception:
            pass


def remove_com(osys, pos, vel, nticom, nscm, step):
    """RemoveCOMMotion::apply at md step `step` on copies of pos / vel; returns
    the new velocities"""
    x, v = np.array(pos, np.float64), np.array(vel, np.float64)
    m = c64(osys.s.mass)
    lib().or_remove_com(len(m), ptr(m), ptr(x), ptr(v), C.byref(osys.pbc), osys.s.dt, int(nticom), int(nscm),
                        int(step))
    return v.reshape(-1, 3)


def nosehoover_step(zeta, ekin, t_ref, tau, dof, dt):
    """one NoseHooverThermostat::apply on the chain state `zeta` (updated in
    place, float64 array); returns the velocity scale"""
    assert zeta.dtype == np.float64 and zeta.flags.c_contiguous
    return lib().or_nosehoover_step(zeta.ctypes.data_as(C.c_void_p), len(zeta), ekin, t_ref, tau, dof, dt)


def berendsen_scale(mass, vel, t_ref, tau, dof, dt):
    m = c64(mass)
    v = c64(vel)
    return lib().or_berendsen_scale(len(m), ptr(m), ptr(v), t_ref, tau, dof, dt)


def set_threads(n):
    lib().or_set_num_threads(int(n))


def num_threads():
    return lib().or_num_threads()
