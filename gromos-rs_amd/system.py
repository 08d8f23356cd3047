"""Host-side description of one MD system: the topology essentials the nonbonded
path reads, plus coordinates, box and run parameters.

The fields mirror what gromos-rs keeps on `Topology` for this path
(crates/gromos-core/src/topology.rs:619-712): flat per-atom arrays (charge,
iac, mass), contiguous charge groups (`chargegroups`, `num_solute_chargegroups`),
sorted symmetric exclusion lists (`exclusions`), the separate 1-4 lists
(`one_four_pairs`), the LJ matrix (`lj_parameters`, stored flat as
[ti * n_types + tj] x (c6, c12, cs6, cs12) like `LJParamMatrix`,
crates/gromos-forces/src/nonbonded/params.rs:26-58), the solvent template size
and its constraint lengths.  Lists are stored as CSR (offsets + indices) so they
can be copied to the device as two flat arrays.
"""
from dataclasses import dataclass, field
from typing import Optional

import numpy as np

FOUR_PI_EPS_I = 138.9354          # crates/gromos-core/src/units.rs:47-53
K_BOLTZMANN = 0.00831441          # crates/gromos-core/src/units.rs:42-46


def csr_from_lists(lists, n):
    off = np.zeros(n + 1, dtype=np.int32)
    for i in range(min(n, len(lists))):
        off[i + 1] = len(lists[i])
    off = np.cumsum(off).astype(np.int32)
    idx = np.zeros(int(off[-1]), dtype=np.int32)
    for i in range(min(n, len(lists))):
        if lists[i]:
            idx[off[i]:off[i + 1]] = lists[i]
    return off, idx


@dataclass
class Perturbation:
    """Per-atom dual-topology data (PertAtomInfo, nonbonded/perturbed.rs:271-280);
    alphas already multiplied by the global ALPHLJ / ALPHC
    (crates/gromos-integrators/src/algorithms/forcefield.rs:175-199)."""
    flag: np.ndarray          # u8 [n]
    a_iac: np.ndarray         # i32 [n]
    b_iac: np.ndarray
    a_q: np.ndarray           # f64 [n]
    b_q: np.ndarray
    alpha_lj: np.ndarray
    alpha_crf: np.ndarray
    lam: float = 0.0
    nlam: int = 1
    atom_pairs: np.ndarray = field(default_factory=lambda: np.zeros((0, 4), np.int32))

    @staticmethod
    def empty(n):
        z = np.zeros(n)
        return Perturbation(np.zeros(n, np.uint8), np.zeros(n, np.int32), np.zeros(n, np.int32),
                            z.copy(), z.copy(), z.copy(), z.copy())


@dataclass
class Bonded:
    """Covalent terms of the solute (crates/gromos-forces/src/bonded): index
    arrays with a 0-based type per term, and the type tables as the reference
    stores them (crates/gromos-io/src/topology.rs:373-710)."""
    bonds: np.ndarray            # i32 [nb, 3]  i, j, type
    bond_types: np.ndarray       # f64 [tb, 3]  k_quartic (CB), k_harmonic (CHB), r0
    angles: np.ndarray           # i32 [na, 4]  i, j, k, type
    angle_types: np.ndarray      # f64 [ta, 3]  k_cosine (CT), k_harmonic (CHT), theta0 (rad)
    impropers: np.ndarray        # i32 [ni, 5]  i, j, k, l, type
    improper_types: np.ndarray   # f64 [ti, 2]  k (CQ), q0 (rad)
    dihedrals: np.ndarray        # i32 [nd, 5]  i, j, k, l, type
    dihedral_types: np.ndarray   # f64 [td, 3]  k (CP), cos(PD), multiplicity
    # FORCE block NTF(1..4): bonds, angles, impropers, dihedrals switched on (md.rs:331-335)
    ntf: tuple = (1, 1, 1, 1)

    @staticmethod
    def of(bonds, bond_types, angles, angle_types, impropers, improper_types, dihedrals, dihedral_types,
           ntf=(1, 1, 1, 1)):
        def ia(a, w):
            return np.asarray(a, np.int32).reshape(-1, w)

        def fa(a, w):
            return np.asarray(a, float).reshape(-1, w)

        return Bonded(ia(bonds, 3), fa(bond_types, 3), ia(angles, 4), fa(angle_types, 3), ia(impropers, 5),
                      fa(improper_types, 2), ia(dihedrals, 5), fa(dihedral_types, 3),
                      tuple(int(bool(x)) for x in ntf))

    @property
    def n_terms(self):
        return len(self.bonds) + len(self.angles) + len(self.impropers) + len(self.dihedrals)


@dataclass
class System:
    n_atoms: int
    n_solute_atoms: int
    cg_start: np.ndarray        # i32 [n_cg + 1]
    n_solute_cg: int
    charge: np.ndarray          # f64 [n]
    iac: np.ndarray             # i32 [n], 0-based
    mass: np.ndarray            # f64 [n]
    n_types: int
    lj: np.ndarray              # f64 [n_types * n_types * 4]
    excl_off: np.ndarray        # i32 [n + 1]
    excl: np.ndarray            # i32
    o14_off: np.ndarray         # i32 [n + 1]
    o14: np.ndarray             # i32
    atoms_per_solvent: int
    pos: np.ndarray             # f64 [n, 3]
    vel: np.ndarray             # f64 [n, 3]
    box: np.ndarray             # f64 [3]; zeros for vacuum; the diagonal for triclinic
    boundary: int = 1           # 0 vacuum, 1 rectangular, 2 triclinic
    # run parameters (IMD PAIRLIST / NONBONDED / STEP blocks, crates/gromos-io/src/imd.rs:53-176)
    rcutp: float = 1.4
    rcutl: float = 1.4
    skin: float = 0.0
    nsnb: int = 5
    algorithm: int = 0
    epsrf: float = 61.0
    kappa: float = 0.0
    fpepsi: float = FOUR_PI_EPS_I
    dt: float = 0.002
    dist_oh: float = 0.1
    dist_hh: float = 0.1632990
    pert: Optional[Perturbation] = None
    thermostat: Optional[tuple] = None   # one bath: (T_ref K, tau ps, degrees of freedom)
    thermostat_nhc: int = 0              # 0: Berendsen; N >= 1: Nose-Hoover chain of length N
    # centre-of-mass motion removal (remove_com_motion.rs): NTICOM 1 / 2 removes the
    # translation / translation and rotation at step 0; NSCM > 0 the translation
    # every NSCM steps, NSCM < 0 translation and (vacuum only) rotation every |NSCM|
    nticom: int = 0
    nscm: int = 0
    # pressure coupling (berendsen_barostat.rs, pressure_calculation.rs), isotropic,
    # rectangular or triclinic boxes: (P0 kJ mol^-1 nm^-3, compressibility, tau_p ps, virial
    # 0 none / 1 atomic / 2 molecular)
    barostat: Optional[tuple] = None
    # SHAKE (constraints.rs:213-570): solvent_shake constrains every solvent
    # molecule by the template solvent_constraints ((i, j, length) within the
    # molecule) instead of SETTLE; solute_constraints (i, j, length) are the
    # solute's (NTC 2 / 3); tolerance NTCP0 / NTCS0
    solvent_shake: bool = False
    solvent_constraints: Optional[np.ndarray] = None
    solute_constraints: Optional[np.ndarray] = None
    shake_tolerance: float = 1e-4
    # LINCS (constraints.rs:1168-1445): an expansion order above 0 constrains
    # the solute_constraints (solute_lincs) or every solvent molecule by the
    # template solvent_constraints (solvent_lincs, instead of SHAKE or SETTLE)
    # by LINCS
    solute_lincs: int = 0
    solvent_lincs: int = 0
    # last atom of each solute pressure group (topology PRESSUREGROUPS); every
    # solvent molecule is a group of its own
    pressure_groups: Optional[np.ndarray] = None
    name: str = "system"
    # triclinic box (boundary 2): f64 [3, 3], columns a, b, c in the lower-triangular
    # GROMOS convention a = (ax, 0, 0), b = (bx, by, 0), c = (cx, cy, cz)
    # (crates/gromos-core/src/math.rs:75-92)
    box_matrix: Optional[np.ndarray] = None
    bonded: Optional[Bonded] = None      # covalent terms of the solute (none: nonbonded only)

    def box_values(self):
        """the box as the C ABI takes it: 3 lengths, or a, b, c (9 values) for triclinic"""
        if self.boundary == 2:
            return np.ascontiguousarray(np.asarray(self.box_matrix, float).T.reshape(-1))
        return np.ascontiguousarray(self.box, dtype=float)

    @property
    def n_cg(self):
        return len(self.cg_start) - 1

    @property
    def nsm(self):
        if self.atoms_per_solvent == 0:
            return 0
        return (self.n_atoms - self.n_solute_atoms) // self.atoms_per_solvent

    def copy(self):
        import copy
        return copy.deepcopy(self)

    def pressure_group_last(self):
        """the last atom of every pressure group: the solute's groups, then each
        solvent molecule (core topology.rs:1242-1263)"""
        sol = [] if self.pressure_groups is None else [int(x) for x in self.pressure_groups]
        aps = self.atoms_per_solvent
        solv = list(range(self.n_solute_atoms + aps - 1, self.n_atoms, aps)) if aps > 0 else []
        return np.array(sol + solv, np.int32)

    def validate(self):
        n = self.n_atoms
        assert self.pos.shape == (n, 3) and self.vel.shape == (n, 3)
        assert self.charge.shape == (n,) and self.iac.shape == (n,) and self.mass.shape == (n,)
        assert self.cg_start[0] == 0 and self.cg_start[-1] == n
        assert np.all(np.diff(self.cg_start) > 0)
        assert self.lj.shape == (self.n_types * self.n_types * 4,)
        assert self.iac.min(initial=0) >= 0 and self.iac.max(initial=0) < self.n_types
        assert len(self.excl_off) == n + 1 and len(self.o14_off) == n + 1
        if self.boundary == 1:
            assert np.all(self.box > 0)
        if self.boundary == 2:
            m = np.asarray(self.box_matrix, float)
            assert m.shape == (3, 3) and m[1, 0] == 0 and m[2, 0] == 0 and m[2, 1] == 0
            assert m[0, 0] > 0 and m[1, 1] > 0 and m[2, 2] > 0
            assert np.allclose(self.box, np.diag(m))
        if self.pert is not None:
            p = self.pert
            assert p.a_iac.max(initial=0) < self.n_types and p.b_iac.max(initial=0) < self.n_types
        if self.bonded is not None:
            b = self.bonded
            for idx, w in ((b.bonds, 2), (b.angles, 3), (b.impropers, 4), (b.dihedrals, 4)):
                if len(idx):
                    assert idx[:, :w].min() >= 0 and idx[:, :w].max() < self.n_solute_atoms
                    assert idx[:, w].min() >= 0
        return self


def lj_table_from_types(c6, c12, cs6=None, cs12=None):
    """Flat (c6, c12, cs6, cs12) table from full per-pair matrices."""
    c6 = np.asarray(c6, float)
    c12 = np.asarray(c12, float)
    cs6 = c6 if cs6 is None else np.asarray(cs6, float)
    cs12 = c12 if cs12 is None else np.asarray(cs12, float)
    nt = c6.shape[0]
    t = np.zeros((nt, nt, 4))
    t[..., 0], t[..., 1], t[..., 2], t[..., 3] = c6, c12, cs6, cs12
    return t.reshape(-1)
