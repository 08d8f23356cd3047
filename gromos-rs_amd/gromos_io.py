"""GROMOS file formats: topology, configuration, input parameters and
perturbation topology in; energy / force trajectories and configurations out.

Readers follow the block semantics of gromos-rs's readers (crates/gromos-io):

  * topology (topology.rs:89-236, build_topology :1035-1300): SOLUTEATOM with its
    exclusions on the atom line and the 1-4 partners on the next line (:263-370),
    0-based type indices, charge groups closed at CGC = 1 (:1093-1110), bonds and
    angles adding 1-2 / 1-3 exclusions while the 1-4 lists stay separate and are
    not mirrored (:1141-1183), sorted de-duplicated exclusion lists (:1294-1297),
    LJPARAMETERS / CGPARAMETERS as (c12, c6, cs12, cs6) per type pair (:711-780),
    bonded type tables (BONDSTRETCHTYPE CB CHB B0, BONDANGLEBENDTYPE CT CHT T0,
    IMPDIHEDRALTYPE CQ Q0, TORSDIHEDRALTYPE CP PD NP) and their index blocks;
    every solvent molecule is its own charge group and excludes all of its own
    atom pairs (crates/gromos-core/src/topology.rs:1224-1263);
  * configuration (coordinate.rs:49-200): POSITION / POSITIONRED, VELOCITY /
    VELOCITYRED, GENBOX or BOX;
  * input (imd.rs:53-176): SYSTEM, STEP, BOUNDCOND, PAIRLIST, NONBONDED,
    CONSTRAINT, MULTIBATH, PERTURBATION, WRITETRAJ, COMTRANSROT;
  * perturbation topology (ptp.rs): PERTATOMPARAM, PERTATOMPAIR.

NTB = -1 (truncated octahedron) converts the cube edge of the box block into
the lower-triangular triclinic box and rotates positions and velocities into
that frame (crates/gromos-md/src/bin/md.rs:530-545, math.rs
truncoct_triclinic_box / truncoct_triclinic).

Writers produce the layouts gromos-rs writes: the energy trajectory's ENERTRJ
block (energy.rs:190-275, 17 values per frame), the force trajectory's TIMESTEP
+ FREEFORCERED blocks (force.rs, 9 decimals in 18 columns) and a configuration
with POSITION, VELOCITY and GENBOX.

`run_md` drives a whole md run from these files on the device and writes the
trajectories, the end-to-end path of the gromos-rs `md` binary for the
nonbonded + SETTLE systems this library covers.
"""
import math
import os
from dataclasses import dataclass, field
from typing import Dict, List, Optional, Tuple

import numpy as np

from .system import FOUR_PI_EPS_I, K_BOLTZMANN, Bonded, Perturbation, System, csr_from_lists


class GromosFormatError(ValueError):
    pass


# ---------------------------------------------------------------------------
# blocks
# ---------------------------------------------------------------------------
def read_blocks(path):
    """A GROMOS block file as an ordered list of (NAME, lines between NAME and END)."""
    out = []
    name, buf = None, []
    with open(path) as fh:
        for raw in fh:
            line = raw.rstrip("\n")
            s = line.strip()
            if name is None:
                if s and not s.startswith("#"):
                    name, buf = s.split()[0], []
                continue
            if s == "END":
                out.append((name, buf))
                name = None
                continue
            buf.append(line)
    if name is not None:
        raise GromosFormatError(f"{path}: block {name} has no END")
    return out


def _first(blocks, *names):
    for n, lines in blocks:
        if n in names:
            return n, lines
    return None, None


def _data(lines):
    """non-empty lines without comments (a '#' starts a comment anywhere)"""
    out = []
    for l in lines or []:
        s = l.split("#", 1)[0].strip()
        if s:
            out.append(s)
    return out


def _tokens(lines):
    return [t for l in _data(lines) for t in l.split()]


# ---------------------------------------------------------------------------
# topology
# ---------------------------------------------------------------------------
@dataclass
class GromosTopology:
    fpepsi: float = FOUR_PI_EPS_I
    hbar: float = 0.0635078
    boltz: float = K_BOLTZMANN
    atom_type_names: List[str] = field(default_factory=list)
    # solute atoms (0-based types and indices)
    atom_names: List[str] = field(default_factory=list)
    residue: List[int] = field(default_factory=list)
    iac: List[int] = field(default_factory=list)
    mass: List[float] = field(default_factory=list)
    charge: List[float] = field(default_factory=list)
    cgc: List[int] = field(default_factory=list)
    excl: List[List[int]] = field(default_factory=list)
    o14: List[List[int]] = field(default_factory=list)
    # bonded terms: type tables and (atoms..., type) entries, H blocks first as read
    bond_types: List[Tuple[float, float, float]] = field(default_factory=list)      # CB, CHB, B0
    bonds: List[Tuple[int, int, int]] = field(default_factory=list)
    angle_types: List[Tuple[float, float, float]] = field(default_factory=list)     # CT, CHT, T0 (deg)
    angles: List[Tuple[int, int, int, int]] = field(default_factory=list)
    improper_types: List[Tuple[float, float]] = field(default_factory=list)         # CQ, Q0 (deg)
    impropers: List[Tuple[int, int, int, int, int]] = field(default_factory=list)
    dihedral_types: List[Tuple[float, float, int]] = field(default_factory=list)    # CP, PD, NP
    dihedrals: List[Tuple[int, int, int, int, int]] = field(default_factory=list)
    lj: Dict[Tuple[int, int], Tuple[float, float, float, float]] = field(default_factory=dict)  # c6, c12, cs6, cs12
    # solvent template
    solvent: List[Tuple[str, int, float, float]] = field(default_factory=list)      # name, iac, mass, charge
    solvent_constr: List[Tuple[int, int, float]] = field(default_factory=list)
    temperature_groups: List[int] = field(default_factory=list)
    pressure_groups: List[int] = field(default_factory=list)

    @property
    def n_solute(self):
        return len(self.mass)


def _int_index_block(lines, n_atoms_per_entry):
    """a bonded index block: count, then entries of n atoms (1-based) and a type (1-based)"""
    toks = _tokens(lines)
    if not toks:
        return []
    n = int(toks[0])
    w = n_atoms_per_entry + 1
    out = []
    for k in range(n):
        e = toks[1 + k * w: 1 + (k + 1) * w]
        if len(e) < w:
            raise GromosFormatError("truncated bonded block")
        out.append(tuple(int(x) - 1 for x in e))
    return out


def read_topology(path):
    blocks = read_blocks(path)
    t = GromosTopology()
    _, pc = _first(blocks, "PHYSICALCONSTANTS")
    if pc:
        v = [float(x) for x in _tokens(pc)]
        t.fpepsi = v[0]
        if len(v) > 1:
            t.hbar = v[1]
        if len(v) > 3:
            t.boltz = v[3]
    _, an = _first(blocks, "ATOMTYPENAME")
    if an:
        d = _data(an)
        t.atom_type_names = [l.split()[0] for l in d[1:1 + int(d[0].split()[0])]]
    _, sa = _first(blocks, "SOLUTEATOM")
    if sa is None:
        raise GromosFormatError(f"{path}: no SOLUTEATOM block")
    d = _data(sa)
    n = int(d[0].split()[0])
    k = 1
    while len(t.mass) < n:
        # ATNM MRES PANM IAC MASS CG CGC INE [entries]: the exclusions are the
        # entries on this line (topology.rs:327-337), the next line is INE14
        # [entries] (topology.rs:339-362)
        p = d[k].split()
        k += 1
        ne = int(p[7])
        ex = [int(x) - 1 for x in p[8:8 + ne]]
        p14 = d[k].split()
        k += 1
        n14 = int(p14[0])
        o14 = [int(x) - 1 for x in p14[1:1 + n14]]
        t.residue.append(int(p[1]) - 1)
        t.atom_names.append(p[2])
        t.iac.append(int(p[3]) - 1)
        t.mass.append(float(p[4]))
        t.charge.append(float(p[5]))
        t.cgc.append(int(p[6]))
        t.excl.append(ex[:ne])
        t.o14.append(o14[:n14])
    for name, lines in blocks:
        if name == "BONDSTRETCHTYPE":
            d = _data(lines)
            t.bond_types = [tuple(float(x) for x in l.split()[:3]) for l in d[1:1 + int(d[0].split()[0])]]
        elif name in ("BONDH", "BOND"):
            t.bonds += _int_index_block(lines, 2)
        elif name == "BONDANGLEBENDTYPE":
            d = _data(lines)
            t.angle_types = [tuple(float(x) for x in l.split()[:3]) for l in d[1:1 + int(d[0].split()[0])]]
        elif name in ("BONDANGLEH", "BONDANGLE"):
            t.angles += _int_index_block(lines, 3)
        elif name == "IMPDIHEDRALTYPE":
            d = _data(lines)
            t.improper_types = [tuple(float(x) for x in l.split()[:2]) for l in d[1:1 + int(d[0].split()[0])]]
        elif name in ("IMPDIHEDRALH", "IMPDIHEDRAL"):
            t.impropers += _int_index_block(lines, 4)
        elif name == "TORSDIHEDRALTYPE":
            d = _data(lines)
            t.dihedral_types = [(float(l.split()[0]), float(l.split()[1]), int(l.split()[2]))
                                for l in d[1:1 + int(d[0].split()[0])]]
        elif name in ("DIHEDRALH", "DIHEDRAL"):
            t.dihedrals += _int_index_block(lines, 4)
        elif name in ("LJPARAMETERS", "CGPARAMETERS") and not t.lj:
            d = _data(lines)
            for l in d[1:]:
                p = l.split()
                i, j = int(p[0]) - 1, int(p[1]) - 1
                c12, c6 = float(p[2]), float(p[3])
                cs12, cs6 = (float(p[4]), float(p[5])) if len(p) >= 6 else (c12, c6)
                t.lj[(i, j)] = t.lj[(j, i)] = (c6, c12, cs6, cs12)
        elif name == "SOLVENTATOM":
            d = _data(lines)
            for l in d[1:1 + int(d[0].split()[0])]:
                p = l.split()
                t.solvent.append((p[1], int(p[2]) - 1, float(p[3]), float(p[4])))
        elif name == "SOLVENTCONSTR":
            d = _data(lines)
            for l in d[1:1 + int(d[0].split()[0])]:
                p = l.split()
                t.solvent_constr.append((int(p[0]) - 1, int(p[1]) - 1, float(p[2])))
        elif name == "TEMPERATUREGROUPS":
            v = [int(x) for x in _tokens(lines)]
            t.temperature_groups = [x - 1 for x in v[1:1 + v[0]]]
        elif name == "PRESSUREGROUPS":
            v = [int(x) for x in _tokens(lines)]
            t.pressure_groups = [x - 1 for x in v[1:1 + v[0]]]
    return t


# ---------------------------------------------------------------------------
# configuration
# ---------------------------------------------------------------------------
@dataclass
class GromosConfiguration:
    title: str = ""
    pos: Optional[np.ndarray] = None       # (n, 3)
    vel: Optional[np.ndarray] = None
    ntb: int = 0                           # GENBOX box type (BOX: 1)
    box: Tuple[float, float, float] = (0.0, 0.0, 0.0)
    angles: Tuple[float, float, float] = (90.0, 90.0, 90.0)


def _vectors(lines):
    return np.array([[float(x) for x in l.split()[-3:]] for l in _data(lines)], float)


def read_configuration(path):
    blocks = read_blocks(path)
    c = GromosConfiguration()
    _, ti = _first(blocks, "TITLE")
    c.title = "\n".join(l.strip() for l in (ti or []))
    _, p = _first(blocks, "POSITION", "POSITIONRED")
    if p is None:
        raise GromosFormatError(f"{path}: no POSITION / POSITIONRED block")
    c.pos = _vectors(p)
    _, v = _first(blocks, "VELOCITY", "VELOCITYRED")
    if v is not None:
        c.vel = _vectors(v)
    _, g = _first(blocks, "GENBOX")
    if g is not None:
        d = _data(g)
        c.ntb = int(d[0].split()[0])
        c.box = tuple(float(x) for x in d[1].split()[:3])
        if len(d) > 2:
            c.angles = tuple(float(x) for x in d[2].split()[:3])
    else:
        _, b = _first(blocks, "BOX")
        if b is not None:
            c.box = tuple(float(x) for x in _data(b)[0].split()[:3])
            c.ntb = 1
    return c


# ---------------------------------------------------------------------------
# input parameters
# ---------------------------------------------------------------------------
@dataclass
class GromosInput:
    npm: int = 1
    nsm: int = 0
    nstlim: int = 0
    t0: float = 0.0
    dt: float = 0.002
    ntb: int = 0
    ndfmin: int = 0
    algorithm: int = 0                     # 0 standard, 1 grid, 2 grid_cell
    nsnb: int = 5
    rcutp: float = 0.8
    rcutl: float = 1.4
    size: float = 0.4
    appak: float = 0.0
    rcrf: float = 1.4
    epsrf: float = 61.0
    ntc: int = 1
    ntcs: str = "shake"                    # solvent constraints algorithm
    ntcp: str = "shake"                    # solute constraints algorithm
    shake_tol: float = 1e-4                # NTCP0 when NTCP is shake
    lincs_order_solute: int = 4            # NTCP0 when NTCP is lincs
    lincs_order_solvent: int = 4           # NTCS0 when NTCS is lincs
    ntf: Tuple[int, int, int, int] = (1, 1, 1, 1)   # FORCE: bonds, angles, impropers, dihedrals on
    baths: List[Tuple[float, float]] = field(default_factory=list)   # (TEMP0, TAU)
    coupling: str = ""                     # weak-coupling, nose-hoover, nose-hoover-chains
    nhc: int = 1                           # nose-hoover-chains: the chain length
    ntg: int = 0                           # perturbation on
    rlam: float = 0.0
    nlam: int = 1
    alphlj: float = 1.0
    alphc: float = 1.0
    ntwx: int = 0
    ntwe: int = 0
    ntwv: int = 0
    ntwf: int = 0
    nscm: int = 0
    nticom: int = 0                        # INITIALISE: centre-of-mass motion removal at step 0
    ntishk: int = 0                        # INITIALISE: SHAKE the initial positions (1) and velocities (2)
    couple: str = "off"                    # PRESSURESCALE COUPLE: off, calc, scale
    scale: str = "iso"
    compressibility: float = 4.575e-4
    tau_p: float = 0.5
    virial: int = 0                        # 0 none, 1 atomic, 2 molecular
    p_ref: float = 0.06102                 # PRESSURESCALE reference pressure [0][0]


_ALG = {"standard": 0, "0": 0, "grid": 1, "1": 1, "grid_cell": 2, "2": 2}


def read_imd(path):
    blocks = dict((n, l) for n, l in reversed(read_blocks(path)))   # first occurrence wins
    m = GromosInput()
    if "SYSTEM" in blocks:
        v = _tokens(blocks["SYSTEM"])
        m.npm, m.nsm = int(v[0]), int(v[1])
    if "STEP" in blocks:
        v = _tokens(blocks["STEP"])
        m.nstlim, m.t0, m.dt = int(v[0]), float(v[1]), float(v[2])
    if "BOUNDCOND" in blocks:
        v = _tokens(blocks["BOUNDCOND"])
        m.ntb, m.ndfmin = int(v[0]), int(v[1]) if len(v) > 1 else 0
    if "PAIRLIST" in blocks:
        v = _tokens(blocks["PAIRLIST"])
        m.algorithm = _ALG.get(v[0].lower(), 0)
        m.nsnb, m.rcutp, m.rcutl = int(v[1]), float(v[2]), float(v[3])
        if len(v) > 4:
            try:
                m.size = float(v[4])
            except ValueError:
                pass
    if "NONBONDED" in blocks:
        d = _data(blocks["NONBONDED"])
        p = d[1].split()
        m.appak, m.rcrf, m.epsrf = float(p[0]), float(p[1]), float(p[2])
    if "FORCE" in blocks:
        # md.rs:331-335: NTF(1..4) switch bonds, angles, impropers, dihedrals
        m.ntf = tuple(int(x) for x in _data(blocks["FORCE"])[0].split()[:4])
    if "CONSTRAINT" in blocks:
        d = _data(blocks["CONSTRAINT"])
        m.ntc = int(d[0].split()[0])
        # imd.rs:586-639: NTC; NTCP (a name or 1 shake, 2 lincs, 3 settle); NTCP0, the
        # one SHAKE tolerance, or the solute's LINCS order when NTCP is lincs; NTCS
        # (as NTCP, 0 none); NTCS0, read only as the solvent's LINCS order
        names = {"1": "shake", "2": "lincs", "3": "settle", "0": "none"}
        if len(d) > 1:
            w = d[1].split()[0].lower()
            m.ntcp = names.get(w, w)
        if len(d) > 2 and m.ntcp != "lincs":
            m.shake_tol = float(d[2].split()[0])
        elif len(d) > 2:
            m.lincs_order_solute = int(float(d[2].split()[0]))
        if len(d) > 3:
            w = d[3].split()[0].lower()
            m.ntcs = names.get(w, w)
        if len(d) > 4 and m.ntcs == "lincs":
            m.lincs_order_solvent = int(float(d[4].split()[0]))
    if "MULTIBATH" in blocks:
        d = _data(blocks["MULTIBATH"])
        k = 0
        if not d[0].split()[0].lstrip("-").isdigit():
            w = d[0].split()
            m.coupling = w[0].lower()
            if m.coupling == "nose-hoover-chains" and len(w) > 1:
                m.nhc = int(w[1])
            k = 1
        nb = int(d[k].split()[0])
        vals = " ".join(d[k + 1:]).split()
        m.baths = [(float(vals[2 * b]), float(vals[2 * b + 1])) for b in range(nb)]
    if "PERTURBATION" in blocks:
        d = _data(blocks["PERTURBATION"])
        a, c = d[0].split(), d[1].split()
        m.ntg, m.rlam = int(a[0]), float(a[2])
        m.alphlj, m.alphc, m.nlam = float(c[0]), float(c[1]), int(c[2])
    if "WRITETRAJ" in blocks:
        v = [int(x) for x in _tokens(blocks["WRITETRAJ"])]
        m.ntwx = v[0]
        if len(v) >= 7:
            m.ntwv, m.ntwf, m.ntwe = v[2], v[3], v[4]
    if "PRESSURESCALE" in blocks:
        d = _data(blocks["PRESSURESCALE"])
        w = d[0].split()
        m.couple = {"0": "off", "1": "calc", "2": "scale"}.get(w[0], w[0].lower())
        m.scale = {"1": "iso", "2": "aniso", "3": "full-aniso", "4": "semi-aniso"}.get(w[1], w[1].lower())
        m.compressibility, m.tau_p = float(w[2]), float(w[3])
        m.virial = {"none": 0, "atomic": 1, "molecular": 2}.get(w[4].lower(), None)
        if m.virial is None:
            m.virial = int(w[4])
        m.p_ref = float(" ".join(d[2:5]).split()[0])
    if "COMTRANSROT" in blocks:
        m.nscm = int(_tokens(blocks["COMTRANSROT"])[0])
    if "INITIALISE" in blocks:
        v = _tokens(blocks["INITIALISE"])
        if len(v) > 1:
            m.ntishk = int(v[1])
        if len(v) > 6:
            m.nticom = int(v[6])     # NTIVEL NTISHK NTINHT NTINHB NTISHI NTIRTC NTICOM
    return m


# ---------------------------------------------------------------------------
# perturbation topology
# ---------------------------------------------------------------------------
def read_ptp(path):
    """PERTATOMPARAM (seq, A/B type and charge, soft-core alphas) and PERTATOMPAIR
    (i, j, A and B interaction types; B = 0 absent) -> (atoms, atom_pairs), 0-based."""
    blocks = dict((n, l) for n, l in reversed(read_blocks(path)))
    d = _data(blocks.get("PERTATOMPARAM"))
    if not d:
        raise GromosFormatError(f"{path}: no PERTATOMPARAM block")
    atoms = []
    for l in d[1:1 + int(d[0].split()[0])]:
        p = l.split()
        atoms.append(dict(seq=int(p[0]) - 1, a_iac=int(p[3]) - 1, a_q=float(p[5]), b_iac=int(p[6]) - 1,
                          b_q=float(p[8]), alj=float(p[9]), acrf=float(p[10])))
    pairs = []
    d = _data(blocks.get("PERTATOMPAIR"))
    if d:
        for l in d[1:1 + int(d[0].split()[0])]:
            p = l.split()
            bt = int(p[3]) - 1
            pairs.append((int(p[0]) - 1, int(p[1]) - 1, int(p[2]) - 1, bt if bt >= 0 else -1))
    return atoms, pairs


# ---------------------------------------------------------------------------
# truncated octahedron
# ---------------------------------------------------------------------------
def truncoct_triclinic(L, pos, vel=None):
    """NTB = -1: the cube edge L becomes the lower-triangular truncated-octahedron
    box (columns a, b, c) and positions / velocities are rotated into its frame
    by the transpose of the rotation with columns (1, 1, 1)/sqrt3,
    (-2, 1, 1)/sqrt6, (0, -1, 1)/sqrt2, the columns of the applied matrix summed
    x, y, z in that order (glam's Mat3 * Vec3)."""
    d = 0.5 * math.sqrt(3.0) * L
    M = np.array([[d, d / 3.0, -d / 3.0],
                  [0.0, 2.0 * math.sqrt(2.0) / 3.0 * d, math.sqrt(2.0) / 3.0 * d],
                  [0.0, 0.0, math.sqrt(6.0) / 3.0 * d]])
    sq3i, sq2i = 1.0 / math.sqrt(3.0), 1.0 / math.sqrt(2.0)
    mx = (sq3i, -2.0 * sq2i * sq3i, 0.0)
    my = (sq3i, sq3i * sq2i, -sq2i)
    mz = (sq3i, sq2i * sq3i, sq2i)

    def rot(v):
        v = np.asarray(v, float)
        out = np.empty_like(v)
        for k in range(3):
            out[:, k] = mx[k] * v[:, 0] + my[k] * v[:, 1] + mz[k] * v[:, 2]
        return out

    return M, rot(pos), (rot(vel) if vel is not None else None)


# ---------------------------------------------------------------------------
# System from files
# ---------------------------------------------------------------------------
def build_system(topo, conf, imd, ptp=None, name="gromos"):
    """The device path's System from a read topology, configuration and input
    (and optional perturbation topology), the way gromos-rs md builds its
    Topology and Forcefield from them."""
    ns = topo.n_solute
    aps = len(topo.solvent)
    nsm = imd.nsm
    n = ns + nsm * aps
    if conf.pos.shape != (n, 3):
        raise GromosFormatError(f"configuration holds {len(conf.pos)} positions, topology and SYSTEM need {n}")
    # charge groups: solute groups close at CGC = 1; each solvent molecule is one
    cg = [0]
    for i in range(ns):
        if topo.cgc[i] == 1:
            cg.append(i + 1)
    if cg[-1] != ns:
        cg.append(ns)
    n_solute_cg = len(cg) - 1
    cg += [ns + (m + 1) * aps for m in range(nsm)]
    # exclusions: atom lines + bonds (1-2) + angles (1-3), symmetric, sorted; solvent all-pairs
    ex = [set() for _ in range(n)]
    for i, lst in enumerate(topo.excl):
        for j in lst:
            ex[i].add(j)
            ex[j].add(i)
    for b in topo.bonds:
        ex[b[0]].add(b[1])
        ex[b[1]].add(b[0])
    for a in topo.angles:
        ex[a[0]].add(a[2])
        ex[a[2]].add(a[0])
    for m in range(nsm):
        base = ns + m * aps
        for a in range(aps):
            for b in range(a + 1, aps):
                ex[base + a].add(base + b)
                ex[base + b].add(base + a)
    excl_off, excl = csr_from_lists([sorted(s) for s in ex], n)
    o14_off, o14 = csr_from_lists([list(l) for l in topo.o14], n)
    iac = np.array(list(topo.iac) + [s[1] for _ in range(nsm) for s in topo.solvent], np.int32)
    mass = np.array(list(topo.mass) + [s[2] for _ in range(nsm) for s in topo.solvent], float)
    charge = np.array(list(topo.charge) + [s[3] for _ in range(nsm) for s in topo.solvent], float)
    pert_atoms, pert_pairs = ptp if ptp is not None else ([], [])
    # the LJ table over the types in use (0-based, compacted)
    used = sorted(set(iac.tolist()) | {a["a_iac"] for a in pert_atoms} | {a["b_iac"] for a in pert_atoms})
    remap = {old: k for k, old in enumerate(used)}
    nt = len(used)
    lj = np.zeros((nt, nt, 4))
    for a, ta in enumerate(used):
        for b, tb in enumerate(used):
            lj[a, b] = topo.lj.get((ta, tb), (0.0, 0.0, 0.0, 0.0))
    iac = np.array([remap[x] for x in iac], np.int32)
    pos = np.asarray(conf.pos, float)
    vel = np.asarray(conf.vel, float) if conf.vel is not None else np.zeros_like(pos)
    ntb = imd.ntb
    box = np.array(conf.box, float)
    boundary, box_matrix = 0, None
    if ntb == 1:
        boundary = 1
    elif ntb == -1:
        box_matrix, pos, vel = truncoct_triclinic(box[0], pos, vel)
        boundary, box = 2, np.diag(box_matrix).copy()
    elif ntb != 0:
        raise GromosFormatError(f"NTB = {ntb} is not supported (0 vacuum, 1 rectangular, -1 truncated octahedron)")
    if boundary == 0:
        box = np.zeros(3)
    d_oh, d_hh = 0.1, 0.163299
    for i, j, l in topo.solvent_constr:
        if {i, j} == {1, 2}:
            d_hh = l
        else:
            d_oh = l
    pert = None
    if pert_atoms:
        pert = Perturbation.empty(n)
        for a in pert_atoms:
            k = a["seq"]
            pert.flag[k] = 1
            pert.a_iac[k], pert.b_iac[k] = remap[a["a_iac"]], remap[a["b_iac"]]
            pert.a_q[k], pert.b_q[k] = a["a_q"], a["b_q"]
            pert.alpha_lj[k] = a["alj"] * imd.alphlj
            pert.alpha_crf[k] = a["acrf"] * imd.alphc
        pert.lam, pert.nlam = imd.rlam, imd.nlam
        pert.atom_pairs = np.array(pert_pairs, np.int32).reshape(-1, 4)
    bonded = None
    top = topo
    if top.bonds or top.angles or top.impropers or top.dihedrals:
        # type tables as the reference stores them: theta0 and Q0 in radians
        # (topology.rs:505-513, :699-703), cos(PD) (:661-664)
        bonded = Bonded.of(top.bonds, [list(t) for t in top.bond_types], top.angles,
                           [[ct, cht, t0 * (math.pi / 180.0)] for ct, cht, t0 in top.angle_types],
                           top.impropers, [[cq * 180.0 * 180.0 / (math.pi * math.pi), q0 * math.pi / 180.0]
                                            for cq, q0 in top.improper_types],
                           top.dihedrals, [[cp, math.cos(pd * math.pi / 180.0), npd]
                                           for cp, pd, npd in top.dihedral_types], imd.ntf)
    s = System(n_atoms=n, n_solute_atoms=ns, cg_start=np.array(cg, np.int32), n_solute_cg=n_solute_cg,
               charge=charge, iac=iac, mass=mass, n_types=nt, lj=lj.reshape(-1), excl_off=excl_off, excl=excl,
               o14_off=o14_off, o14=o14, atoms_per_solvent=aps, pos=pos, vel=vel, box=box, boundary=boundary,
               rcutp=imd.rcutp, rcutl=imd.rcutl, skin=0.0, nsnb=imd.nsnb, algorithm=imd.algorithm,
               epsrf=imd.epsrf, kappa=imd.appak, fpepsi=topo.fpepsi, dt=imd.dt, dist_oh=d_oh, dist_hh=d_hh,
               pert=pert, name=name, box_matrix=box_matrix, bonded=bonded)
    if topo.pressure_groups:
        s.pressure_groups = np.array(topo.pressure_groups, np.int32)
    return s.validate()


def load(topology, configuration, imd, ptp=None, name=None):
    """read the files and build the System"""
    top = read_topology(topology)
    conf = read_configuration(configuration)
    inp = read_imd(imd)
    pt = read_ptp(ptp) if ptp else None
    return build_system(top, conf, inp, pt, name=name or os.path.splitext(os.path.basename(imd))[0]), top, inp


# ---------------------------------------------------------------------------
# writers
# ---------------------------------------------------------------------------
ENERTRJ_FIELDS = ("time", "kinetic", "potential", "total", "temperature", "volume", "pressure", "bond", "angle",
                  "improper", "dihedral", "lj", "coul_real", "coul_recip", "coul_self", "shake", "restraint")


class EnergyTrajectoryWriter:
    """ENERTRJ energy trajectory (.tre) as gromos-rs writes it: a header naming
    the 17 values, then one line per frame in 20.12e."""

    def __init__(self, path, title="Energy trajectory"):
        self.fh = open(path, "w")
        self.fh.write(f"TITLE\n  {title}\nEND\nENERTRJ\n")
        self.fh.write("# Block definitions (ENE format):\n")
        for k, f in enumerate(ENERTRJ_FIELDS):
            self.fh.write(f"# {k:>3} = {f}\n")
        self.frames = 0

    def write(self, **values):
        row = [float(values.get(f, 0.0)) for f in ENERTRJ_FIELDS]
        self.fh.write(" ".join(f"{v:20.12e}" for v in row) + "\n")
        self.frames += 1

    def close(self):
        if self.fh:
            self.fh.write("END\n")
            self.fh.close()
            self.fh = None


def read_energy_trajectory(path):
    """an ENERTRJ file -> (frames, 17) array (columns ENERTRJ_FIELDS)"""
    rows = []
    for name, lines in read_blocks(path):
        if name == "ENERTRJ":
            rows += [[float(x) for x in l.split()] for l in _data(lines)]
    return np.array(rows, float).reshape(-1, len(ENERTRJ_FIELDS))


class ForceTrajectoryWriter:
    """Force trajectory (.trf): TIMESTEP (step, time) and FREEFORCERED per frame,
    9 decimals in 18 columns, a comment line every 10 atoms."""

    def __init__(self, path, title="Force trajectory"):
        self.fh = open(path, "w")
        self.fh.write(f"TITLE\n  {title}\nEND\n")

    def write(self, step, time, forces):
        fh = self.fh
        fh.write(f"TIMESTEP\n{step:15d}{time:15.6f}\nEND\nFREEFORCERED\n")
        for i, f in enumerate(np.asarray(forces, float)):
            fh.write(f"{f[0]:18.9f}{f[1]:18.9f}{f[2]:18.9f}\n")
            if (i + 1) % 10 == 0:
                fh.write(f"#{i + 1:10d}\n")
        fh.write("END\n")

    def close(self):
        if self.fh:
            self.fh.close()
            self.fh = None


def read_force_trajectory(path):
    """a .trf file -> list of (step, (n, 3) forces)"""
    out, step = [], None
    for name, lines in read_blocks(path):
        if name == "TIMESTEP":
            step = int(_data(lines)[0].split()[0])
        elif name == "FREEFORCERED":
            out.append((step, np.array([[float(x) for x in l.split()[:3]] for l in _data(lines)], float)))
    return out


def write_configuration(path, system, pos, vel, title="final configuration", ntb=None, box=None):
    """POSITION, VELOCITY (atom numbers and 15.9 columns) and GENBOX (the
    system's box, or `box` edge lengths: a barostat's final box)."""
    with open(path, "w") as fh:
        fh.write(f"TITLE\n  {title}\nEND\nPOSITION\n")
        for i, p in enumerate(np.asarray(pos, float)):
            fh.write(f"{1:5d} {'RES':<5} {'AT':<5}{i + 1:7d}{p[0]:15.9f}{p[1]:15.9f}{p[2]:15.9f}\n")
        fh.write("END\nVELOCITY\n")
        for i, v in enumerate(np.asarray(vel, float)):
            fh.write(f"{1:5d} {'RES':<5} {'AT':<5}{i + 1:7d}{v[0]:15.9f}{v[1]:15.9f}{v[2]:15.9f}\n")
        b = np.asarray(system.box if box is None else box, float)
        kind = system.boundary if ntb is None else ntb
        fh.write(f"END\nGENBOX\n{kind:5d}\n{b[0]:15.9f}{b[1]:15.9f}{b[2]:15.9f}\n"
                 f"{90.0:15.9f}{90.0:15.9f}{90.0:15.9f}\n{0.0:15.9f}{0.0:15.9f}{0.0:15.9f}\n"
                 f"{0.0:15.9f}{0.0:15.9f}{0.0:15.9f}\nEND\n")


# ---------------------------------------------------------------------------
# an md run from GROMOS files on the device
# ---------------------------------------------------------------------------
def apply_constraints(s, top, inp):
    """the CONSTRAINT block on a System (md.rs:337-352, 1208-1236; constraints.rs:36-60):
    NTC 2 constrains the solute bonds with a hydrogen (an atom of mass 2 or less),
    NTC 3 all of them, each of length B0 (bonds with B0 below 1e-10 are skipped), in
    the topology's order, by SHAKE or LINCS (NTCP, order NTCP0); the solvent by
    SHAKE or LINCS (NTCS shake / lincs, order NTCS0: every molecule by the
    SOLVENTCONSTR template) or SETTLE; one SHAKE tolerance, NTCP0."""
    if inp.ntc > 1 and top.bonds:
        if inp.ntcp not in ("shake", "lincs"):
            raise GromosFormatError(f"solute constraints by {inp.ntcp!r} are not provided (SHAKE or LINCS)")
        if inp.ntcp == "lincs":
            s.solute_lincs = inp.lincs_order_solute
        cons = []
        for i, j, t in top.bonds:
            r0 = top.bond_types[t][2]
            if r0 < 1e-10 or (inp.ntc == 2 and s.mass[i] > 2.0 and s.mass[j] > 2.0):
                continue
            cons.append((i, j, r0))
        s.solute_constraints = np.array(cons, float).reshape(-1, 3)
    if s.nsm > 0 and inp.ntcs in ("shake", "lincs"):
        s.solvent_constraints = np.array(top.solvent_constr, float).reshape(-1, 3)
        if inp.ntcs == "shake":
            s.solvent_shake = True
        else:
            s.solvent_lincs = inp.lincs_order_solvent
    elif s.nsm > 0 and inp.ntcs != "settle":
        raise GromosFormatError(f"solvent constraints by {inp.ntcs!r} are not provided (SHAKE, LINCS or SETTLE)")
    s.shake_tolerance = inp.shake_tol
    return s


def run_md(topology, configuration, imd, out_prefix, precision=None, ptp=None, nsteps=None):
    """gromos-rs md for the systems this library covers (nonbonded forces, the
    solute's covalent terms, SHAKE or SETTLE constraints, a Berendsen or Nose-Hoover
    bath, centre-of-mass motion removal, isotropic Berendsen pressure scaling and
    perturbation of the nonbonded terms): reads the files, runs NSTLIM (or
    nsteps) md steps on the device, and writes out_prefix.tre (every step, or
    every NTWE), out_prefix.trf (the step-0 forces) and out_prefix.cnf.
    Returns the System and the per-step energies (e_lj, e_crf, e_pot, e_kin, dhdl)."""
    from . import runtime as rt
    s, top, inp = load(topology, configuration, imd, ptp)
    if inp.ntishk >= 1:
        # md.rs:1227-1235, shake_positions / shake_velocities (constraints.rs:575-650)
        raise GromosFormatError("INITIALISE NTISHK >= 1 (SHAKE of the initial configuration) is not provided")
    apply_constraints(s, top, inp)
    precision = rt.FP64 if precision is None else precision
    n = inp.nstlim if nsteps is None else int(nsteps)
    if inp.baths and inp.baths[0][1] > 0.0:
        # md.rs:1143-1206: weak-coupling -> Berendsen, nose-hoover -> one zeta,
        # nose-hoover-chains N -> a chain of N; one bath over all atoms
        dof = 3 * s.n_atoms - 3 * s.nsm - inp.ndfmin
        s.thermostat = (inp.baths[0][0], inp.baths[0][1], dof)
        if inp.coupling.startswith("nose-hoover"):
            s.thermostat_nhc = inp.nhc if inp.coupling == "nose-hoover-chains" else 1
        elif inp.coupling not in ("", "weak-coupling", "berendsen"):
            raise GromosFormatError(f"MULTIBATH coupling {inp.coupling!r} is not provided")
    s.nticom, s.nscm = inp.nticom, inp.nscm      # md.rs:1095-1097
    if inp.couple == "scale":
        # md.rs:1116-1129, 1259-1277: PressureCalculation + BerendsenBarostat
        if inp.scale != "iso":
            raise GromosFormatError(f"PRESSURESCALE scaling {inp.scale!r} is not provided (isotropic only)")
        s.barostat = (inp.p_ref, inp.compressibility, inp.tau_p, inp.virial)
    elif inp.couple != "off":
        raise GromosFormatError(f"PRESSURESCALE coupling {inp.couple!r} is not provided")
    # step-0 forces (a separate single point: the md run keeps its own rebuild cadence)
    ds = rt.DeviceSystem(s, precision=precision)
    ds.pairlist_update()
    sp = ds.forcefield_apply()
    ds.close()
    fw = ForceTrajectoryWriter(out_prefix + ".trf")
    fw.write(0, inp.t0, sp["forces"])
    fw.close()
    ds = rt.DeviceSystem(s, precision=precision)
    ds.pairlist_update()
    every = inp.ntwe if inp.ntwe > 0 else 1
    volume = {}
    pressure = {}
    if s.barostat is None or n == 0:
        e, cov = ds.md(n, covalent=True) if n > 0 else (np.zeros((0, 5)), np.zeros((0, 4)))
    else:
        # the box changes every step: each written frame's volume (its step's
        # box) and pressure (after its step), the steps between in one call
        parts = []
        for k in range(0, n, every):
            volume[k] = float(np.prod(ds.box()[0]))
            parts.append(ds.md(1, covalent=True))
            pressure[k] = ds.box()[1]
            if min(every, n - k) > 1:
                parts.append(ds.md(min(every, n - k) - 1, covalent=True))
        e = np.concatenate([p[0] for p in parts])
        cov = np.concatenate([p[1] for p in parts])
    x, v = ds.get_state()
    final_box = ds.box()[0]
    ds.close()
    ew = EnergyTrajectoryWriter(out_prefix + ".tre")
    dof = 3 * s.n_atoms - 3 * s.nsm - inp.ndfmin
    for k in range(0, len(e), every):
        ew.write(time=inp.t0 + k * s.dt, kinetic=e[k, 3], potential=e[k, 2], total=e[k, 2] + e[k, 3],
                 temperature=2.0 * e[k, 3] / (dof * K_BOLTZMANN) if dof > 0 else 0.0,
                 volume=volume.get(k, float(np.prod(s.box)) if s.boundary else 0.0),
                 pressure=pressure.get(k, 0.0), lj=e[k, 0], coul_real=e[k, 1],
                 bond=cov[k, 0], angle=cov[k, 1], dihedral=cov[k, 2], improper=cov[k, 3])
    ew.close()
    out_pos, out_vel = x, v
    if inp.ntb == -1:
        # back to the truncated-octahedron frame (the rotation's inverse)
        R = _truncoct_rotation()
        out_pos, out_vel = x @ R, v @ R
    write_configuration(out_prefix + ".cnf", s, out_pos, out_vel, ntb=inp.ntb,
                        box=final_box if s.barostat is not None else None)
    return s, e


def _truncoct_rotation():
    """the matrix truncoct_triclinic applies (rows: the applied columns' components)"""
    _, r, _ = truncoct_triclinic(1.0, np.eye(3))
    return r.T
