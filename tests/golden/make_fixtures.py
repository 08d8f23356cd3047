"""Extract small golden vectors from the gromosXX double-precision fixtures that
gromos-rs ships (crates/gromos-md/tests/gromosXX_references/), as JSON.

Run once in a container that has the reference checked out:

    python tests/golden/make_fixtures.py /root/reference

Only data is written: topology essentials (charges, types, charge groups,
exclusions, 1-4 lists, the LJ table restricted to the types in use),
positions / velocities / box, run parameters, and the expected gromosXX
energies, forces and positions.  The parsing follows the reference's own
readers so that the topology built here is the one gromos-rs builds:

  * SOLUTEATOM: exclusions only from the atom line, 1-4 partners from the next
    non-comment line (crates/gromos-io/src/topology.rs:270-370);
  * IAC / type indices are stored 0-based (topology.rs:300-306, :738-745);
  * bonds and angles add 1-2 / 1-3 exclusions, the 1-4 list stays separate
    and is not mirrored (topology.rs:1141-1183); exclusion lists are sorted and
    de-duplicated (topology.rs:1294-1297);
  * charge groups close at CGC == 1 (topology.rs:1093-1110); every solvent
    molecule is its own charge group and excludes all of its own atom pairs
    (crates/gromos-core/src/topology.rs:1224-1263);
  * perturbed alphas are multiplied by the global ALPHLJ / ALPHC
    (crates/gromos-integrators/src/algorithms/forcefield.rs:175-199).

The fixtures are gromosXX/gromos-rs data (GPL-2.0-or-later).
"""
import json
import os
import sys

FIXDIR = "crates/gromos-md/tests/gromosXX_references"

SYSTEMS = [
    # name, topology, configuration, input, pttopo, frames of forces to keep
    ("pair_lj", None, None, None, None, "all"),
    ("pair_lj_mixed", None, None, None, None, "all"),
    ("nacl_pair", None, None, None, None, "all"),
    ("nacl_pair_box", None, None, None, None, "all"),
    ("water_3_box", None, None, None, None, "all"),
    ("water_3_box_twinrange", None, None, None, None, "all"),
    ("water_10_box", None, None, None, None, "all"),
    ("nacl_3water_box", None, None, None, None, "all"),
    ("nacl_3water_cutoff", None, None, None, None, "all"),
    ("nacl_water_box", None, None, None, None, "all"),
    ("nacl_water_box_shifted", None, None, None, None, "all"),
    ("nacl_1water_settle", None, None, None, None, "all"),
    ("nacl_1water_lincs", None, None, None, None, "all"),
    ("aladip_vacuum_lincs", "../shared/aladip.topo", None, None, None, "all"),
    ("aladip_solvated", "../shared/aladip.topo", "../shared/aladip.conf", None, None, "first"),
    ("water_216_box", None, None, None, None, "first"),
    ("aladip_trunc_oct", "../shared/aladip.topo", "../shared/aladip.conf", None, None, "first"),
    ("aladip_vacuum", "../shared/aladip.topo", None, None, None, "all"),
    ("butane_vacuum", None, None, None, None, "all"),
    ("benzene_vacuum", None, None, None, None, "all"),
    ("water_1000_spc_gridcell", None, None, None, None, "first"),
    ("water_216_nvt", "../water_216_box/water_216_box.topo", "../water_216_box/water_216_box.conf", None, None,
     "all"),
    ("water_216_nvt_nosehoover", "../water_216_box/water_216_box.topo", "../water_216_box/water_216_box.conf",
     None, None, "all"),
    ("water_216_nvt_nhc_chain", "../water_216_box/water_216_box.topo", "../water_216_box/water_216_box.conf",
     None, None, "all"),
    ("water_216_box_com", "../water_216_box/water_216_box.topo", "../water_216_box/water_216_box.conf",
     None, None, "all"),
    ("water_216_box_com_rot", "../water_216_box/water_216_box.topo", "../water_216_box/water_216_box.conf",
     None, None, "all"),
    ("water_216_npt", "../water_216_box/water_216_box.topo", "../water_216_box/water_216_box.conf",
     None, None, "all"),
    ("ch4_water_fep", "../shared/ch4_spc.top", "../shared/ch4_spc.cnf", None,
     "../shared/ch4_spc_dummy.ptp", "first"),
]


def blocks(path):
    """GROMOS block file -> {NAME: [lines]} (first occurrence) plus an ordered list."""
    out, order = {}, []
    name, buf = None, []
    with open(path) as fh:
        for raw in fh:
            line = raw.rstrip("\n")
            s = line.strip()
            if name is None:
                if s and not s.startswith("#"):
                    name, buf = s, []
                continue
            if s == "END":
                order.append((name, buf))
                out.setdefault(name, buf)
                name = None
                continue
            buf.append(line)
    return out, order


def data_lines(lines):
    return [l for l in lines if l.strip() and not l.strip().startswith("#")]


def parse_topology(path):
    b, _ = blocks(path)
    t = {}
    fp = [float(x.split()[0]) for x in data_lines(b.get("PHYSICALCONSTANTS", []))]
    t["fpepsi"] = fp[0] if fp else 138.9354
    # PRESSUREGROUPS (gromos-io topology.rs:822-851): NSVM, then the last atom
    # (1-based) of each solute pressure group
    pg = " ".join(data_lines(b.get("PRESSUREGROUPS", []))).split()
    t["pressure_groups"] = [int(x) - 1 for x in pg[1:1 + int(pg[0])]] if pg else []
    # SOLUTEATOM (topology.rs:270-370)
    lines = data_lines(b.get("SOLUTEATOM", []))
    n = int(lines[0].split()[0])
    mass, q, iac, cgc, excl, o14 = [], [], [], [], [], []
    k = 1
    while len(mass) < n:
        p = lines[k].split()
        k += 1
        iac.append(int(p[3]) - 1)
        mass.append(float(p[4]))
        q.append(float(p[5]))
        cgc.append(int(p[6]))
        ne = int(p[7])
        excl.append([int(x) - 1 for x in p[8:8 + ne]])
        p14 = lines[k].split()
        k += 1
        n14 = int(p14[0])
        o14.append([int(x) - 1 for x in p14[1:1 + n14]])
    t.update(n_solute=n, mass=mass, charge=q, iac=iac, cgc=cgc, excl=excl, o14=o14)
    bonds, angles = [], []
    for key in ("BONDH", "BOND"):
        for l in data_lines(b.get(key, []))[1:]:
            p = l.split()
            bonds.append((int(p[0]) - 1, int(p[1]) - 1))
    for key in ("BONDANGLEH", "BONDANGLE"):
        for l in data_lines(b.get(key, []))[1:]:
            p = l.split()
            angles.append((int(p[0]) - 1, int(p[1]) - 1, int(p[2]) - 1))
    t["bonds"], t["angles"] = bonds, angles
    # bonded terms with their types (1-based type in the file), H blocks first
    # (topology.rs:157-180); type tables converted as the reference stores them
    # (topology.rs:499-513 theta0 in radians, :661-664 cos(PD), :699-703 Q0 in radians)
    import math

    def entries(keys, natoms):
        out = []
        for key in keys:
            tok = [x for l in data_lines(b.get(key, [])) for x in l.split()]
            if not tok:
                continue
            n = int(tok[0])
            for k in range(n):
                e = tok[1 + k * (natoms + 1): 1 + (k + 1) * (natoms + 1)]
                out.append([int(x) - 1 for x in e])
        return out

    def table(key, ncol):
        d = data_lines(b.get(key, []))
        if not d:
            return []
        return [[float(x) for x in l.split()[:ncol]] for l in d[1:1 + int(d[0].split()[0])]]

    t["bond_terms"] = entries(("BONDH", "BOND"), 2)
    t["angle_terms"] = entries(("BONDANGLEH", "BONDANGLE"), 3)
    t["improper_terms"] = entries(("IMPDIHEDRALH", "IMPDIHEDRAL"), 4)
    t["dihedral_terms"] = entries(("DIHEDRALH", "DIHEDRAL"), 4)
    t["bond_types"] = [[cb, chb, b0] for cb, chb, b0 in table("BONDSTRETCHTYPE", 3)]
    t["angle_types"] = [[ct, cht, t0 * (math.pi / 180.0)] for ct, cht, t0 in table("BONDANGLEBENDTYPE", 3)]
    # topology.rs:702-705: CQ kJ/(mol deg^2) -> kJ/(mol rad^2), Q0 deg -> rad
    t["improper_types"] = [[cq * 180.0 * 180.0 / (math.pi * math.pi), q0 * math.pi / 180.0]
                           for cq, q0 in table("IMPDIHEDRALTYPE", 2)]
    t["dihedral_types"] = [[cp, math.cos(pd * math.pi / 180.0), int(np_)]
                           for cp, pd, np_ in table("TORSDIHEDRALTYPE", 3)]
    lj = {}
    key = "LJPARAMETERS" if "LJPARAMETERS" in b else "CGPARAMETERS"
    for l in data_lines(b.get(key, []))[1:]:
        p = l.split()
        i, j = int(p[0]) - 1, int(p[1]) - 1
        c12, c6 = float(p[2]), float(p[3])
        cs12, cs6 = (float(p[4]), float(p[5])) if len(p) >= 6 else (c12, c6)
        lj[(i, j)] = lj[(j, i)] = (c6, c12, cs6, cs12)
    t["lj"] = lj
    sv = data_lines(b.get("SOLVENTATOM", []))
    solv = []
    if sv:
        for l in sv[1:1 + int(sv[0].split()[0])]:
            p = l.split()
            solv.append((int(p[2]) - 1, float(p[3]), float(p[4])))
    t["solvent"] = solv
    sc = data_lines(b.get("SOLVENTCONSTR", []))
    cons = []
    if sc:
        for l in sc[1:1 + int(sc[0].split()[0])]:
            p = l.split()
            cons.append((int(p[0]) - 1, int(p[1]) - 1, float(p[2])))
    t["solvent_constr"] = cons
    return t


def parse_conf(path):
    b, _ = blocks(path)
    pos = [[float(x) for x in l.split()[-3:]] for l in data_lines(b["POSITION"])]
    vel = None
    if "VELOCITY" in b:
        vel = [[float(x) for x in l.split()[-3:]] for l in data_lines(b["VELOCITY"])]
    box = [0.0, 0.0, 0.0]
    ntb = 0
    if "GENBOX" in b:
        g = data_lines(b["GENBOX"])
        ntb = int(g[0].split()[0])
        box = [float(x) for x in g[1].split()[:3]]
    elif "BOX" in b:
        box = [float(x) for x in data_lines(b["BOX"])[0].split()[:3]]
        ntb = 1
    return pos, vel, box, ntb


def parse_imd(path):
    b, _ = blocks(path)
    d = {}
    s = data_lines(b["SYSTEM"])[0].split()
    d["npm"], d["nsm"] = int(s[0]), int(s[1])
    st = data_lines(b["STEP"])[0].split()
    d["nstlim"], d["dt"] = int(st[0]), float(st[2])
    bc = data_lines(b["BOUNDCOND"])[0].split()
    d["ntb"] = int(bc[0])
    d["ndfmin"] = int(bc[1]) if len(bc) > 1 else 0
    # PRESSURESCALE (md.rs:1116-1129, 1259-1277): COUPLE SCALE COMP TAUP VIRIAL,
    # the SEMI row, then the 3x3 reference pressure (isotropic: its [0][0])
    ps = data_lines(b.get("PRESSURESCALE", []))
    d["pressure"] = None
    if ps:
        w = ps[0].split()
        ref = [float(x) for x in " ".join(ps[2:5]).split()]
        d["pressure"] = dict(couple=w[0], scale=w[1], comp=float(w[2]), tau_p=float(w[3]),
                             virial={"none": 0, "atomic": 1, "molecular": 2}.get(w[4].lower(), int(w[4]) if
                                                                                w[4].isdigit() else 0),
                             p_ref=ref[0])
    # centre-of-mass motion removal: INITIALISE NTICOM (third value of its second
    # row), COMTRANSROT NSCM
    ini = data_lines(b.get("INITIALISE", []))
    d["nticom"] = int(ini[1].split()[2]) if len(ini) > 1 else 0
    ctr = data_lines(b.get("COMTRANSROT", []))
    d["nscm"] = int(ctr[0].split()[0]) if ctr else 0
    # MULTIBATH (one bath): weak-coupling -> Berendsen, nose-hoover -> a chain of
    # one, nose-hoover-chains N -> a chain of N
    mb = data_lines(b.get("MULTIBATH", []))
    d["thermostat"] = None
    if mb:
        alg = mb[0].split()
        tt = mb[2].split()
        kind = {"weak-coupling": "berendsen", "0": "berendsen", "nose-hoover": "nosehoover", "1": "nosehoover",
                "nose-hoover-chains": "nosehoover"}.get(alg[0], alg[0])
        nhc = int(alg[1]) if alg[0] == "nose-hoover-chains" and len(alg) > 1 else 1
        d["thermostat"] = dict(kind=kind, nhc=nhc, t_ref=float(tt[0]), tau=float(tt[1]), n_baths=int(mb[1].split()[0]))
    pl = data_lines(b["PAIRLIST"])[0].split()
    alg = pl[0]
    d["algorithm"] = {"standard": 0, "0": 0, "grid": 1, "1": 1, "grid_cell": 2, "2": 2}.get(alg, 0)
    d["nsnb"], d["rcutp"], d["rcutl"] = int(pl[1]), float(pl[2]), float(pl[3])
    nb = data_lines(b["NONBONDED"])
    p = nb[1].split()
    d["appak"], d["rcrf"], d["epsrf"] = float(p[0]), float(p[1]), float(p[2])
    frc = data_lines(b.get("FORCE", []))
    d["ntf"] = [int(x) for x in frc[0].split()[:4]] if frc else [1, 1, 1, 1]
    cons = data_lines(b.get("CONSTRAINT", []))
    d["ntcs"] = cons[3].split()[0] if len(cons) > 3 else "shake"
    # NTC, NTCP, NTCP0 (tolerance, or LINCS order), NTCS, NTCS0 as written
    d["constraint"] = dict(ntc=int(cons[0].split()[0]) if cons else 1,
                           ntcp=cons[1].split()[0] if len(cons) > 1 else "shake",
                           ntcp0=float(cons[2].split()[0]) if len(cons) > 2 else 1e-4,
                           ntcs0=float(cons[4].split()[0]) if len(cons) > 4 else 1e-4)
    if "PERTURBATION" in b:
        pr = data_lines(b["PERTURBATION"])
        a = pr[0].split()
        c = pr[1].split()
        d["ntg"], d["rlam"] = int(a[0]), float(a[2])
        d["alphlj"], d["alphc"], d["nlam"] = float(c[0]), float(c[1]), int(c[2])
    return d


def parse_ptp(path):
    b, _ = blocks(path)
    lines = data_lines(b["PERTATOMPARAM"])
    n = int(lines[0].split()[0])
    atoms = []
    for l in lines[1:1 + n]:
        p = l.split()
        atoms.append(dict(seq=int(p[0]) - 1, a_iac=int(p[3]) - 1, a_q=float(p[5]),
                          b_iac=int(p[6]) - 1, b_q=float(p[8]), alj=float(p[9]), acrf=float(p[10])))
    pairs = []
    pl = data_lines(b.get("PERTATOMPAIR", []))
    if pl:
        for l in pl[1:1 + int(pl[0].split()[0])]:
            p = l.split()
            bt = int(p[3]) - 1
            pairs.append([int(p[0]) - 1, int(p[1]) - 1, int(p[2]) - 1, bt if bt >= 0 else -1])
    return atoms, pairs


def parse_frames(path, block):
    out = []
    _, order = blocks(path)
    step = None
    for name, lines in order:
        if name == "TIMESTEP":
            step = int(data_lines(lines)[0].split()[0])
        elif name == block:
            out.append((step, lines))
    return out


def energies(path):
    frames = []
    for step, lines in parse_frames(path, "ENERGY03"):
        vals = []
        for l in lines:
            s = l.strip()
            if s.startswith("# baths"):
                break
            if s and not s.startswith("#"):
                vals.append(float(s.split()[0]))
        frames.append(dict(step=step, e_tot=vals[0], e_kin=vals[1], e_pot=vals[2], e_covalent=vals[3],
                           e_bond=vals[4], e_angle=vals[5], e_improper=vals[6], e_dihedral=vals[7],
                           e_nonbonded=vals[9], e_lj=vals[10], e_crf=vals[11]))
    return frames


def forces(path):
    out = []
    for step, lines in parse_frames(path, "FREEFORCERED"):
        f = [float(x) for l in data_lines(lines) for x in l.split()[:3]]
        out.append(dict(step=step, f=f))
    return out


def positions(path):
    boxes = {step: [float(x) for x in data_lines(lines)[1].split()[:3]]
             for step, lines in parse_frames(path, "GENBOX")}
    out = []
    for step, lines in parse_frames(path, "POSITIONRED"):
        fr = dict(step=step, x=[float(x) for l in data_lines(lines) for x in l.split()[:3]])
        if step in boxes:
            fr["box"] = boxes[step]
        out.append(fr)
    return out


def build(ref, name, topo_rel, conf_rel, imd_rel, ptp_rel, keep):
    d = os.path.join(ref, FIXDIR, name)
    files = os.listdir(d)
    topo_p = os.path.join(d, topo_rel) if topo_rel else os.path.join(
        d, [f for f in files if f.endswith((".topo", ".top"))][0])
    conf_p = os.path.join(d, conf_rel) if conf_rel else os.path.join(
        d, [f for f in files if f.endswith((".conf", ".cnf"))][0])
    imd_p = os.path.join(d, imd_rel) if imd_rel else os.path.join(
        d, [f for f in files if f.endswith(".in")][0])
    t = parse_topology(topo_p)
    imd = parse_imd(imd_p)
    pos, vel, box, ntb_conf = parse_conf(conf_p)
    nsm = imd["nsm"]
    ns = t["n_solute"]
    aps = len(t["solvent"])
    n = ns + nsm * aps
    assert len(pos) == n, (name, len(pos), n)
    # charge groups (topology.rs:1093-1110) + solvent molecules
    cg_start, cur = [0], 0
    for i in range(ns):
        if t["cgc"][i] == 1:
            cg_start.append(i + 1)
    if cg_start[-1] != ns:
        cg_start.append(ns)
    n_solute_cg = len(cg_start) - 1
    for m in range(nsm):
        cg_start.append(ns + (m + 1) * aps)
    # exclusions (topology.rs:1141-1183, 1294-1297; core topology.rs:1242-1263)
    ex = [set() for _ in range(n)]
    for i, lst in enumerate(t["excl"]):
        for j in lst:
            ex[i].add(j)
            ex[j].add(i)
    for i, j in t["bonds"]:
        ex[i].add(j); ex[j].add(i)
    for i, _, k in t["angles"]:
        ex[i].add(k); ex[k].add(i)
    for m in range(nsm):
        base = ns + m * aps
        for a in range(aps):
            for b2 in range(a + 1, aps):
                ex[base + a].add(base + b2); ex[base + b2].add(base + a)
    excl = [sorted(s) for s in ex]
    o14 = [list(l) for l in t["o14"]] + [[] for _ in range(nsm * aps)]
    iac = list(t["iac"]) + [s[0] for _ in range(nsm) for s in t["solvent"]]
    mass = list(t["mass"]) + [s[1] for _ in range(nsm) for s in t["solvent"]]
    charge = list(t["charge"]) + [s[2] for _ in range(nsm) for s in t["solvent"]]
    pert = None
    if ptp_rel:
        atoms, apairs = parse_ptp(os.path.join(d, ptp_rel))
        pert = dict(atoms=atoms, atom_pairs=apairs, rlam=imd.get("rlam", 0.0),
                    nlam=imd.get("nlam", 1), alphlj=imd.get("alphlj", 1.0), alphc=imd.get("alphc", 1.0))
    used = sorted(set(iac) | ({a["a_iac"] for a in pert["atoms"]} | {a["b_iac"] for a in pert["atoms"]}
                              if pert else set()))
    remap = {old: k for k, old in enumerate(used)}
    nt = len(used)
    lj = []
    for a in used:
        for b2 in used:
            lj.extend(t["lj"].get((a, b2), (0.0, 0.0, 0.0, 0.0)))
    if pert:
        for a in pert["atoms"]:
            a["a_iac"], a["b_iac"] = remap[a["a_iac"]], remap[a["b_iac"]]
    ntb = imd["ntb"]
    box_matrix = None
    if ntb == -1:
        (a, b3, c), pos, vel = truncoct_triclinic(box[0], pos, vel)
        box_matrix = a + b3 + c            # columns a, b, c
        box = [a[0], b3[1], c[2]]          # the diagonal
    fx = dict(
        name=name, source=os.path.join(FIXDIR, name),
        n_atoms=n, n_solute_atoms=ns, n_cg=len(cg_start) - 1, n_solute_cg=n_solute_cg,
        atoms_per_solvent=aps, nsm=nsm, cg_start=cg_start,
        charge=charge, iac=[remap[x] for x in iac], mass=mass,
        n_types=nt, lj=lj, excl=excl, o14=o14,
        pos=[x for p in pos for x in p], vel=[x for v in vel for x in v] if vel else None,
        box=box if ntb != 0 else [0.0, 0.0, 0.0], ntb=2 if ntb == -1 else (1 if ntb != 0 else 0),
        box_matrix=box_matrix,
        fpepsi=t["fpepsi"], rcutp=imd["rcutp"], rcutl=imd["rcutl"], nsnb=imd["nsnb"],
        algorithm=imd["algorithm"], epsrf=imd["epsrf"], appak=imd["appak"],
        dt=imd["dt"], nstlim=imd["nstlim"], ntcs=imd["ntcs"], constraint=imd["constraint"], ndfmin=imd["ndfmin"], thermostat=imd["thermostat"],
        nticom=imd["nticom"], nscm=imd["nscm"], pressure=imd["pressure"],
        pressure_groups=t["pressure_groups"],
        solvent_constr=t["solvent_constr"], pert=pert,
        has_bonded_terms=bool(t["bonds"] or t["angles"]),
        bonded=dict(bonds=t["bond_terms"], angles=t["angle_terms"], impropers=t["improper_terms"],
                    dihedrals=t["dihedral_terms"], bond_types=t["bond_types"], angle_types=t["angle_types"],
                    improper_types=t["improper_types"], dihedral_types=t["dihedral_types"], ntf=imd["ntf"]),
    )
    ex_dir = os.path.join(d, "expected")
    fx["expected_energies"] = energies(os.path.join(ex_dir, "energies.tre"))
    fr = forces(os.path.join(ex_dir, "forces.trf"))
    fx["expected_forces"] = fr if keep == "all" else fr[:1]
    trc = os.path.join(ex_dir, "trajectory.trc")
    if keep == "all" and os.path.exists(trc):
        fx["expected_positions"] = positions(trc)
    if imd.get("pressure"):
        # the box after the run (final.conf GENBOX edge lengths)
        b, _ = blocks(os.path.join(ex_dir, "final.conf"))
        fx["final_box"] = [float(x) for x in data_lines(b["GENBOX"])[1].split()[:3]]
    return fx


def truncoct_triclinic(L, pos, vel):
    """NTB = -1 (crates/gromos-md/src/bin/md.rs:530-545): the cube edge L of the
    BOX block becomes the lower-triangular truncated-octahedron box
    (crates/gromos-core/src/math.rs truncoct_triclinic_box, forward), and every
    position and velocity is rotated into that frame by the transpose of the
    rotation whose columns are (1, 1, 1)/sqrt3, (-2, 1, 1)/sqrt6, (0, -1, 1)/sqrt2
    (truncoct_triclinic_rotmat(true)); glam's Mat3 * Vec3 sums the columns
    times x, y, z left to right.  Returns the box vectors a, b, c and the
    rotated vectors."""
    import math
    d = 0.5 * math.sqrt(3.0) * L
    a = [d, 0.0, 0.0]
    b = [d / 3.0, 2.0 * math.sqrt(2.0) / 3.0 * d, 0.0]
    c = [-d / 3.0, math.sqrt(2.0) / 3.0 * d, math.sqrt(6.0) / 3.0 * d]
    sq3i = 1.0 / math.sqrt(3.0)
    sq2i = 1.0 / math.sqrt(2.0)
    # columns of the applied matrix (the rows of the rotation)
    mx = [sq3i, -2.0 * sq2i * sq3i, 0.0]
    my = [sq3i, sq3i * sq2i, -sq2i]
    mz = [sq3i, sq2i * sq3i, sq2i]

    def rot(v):
        return [mx[k] * v[0] + my[k] * v[1] + mz[k] * v[2] for k in range(3)]

    return (a, b, c), [rot(v) for v in pos], ([rot(v) for v in vel] if vel else vel)


# input files copied as data for the product's GROMOS readers (name: files
# relative to the fixture directory)
INPUT_SETS = {
    "water_10_box": ["water_10_box.topo", "water_10_box.conf", "water_10_box.in"],
    "nacl_water_box": ["nacl_water_box.topo", "nacl_water_box.conf", "nacl_water_box.in"],
    "nacl_1water_settle": ["nacl_1water_settle.topo", "nacl_1water_settle.conf", "nacl_1water_settle.in"],
    "nacl_3water_cutoff": ["nacl_3water_cutoff.topo", "nacl_3water_cutoff.conf", "nacl_3water_cutoff.in"],
    "water_3_box_twinrange": ["water_3_box_twinrange.topo", "water_3_box_twinrange.conf",
                              "water_3_box_twinrange.in"],
    "aladip_solvated": ["../shared/aladip.topo", "../shared/aladip.conf", "aladip_solvated.in"],
    "aladip_trunc_oct": ["../shared/aladip.topo", "../shared/aladip.conf", "aladip_trunc_oct.in"],
    "ch4_water_fep": ["../shared/ch4_spc.top", "../shared/ch4_spc.cnf", "ch4_water_fep.in",
                      "../shared/ch4_spc_dummy.ptp"],
    "aladip_vacuum": ["../shared/aladip.topo", "aladip_vacuum.conf", "aladip_vacuum.in"],
    "nacl_1water_lincs": ["nacl_1water_lincs.topo", "nacl_1water_lincs.conf", "nacl_1water_lincs.in"],
    "aladip_vacuum_lincs": ["../shared/aladip.topo", "aladip_vacuum_lincs.conf", "aladip_vacuum_lincs.in"],
    "butane_vacuum": ["butane_vacuum.topo", "butane_vacuum.conf", "butane_vacuum.in"],
    "benzene_vacuum": ["benzene_vacuum.topo", "benzene_vacuum.conf", "benzene_vacuum.in"],
    "water_216_nvt_nosehoover": ["../water_216_box/water_216_box.topo", "../water_216_box/water_216_box.conf",
                                 "water_216_nvt_nosehoover.in"],
    "water_216_nvt_nhc_chain": ["../water_216_box/water_216_box.topo", "../water_216_box/water_216_box.conf",
                                "water_216_nvt_nhc_chain.in"],
    "water_216_box_com_rot": ["../water_216_box/water_216_box.topo", "../water_216_box/water_216_box.conf",
                              "water_216_box_com_rot.in"],
    "water_216_npt": ["../water_216_box/water_216_box.topo", "../water_216_box/water_216_box.conf",
                      "water_216_npt.in"],
}


def copy_inputs(ref, here):
    """tests/golden/gromos/<name>/ <- the fixture's input files (data)"""
    import shutil
    for name, files in INPUT_SETS.items():
        dst = os.path.join(here, "gromos", name)
        os.makedirs(dst, exist_ok=True)
        for f in files:
            shutil.copyfile(os.path.join(ref, FIXDIR, name, f), os.path.join(dst, os.path.basename(f)))
        print(f"inputs of {name}: {', '.join(os.path.basename(f) for f in files)}")


def main():
    ref = sys.argv[1] if len(sys.argv) > 1 else "/root/reference"
    if len(sys.argv) > 2 and sys.argv[2] == "--inputs":
        copy_inputs(ref, os.path.dirname(os.path.abspath(__file__)))
        return
    only = set(sys.argv[2:])   # names to (re)build; all when empty
    here = os.path.dirname(os.path.abspath(__file__))
    for spec in SYSTEMS:
        if only and spec[0] not in only:
            continue
        fx = build(ref, *spec)
        out = os.path.join(here, spec[0] + ".json")
        with open(out, "w") as fh:
            json.dump(fx, fh, separators=(",", ":"))
        print(f"{spec[0]}: {fx['n_atoms']} atoms, {len(fx['expected_energies'])} energy frames "
              f"-> {os.path.getsize(out) // 1024} KiB")


if __name__ == "__main__":
    main()
