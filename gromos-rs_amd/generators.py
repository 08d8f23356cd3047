"""Seeded synthetic systems for the benchmark and parity workloads.

SPC water (all workloads): O -0.82 e / H +0.41 e, masses 15.9994 / 1.008,
O-H 0.1 nm, H-O-H 109.47 deg, LJ on O-O only (C6 = 2.6173456e-3,
C12 = 2.634129e-6 kJ mol^-1 nm^6 / nm^12), every water its own charge group
and solvent molecule (crates/gromos-core/src/topology.rs:1204-1263), reaction
field eps_rf = 61, kappa = 0, four_pi_eps_i = 138.9354.

Waters sit on a cubic lattice with uniform +-jitter per coordinate, uniformly
random rigid orientations and Maxwell-Boltzmann velocities projected onto
rigid-body motion (so SETTLE starts from consistent velocities).
"""
import math

import numpy as np

from .system import FOUR_PI_EPS_I, K_BOLTZMANN, Bonded, Perturbation, System, csr_from_lists

SPC_C6 = 2.6173456e-3
SPC_C12 = 2.634129e-6
M_O, M_H = 15.9994, 1.008
Q_O, Q_H = -0.82, 0.41
D_OH = 0.1
THETA_HOH = math.radians(109.47)
D_HH = 2.0 * D_OH * math.sin(0.5 * THETA_HOH)
AVOGADRO = 6.02214076e23


def box_for_density(n_waters, density_g_cm3=0.997):
    m = n_waters * (M_O + 2 * M_H) / AVOGADRO        # g
    vol_nm3 = m / density_g_cm3 * 1e21
    return vol_nm3 ** (1.0 / 3.0)


def random_rotations(rng, n):
    """Uniform random rotation matrices (Shoemake quaternions)."""
    u1, u2, u3 = rng.random(n), rng.random(n), rng.random(n)
    q = np.stack([np.sqrt(1 - u1) * np.sin(2 * np.pi * u2), np.sqrt(1 - u1) * np.cos(2 * np.pi * u2),
                  np.sqrt(u1) * np.sin(2 * np.pi * u3), np.sqrt(u1) * np.cos(2 * np.pi * u3)], axis=1)
    x, y, z, w = q.T
    R = np.empty((n, 3, 3))
    R[:, 0, 0] = 1 - 2 * (y * y + z * z); R[:, 0, 1] = 2 * (x * y - z * w); R[:, 0, 2] = 2 * (x * z + y * w)
    R[:, 1, 0] = 2 * (x * y + z * w); R[:, 1, 1] = 1 - 2 * (x * x + z * z); R[:, 1, 2] = 2 * (y * z - x * w)
    R[:, 2, 0] = 2 * (x * z - y * w); R[:, 2, 1] = 2 * (y * z + x * w); R[:, 2, 2] = 1 - 2 * (x * x + y * y)
    return R


def water_template():
    s, c = math.sin(0.5 * THETA_HOH), math.cos(0.5 * THETA_HOH)
    return np.array([[0.0, 0.0, 0.0], [D_OH * s, D_OH * c, 0.0], [-D_OH * s, D_OH * c, 0.0]])


def rigid_mb_velocities(rng, pos, masses, temperature):
    """Per-atom Maxwell-Boltzmann velocities projected on rigid-body motion per water."""
    n = len(pos) // 3
    m = np.asarray(masses[:3])
    sig = np.sqrt(K_BOLTZMANN * temperature / np.tile(m, n))[:, None]
    v = rng.normal(size=(3 * n, 3)) * sig
    x = pos.reshape(n, 3, 3)
    v = v.reshape(n, 3, 3)
    M = m.sum()
    com = (x * m[None, :, None]).sum(1) / M
    vcom = (v * m[None, :, None]).sum(1) / M
    r = x - com[:, None, :]
    L = (np.cross(r, v) * m[None, :, None]).sum(1)
    I = np.zeros((n, 3, 3))
    for a in range(3):
        rr = r[:, a, :]
        I += m[a] * (np.einsum("ni,ni->n", rr, rr)[:, None, None] * np.eye(3)[None] - rr[:, :, None] * rr[:, None, :])
    omega = np.linalg.solve(I, L[..., None])[..., 0]
    vr = vcom[:, None, :] + np.cross(omega[:, None, :], r)
    return vr.reshape(-1, 3)


def spc_water_box(n_waters, box=None, density=0.997, jitter=0.02, seed=0, temperature=300.0,
                  rcutp=1.4, rcutl=1.4, nsnb=5, epsrf=61.0, dt=0.002, name="spc"):
    rng = np.random.default_rng(seed)
    L = box_for_density(n_waters, density) if box is None else float(box)
    side = int(math.ceil(round(n_waters ** (1.0 / 3.0), 9)))
    while side ** 3 < n_waters:
        side += 1
    sites = np.stack(np.meshgrid(np.arange(side), np.arange(side), np.arange(side), indexing="ij"),
                     axis=-1).reshape(-1, 3)[:, ::-1]
    if len(sites) > n_waters:
        sites = sites[np.sort(rng.choice(len(sites), n_waters, replace=False))]
    spacing = L / side
    o = (sites + 0.5) * spacing + rng.uniform(-jitter, jitter, size=(n_waters, 3))
    R = random_rotations(rng, n_waters)
    tmpl = water_template()
    pos = (o[:, None, :] + np.einsum("nij,aj->nai", R, tmpl)).reshape(-1, 3)
    mass = np.tile([M_O, M_H, M_H], n_waters)
    vel = rigid_mb_velocities(rng, pos, mass, temperature) if temperature > 0 else np.zeros_like(pos)
    n = 3 * n_waters
    charge = np.tile([Q_O, Q_H, Q_H], n_waters)
    iac = np.tile(np.array([0, 1, 1], np.int32), n_waters)
    lj = np.zeros((2, 2, 4))
    lj[0, 0] = (SPC_C6, SPC_C12, SPC_C6, SPC_C12)
    excl = []
    for w in range(n_waters):
        b = 3 * w
        excl += [[b + 1, b + 2], [b, b + 2], [b, b + 1]]
    excl_off, excl_idx = csr_from_lists(excl, n)
    o14_off = np.zeros(n + 1, np.int32)
    return System(n_atoms=n, n_solute_atoms=0, cg_start=np.arange(0, n + 1, 3, dtype=np.int32), n_solute_cg=0,
                  charge=charge, iac=iac, mass=mass, n_types=2, lj=lj.reshape(-1), excl_off=excl_off,
                  excl=excl_idx, o14_off=o14_off, o14=np.zeros(0, np.int32), atoms_per_solvent=3, pos=pos,
                  vel=vel, box=np.array([L, L, L]), boundary=1, rcutp=rcutp, rcutl=rcutl, nsnb=nsnb,
                  epsrf=epsrf, kappa=0.0, fpepsi=FOUR_PI_EPS_I, dt=dt, dist_oh=D_OH, dist_hh=D_HH,
                  name=name).validate()


def truncoct_box(L):
    """The truncated-octahedron box of cube edge L as lower-triangular vectors
    a, b, c (columns; crates/gromos-core/src/math.rs truncoct_triclinic_box)."""
    d = 0.5 * math.sqrt(3.0) * L
    a = [d, 0.0, 0.0]
    b = [d / 3.0, 2.0 * math.sqrt(2.0) / 3.0 * d, 0.0]
    c = [-d / 3.0, math.sqrt(2.0) / 3.0 * d, math.sqrt(6.0) / 3.0 * d]
    return np.array([a, b, c]).T


def water_truncoct(n_side, density=0.997, jitter=0.02, seed=0, temperature=300.0, rc=1.2, nsnb=5, name=None):
    """n_side^3 SPC waters in a truncated-octahedron box (triclinic boundary)
    sized for the density: a lattice n_side^3 in fractional coordinates of the
    box vectors, random rigid orientations, Maxwell-Boltzmann velocities."""
    rng = np.random.default_rng(seed)
    n_waters = n_side ** 3
    vol = n_waters * 18.01528 / (density * 602.214076)          # nm^3
    L = (2.0 * vol) ** (1.0 / 3.0)                              # the octahedron is half the cube
    M = truncoct_box(L)
    f = (np.stack(np.meshgrid(*[np.arange(n_side)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) / n_side
    o = f @ M.T + rng.uniform(-jitter, jitter, size=(n_waters, 3))
    s = spc_water_box(n_waters, box=L, jitter=0.0, seed=seed, temperature=0.0, rcutp=rc, rcutl=rc, nsnb=nsnb,
                      name=name or f"truncoct_{n_waters}")
    R = random_rotations(rng, n_waters)
    pos = (o[:, None, :] + np.einsum("nij,aj->nai", R, water_template())).reshape(-1, 3)
    s.pos = pos
    s.vel = rigid_mb_velocities(rng, pos, s.mass, temperature) if temperature > 0 else np.zeros_like(pos)
    s.boundary = 2
    s.box_matrix = M
    s.box = np.diag(M).copy()
    return s.validate()


def water_216(seed=1):
    """216 SPC on a 6^3 lattice, L = 1.8621 nm, twin range 0.8 / 1.4 (cutoff > L/2)."""
    return spc_water_box(216, box=1.8621, seed=seed, rcutp=0.8, rcutl=1.4, nsnb=5, name="water_216")


def water_32768(seed=2):
    """32,768 SPC at 0.997 g/cm^3 (L ~ 9.94 nm), single range 1.4 nm."""
    return spc_water_box(32768, seed=seed, rcutp=1.4, rcutl=1.4, nsnb=5, name="water_32768")


def water_replica(seed, temperature, n_waters=4000, rc=1.2):
    """12,000-atom SPC replica box (L ~ 4.93 nm), cutoff 1.2 nm."""
    return spc_water_box(n_waters, seed=seed, temperature=temperature, rcutp=rc, rcutl=rc, nsnb=5,
                         name=f"replica_{temperature:.1f}K")


# ---------------------------------------------------------------------------
# heterogeneous solute + water, FEP jobs
# ---------------------------------------------------------------------------
def lj_type_table(rng, n_types=20):
    """C6 log-uniform in [1e-4, 1e-2], C12 log-uniform in [1e-7, 1e-4] per type;
    geometric mixing (crates/gromos-core/src/topology.rs:1141-1152); 1-4 table
    scaled by 1/2."""
    c6 = np.exp(rng.uniform(np.log(1e-4), np.log(1e-2), n_types))
    c12 = np.exp(rng.uniform(np.log(1e-7), np.log(1e-4), n_types))
    return c6, c12


def _mixed_table(c6s, c12s):
    """types: solute types ..., water O, water H; returns flat (c6, c12, cs6, cs12)."""
    c6 = np.concatenate([c6s, [SPC_C6, 0.0]])
    c12 = np.concatenate([c12s, [SPC_C12, 0.0]])
    nt = len(c6)
    C6 = np.sqrt(np.outer(c6, c6))
    C12 = np.sqrt(np.outer(c12, c12))
    t = np.zeros((nt, nt, 4))
    t[..., 0], t[..., 1] = C6, C12
    t[..., 2], t[..., 3] = 0.5 * C6, 0.5 * C12
    return t.reshape(-1), nt


def _periodic_min_dist2(a, b, L):
    d = a[:, None, :] - b[None, :, :]
    d -= L * np.round(d / L)
    return (d * d).sum(-1)


def _random_walk(rng, n, bond=0.15, start=None, min_dist=0.22, L=None):
    """Self-avoiding-ish chain: successive bonds at >= 70 degrees turning
    freedom, new atoms at least min_dist from all but their two predecessors
    (where 200 tries find no such place, the try farthest from the others)."""
    pos = np.zeros((n, 3))
    pos[0] = start if start is not None else rng.uniform(0, 1, 3)
    prev_dir = None
    for i in range(1, n):
        best, best_d2, best_dir = None, -1.0, None
        for _ in range(200):
            d = rng.normal(size=3)
            d /= np.linalg.norm(d)
            if prev_dir is not None and np.dot(d, prev_dir) > 0.34:   # angle to previous bond > ~110 deg apart
                continue
            cand = pos[i - 1] + bond * d
            d2 = np.inf
            if i > 2:
                dd = pos[: i - 2] - cand
                if L is not None:
                    dd -= L * np.round(dd / L)
                d2 = (dd * dd).sum(1).min()
            if d2 > best_d2:
                best, best_d2, best_dir = cand, d2, d
            if d2 >= min_dist * min_dist:
                break
        pos[i] = best
        prev_dir = -best_dir
    return pos


def _dihedral_angle(a, b, c, d):
    """signed angle between the planes (a, b, c) and (b, c, d), as the improper
    term measures it (improper.rs:27-47)"""
    rij, rkj, rkl = a - b, c - b, c - d
    m, n = np.cross(rij, rkj), np.cross(rkj, rkl)
    cs = np.clip((m * n).sum(1) / (np.linalg.norm(m, axis=1) * np.linalg.norm(n, axis=1)), -1.0, 1.0)
    z = np.arccos(cs)
    return np.where((rij * n).sum(1) < 0.0, -z, z)


def chain_terms(rng, x, off=0):
    """Covalent terms along a chain at positions x (first atom index off): a
    bond per neighbour pair, an angle per triple, a proper dihedral per quadruple
    and an improper on every fifth quadruple.  Bonds, angles and impropers get a
    type each, whose equilibrium value is the present one with a little noise
    (r0 within 1 %, theta0 within 2 degrees, q0 within 0.05 rad) and GROMOS-like
    constants (quartic 1.0-1.4e7 kJ/mol/nm^4, cosine 400-600 kJ/mol, improper
    100-300 kJ/mol/rad^2); dihedrals draw from seven types, multiplicities 0-6,
    cos(delta) = +-1, K 1-6 kJ/mol."""
    n = len(x)
    i = np.arange(n, dtype=np.int32)
    r = np.linalg.norm(x[1:] - x[:-1], axis=1)
    bt = np.stack([rng.uniform(1.0e7, 1.4e7, n - 1), np.zeros(n - 1), r * rng.uniform(0.99, 1.01, n - 1)], 1)
    u, v = x[:-2] - x[1:-1], x[2:] - x[1:-1]
    th = np.arccos(np.clip((u * v).sum(1) / (np.linalg.norm(u, axis=1) * np.linalg.norm(v, axis=1)), -1, 1))
    at = np.stack([rng.uniform(400, 600, n - 2), np.zeros(n - 2), th + np.radians(rng.uniform(-2, 2, n - 2))], 1)
    dt = np.stack([rng.uniform(1.0, 6.0, 7), rng.choice([-1.0, 1.0], 7), np.arange(7, dtype=float)], 1)
    q = np.arange(0, n - 3, 5)
    zeta = _dihedral_angle(x[q], x[q + 1], x[q + 2], x[q + 3])
    it = np.stack([rng.uniform(100, 300, len(q)), zeta + rng.uniform(-0.05, 0.05, len(q))], 1)
    bonds = np.stack([i[:-1] + off, i[1:] + off, np.arange(n - 1)], 1)
    angles = np.stack([i[:-2] + off, i[1:-1] + off, i[2:] + off, np.arange(n - 2)], 1)
    dih = np.stack([i[:-3] + off, i[1:-2] + off, i[2:-1] + off, i[3:] + off, rng.integers(0, 7, n - 3)], 1)
    imp = np.stack([q + off, q + 1 + off, q + 2 + off, q + 3 + off, np.arange(len(q))], 1)
    return Bonded.of(bonds, bt, angles, at, imp, it, dih, dt)


def hetero_system(n_chain=6000, n_waters=18000, box=8.4, seed=5, rc=1.4, nsnb=5, n_types=20, pert_every=0,
                  lam=0.5, name="hetero", bonded=False):
    """Heterogeneous single-point system: a random-walk solute chain (bond 0.15 nm)
    in charge groups of 1-4 atoms with N(0, 0.3^2) charges shifted to zero per
    group, 20 LJ types, 1-2 / 1-3 exclusions, 1-4 pairs with the scaled table,
    plus SPC water.  pert_every > 0 perturbs every pert_every-th chain atom
    (types and charges of state B drawn the same way).  bonded: the chain keeps
    its covalent terms (chain_terms) and stays whole (the covalent terms take
    raw coordinate differences); otherwise each group is moved into the box."""
    rng = np.random.default_rng(seed)
    L = float(box)
    chain = _random_walk(rng, n_chain, start=rng.uniform(0, L, 3), L=L)
    # charge groups of 1-4 consecutive atoms, each moved by lattice vectors so that
    # its first atom lies in the box (the reference takes raw COGs)
    sizes = []
    left = n_chain
    while left > 0:
        k = int(min(left, rng.integers(1, 5)))
        sizes.append(k)
        left -= k
    starts = np.concatenate([[0], np.cumsum(sizes)]).astype(np.int32)
    if not bonded:
        for c in range(len(sizes)):
            a, b = starts[c], starts[c + 1]
            chain[a:b] -= L * np.floor(chain[a] / L)
    q = np.zeros(n_chain)
    for c in range(len(sizes)):
        a, b = starts[c], starts[c + 1]
        v = rng.normal(0, 0.3, b - a)
        q[a:b] = v - v.mean()
    c6s, c12s = lj_type_table(rng, n_types)
    lj, nt = _mixed_table(c6s, c12s)
    t_o, t_h = n_types, n_types + 1
    iac_chain = rng.integers(0, n_types, n_chain).astype(np.int32)
    # water on a lattice, sites too close to the chain removed; the lattice is
    # refined until at least n_waters sites remain
    cw = chain - L * np.floor(chain / L)
    cell = 0.6
    nc = max(1, int(L // cell))
    ci = np.floor(cw / (L / nc)).astype(int) % nc
    buckets = {}
    for idx, c in enumerate(map(tuple, ci)):
        buckets.setdefault(c, []).append(idx)
    side = int(np.ceil((n_waters * 1.25) ** (1 / 3)))
    while True:
        g = (np.stack(np.meshgrid(*[np.arange(side)] * 3, indexing="ij"), -1).reshape(-1, 3) + 0.5) * (L / side)
        g += rng.uniform(-0.02, 0.02, g.shape)
        keep = np.ones(len(g), bool)
        gi = np.floor(g / (L / nc)).astype(int) % nc
        for s_idx in range(len(g)):
            cx, cy, cz = gi[s_idx]
            near = []
            for dx in (-1, 0, 1):
                for dy in (-1, 0, 1):
                    for dz in (-1, 0, 1):
                        near += buckets.get(((cx + dx) % nc, (cy + dy) % nc, (cz + dz) % nc), [])
            if near:
                d = cw[near] - g[s_idx]
                d -= L * np.round(d / L)
                if ((d * d).sum(1) < 0.28 ** 2).any():
                    keep[s_idx] = False
        if keep.sum() >= n_waters:
            break
        side += 1
    g = g[keep]
    if len(g) > n_waters:
        g = g[np.sort(rng.choice(len(g), n_waters, replace=False))]
    nw = len(g)
    R = random_rotations(rng, nw)
    wpos = (g[:, None, :] + np.einsum("nij,aj->nai", R, water_template())).reshape(-1, 3)
    n = n_chain + 3 * nw
    pos = np.concatenate([chain, wpos])
    mass = np.concatenate([np.full(n_chain, 13.019), np.tile([M_O, M_H, M_H], nw)])
    charge = np.concatenate([q, np.tile([Q_O, Q_H, Q_H], nw)])
    iac = np.concatenate([iac_chain, np.tile(np.array([t_o, t_h, t_h], np.int32), nw)]).astype(np.int32)
    excl = [[] for _ in range(n)]
    o14 = [[] for _ in range(n)]
    for i in range(n_chain):
        for d in (1, 2):
            if i + d < n_chain:
                excl[i].append(i + d)
                excl[i + d].append(i)
        if i + 3 < n_chain:
            o14[i].append(i + 3)
    for w in range(nw):
        b = n_chain + 3 * w
        excl[b] += [b + 1, b + 2]
        excl[b + 1] += [b, b + 2]
        excl[b + 2] += [b, b + 1]
    excl = [sorted(e) for e in excl]
    excl_off, excl_idx = csr_from_lists(excl, n)
    o14_off, o14_idx = csr_from_lists(o14, n)
    cg_start = np.concatenate([starts, n_chain + 3 * np.arange(1, nw + 1)]).astype(np.int32)
    pert = None
    if pert_every:
        pert = Perturbation.empty(n)
        for i in range(0, n_chain, pert_every):
            pert.flag[i] = 1
            pert.a_iac[i], pert.a_q[i] = iac[i], charge[i]
            pert.b_iac[i] = rng.integers(0, n_types)
            pert.b_q[i] = rng.uniform(-0.5, 0.5)
            pert.alpha_lj[i], pert.alpha_crf[i] = 0.5, 0.5
        pert.lam, pert.nlam = lam, 1
    vel = np.zeros((n, 3))
    return System(n_atoms=n, n_solute_atoms=n_chain, cg_start=cg_start, n_solute_cg=len(sizes), charge=charge,
                  iac=iac, mass=mass, n_types=nt, lj=lj, excl_off=excl_off, excl=excl_idx, o14_off=o14_off,
                  o14=o14_idx, atoms_per_solvent=3, pos=pos, vel=vel, box=np.array([L, L, L]), boundary=1,
                  rcutp=rc, rcutl=rc, nsnb=nsnb, epsrf=61.0, kappa=0.0, fpepsi=FOUR_PI_EPS_I, dt=0.002,
                  dist_oh=D_OH, dist_hh=D_HH, pert=pert, name=name,
                  bonded=chain_terms(rng, chain) if bonded else None).validate()


def fep_job(n_atoms, n_pert, lam, seed, rc=1.4, nsnb=5, alpha_lj=0.5, alpha_crf=0.5, n_types=20,
            type_seed=5, temperature=300.0, name=None, min_sigma_ox=0.28):
    """One FEP lambda-window job: n_pert perturbed solute atoms (each its own
    charge group, no bonds, >= 0.4 nm apart) with A/B charges uniform in
    [-0.5, 0.5] e and A/B types from the 20-type table (those whose LJ diameter
    against the water oxygen is at least min_sigma_ox nm), in SPC water at
    0.997 g/cm^3 in the box sized for n_atoms, the waters of each site's first
    shell removed (so the job holds somewhat fewer than n_atoms atoms);
    soft-core alphas 0.5 / 0.5 nm^2 (global ALPHLJ = ALPHC = 1)."""
    rng = np.random.default_rng(seed)
    n_w = max(1, (n_atoms - n_pert) // 3)
    L = box_for_density(n_w + n_pert // 3 + 1)
    base = spc_water_box(n_w + n_pert // 3 + 1, box=L, seed=seed, temperature=temperature, rcutp=rc, rcutl=rc,
                         nsnb=nsnb)
    # solute sites
    sites = []
    while len(sites) < n_pert:
        c = rng.uniform(0, L, 3)
        if all(((np.asarray(sv) - c - L * np.round((np.asarray(sv) - c) / L)) ** 2).sum() >= 0.16 for sv in sites):
            sites.append(c)
    sites = np.array(sites)
    # clear the first shell of every site: no O within 0.3 nm, no H within 0.22 nm
    # (H carries no LJ, so nothing else would keep it off a charged site)
    wp = base.pos.reshape(-1, 3, 3)
    d_o = _periodic_min_dist2(wp[:, 0], sites, L).min(1)
    d_h = np.minimum(_periodic_min_dist2(wp[:, 1], sites, L).min(1), _periodic_min_dist2(wp[:, 2], sites, L).min(1))
    ok = np.where((d_o >= 0.09) & (d_h >= 0.22 ** 2))[0]
    keepw = np.sort(ok[np.argsort(-d_o[ok], kind="stable")[:n_w]])
    n_w = len(keepw)
    wpos = base.pos.reshape(-1, 3, 3)[keepw].reshape(-1, 3)
    wvel = base.vel.reshape(-1, 3, 3)[keepw].reshape(-1, 3)
    trng = np.random.default_rng(type_seed)
    c6s, c12s = lj_type_table(trng, n_types)
    lj, nt = _mixed_table(c6s, c12s)
    t_o, t_h = n_types, n_types + 1
    n = n_pert + 3 * n_w
    # A/B types from the table's types with an oxygen-site LJ diameter
    # (C12/C6)^(1/6) >= min_sigma_ox: with charges up to 0.5 e, a smaller site lets
    # a water hydrogen (no LJ of its own) collapse onto it at the unsoftened end
    # states and the job fails (SETTLE)
    ljt = np.asarray(lj).reshape(nt, nt, 4)
    c6o, c12o = ljt[t_o, :n_types, 0], ljt[t_o, :n_types, 1]
    sig = np.where(c6o > 0, (c12o / np.where(c6o > 0, c6o, 1.0)) ** (1.0 / 6.0), 0.0)
    allowed = np.where(sig >= min_sigma_ox)[0]
    a_iac = allowed[rng.integers(0, len(allowed), n_pert)].astype(np.int32)
    b_iac = allowed[rng.integers(0, len(allowed), n_pert)].astype(np.int32)
    a_q = rng.uniform(-0.5, 0.5, n_pert)
    b_q = rng.uniform(-0.5, 0.5, n_pert)
    pos = np.concatenate([sites, wpos])
    mass = np.concatenate([np.full(n_pert, 15.035), np.tile([M_O, M_H, M_H], n_w)])
    vs = np.sqrt(K_BOLTZMANN * temperature / 15.035) * rng.normal(size=(n_pert, 3)) if temperature > 0 \
        else np.zeros((n_pert, 3))
    vel = np.concatenate([vs, wvel])
    charge = np.concatenate([a_q, np.tile([Q_O, Q_H, Q_H], n_w)])
    iac = np.concatenate([a_iac, np.tile(np.array([t_o, t_h, t_h], np.int32), n_w)]).astype(np.int32)
    excl = [[] for _ in range(n)]
    for w in range(n_w):
        b = n_pert + 3 * w
        excl[b] += [b + 1, b + 2]
        excl[b + 1] += [b, b + 2]
        excl[b + 2] += [b, b + 1]
    excl_off, excl_idx = csr_from_lists(excl, n)
    pert = Perturbation.empty(n)
    pert.flag[:n_pert] = 1
    pert.a_iac[:n_pert], pert.b_iac[:n_pert] = a_iac, b_iac
    pert.a_q[:n_pert], pert.b_q[:n_pert] = a_q, b_q
    pert.alpha_lj[:n_pert], pert.alpha_crf[:n_pert] = alpha_lj, alpha_crf
    pert.lam, pert.nlam = float(lam), 1
    cg_start = np.concatenate([np.arange(n_pert + 1), n_pert + 3 * np.arange(1, n_w + 1)]).astype(np.int32)
    return System(n_atoms=n, n_solute_atoms=n_pert, cg_start=cg_start, n_solute_cg=n_pert, charge=charge, iac=iac,
                  mass=mass, n_types=nt, lj=lj, excl_off=excl_off, excl=excl_idx,
                  o14_off=np.zeros(n + 1, np.int32), o14=np.zeros(0, np.int32), atoms_per_solvent=3, pos=pos,
                  vel=vel, box=np.array([L, L, L]), boundary=1, rcutp=rc, rcutl=rc, nsnb=nsnb, epsrf=61.0,
                  kappa=0.0, fpepsi=FOUR_PI_EPS_I, dt=0.002, dist_oh=D_OH, dist_hh=D_HH, pert=pert,
                  name=name or f"fep_{n_atoms}_l{lam:.2f}").validate()


def fep_mix(n_jobs=64, seed=3, rate_per_step=0.05, lo=1000, hi=30000, n_windows=21):
    """The FEP serving mix: n_jobs jobs with log-uniform atom counts in [lo, hi],
    3-30 perturbed atoms, lambda from n_windows evenly spaced windows, and seeded
    Poisson arrival steps (exponential gaps with the given rate per MD step)."""
    rng = np.random.default_rng(seed)
    sizes = np.exp(rng.uniform(np.log(lo), np.log(hi), n_jobs)).astype(int)
    npert = rng.integers(3, 31, n_jobs)
    lams = rng.integers(0, n_windows, n_jobs) / (n_windows - 1)
    gaps = rng.exponential(1.0 / rate_per_step, n_jobs)
    arrivals = np.floor(np.cumsum(gaps) - gaps[0]).astype(int)
    return [dict(n_atoms=int(sizes[k]), n_pert=int(npert[k]), lam=float(lams[k]), seed=int(seed * 1000 + k),
                 arrival=int(arrivals[k])) for k in range(n_jobs)]
