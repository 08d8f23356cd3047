"""Pinning of the oracle's covalent terms (calculate_bonded_forces_ntf,
crates/gromos-forces/src/bonded/mod.rs:141-185, restated in
oracle/gromos_oracle.c).

  * gromosXX fixtures: step-0 bond / angle / improper / dihedral energies and
    full forces of aladip (vacuum, solvated, truncated octahedron), butane and
    benzene (NTF switched off) at the files' printed precision, and the first 10
    md steps of the three vacuum systems (positions, forces, E_pot, E_kin);
  * the mathematics: each kind's forces are minus the gradient of its energy
    (central differences), sum to zero, leave the energy unchanged under rigid
    motion, give the virial sum_i x_i f_i^T; the energies equal textbook forms
    computed another way (the bond length, the angle by arccos, the dihedral by
    atan2 for every multiplicity 0-6, the improper by the signed plane angle);
  * the NTF switches and out-of-range types drop exactly their terms."""
import importlib

import numpy as np
import pytest

import golden_io
import oracle

gamd = importlib.import_module("gromos-rs_amd")
Bonded = gamd.system.Bonded

FIXTURES = ["aladip_vacuum", "butane_vacuum", "benzene_vacuum", "aladip_solvated", "aladip_trunc_oct"]
VACUUM = ["aladip_vacuum", "butane_vacuum", "benzene_vacuum"]
KINDS = ("e_bond", "e_angle", "e_improper", "e_dihedral")


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


@pytest.mark.parametrize("name", FIXTURES)
def test_fixture_step0_energies_and_forces(name):
    s, d = golden_io.load(name)
    r = oracle.single_point(oracle.OracleSystem(s))
    e0 = d["expected_energies"][0]
    for k in KINDS + ("e_lj", "e_crf"):
        if e0[k] == 0.0:
            assert r[k] == 0.0, k
        else:
            assert rel(r[k], e0[k]) < 6e-10, (k, r[k], e0[k])
    f = np.array(d["expected_forces"][0]["f"]).reshape(-1, 3)
    if s.boundary == 2:
        # the fixture prints forces in the original frame; the run is in the
        # rotated one (NTB = -1 conversion)
        gio = importlib.import_module("gromos-rs_amd.gromos_io")
        f = gio.truncoct_triclinic(1.0, f)[1]
        assert np.abs(r["forces"] - f).max() < 1e-9
    else:
        assert np.abs(r["forces"] - f).max() <= 5.01e-10


@pytest.mark.parametrize("name", VACUUM)
def test_vacuum_md_follows_gromosxx(name):
    s, d = golden_io.load(name)
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    for k in range(10):
        x, _, _ = md.state()
        assert np.abs(x - np.array(d["expected_positions"][k]["x"]).reshape(-1, 3)).max() <= 5.01e-10, k
        e = md.run(1)[0]
        _, _, f = md.state()
        assert np.abs(f - np.array(d["expected_forces"][k]["f"]).reshape(-1, 3)).max() <= 5.01e-10, k
        assert rel(e[2], d["expected_energies"][k]["e_pot"]) < 6e-10
        assert rel(e[3], d["expected_energies"][k]["e_kin"]) < 6e-10


# ---------------------------------------------------------------------------
# the mathematics
# ---------------------------------------------------------------------------

def one_kind(kind, rng, n_terms=6, mult=None):
    """Random terms of one kind among 12 atoms (types one per term)."""
    z3, z4, z5 = np.zeros((0, 3)), np.zeros((0, 4)), np.zeros((0, 5))
    idx = np.array([rng.choice(12, {"bond": 2, "angle": 3}.get(kind, 4), replace=False) for _ in range(n_terms)])
    t = np.arange(n_terms)[:, None]
    if kind == "bond":
        ty = np.stack([rng.uniform(1e6, 1e7, n_terms), np.zeros(n_terms), rng.uniform(0.1, 0.3, n_terms)], 1)
        return Bonded.of(np.hstack([idx, t]), ty, z4, z3, z5, np.zeros((0, 2)), z5, z3)
    if kind == "angle":
        ty = np.stack([rng.uniform(100, 600, n_terms), np.zeros(n_terms), rng.uniform(1.5, 2.5, n_terms)], 1)
        return Bonded.of(np.zeros((0, 3)), z3, np.hstack([idx, t]), ty, z5, np.zeros((0, 2)), z5, z3)
    if kind == "dihedral":
        m = rng.integers(0, 7, n_terms) if mult is None else np.full(n_terms, mult)
        ty = np.stack([rng.uniform(1, 6, n_terms), rng.choice([-1.0, 1.0], n_terms), m.astype(float)], 1)
        return Bonded.of(np.zeros((0, 3)), z3, z4, z3, z5, np.zeros((0, 2)), np.hstack([idx, t]), ty)
    ty = np.stack([rng.uniform(50, 300, n_terms), rng.uniform(-1.0, 1.0, n_terms)], 1)
    return Bonded.of(np.zeros((0, 3)), z3, z4, z3, np.hstack([idx, t]), ty, z5, z3)


def bforces(b, x):
    return oracle.bonded_forces(oracle.BondedOnly(x, b), x)


def total(r):
    return sum(r[k] for k in KINDS)


@pytest.mark.parametrize("kind", ["bond", "angle", "dihedral", "improper"])
def test_forces_are_minus_the_energy_gradient(kind):
    rng = np.random.default_rng({"bond": 1, "angle": 2, "dihedral": 3, "improper": 4}[kind])
    x = rng.uniform(0, 0.8, (12, 3))
    b = one_kind(kind, rng)
    r = bforces(b, x)
    h = 1e-6
    g = np.zeros_like(x)
    for i in range(12):
        for a in range(3):
            xp, xm = x.copy(), x.copy()
            xp[i, a] += h
            xm[i, a] -= h
            g[i, a] = (total(bforces(b, xp)) - total(bforces(b, xm))) / (2 * h)
    scale = max(np.abs(r["forces"]).max(), 1.0)
    assert np.abs(r["forces"] + g).max() < 1e-5 * scale
    # no net force, energy unchanged by a rigid rotation + translation
    assert np.abs(r["forces"].sum(0)).max() < 1e-9 * scale
    q, _ = np.linalg.qr(rng.normal(size=(3, 3)))
    y = x @ q.T + rng.normal(size=3)
    assert abs(total(bforces(b, y)) - total(r)) < 1e-9 * max(abs(total(r)), 1.0)
    # virial = sum_i x_i f_i^T (the terms' forces sum to zero, so relative
    # vectors and absolute positions give the same tensor)
    v = x.T @ r["forces"]
    assert np.abs(r["virial"] - v).max() < 1e-9 * max(np.abs(v).max(), 1.0)


def test_bond_and_angle_energies_closed_form():
    rng = np.random.default_rng(5)
    x = rng.uniform(0, 0.8, (12, 3))
    b = one_kind("bond", rng)
    e = sum(0.25 * t[0] * (np.sum((x[i] - x[j]) ** 2) - t[2] ** 2) ** 2
            for (i, j, k), t in zip(b.bonds, b.bond_types))
    assert rel(bforces(b, x)["e_bond"], e) < 1e-12
    b = one_kind("angle", rng)
    e = 0.0
    for (i, j, k, ty), t in zip(b.angles, b.angle_types):
        u, v = x[i] - x[j], x[k] - x[j]
        th = np.arccos(u @ v / np.linalg.norm(u) / np.linalg.norm(v))
        e += 0.5 * t[0] * (np.cos(th) - np.cos(t[2])) ** 2
    assert rel(bforces(b, x)["e_angle"], e) < 1e-10


def phi_atan2(a, b, c, d):
    """IUPAC dihedral angle by atan2 (independent of the Chebyshev path):
    atan2(b2^ . (n1 x n2), n1 . n2), n1 = b1 x b2, n2 = b2 x b3"""
    b1, b2, b3 = b - a, c - b, d - c
    n1, n2 = np.cross(b1, b2), np.cross(b2, b3)
    m = np.cross(b2 / np.linalg.norm(b2), n1)
    return np.arctan2(m @ n2, n1 @ n2)


@pytest.mark.parametrize("m", range(7))
def test_dihedral_energy_is_k_1_plus_cosd_cos_m_phi(m):
    rng = np.random.default_rng(10 + m)
    x = rng.uniform(0, 0.8, (12, 3))
    b = one_kind("dihedral", rng, mult=m)
    e = 0.0
    for (i, j, k, l, ty) in b.dihedrals:
        K, cd, _ = b.dihedral_types[ty]
        # cos(m phi) for m = 0 is taken as 0 by the reference's table (energy K)
        cm = np.cos(m * phi_atan2(x[i], x[j], x[k], x[l])) if m > 0 else 0.0
        e += K * (1.0 + cd * cm)
    assert rel(bforces(b, x)["e_dihedral"], e) < 1e-9


def test_improper_energy_from_signed_plane_angle():
    rng = np.random.default_rng(20)
    x = rng.uniform(0, 0.8, (12, 3))
    b = one_kind("improper", rng)
    e = 0.0
    for (i, j, k, l, ty) in b.impropers:
        K, q0 = b.improper_types[ty]
        # the improper's zeta is the IUPAC angle of (i, j, k, l) (same planes,
        # sign of rij . (rkj x rkl))
        z = phi_atan2(x[i], x[j], x[k], x[l])
        dz = (z - q0 + np.pi) % (2 * np.pi) - np.pi
        e += 0.5 * K * dz * dz
    assert rel(bforces(b, x)["e_improper"], e) < 1e-9


def test_switches_and_out_of_range_types_drop_their_terms():
    rng = np.random.default_rng(30)
    x = rng.uniform(0, 0.8, (12, 3))
    parts = {k: one_kind(k, rng) for k in ("bond", "angle", "dihedral", "improper")}
    b = Bonded.of(parts["bond"].bonds, parts["bond"].bond_types, parts["angle"].angles, parts["angle"].angle_types,
                  parts["improper"].impropers, parts["improper"].improper_types, parts["dihedral"].dihedrals,
                  parts["dihedral"].dihedral_types)
    full = bforces(b, x)
    alone = {k: bforces(v, x) for k, v in parts.items()}
    for k, name in zip(("bond", "angle", "dihedral", "improper"), ("e_bond", "e_angle", "e_dihedral", "e_improper")):
        assert full[name] == alone[k][name]
    # NTF order: bonds, angles, impropers, dihedrals; the rest stays bit-equal
    # to a sum in the reference's kind order
    for ntf, kept in (((1, 0, 1, 1), "angle"), ((1, 1, 0, 1), "improper"), ((1, 1, 1, 0), "dihedral"),
                      ((0, 1, 1, 1), "bond")):
        off = bforces(Bonded.of(b.bonds, b.bond_types, b.angles, b.angle_types, b.impropers, b.improper_types,
                                b.dihedrals, b.dihedral_types, ntf), x)
        assert off["e_" + kept] == 0.0
        ref = np.zeros_like(x)
        for k in ("bond", "angle", "dihedral", "improper"):
            if k != kept:
                ref = ref + alone[k]["forces"]
        assert np.array_equal(off["forces"], ref)
    # a type beyond the table: the term is skipped
    bb = parts["bond"]
    cut = Bonded.of(bb.bonds, bb.bond_types[:3], b.angles[:0], b.angle_types, b.impropers[:0], b.improper_types,
                    b.dihedrals[:0], b.dihedral_types)
    kept_terms = Bonded.of(bb.bonds[bb.bonds[:, 2] < 3], bb.bond_types[:3], b.angles[:0], b.angle_types,
                           b.impropers[:0], b.improper_types, b.dihedrals[:0], b.dihedral_types)
    assert bforces(cut, x)["e_bond"] == bforces(kept_terms, x)["e_bond"]


def test_oracle_catches_a_planted_sign_error():
    """A wrong sign in one force component is caught by the gradient check."""
    rng = np.random.default_rng(40)
    x = rng.uniform(0, 0.8, (12, 3))
    b = one_kind("angle", rng)
    r = bforces(b, x)
    bad = r["forces"].copy()
    bad[:, 1] *= -1.0
    h = 1e-6
    xp, xm = x.copy(), x.copy()
    i = int(b.angles[0, 0])
    xp[i, 1] += h
    xm[i, 1] -= h
    g = (total(bforces(b, xp)) - total(bforces(b, xm))) / (2 * h)
    assert abs(r["forces"][i, 1] + g) < 1e-5 * max(abs(g), 1.0)
    assert abs(bad[i, 1] + g) > 1e-3 * max(abs(g), 1.0)
