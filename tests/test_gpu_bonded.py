"""GPU parity of the covalent terms (calculate_bonded_forces_ntf,
crates/gromos-forces/src/bonded/mod.rs:141-185) against the oracle, itself
pinned by the gromosXX aladip / butane / benzene fixtures.

Bars: FP64 covalent forces within 1e-12 of the oracle's per atom (the device
sums an atom's contributions in the reference's order; sqrt, division and the
host-side cos are correctly rounded, the improper's acos may differ by an ulp),
energies within 1e-10, whole force fields and md energies within the FP64 bar
of 1e-10; the mixed precision path keeps the covalent terms in FP64, so its
force error stays at the nonbonded tolerance."""
import importlib
import os

import numpy as np
import pytest

import golden_io
import oracle

gamd = importlib.import_module("gromos-rs_amd")
rt = gamd.runtime
gen = importlib.import_module("gromos-rs_amd.generators")

pytestmark = pytest.mark.gpu

FIXTURES = ["aladip_vacuum", "butane_vacuum", "benzene_vacuum", "aladip_solvated", "aladip_trunc_oct"]
VACUUM = ["aladip_vacuum", "butane_vacuum", "benzene_vacuum"]
KINDS = ("e_bond", "e_angle", "e_dihedral", "e_improper")
E_TOL = 1e-10
F_TOL = 1e-10
MX_F_TOL = 1e-4
MX_E_TOL = 1e-5
HERE = os.path.dirname(os.path.abspath(__file__))


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def force_err(f, ref):
    d = np.linalg.norm(f - ref, axis=1)
    return float((d / np.maximum(np.linalg.norm(ref, axis=1), 1.0)).max())


def chain(seed=7, n_chain=600, n_waters=900, box=3.6):
    return gen.hetero_system(n_chain=n_chain, n_waters=n_waters, box=box, seed=seed, rc=1.0, bonded=True,
                             name=f"chain{seed}")


@pytest.mark.parametrize("name", FIXTURES + ["chain"])
def test_plan_matches_oracle(name):
    import torch
    bd = importlib.import_module("gromos-rs_amd.bonded")
    s = chain() if name == "chain" else golden_io.load(name)[0]
    if s.bonded is None:
        pytest.skip("no covalent terms")
    plan = bd.BondedPlan(s.bonded, s.n_atoms)
    pos = torch.tensor(s.pos, dtype=torch.float64, device="cuda")
    fe = bd.calculate_bonded_forces_ntf(plan, pos)
    ref = oracle.bonded_forces(oracle.BondedOnly(s.pos, s.bonded), s.pos)
    f = fe.forces.cpu().numpy()
    assert np.abs(f - ref["forces"]).max() <= 1e-12 * max(np.abs(ref["forces"]).max(), 1.0)
    for k in KINDS:
        if ref[k] == 0.0:
            assert getattr(fe, k) == 0.0, k
        else:
            assert rel(getattr(fe, k), ref[k]) < E_TOL, (k, getattr(fe, k), ref[k])
    assert np.abs(fe.virial - ref["virial"]).max() <= 1e-10 * max(np.abs(ref["virial"]).max(), 1.0)
    # every kind switched off but the dihedrals: the plan built with those switches
    fd = bd.calculate_bonded_forces_ntf(plan, pos, ntf_bond=False, ntf_angle=False, ntf_improper=False)
    assert fd.e_bond == 0.0 and fd.e_angle == 0.0 and fd.e_improper == 0.0
    assert rel(fd.e_dihedral, ref["e_dihedral"]) < E_TOL or ref["e_dihedral"] == fd.e_dihedral == 0.0
    plan.close()


@pytest.mark.parametrize("name", FIXTURES)
def test_resident_single_point_with_covalent_terms(name):
    s, d = golden_io.load(name)
    ds = rt.DeviceSystem(s)
    ds.pairlist_update()
    g = ds.forcefield_apply()
    ds.close()
    r = oracle.single_point(oracle.OracleSystem(s))
    for k in KINDS + ("e_lj", "e_crf"):
        if r[k] == 0.0:
            assert abs(g[k]) < 1e-12, k
        else:
            assert rel(g[k], r[k]) < E_TOL, (k, g[k], r[k])
    assert force_err(g["forces"], r["forces"]) < F_TOL
    e0 = d["expected_energies"][0]
    for k in KINDS:
        if e0[k] != 0.0:
            assert rel(g[k], e0[k]) < 6e-10, k


@pytest.mark.parametrize("name", VACUUM)
def test_vacuum_md_matches_oracle_and_gromosxx(name):
    s, d = golden_io.load(name)
    ds = rt.DeviceSystem(s)
    ds.pairlist_update()
    e, cov = ds.md(10, covalent=True)
    x, _ = ds.get_state()
    ds.close()
    oe = oracle.MD(oracle.OracleSystem(s), use_settle=False).run(10)
    for k in range(10):
        for c in (0, 1, 2, 3):
            if oe[k, c] != 0.0:
                assert rel(e[k, c], oe[k, c]) < E_TOL, (k, c, e[k, c], oe[k, c])
        assert rel(e[k, 2], d["expected_energies"][k]["e_pot"]) < 6e-10
        assert rel(e[k, 3], d["expected_energies"][k]["e_kin"]) < 6e-10
        fx = d["expected_energies"][k]
        for c, key in enumerate(("e_bond", "e_angle", "e_dihedral", "e_improper")):
            if fx[key] != 0.0:
                assert rel(cov[k, c], fx[key]) < 6e-10, (k, key)
    # e_pot = ((covalent) + e_lj) + e_crf in the reference's order
    eb = ((cov[:, 0] + cov[:, 1]) + cov[:, 2]) + cov[:, 3]
    assert np.array_equal(e[:, 2], (eb + e[:, 0]) + e[:, 1])


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_solvated_chain(prec):
    s = chain()
    o = oracle.OracleSystem(s)
    r = oracle.single_point(o, parallel=True)
    ds = rt.DeviceSystem(s, precision=prec)
    ds.pairlist_update()
    g = ds.forcefield_apply()
    if prec == rt.FP64:
        for k in KINDS + ("e_lj", "e_crf"):
            assert rel(g[k], r[k]) < E_TOL, k
        assert force_err(g["forces"], r["forces"]) < F_TOL
    else:
        for k in KINDS:
            assert rel(g[k], r[k]) < E_TOL, k   # FP64 in both precisions
        assert force_err(g["forces"], r["forces"]) < MX_F_TOL
    ds.close()
    if prec == rt.FP64:
        ds = rt.DeviceSystem(s)
        ds.pairlist_update()
        e = ds.md(12)
        ds.close()
        oe = oracle.MD(o, parallel=True).run(12)
        for k in range(6):
            for c in range(4):
                assert rel(e[k, c], oe[k, c]) < E_TOL, (k, c, e[k, c], oe[k, c])


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_covalent_job_in_a_batch_matches_alone(prec):
    """Jobs with covalent terms batched with plain water give the bits they give alone."""
    specs = [golden_io.load("aladip_solvated")[0], chain(seed=9, n_chain=300, n_waters=700, box=3.2),
             gen.spc_water_box(500, seed=9), golden_io.load("aladip_vacuum")[0]]
    for s in specs:
        s.nsnb = 5
    ref = []
    for s in specs:
        ds = rt.DeviceSystem(s, precision=prec)
        ds.pairlist_update()
        e = ds.md(11)
        ref.append((e, *ds.get_state()))
        ds.close()
    sess = rt.Session(precision=prec, nsnb=5)
    jobs = [sess.admit(s, 11) for s in specs]
    sess.step(11)
    for j, (e, x, v) in zip(jobs, ref):
        assert sess.query(j) == (rt.DONE, 11)
        assert np.array_equal(sess.energies(j), e)
        xx, vv = sess.fetch(j)
        assert np.array_equal(xx, x) and np.array_equal(vv, v)
    sess.close()


def test_run_md_from_files_with_covalent_terms(tmp_path):
    gio = importlib.import_module("gromos-rs_amd.gromos_io")
    d = os.path.join(HERE, "golden", "gromos", "aladip_vacuum")
    fs = {f.rsplit(".", 1)[1]: os.path.join(d, f) for f in os.listdir(d)}
    out = str(tmp_path / "run")
    s, e = gio.run_md(fs["topo"], fs["conf"], fs["in"], out, nsteps=10)
    _, fx = golden_io.load("aladip_vacuum")
    tre = gio.read_energy_trajectory(out + ".tre")
    col = {f: k for k, f in enumerate(gio.ENERTRJ_FIELDS)}
    for k in range(10):
        ref = fx["expected_energies"][k]
        assert rel(e[k, 2], ref["e_pot"]) < 6e-10
        for f, key in (("bond", "e_bond"), ("angle", "e_angle"), ("improper", "e_improper"),
                       ("dihedral", "e_dihedral")):
            assert rel(tre[k, col[f]], ref[key]) < 1e-9, (k, f)
    frames = gio.read_force_trajectory(out + ".trf")
    f0 = np.array(fx["expected_forces"][0]["f"]).reshape(-1, 3)
    assert np.abs(frames[0][1] - f0).max() <= 1e-9
