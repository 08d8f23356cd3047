"""GPU tests of LINCS in the device step (constraints.rs:1168-1445,
md.rs:1244-1257): the solute's constraints (one thread per job after the
integrator, the solute SHAKE kernel's LINCS branch) and every solvent molecule
by the template (a lane per molecule after the integrator) against the oracle
and the gromosXX LINCS frames; LINCS jobs alone and in a batch; the input
files through run_md."""
import importlib
import os

import numpy as np
import pytest

import oracle
import test_oracle_lincs as tl
import test_oracle_shake as ts

gamd = importlib.import_module("gromos-rs_amd")
rt = gamd.runtime
gen = importlib.import_module("gromos-rs_amd.generators")

pytestmark = pytest.mark.gpu


def rel(a, b):
    return np.abs(a - b) / np.maximum(np.abs(b), 1e-300)


def device_md(s, n, prec=rt.FP64):
    ds = rt.DeviceSystem(s, precision=prec)
    ds.pairlist_update()
    e = ds.md(n)
    x, v = ds.get_state()
    ds.close()
    return e, x, v


def oracle_md(s, n):
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    e = md.run(n)
    x, v, _ = md.state()
    return e, x, v


def assert_matches_oracle(s, n):
    e_ref, x_ref, v_ref = oracle_md(s, n)
    e, x, v = device_md(s, n)
    for c in range(4):
        assert rel(e[:, c], e_ref[:, c]).max() < 1e-10, (c, rel(e[:, c], e_ref[:, c]).max())
    assert np.abs(x - x_ref).max() < 1e-10 and np.abs(v - v_ref).max() < 1e-9
    return e, x, v


@pytest.mark.parametrize("name", tl.FIXTURES)
def test_lincs_matches_oracle_and_gromosxx(name):
    s, d = tl.with_lincs(name)
    e, x, v = assert_matches_oracle(s, d["nstlim"])
    for f in d["expected_energies"]:
        k = f["step"]
        assert rel(e[k, 2], f["e_pot"]) < 6e-10 and rel(e[k, 3], f["e_kin"]) < 6e-10, k


@pytest.mark.parametrize("order", [1, 4, 8])
def test_solvent_lincs_in_a_water_box_matches_oracle(order):
    """nacl in SPC water, every molecule by LINCS of the given order"""
    s, d = ts.with_shake("nacl_water_box")
    s.solvent_shake = False
    s.solvent_lincs = order
    e, x, v = assert_matches_oracle(s, 20)
    aps = s.atoms_per_solvent
    tol = 5e-3 if order == 1 else 1e-3   # the truncated expansion's residual (1.4e-3 seen at order 1)
    for m in range(s.nsm):
        base = s.n_solute_atoms + m * aps
        for i, j, l in d["solvent_constr"]:
            assert abs(np.linalg.norm(x[base + int(i)] - x[base + int(j)]) - l) < tol * l


@pytest.mark.parametrize("ntc", [3, 2])
def test_solute_and_solvent_lincs_match_oracle(ntc):
    """aladip in 20 waters: the solute's bonds and the solvent by LINCS"""
    s, d = ts.with_shake("aladip_solvated")
    s.solvent_shake = False
    s.solvent_lincs = 4
    s.solute_constraints = tl.solute_bonds(s, d, ntc)
    s.solute_lincs = 4
    assert_matches_oracle(s, 30)


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_lincs_jobs_in_a_batch_match_alone(prec):
    a, _ = tl.with_lincs("nacl_1water_lincs")
    b, db = ts.with_shake("aladip_solvated")
    b.solvent_shake, b.solvent_lincs = False, 4
    b.solute_constraints, b.solute_lincs = tl.solute_bonds(b, db, 3), 4
    a.nsnb = b.nsnb = 5
    c = gen.spc_water_box(300, seed=9)
    refs = [device_md(q, n, prec) for q, n in ((a, 12), (b, 15), (c, 8))]
    sess = rt.Session(precision=prec, nsnb=5)
    ja = sess.admit(a, 12)
    sess.step(3)
    jb = sess.admit(b, 15)
    jc = sess.admit(c, 8)
    k = 0
    while not all(sess.query(j)[0] == rt.DONE for j in (ja, jb, jc)):
        sess.step(2)
        k += 1
        assert k < 40
    for j, (e0, x0, v0) in zip((ja, jb, jc), refs):
        assert np.array_equal(sess.energies(j), e0)
        x, v = sess.fetch(j)
        assert np.array_equal(x, x0) and np.array_equal(v, v0)
    sess.close()


@pytest.mark.parametrize("name", tl.FIXTURES)
def test_run_md_from_files_with_lincs(name, tmp_path):
    """the gromosXX input files (NTCP or NTCS lincs) through the readers, the
    device md and the .tre writer: every written frame matches gromosXX"""
    gio = importlib.import_module("gromos-rs_amd.gromos_io")
    d = os.path.join(os.path.dirname(__file__), "golden", "gromos", name)
    fs = {f.rsplit(".", 1)[1]: os.path.join(d, f) for f in os.listdir(d)}
    out = str(tmp_path / "run")
    s, e = gio.run_md(fs["topo"], fs["conf"], fs["in"], out)
    assert s.solute_lincs == 4 or s.solvent_lincs == 4
    _, fx = tl.with_lincs(name)
    tre = gio.read_energy_trajectory(out + ".tre")
    col = {f: k for k, f in enumerate(gio.ENERTRJ_FIELDS)}
    for k, ref in enumerate(fx["expected_energies"]):
        assert rel(tre[k, col["kinetic"]], ref["e_kin"]) < 6e-10, k
        assert rel(tre[k, col["potential"]], ref["e_pot"]) < 6e-10, k
