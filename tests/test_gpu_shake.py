"""GPU tests of SHAKE in the device step (constraints.rs:213-570, md.rs:1208-1236):
solvent molecules constrained by the template (in the integrator's lane per
molecule) against the oracle and the gromosXX frames; the solute's constraints
(one sweep per job after the integrator) against the oracle; SHAKE jobs alone
and in a batch."""
import importlib
import os

import numpy as np
import pytest

import oracle
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


def solute_constraints(s, d, ntc=3):
    """the solute bonds as constraints (md.rs:1208-1236; NTC 3: all bonds, 2:
    those with a hydrogen, mass < 2), the topology's order"""
    b = d["bonded"]
    out = []
    for i, j, t in b["bonds"]:
        r0 = b["bond_types"][t][2]
        if r0 < 1e-10 or (ntc == 2 and s.mass[i] > 2.0 and s.mass[j] > 2.0):
            continue
        out.append((i, j, r0))
    return np.array(out, float)


@pytest.mark.parametrize("name", ts.FIXTURES)
def test_solvent_shake_matches_oracle_and_gromosxx(name):
    s, d = ts.with_shake(name)
    n = d["nstlim"]
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    e_ref = md.run(n)
    x_ref, v_ref, _ = md.state()
    e, x, v = device_md(s, n)
    for c in range(4):
        assert rel(e[:, c], e_ref[:, c]).max() < 1e-10, (c, rel(e[:, c], e_ref[:, c]).max())
    assert np.abs(x - x_ref).max() < 1e-10 and np.abs(v - v_ref).max() < 1e-9
    for f in d["expected_energies"]:
        k = f["step"]
        assert rel(e[k, 2], f["e_pot"]) < 6e-10 and rel(e[k, 3], f["e_kin"]) < 6e-10, k


@pytest.mark.parametrize("ntc", [3, 2])
def test_solute_and_solvent_shake_match_oracle(ntc):
    """aladip in 20 SHAKE waters, the solute's bonds constrained too"""
    s, d = ts.with_shake("aladip_solvated")
    s.solute_constraints = solute_constraints(s, d, ntc)
    assert len(s.solute_constraints) > 0
    n = 30
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    e_ref = md.run(n)
    x_ref, v_ref, _ = md.state()
    e, x, v = device_md(s, n)
    for c in range(4):
        assert rel(e[:, c], e_ref[:, c]).max() < 1e-10, (c, rel(e[:, c], e_ref[:, c]).max())
    assert np.abs(x - x_ref).max() < 1e-10 and np.abs(v - v_ref).max() < 1e-9
    for i, j, l in s.solute_constraints:
        r = np.linalg.norm(x[int(i)] - x[int(j)])
        assert abs(r * r - l * l) < 2.0 * s.shake_tolerance * l * l


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_shake_jobs_in_a_batch_match_alone(prec):
    a, da = ts.with_shake("water_10_box")
    b, db = ts.with_shake("aladip_solvated")
    b.solute_constraints = solute_constraints(b, db)
    a.nsnb = b.nsnb = 5   # one pair-list cadence per session
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


@pytest.mark.parametrize("name", ["water_10_box", "nacl_water_box"])
def test_run_md_from_files_with_solvent_shake(name, tmp_path):
    """the gromosXX input files (NTCS shake) through the readers, the device md
    and the .tre writer: every written frame matches the gromosXX energies"""
    gio = importlib.import_module("gromos-rs_amd.gromos_io")
    d = os.path.join(os.path.dirname(__file__), "golden", "gromos", name)
    fs = {f.rsplit(".", 1)[1]: os.path.join(d, f) for f in os.listdir(d)}
    out = str(tmp_path / "run")
    s, e = gio.run_md(fs["topo"], fs["conf"], fs["in"], out)
    assert s.solvent_shake
    _, fx = ts.with_shake(name)
    tre = gio.read_energy_trajectory(out + ".tre")
    col = {f: k for k, f in enumerate(gio.ENERTRJ_FIELDS)}
    for k, ref in enumerate(fx["expected_energies"]):
        assert rel(tre[k, col["kinetic"]], ref["e_kin"]) < 6e-10, k
        assert rel(tre[k, col["potential"]], ref["e_pot"]) < 6e-10, k


def test_run_md_from_files_with_solute_shake(tmp_path):
    """aladip in SHAKE water with NTC 3 from the files: run_md's constraint
    selection and device run equal the oracle's md over the fixture's bonds"""
    gio = importlib.import_module("gromos-rs_amd.gromos_io")
    d = os.path.join(os.path.dirname(__file__), "golden", "gromos", "aladip_solvated")
    fs = {f.rsplit(".", 1)[1]: os.path.join(d, f) for f in os.listdir(d)}
    text = open(fs["in"]).read().split("\n")
    k = next(i for i, l in enumerate(text) if l.startswith("CONSTRAINT"))
    k = next(i for i in range(k + 1, len(text)) if not text[i].lstrip().startswith("#"))
    assert text[k].split()[0] == "1"
    text[k] = "3"
    imd = str(tmp_path / "ntc3.in")
    open(imd, "w").write("\n".join(text))
    n = 20
    s, e = gio.run_md(fs["topo"], fs["conf"], imd, str(tmp_path / "run"), nsteps=n)
    assert len(s.solute_constraints) == 11
    r, dd = ts.with_shake("aladip_solvated")
    r.solute_constraints = solute_constraints(r, dd, 3)
    assert (s.nticom, s.nscm) == (1, 10)           # the input's centre-of-mass removal
    r.nticom, r.nscm = s.nticom, s.nscm
    md = oracle.MD(oracle.OracleSystem(r), use_settle=False)
    e_ref = md.run(n)
    for c in range(4):
        assert rel(e[:, c], e_ref[:, c]).max() < 1e-10, (c, rel(e[:, c], e_ref[:, c]).max())
