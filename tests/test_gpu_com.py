"""GPU tests of the centre-of-mass motion removal at the start of the device
step (RemoveCOMMotion, remove_com_motion.rs; md.rs:1095-1097): the gromosXX
water_216_box_com / _com_rot runs against the oracle and the fixture frames,
and a job's removal steps counted from its own step 0 in a session."""
import importlib
import os

import numpy as np
import pytest

import oracle
import test_oracle_com as tc

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


@pytest.mark.parametrize("name", tc.FIXTURES)
def test_com_md_matches_oracle_and_gromosxx(name):
    s, d = tc.with_com(name)
    n = 91
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    e_ref = md.run(n)
    x_ref, v_ref, _ = md.state()
    print(f"{name}: {n} device steps", flush=True)
    e, x, v = device_md(s, n)
    for c in range(4):
        assert rel(e[:, c], e_ref[:, c]).max() < 1e-10, (c, rel(e[:, c], e_ref[:, c]).max())
    assert np.abs(x - x_ref).max() < 1e-10 and np.abs(v - v_ref).max() < 1e-9
    for f in d["expected_energies"]:
        k = f["step"]
        assert rel(e[k, 2], f["e_pot"]) < 6e-10 and rel(e[k, 3], f["e_kin"]) < 6e-10, k


def test_device_removal_changes_the_run():
    s, _ = tc.with_com("water_216_box_com_rot")
    v_rot = device_md(s, 3)[2]
    s.nticom = 1
    v_tr = device_md(s, 3)[2]
    s.nticom, s.nscm = 0, 0
    v_none = device_md(s, 3)[2]
    assert not np.array_equal(v_rot, v_tr) and not np.array_equal(v_tr, v_none)


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_removal_steps_follow_each_jobs_own_clock(prec):
    """jobs admitted at different session clocks remove their centre-of-mass
    motion at their own step 0 and every NSCM of their own steps: each ends
    bit-identical to the same job alone"""
    a = gen.spc_water_box(400, seed=2)
    a.nticom, a.nscm = 1, 3
    b = gen.spc_water_box(300, seed=6)
    b.vel = b.vel + np.array([0.2, -0.1, 0.05])      # a drifting box: removal visible
    b.nticom, b.nscm = 2, -4
    c = gen.water_216(seed=3)
    refs = [device_md(q, n, prec) for q, n in ((a, 20), (b, 14), (c, 9))]
    sess = rt.Session(precision=prec, nsnb=5)
    ja = sess.admit(a, 20)
    sess.step(4)
    jb = sess.admit(b, 14)
    sess.step(3)
    jc = sess.admit(c, 9)
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
    # and the removal happened: job b's momentum (1e3 at the start) is gone; the
    # FP64 forces conserve it, the mixed path's per-side sums nearly so
    e, x, v = refs[1]
    assert np.abs((b.mass[:, None] * v).sum(0)).max() < (1e-9 if prec == rt.FP64 else 1e-2)


def test_run_md_from_files_with_com_removal(tmp_path):
    gio = importlib.import_module("gromos-rs_amd.gromos_io")
    d = os.path.join(os.path.dirname(__file__), "golden", "gromos", "water_216_box_com_rot")
    fs = {f.rsplit(".", 1)[1]: os.path.join(d, f) for f in os.listdir(d)}
    out = str(tmp_path / "run")
    s, e = gio.run_md(fs["topo"], fs["conf"], fs["in"], out, nsteps=31)
    assert (s.nticom, s.nscm) == (2, -10)
    _, fx = tc.with_com("water_216_box_com_rot")
    for ref in fx["expected_energies"][:4]:
        k = ref["step"]
        assert rel(e[k, 3], ref["e_kin"]) < 6e-10 and rel(e[k, 2], ref["e_pot"]) < 6e-10, k
