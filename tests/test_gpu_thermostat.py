"""GPU tests of the bath in the device step (md.rs:1143-1206): Berendsen weak
coupling and the Nose-Hoover thermostat, single and chained
(nosehoover_thermostat.rs), against the oracle and the gromosXX frames; a
Nose-Hoover chain's state carried through session layouts."""
import importlib
import os

import numpy as np
import pytest

import oracle
import test_oracle_thermostat as tt

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


@pytest.mark.parametrize("name", tt.FIXTURES)
def test_bath_md_matches_oracle_and_gromosxx(name):
    s, d = tt.with_bath(name)
    n = d["nstlim"]
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    e_ref = md.run(n)
    x_ref, v_ref, _ = md.state()
    print(f"{name}: {n} device steps", flush=True)
    e, x, v = device_md(s, n)
    for c in range(4):   # e_lj, e_crf, e_pot, e_kin
        assert rel(e[:, c], e_ref[:, c]).max() < 1e-10, (c, rel(e[:, c], e_ref[:, c]).max())
    assert np.abs(x - x_ref).max() < 1e-10
    assert np.abs(v - v_ref).max() < 1e-9
    for f in d["expected_energies"]:
        k = f["step"]
        assert rel(e[k, 2], f["e_pot"]) < 6e-10 and rel(e[k, 3], f["e_kin"]) < 6e-10, (k, e[k, 2:4])


def test_bath_kinds_differ_on_the_device():
    """a chain of three, a single Nose-Hoover zeta and Berendsen give different
    velocities from the same start: the device honours the kind and length"""
    s, _ = tt.with_bath("water_216_nvt_nhc_chain")
    out = []
    for nhc in (3, 1, 0):
        s.thermostat_nhc = nhc
        out.append(device_md(s, 12)[2])
    assert not np.array_equal(out[0], out[1]) and not np.array_equal(out[1], out[2])


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_chain_state_survives_layouts(prec):
    """a Nose-Hoover chain job that sees other jobs join and leave (new layouts)
    ends bit-identical to the same job alone"""
    s, _ = tt.with_bath("water_216_nvt_nhc_chain")
    b = gen.spc_water_box(300, seed=9)
    c = gen.spc_water_box(500, seed=4)
    c.thermostat, c.thermostat_nhc = (320.0, 0.2, 3.0 * c.n_atoms - 3 * c.nsm), 2
    e0, x0, v0 = device_md(s, 30, prec)
    ec, xc, vc = device_md(c, 12, prec)
    sess = rt.Session(precision=prec, nsnb=5)
    ja = sess.admit(s, 30)
    sess.step(7)
    jb = sess.admit(b, 6)
    sess.step(4)
    jc = sess.admit(c, 12)
    k = 0
    while not all(sess.query(j)[0] == rt.DONE for j in (ja, jb, jc)):
        sess.step(3)
        k += 1
        assert k < 40
    assert sess.info()["layouts"] >= 4
    assert np.array_equal(sess.energies(ja), e0)
    x, v = sess.fetch(ja)
    assert np.array_equal(x, x0) and np.array_equal(v, v0)
    assert np.array_equal(sess.energies(jc), ec)
    assert np.array_equal(sess.fetch(jc)[1], vc)
    sess.close()


@pytest.mark.parametrize("name", ["water_216_nvt_nosehoover", "water_216_nvt_nhc_chain"])
def test_run_md_from_files_with_nose_hoover(name, tmp_path):
    """the gromosXX input files through the readers, the device md and the
    .tre writer: every written frame matches the gromosXX energies"""
    gio = importlib.import_module("gromos-rs_amd.gromos_io")
    d = os.path.join(os.path.dirname(__file__), "golden", "gromos", name)
    fs = {f.rsplit(".", 1)[1]: os.path.join(d, f) for f in os.listdir(d)}
    out = str(tmp_path / "run")
    s, e = gio.run_md(fs["topo"], fs["conf"], fs["in"], out)
    assert s.thermostat_nhc == (3 if "chain" in name else 1)
    _, fx = tt.with_bath(name)
    tre = gio.read_energy_trajectory(out + ".tre")
    col = {f: k for k, f in enumerate(gio.ENERTRJ_FIELDS)}
    for k, ref in enumerate(fx["expected_energies"]):
        assert rel(tre[k, col["kinetic"]], ref["e_kin"]) < 1e-9, (k, tre[k, col["kinetic"]], ref["e_kin"])
        assert rel(tre[k, col["total"]], ref["e_tot"]) < 1e-9, k
