"""GPU tests of pressure coupling in the device step (md.rs:1116-1129,
1259-1277): the kinetic tensor, the molecular virial, the pressure and the
isotropic Berendsen barostat against the oracle and the gromosXX water_216_npt
frames; a barostat job's box carried through session layouts."""
import importlib
import os

import numpy as np
import pytest

import oracle
import test_oracle_npt as tn

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
    box, p = ds.box()
    ds.close()
    return e, x, v, box, p


@pytest.mark.parametrize("vtype", [2, 1])
def test_npt_md_matches_oracle(vtype):
    s, d = tn.with_npt(vtype=vtype)
    n = d["nstlim"]
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    e_ref = md.run(n)
    x_ref, v_ref, _ = md.state()
    print(f"water_216_npt virial {vtype}: {n} device steps", flush=True)
    e, x, v, box, p = device_md(s, n)
    for c in range(4):
        assert rel(e[:, c], e_ref[:, c]).max() < 1e-10, (c, rel(e[:, c], e_ref[:, c]).max())
    assert np.abs(x - x_ref).max() < 1e-10 and np.abs(v - v_ref).max() < 1e-9
    assert rel(box, md.box()).max() < 1e-13
    assert rel(p, md.pressure()) < 1e-9, (p, md.pressure())
    if vtype == 2:
        for f in d["expected_energies"]:
            k = f["step"]
            assert rel(e[k, 2], f["e_pot"]) < 6e-10 and rel(e[k, 3], f["e_kin"]) < 6e-10, k
        assert np.abs(box - np.array(d["final_box"])).max() <= 5.01e-10


@pytest.mark.parametrize("vtype", [2, 1])
def test_npt_settle_water_matches_oracle(vtype):
    """rigid SETTLE waters under the barostat: SETTLE's constraint virial joins
    the pressure (constraints.rs:978-1010); each molecule its own pressure group"""
    w = gen.spc_water_box(400, seed=12)
    w.barostat = (0.06102, 4.575e-4, 0.5, vtype)
    n = 15
    md = oracle.MD(oracle.OracleSystem(w), use_settle=True)
    e_ref = md.run(n)
    x_ref, _, _ = md.state()
    e, x, v, box, p = device_md(w, n)
    for c in range(4):
        assert rel(e[:, c], e_ref[:, c]).max() < 1e-10, (c, rel(e[:, c], e_ref[:, c]).max())
    assert np.abs(x - x_ref).max() < 1e-10
    assert rel(box, md.box()).max() < 1e-13 and rel(p, md.pressure()) < 1e-9, (p, md.pressure())
    assert box[0] != w.box[0]


@pytest.mark.parametrize("method", ["shake", "lincs"])
def test_npt_constrained_water_matches_oracle(method):
    """waters constrained by SHAKE (its constraint virial joins the pressure,
    constraints.rs:213-389) or by LINCS (no constraint virial, as the
    reference's) under the barostat, molecular virial"""
    w = gen.spc_water_box(400, seed=12)
    w.barostat = (0.06102, 4.575e-4, 0.5, 2)
    hh = 2.0 * 0.1 * np.sin(np.radians(109.47) / 2.0)
    w.solvent_constraints = np.array([[0, 1, 0.1], [0, 2, 0.1], [1, 2, hh]])
    if method == "shake":
        w.solvent_shake = True
    else:
        w.solvent_lincs = 4
    n = 15
    md = oracle.MD(oracle.OracleSystem(w), use_settle=False)
    e_ref = md.run(n)
    x_ref, _, _ = md.state()
    e, x, v, box, p = device_md(w, n)
    for c in range(4):
        assert rel(e[:, c], e_ref[:, c]).max() < 1e-10, (c, rel(e[:, c], e_ref[:, c]).max())
    assert np.abs(x - x_ref).max() < 1e-10
    assert rel(box, md.box()).max() < 1e-13 and rel(p, md.pressure()) < 1e-9, (p, md.pressure())
    assert box[0] != w.box[0]


@pytest.mark.parametrize("vtype", [2, 1])
def test_npt_truncated_octahedron_matches_oracle(vtype):
    """a truncated-octahedron SETTLE water box under the barostat: every box
    vector scaled (berendsen_barostat.rs:40-48), the volume the box matrix's
    determinant (configuration.rs:86-89)"""
    w = gen.water_truncoct(5, seed=3)
    w.barostat = (0.06102, 4.575e-4, 0.5, vtype)
    n = 15
    md = oracle.MD(oracle.OracleSystem(w), use_settle=True)
    e_ref = md.run(n)
    x_ref, _, _ = md.state()
    e, x, v, box, p = device_md(w, n)
    for c in range(4):
        assert rel(e[:, c], e_ref[:, c]).max() < 1e-10, (c, rel(e[:, c], e_ref[:, c]).max())
    assert np.abs(x - x_ref).max() < 1e-10
    assert rel(box, md.box()).max() < 1e-13 and rel(p, md.pressure()) < 1e-9, (p, md.pressure())
    assert box[0] != np.asarray(w.box_matrix, float).reshape(3, 3)[0, 0]


def test_npt_mixed_precision_follows_fp64():
    s, d = tn.with_npt()
    _, _, _, b64, p64 = device_md(s, 20)
    e, _, _, bmx, pmx = device_md(s, 20, rt.MIXED)
    assert np.isfinite(e).all()
    assert rel(bmx, b64).max() < 1e-6 and abs(pmx - p64) < 1e-2 * max(abs(p64), 1.0)


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_barostat_box_survives_layouts(prec):
    """an NPT job whose batch changes around it ends bit-identical to the same
    job alone, box and pressure included"""
    s, _ = tn.with_npt()
    b = gen.spc_water_box(300, seed=9)
    w = gen.spc_water_box(400, seed=12)
    w.barostat = (0.06102, 4.575e-4, 0.5, 2)        # solvent only: each molecule a group
    ref_s = device_md(s, 30, prec)
    ref_w = device_md(w, 12, prec)
    sess = rt.Session(precision=prec, nsnb=5)
    js = sess.admit(s, 30)
    sess.step(6)
    jb = sess.admit(b, 7)
    sess.step(4)
    jw = sess.admit(w, 12)
    k = 0
    while not all(sess.query(j)[0] == rt.DONE for j in (js, jb, jw)):
        sess.step(3)
        k += 1
        assert k < 40
    assert sess.info()["layouts"] >= 4
    for j, (e0, x0, v0, box0, p0) in ((js, ref_s), (jw, ref_w)):
        assert np.array_equal(sess.energies(j), e0)
        x, v = sess.fetch(j)
        assert np.array_equal(x, x0) and np.array_equal(v, v0)
        box, p = sess.box(j)
        assert np.array_equal(box, box0) and p == p0
    assert not np.array_equal(ref_w[3], w.box)
    sess.close()


def test_run_md_from_files_with_barostat(tmp_path):
    gio = importlib.import_module("gromos-rs_amd.gromos_io")
    d = os.path.join(os.path.dirname(__file__), "golden", "gromos", "water_216_npt")
    fs = {f.rsplit(".", 1)[1]: os.path.join(d, f) for f in os.listdir(d)}
    out = str(tmp_path / "run")
    s, e = gio.run_md(fs["topo"], fs["conf"], fs["in"], out)
    assert s.barostat is not None and s.barostat[3] == 2
    _, fx = tn.with_npt()
    for ref in fx["expected_energies"]:
        k = ref["step"]
        assert rel(e[k, 3], ref["e_kin"]) < 6e-10 and rel(e[k, 2], ref["e_pot"]) < 6e-10, k
    final = gio.read_configuration(out + ".cnf")
    assert np.abs(np.asarray(final.box)[:3] - np.array(fx["final_box"])).max() <= 5.01e-10
    tre = gio.read_energy_trajectory(out + ".tre")
    col = {f: k for k, f in enumerate(gio.ENERTRJ_FIELDS)}
    vols = [np.prod(p["box"]) for p in fx["expected_positions"]]
    assert rel(tre[:len(vols), col["volume"]], np.array(vols)).max() < 1e-8
