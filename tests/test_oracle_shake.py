"""Pinning of the oracle's SHAKE (constraints.rs:213-570, md.rs:1208-1236),
restated in oracle/gromos_oracle.c.

  * gromosXX fixtures whose solvent gromosXX constrains by SHAKE (NTCS shake,
    tolerance 1e-4): water_3_box, water_10_box, nacl_3water_box, nacl_water_box
    (every step of 10: positions, forces, E_pot, E_kin); the frames reject
    SETTLE in its place;
  * the mathematics: every constraint within the tolerance after the step,
    the total momentum conserved by the constraint forces, velocities the
    constrained displacement over dt, and a bond already at its length left
    alone."""
import numpy as np
import pytest

import golden_io
import oracle

FIXTURES = ["water_3_box", "water_10_box", "nacl_3water_box", "nacl_water_box"]


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def with_shake(name):
    s, d = golden_io.load(name)
    assert d["ntcs"] == "shake"
    s.solvent_shake = True
    s.solvent_constraints = np.array(d["solvent_constr"], float)
    return s, d


def follow(s, d, use_settle=False):
    md = oracle.MD(oracle.OracleSystem(s), use_settle=use_settle)
    frames = {e["step"]: e for e in d["expected_energies"]}
    pos = {p["step"]: p for p in d["expected_positions"]}
    frc = {f["step"]: f for f in d["expected_forces"]}
    dev = dict(x=0.0, f=0.0, epot=0.0, ekin=0.0)
    for k in range(max(frames) + 1):
        if k in pos:
            x, _, _ = md.state()
            dev["x"] = max(dev["x"], np.abs(x - np.array(pos[k]["x"]).reshape(-1, 3)).max())
        e = md.run(1)[0]
        if k in frames:
            _, _, f = md.state()
            if k in frc:
                dev["f"] = max(dev["f"], np.abs(f - np.array(frc[k]["f"]).reshape(-1, 3)).max())
            dev["epot"] = max(dev["epot"], rel(e[2], frames[k]["e_pot"]))
            dev["ekin"] = max(dev["ekin"], rel(e[3], frames[k]["e_kin"]))
    return dev


@pytest.mark.parametrize("name", FIXTURES)
def test_md_with_solvent_shake_follows_gromosxx(name):
    s, d = with_shake(name)
    dev = follow(s, d)
    assert dev["x"] <= 5.01e-10 and dev["f"] <= 1e-9, dev
    assert dev["epot"] < 6e-10 and dev["ekin"] < 6e-10, dev


def test_fixture_rejects_settle():
    s, d = with_shake("water_10_box")
    s.solvent_shake = False
    dev = follow(s, d, use_settle=True)
    assert dev["x"] > 1e-8 or dev["ekin"] > 1e-8, dev


def test_constraints_hold_and_momentum_is_kept():
    s, d = with_shake("nacl_water_box")
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    x0, v0, _ = md.state()
    md.run(1)
    x1, v1, f1 = md.state()
    aps = s.atoms_per_solvent
    tol = s.shake_tolerance
    for m in range(s.nsm):
        base = s.n_solute_atoms + m * aps
        for i, j, l in d["solvent_constr"]:
            r2 = np.sum((x1[base + int(i)] - x1[base + int(j)]) ** 2)
            assert abs(l * l - r2) < 2.0 * tol * l * l
    # constraint forces are internal: the momentum changes by the step's
    # (unconstrained) forces only, which sum to ~zero here
    p0 = (s.mass[:, None] * v0).sum(0)
    p1 = (s.mass[:, None] * v1).sum(0)
    assert np.abs(p1 - p0).max() < 1e-9 * np.abs(s.mass[:, None] * v1).sum()


def test_shake_kernel_closed_forms():
    """a lone bond: one sweep puts it at the target length to first order and
    moves the atoms by inverse mass along the old bond; a bond at its length
    is left alone"""
    inv_m = np.array([1.0 / 16.0, 1.0 / 1.008])
    x_old = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0]])
    x_new = np.array([[0.001, 0.002, 0.0], [0.1021, -0.001, 0.0003]])
    v = np.zeros((2, 3))
    ci, cj, ln = np.array([0], np.int32), np.array([1], np.int32), np.array([0.1])
    xn = x_new.copy()
    it = oracle.lib().or_shake(2, oracle.ptr(inv_m), 1, oracle.ptr(ci), oracle.ptr(cj), oracle.ptr(ln), 1e-10,
                               1000, 0.002, oracle.ptr(x_old), oracle.ptr(xn), oracle.ptr(v), None)
    assert it > 1
    assert abs(np.linalg.norm(xn[0] - xn[1]) - 0.1) < 1e-10
    d0, d1 = xn[0] - x_new[0], xn[1] - x_new[1]
    # both moved along the old bond, in the inverse-mass ratio, opposite ways
    b = x_old[0] - x_old[1]
    assert np.linalg.norm(np.cross(d0, b)) < 1e-15 and np.linalg.norm(np.cross(d1, b)) < 1e-15
    assert np.allclose(d0 / inv_m[0], -d1 / inv_m[1], atol=1e-15)
    assert np.allclose(v, (xn - x_old) / 0.002, rtol=0, atol=1e-12)
    xs = x_old.copy()
    it = oracle.lib().or_shake(2, oracle.ptr(inv_m), 1, oracle.ptr(ci), oracle.ptr(cj), oracle.ptr(ln), 1e-4,
                               1000, 0.002, oracle.ptr(x_old), oracle.ptr(xs), oracle.ptr(v), None)
    assert it == 1 and np.array_equal(xs, x_old)
