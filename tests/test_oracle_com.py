"""Pinning of the oracle's centre-of-mass motion removal (RemoveCOMMotion,
remove_com_motion.rs, first in the md sequence, md.rs:1095-1097), restated in
oracle/gromos_oracle.c.

  * gromosXX fixtures water_216_box_com (NTICOM 1, NSCM 10: translation at
    step 0 and every 10 steps) and water_216_box_com_rot (NTICOM 2, NSCM -10:
    rotation too at step 0; a periodic box gets no periodic rotation removal):
    every printed frame of positions, forces, E_pot and E_kin over 100 steps;
    the frames reject the run without removal, and the rotation fixture the
    run that removes the translation only;
  * the mathematics: after the translation step the total momentum vanishes;
    after the rotation step so does the angular momentum about the centre of
    mass (of the positions taken back half a step, in the box's (-L/2, L/2]
    frame); a rigid rotation plus translation is removed entirely; vacuum
    systems get the periodic rotation removal, periodic ones do not."""
import importlib

import numpy as np
import pytest

import golden_io
import oracle

FIXTURES = ["water_216_box_com", "water_216_box_com_rot"]


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def with_com(name):
    s, d = golden_io.load(name)
    s.nticom, s.nscm = d["nticom"], d["nscm"]
    return s, d


def follow(s, d, steps=None):
    """max deviations from the fixture frames: positions, forces, E_pot, E_kin"""
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    frames = {e["step"]: e for e in d["expected_energies"]}
    pos = {p["step"]: p for p in d["expected_positions"]}
    frc = {f["step"]: f for f in d["expected_forces"]}
    last = max(frames) if steps is None else steps
    dx = df = dp = dk = 0.0
    for k in range(last + 1):
        if k in pos:
            x, _, _ = md.state()
            dx = max(dx, np.abs(x - np.array(pos[k]["x"]).reshape(-1, 3)).max())
        e = md.run(1)[0]
        if k in frames:
            _, _, f = md.state()
            df = max(df, np.abs(f - np.array(frc[k]["f"]).reshape(-1, 3)).max())
            dp = max(dp, rel(e[2], frames[k]["e_pot"]))
            dk = max(dk, rel(e[3], frames[k]["e_kin"]))
    return dx, df, dp, dk


@pytest.mark.parametrize("name", FIXTURES)
def test_md_with_com_removal_follows_gromosxx(name):
    s, d = with_com(name)
    dx, df, dp, dk = follow(s, d)
    # positions at their 9 printed decimals; forces (to ~1,800 kJ/mol/nm, 9
    # decimals) within 5e-9 as the two runs' rounding drifts apart over 100 steps
    assert dx <= 5.01e-10 and df <= 5e-9, (dx, df)
    assert dp < 6e-10 and dk < 6e-10, (dp, dk)


@pytest.mark.parametrize("name,nticom,nscm", [("water_216_box_com", 0, 0), ("water_216_box_com_rot", 0, 0),
                                             ("water_216_box_com_rot", 1, -10)])
def test_fixtures_reject_other_removals(name, nticom, nscm):
    s, d = with_com(name)
    s.nticom, s.nscm = nticom, nscm
    _, _, _, dk = follow(s, d, steps=10)
    assert dk > 1e-8, dk


def momenta(s, x, v):
    m = s.mass[:, None]
    p = (m * v).sum(0)
    return p


def test_translation_removal_zeroes_momentum():
    s, _ = with_com("water_216_box_com")
    osys = oracle.OracleSystem(s)
    rng = np.random.default_rng(2)
    v = s.vel + rng.normal(0, 0.3, 3)[None, :]
    v1 = oracle.remove_com(osys, s.pos, v, 1, 0, 0)
    assert np.abs(momenta(s, s.pos, v1)).max() < 1e-11
    # v - v1 is one common vector
    dv = v - v1
    assert np.abs(dv - dv[0]).max() < 1e-14


def test_rotation_removal_zeroes_angular_momentum():
    s, _ = with_com("water_216_box_com_rot")
    osys = oracle.OracleSystem(s)
    L, dt = s.box, s.dt
    rng = np.random.default_rng(3)
    v = s.vel.copy()
    v1 = oracle.remove_com(osys, s.pos, v, 2, 0, 0)
    m = s.mass[:, None]
    r = s.pos - np.round(s.pos / L) * L - 0.5 * dt * v1
    # the centre of mass the removal used (of the velocities after translation removal)
    vt = v - (m * v).sum(0) / m.sum()
    rc = (m * (s.pos - np.round(s.pos / L) * L) - 0.5 * m * vt * dt).sum(0) / m.sum()
    r = s.pos - np.round(s.pos / L) * L - 0.5 * dt * vt - rc
    ang = (m * np.cross(r, v1)).sum(0)
    scale = (m * np.abs(np.cross(r, vt))).sum()
    assert np.abs(ang).max() < 1e-12 * scale
    assert np.abs((m * v1).sum(0)).max() < 1e-11
    # a pure rigid motion of the half-step positions is removed entirely
    w, c = rng.normal(0, 2.0, 3), rng.normal(0, 0.5, 3)
    x0 = s.pos - np.round(s.pos / L) * L
    # velocities whose half-step-back positions rotate rigidly: solve v = c + w x (x0 - dt/2 v - rc')
    vr = np.tile(c, (s.n_atoms, 1))
    for _ in range(60):   # fixed point (dt/2 |w| << 1)
        rcp = (m * x0 - 0.5 * m * vr * dt).sum(0) / m.sum()
        vr = c + np.cross(w, x0 - 0.5 * dt * vr - rcp)
    out = oracle.remove_com(osys, s.pos, vr, 2, 0, 0)
    assert np.abs(out).max() < 1e-9 * np.abs(vr).max()


def test_periodic_rotation_only_in_vacuum():
    s, _ = with_com("water_216_box_com_rot")
    osys = oracle.OracleSystem(s)
    v = s.vel.copy()
    vt = oracle.remove_com(osys, s.pos, v, 0, 1, 10)       # translation
    vb = oracle.remove_com(osys, s.pos, v, 0, -10, 10)     # periodic box: still translation only
    assert np.array_equal(vt, vb)
    assert np.array_equal(oracle.remove_com(osys, s.pos, v, 2, -10, 5), v)   # not a removal step
    s.boundary = 0
    s.box = np.zeros(3)
    osv = oracle.OracleSystem(s)
    vv = oracle.remove_com(osv, s.pos, v, 0, -10, 10)      # vacuum: rotation too
    assert not np.array_equal(vv, oracle.remove_com(osv, s.pos, v, 0, 10, 10))
