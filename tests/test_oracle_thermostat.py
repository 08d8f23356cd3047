"""Pinning of the oracle's thermostats in the md sequence (md.rs:1143-1206:
LeapFrogVelocity, the bath's velocity scale, LeapFrogPosition):
BerendsenThermostat (berendsen_thermostat.rs:61-100) and NoseHooverThermostat,
single and chained (nosehoover_thermostat.rs:116-227), restated in
oracle/gromos_oracle.c.

  * gromosXX fixtures water_216_nvt (weak coupling), water_216_nvt_nosehoover
    (one zeta) and water_216_nvt_nhc_chain (a chain of three): 216 flexible
    waters, T0 300 K, tau 0.1 ps, every printed frame of positions, forces,
    E_pot and E_kin over the run;
  * the mathematics: a bath exactly at T0 leaves the chain at zero and the
    velocities untouched; without coupling (tau < 0) nothing changes; the single
    chain's first step and a two-link chain's first step from rest are the
    closed forms zeta_0 = dt/tau^2 (T/T0 - 1), zeta_1 = -dt/tau^2."""
import math

import numpy as np
import pytest

import golden_io
import oracle

KB = 0.00831441   # units.rs
FIXTURES = ["water_216_nvt", "water_216_nvt_nosehoover", "water_216_nvt_nhc_chain"]


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def with_bath(name):
    s, d = golden_io.load(name)
    th = d["thermostat"]
    # md.rs:1152-1172: 3N minus the solvent's constraints (none here) minus NDFMIN
    dof = 3 * s.n_atoms - 3 * s.nsm * 0 - d["ndfmin"]
    s.thermostat = (th["t_ref"], th["tau"], float(dof))
    s.thermostat_nhc = th["nhc"] if th["kind"] == "nosehoover" else 0
    return s, d


@pytest.mark.parametrize("name", FIXTURES)
def test_md_with_bath_follows_gromosxx(name):
    s, d = with_bath(name)
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    frames = {e["step"]: e for e in d["expected_energies"]}
    pos = {p["step"]: p for p in d["expected_positions"]}
    frc = {f["step"]: f for f in d["expected_forces"]}
    last = max(frames)
    checked = 0
    for k in range(last + 1):
        if k in pos:
            x, _, _ = md.state()
            assert np.abs(x - np.array(pos[k]["x"]).reshape(-1, 3)).max() <= 5.01e-10, (name, k)
        e = md.run(1)[0]
        if k in frames:
            _, _, f = md.state()
            if k in frc:
                # 9 printed decimals on forces up to ~1,800 kJ/mol/nm
                assert np.abs(f - np.array(frc[k]["f"]).reshape(-1, 3)).max() <= 1e-9, (name, k)
            assert rel(e[2], frames[k]["e_pot"]) < 6e-10, (name, k, e[2], frames[k]["e_pot"])
            assert rel(e[3], frames[k]["e_kin"]) < 6e-10, (name, k, e[3], frames[k]["e_kin"])
            checked += 1
    assert checked == len(frames) and checked >= 2
    if s.thermostat_nhc:
        assert np.all(md.zeta() != 0.0)


@pytest.mark.parametrize("name,wrong", [("water_216_nvt_nhc_chain", 1), ("water_216_nvt_nhc_chain", 0),
                                        ("water_216_nvt_nosehoover", 3), ("water_216_nvt_nosehoover", 0)])
def test_fixtures_reject_another_bath(name, wrong):
    """the frames tell the baths apart: a chain of the wrong length, or
    Berendsen in place of Nose-Hoover, misses E_kin at step 10 by >1e-5"""
    s, d = with_bath(name)
    s.thermostat_nhc = wrong
    e = oracle.MD(oracle.OracleSystem(s), use_settle=False).run(11)
    assert rel(e[10, 3], d["expected_energies"][1]["e_kin"]) > 1e-5


def test_bath_at_reference_temperature_stays_put():
    dof, t0 = 30.0, 300.0
    ekin = 0.5 * dof * KB * t0
    for nhc in (1, 3):
        z = np.zeros(nhc)
        assert oracle.nosehoover_step(z, ekin, t0, 0.1, dof, 0.002) == 1.0
        # a single zeta stays exactly zero; a chain's tail relaxes on its own
        if nhc == 1:
            assert z[0] == 0.0
    z = np.zeros(2)
    assert oracle.nosehoover_step(z, 2.0 * ekin, t0, -1.0, dof, 0.002) == 1.0   # no coupling
    assert np.all(z == 0.0)


@pytest.mark.parametrize("ratio", [0.5, 1.3, 2.0])
def test_first_step_closed_forms(ratio):
    dof, t0, tau, dt = 1944.0, 300.0, 0.1, 0.002
    ekin = 0.5 * dof * KB * t0 * ratio            # T = ratio * T0
    z = np.zeros(1)
    lam = oracle.nosehoover_step(z, ekin, t0, tau, dof, dt)
    z0 = dt / tau ** 2 * (ratio - 1.0)
    assert math.isclose(z[0], z0, rel_tol=1e-13)
    assert math.isclose(lam, 1.0 - z0 * dt, rel_tol=1e-15)
    # a chain of two from rest: the tail feels only -1/dof, the head the bath
    z = np.zeros(2)
    lam = oracle.nosehoover_step(z, ekin, t0, tau, dof, dt)
    assert math.isclose(z[1], -dt / tau ** 2, rel_tol=1e-13)
    assert math.isclose(z[0], z0, rel_tol=1e-13)
    assert math.isclose(lam, 1.0 - z0 * dt, rel_tol=1e-15)


def test_free_particles_follow_the_chain():
    """atoms with no interactions: each md step multiplies every velocity by
    1 - zeta dt with zeta advanced by the temperature of the velocities entering
    the step, so T_(n+1) = T_n (1 - zeta_(n+1) dt)^2"""
    import importlib
    test_oracle = importlib.import_module("test_oracle")
    s = test_oracle._free_particles(n=12, seed=3)
    dof = 3.0 * s.n_atoms
    temp = lambda v: np.sum(0.5 * s.mass[:, None] * v ** 2) * 2.0 / (dof * KB)
    s.thermostat = (300.0, 0.05, dof)
    s.thermostat_nhc = 1
    md = oracle.MD(oracle.OracleSystem(s), use_settle=False)
    zeta = 0.0
    t_prev = temp(s.vel)
    for _ in range(5):
        md.run(1)
        _, v, _ = md.state()
        zeta += s.dt / 0.05 ** 2 * (t_prev / 300.0 - 1.0)
        assert md.zeta()[0] == pytest.approx(zeta, rel=1e-12)
        t = temp(v)
        assert t == pytest.approx(t_prev * (1.0 - zeta * s.dt) ** 2, rel=1e-12)
        t_prev = t
