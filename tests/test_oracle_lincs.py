"""Pinning of the oracle's LINCS (constraints.rs:1168-1445, md.rs:1244-1257),
restated in oracle/gromos_oracle.c.

  * gromosXX fixtures: nacl_1water_lincs (the solvent by LINCS, order 4) and
    aladip_vacuum_lincs (every solute bond by LINCS, NTC 3, order 4), every
    step of 10: positions, forces, E_pot, E_kin; the frames reject SHAKE in
    LINCS's place and another expansion order;
  * the mathematics: a lone constraint (no coupling) is brought to its length
    along the old bond, both atoms moved in the inverse-mass ratio, velocities
    the displacement over dt; a constraint already at its length is left where
    it is; the constraint forces keep the momentum."""
import numpy as np
import pytest

import golden_io
import oracle
import test_oracle_shake as ts

FIXTURES = ["nacl_1water_lincs", "aladip_vacuum_lincs"]


def solute_bonds(s, d, ntc=3):
    """the solute bonds as constraints (NTC 3: all, 2: with a hydrogen, mass
    below 2), the topology's order"""
    b = d["bonded"]
    out = []
    for i, j, t in b["bonds"]:
        r0 = b["bond_types"][t][2]
        if r0 < 1e-10 or (ntc == 2 and s.mass[i] > 2.0 and s.mass[j] > 2.0):
            continue
        out.append((i, j, r0))
    return np.array(out, float).reshape(-1, 3)


def with_lincs(name, order=None):
    s, d = golden_io.load(name)
    c = d["constraint"]
    if c["ntc"] > 1 and c["ntcp"] == "lincs":
        s.solute_constraints = solute_bonds(s, d, c["ntc"])
        s.solute_lincs = int(order or c["ntcp0"])
    if d["ntcs"] == "lincs" and s.nsm > 0:
        s.solvent_constraints = np.array(d["solvent_constr"], float)
        s.solvent_lincs = int(order or c["ntcs0"])
    return s, d


@pytest.mark.parametrize("name", FIXTURES)
def test_md_with_lincs_follows_gromosxx(name):
    s, d = with_lincs(name)
    assert s.solute_lincs == 4 or s.solvent_lincs == 4
    dev = ts.follow(s, d)
    assert dev["x"] <= 5.01e-10 and dev["f"] <= 1e-9, dev
    assert dev["epot"] < 6e-10 and dev["ekin"] < 6e-10, dev


@pytest.mark.parametrize("name", FIXTURES)
def test_fixtures_reject_shake_and_another_order(name):
    s, d = with_lincs(name, order=1)
    dev = ts.follow(s, d)
    assert dev["x"] > 1e-8 or dev["ekin"] > 1e-8, dev
    s, d = with_lincs(name)
    if s.solute_lincs:
        s.solute_lincs = 0            # the same constraints by SHAKE
    else:
        s.solvent_lincs, s.solvent_shake = 0, True
    dev = ts.follow(s, d)
    assert dev["x"] > 1e-8 or dev["ekin"] > 1e-8, dev


def run_lincs(ci, cj, ln, inv_m, x_old, x_new, order=4, dt=0.002):
    n = len(inv_m)
    xn = np.array(x_new, float).copy()
    v = np.zeros((n, 3))
    ci, cj = np.array(ci, np.int32), np.array(cj, np.int32)
    ln, inv_m, x_old = np.array(ln, float), np.array(inv_m, float), np.array(x_old, float)
    err = oracle.lib().or_lincs(len(ci), oracle.ptr(ci), oracle.ptr(cj), oracle.ptr(ln), 0, order, oracle.ptr(inv_m),
                                dt, oracle.ptr(x_old), oracle.ptr(xn), oracle.ptr(v))
    return xn, v, err


def test_lincs_closed_forms():
    inv_m = np.array([1.0 / 16.0, 1.0 / 1.008])
    x_old = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0]])
    # stretched along the bond: one projection step is exact
    x_new = np.array([[0.001, 0.0, 0.0], [0.1032, 0.0, 0.0]])
    xn, v, err = run_lincs([0], [1], [0.1], inv_m, x_old, x_new)
    assert abs(np.linalg.norm(xn[1] - xn[0]) - 0.1) < 1e-15 and err < 1e-14
    d0, d1 = xn[0] - x_new[0], xn[1] - x_new[1]
    assert np.allclose(d0 / inv_m[0], -d1 / inv_m[1], atol=1e-15) and np.all(d0[1:] == 0) and np.all(d1[1:] == 0)
    assert np.allclose(v, (xn - x_old) / 0.002, rtol=0, atol=1e-12)
    # a small rotation: the rotational correction leaves an O(theta^4) error
    th = 0.05
    x_new = np.array([[0.0, 0.0, 0.0], [0.1 * np.cos(th) * 1.01, 0.1 * np.sin(th) * 1.01, 0.0]])
    xn, v, err = run_lincs([0], [1], [0.1], inv_m, x_old, x_new)
    assert abs(np.linalg.norm(xn[1] - xn[0]) - 0.1) < 0.1 * th ** 4
    # already at its length and unmoved: stays
    xn, v, err = run_lincs([0], [1], [0.1], inv_m, x_old, x_old)
    assert np.abs(xn - x_old).max() < 1e-16 and np.abs(v).max() < 1e-13


def test_coupled_lincs_keeps_momentum_and_converges_with_order():
    """a water-like triangle of constraints sharing atoms, perturbed: the
    corrections sum to zero momentum, and the error falls with the order"""
    rng = np.random.default_rng(5)
    inv_m = 1.0 / np.array([15.9994, 1.008, 1.008])
    x_old = np.array([[0.0, 0.0, 0.0], [0.1, 0.0, 0.0], [-0.0333, 0.0943, 0.0]])
    ln = [0.1, 0.1, np.linalg.norm(x_old[1] - x_old[2])]
    x_new = x_old + rng.normal(0, 0.003, (3, 3))
    errs = []
    for order in (1, 2, 4, 8):
        xn, v, err = run_lincs([0, 0, 1], [1, 2, 2], ln, inv_m, x_old, x_new, order=order)
        p = ((xn - x_new) / inv_m[:, None]).sum(0)
        assert np.abs(p).max() < 1e-15
        errs.append(err)
    assert errs[-1] < errs[0] and errs[-1] < 1e-3
