"""GPU parity: the HIP library (through its C ABI) against the CPU oracle.

Bars (FP64 plain path):
  * pair list: identical sets of integer pairs, per list (exact);
  * energies: relative difference <= 1e-10;
  * forces: per atom |dF_i| <= 1e-10 * max(|F_i|, F_floor), F_floor = 1 kJ/mol/nm;
  * single pairs: bit-identical to the oracle (same operation order, correctly
    rounded division and square root).
"""
import importlib

import numpy as np
import pytest

import golden_io
import oracle

gamd = importlib.import_module("gromos-rs_amd")
rt = gamd.runtime
gen = importlib.import_module("gromos-rs_amd.generators")

pytestmark = pytest.mark.gpu

F_FLOOR = 1.0
E_TOL = 1e-10
F_TOL = 1e-10


def torch_cuda():
    import torch
    assert torch.cuda.is_available(), "GPU test needs a visible device"
    return torch


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def force_err(f, ref):
    d = np.linalg.norm(f - ref, axis=1)
    s = np.maximum(np.linalg.norm(ref, axis=1), F_FLOOR)
    return float((d / s).max())


def pair_keys(a):
    a = np.asarray(a, np.uint64).reshape(-1, 2)
    return np.sort(a[:, 0] << np.uint64(32) | a[:, 1])


def assert_same_lists(gpu, ref):
    for k in ("solute_short", "solute_long", "solvent_short", "solvent_long"):
        g, r = pair_keys(gpu[k]), pair_keys(ref[k])
        assert len(g) == len(r), (k, len(g), len(r))
        assert np.array_equal(g, r), k


def test_device_and_library():
    assert rt.device_count() >= 1
    assert b"gfx950" in rt.lib().gromos_amd_version()


# --------------------------------------------------------------------------
# explicit pair-list kernels
# --------------------------------------------------------------------------
def test_single_pair_bitwise():
    torch = torch_cuda()
    nb = importlib.import_module("gromos-rs_amd.nonbonded")
    rng = np.random.default_rng(7)
    crf = rt.crf_new(1.4, 1.0, 61.0, 0.0)
    ocrf = oracle.crf_new(1.4, 1.0, 61.0, 0.0)
    lj = nb.LJParamMatrix(np.array([2.6e-3, 2.6e-6, 2.6e-3, 2.6e-6]), 1)
    for _ in range(200):
        pos = rng.uniform(-1, 1, size=(2, 3)) * rng.uniform(0.05, 1.0)
        q = rng.uniform(-1, 1, size=2)
        st = nb.ForceStorage(2)
        nb.lj_crf_innerloop(torch.tensor(pos), torch.tensor(q), torch.zeros(2, dtype=torch.int32),
                            torch.tensor([[0, 1]], dtype=torch.int32), lj, crf, nb.Vacuum(), 138.9354, st)
        torch.cuda.synchronize()
        r = pos[0] - pos[1]
        fm, elj, ecrf = oracle.lj_crf_interaction(r, 2.6e-3, 2.6e-6, q[0] * q[1] * 138.9354, ocrf)
        f = st.forces.cpu().numpy()
        assert np.array_equal(f[0], r * fm)
        assert st.e_lj == elj and st.e_crf == ecrf


@pytest.mark.parametrize("name", ["aladip_solvated", "water_10_box", "nacl_water_box", "ch4_water_fep",
                                  "water_1000_spc_gridcell"])
def test_explicit_kernels_match_oracle(name):
    torch = torch_cuda()
    nb = importlib.import_module("gromos-rs_amd.nonbonded")
    s, _ = golden_io.load(name)
    o = oracle.OracleSystem(s)
    pl = oracle.pairlist(o)
    crf = rt.crf_new(s.rcutl, 1.0, s.epsrf, s.kappa)
    ocrf = oracle.crf_new(s.rcutl, 1.0, s.epsrf, s.kappa)
    bc = nb.Rectangular(s.box) if s.boundary else nb.Vacuum()
    lj = nb.LJParamMatrix(s.lj, s.n_types)
    pos, q, iac = torch.tensor(s.pos), torch.tensor(s.charge), torch.tensor(s.iac, dtype=torch.int32)
    cases = [("plain", "solute_short"), ("plain", "solute_long"), ("plain", "solvent_long"),
             ("solvent", "solvent_short")]
    for kind, lst in cases:
        if len(pl[lst]) == 0:
            continue
        st = nb.ForceStorage(s.n_atoms)
        if kind == "plain":
            nb.lj_crf_innerloop(pos, q, iac, torch.tensor(pl[lst].astype(np.int32)), lj, crf, bc, s.fpepsi, st)
        else:
            nb.solvent_innerloop(pos, q, iac, torch.tensor(pl[lst].astype(np.int32)), lj, crf, bc,
                                 s.atoms_per_solvent, s.fpepsi, st)
        torch.cuda.synchronize()
        ref = oracle.innerloop(kind, o, pl[lst], ocrf)
        assert rel(st.e_lj, ref.st.e_lj) < E_TOL or abs(st.e_lj - ref.st.e_lj) < 1e-12, (lst, "lj")
        assert rel(st.e_crf, ref.st.e_crf) < E_TOL, (lst, "crf")
        assert force_err(st.forces.cpu().numpy(), ref.forces) < F_TOL, lst
        assert np.allclose(st.virial.cpu().numpy(), ref.virial, rtol=1e-10, atol=1e-9)
    st = nb.ForceStorage(s.n_atoms)
    nb.rf_excluded_interactions(q, torch.tensor(s.excl_off), torch.tensor(s.excl), pos, crf, bc, s.fpepsi, st,
                                s.n_solute_atoms)
    nb.one_four_interaction_loop(torch.tensor(s.o14_off), torch.tensor(s.o14), pos, q, iac, lj, crf, bc,
                                 s.fpepsi, st)
    torch.cuda.synchronize()
    r1 = oracle.rf_excluded(o, ocrf)
    r2 = oracle.one_four(o, ocrf)
    assert rel(st.e_crf, r1.st.e_crf + r2.st.e_crf) < E_TOL
    assert force_err(st.forces.cpu().numpy(), r1.forces + r2.forces) < F_TOL


def test_cg_grouped_kernel_matches_oracle():
    torch = torch_cuda()
    nb = importlib.import_module("gromos-rs_amd.nonbonded")
    s, _ = golden_io.load("water_1000_spc_gridcell")
    o = oracle.OracleSystem(s)
    ff = oracle.ForceField(o)
    ff.initial_pairlist(s.pos)
    pl = ff.pairlist()["solute_short"]
    # CG-pair runs (forcefield.rs:321-358)
    a2cg = np.repeat(np.arange(len(s.cg_start) - 1), np.diff(s.cg_start))
    cgi, cgj = a2cg[pl[:, 0]], a2cg[pl[:, 1]]
    brk = np.flatnonzero((np.diff(cgi) != 0) | (np.diff(cgj) != 0)) + 1
    starts = np.concatenate([[0], brk])
    ends = np.concatenate([brk, [len(pl)]])
    groups = np.stack([s.cg_start[cgi[starts]], s.cg_start[cgj[starts]], starts, ends], axis=1).astype(np.int32)
    crf = rt.crf_new(s.rcutl, 1.0, s.epsrf, s.kappa)
    ocrf = oracle.crf_new(s.rcutl, 1.0, s.epsrf, s.kappa)
    st = nb.ForceStorage(s.n_atoms)
    nb.lj_crf_innerloop_cg_grouped(torch.tensor(s.pos), torch.tensor(s.charge), torch.tensor(s.iac),
                                   torch.tensor(pl.astype(np.int32)), torch.tensor(groups),
                                   nb.LJParamMatrix(s.lj, s.n_types), crf, nb.Rectangular(s.box), s.fpepsi, st)
    torch.cuda.synchronize()
    ref = oracle.innerloop("cg_grouped", o, pl, ocrf, groups=groups)
    assert rel(st.e_lj, ref.st.e_lj) < E_TOL and rel(st.e_crf, ref.st.e_crf) < E_TOL
    assert force_err(st.forces.cpu().numpy(), ref.forces) < F_TOL


# --------------------------------------------------------------------------
# device-resident systems: GPU pair list + Forcefield::apply
# --------------------------------------------------------------------------
RESIDENT = golden_io.names()


@pytest.mark.parametrize("name", RESIDENT)
def test_resident_pairlist_and_forces_fixture(name):
    s, d = golden_io.load(name)
    o = oracle.OracleSystem(s)
    ds = rt.DeviceSystem(s)
    ds.pairlist_update()
    assert_same_lists(ds.pairlist(), oracle.pairlist(o))
    g = ds.forcefield_apply()
    r = oracle.single_point(o)
    for k in ("e_lj", "e_crf", "dhdl"):
        assert rel(g[k], r[k]) < E_TOL or abs(g[k] - r[k]) < 1e-12, (k, g[k], r[k])
    assert force_err(g["forces"], r["forces"]) < F_TOL
    ds.close()


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_vacuum_single_cell(prec):
    """No periodic box: every charge group lies in the job's one cell, more of
    them than a wave holds (the cell sort's large-cell branch); the pair set is
    the oracle's and the forces match (FP64 at the parity tolerance, mixed at
    the mixed-precision one)."""
    s = gen.spc_water_box(150, seed=21)
    s.boundary = 0
    o = oracle.OracleSystem(s)
    ds = rt.DeviceSystem(s, precision=prec)
    ds.pairlist_update()
    assert_same_lists(ds.pairlist(), oracle.pairlist(o))
    g = ds.forcefield_apply()
    r = oracle.single_point(o)
    ds.close()
    if prec == rt.FP64:
        assert rel(g["e_lj"], r["e_lj"]) < E_TOL and rel(g["e_crf"], r["e_crf"]) < E_TOL
        assert force_err(g["forces"], r["forces"]) < F_TOL
    else:
        assert rel(g["e_lj"] + g["e_crf"], r["e_lj"] + r["e_crf"]) < 1e-5
        assert force_err(g["forces"], r["forces"]) < 1e-4


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_search_exact_half_box_ties(prec):
    """An exact lattice (64 waters, L = 2 nm, first atoms on multiples of 0.5 nm,
    cutoff 1.4 nm > L/2): many first-atom separations are exactly L/2 along an
    axis, where x / L = +-0.5 and the nearest image takes its half-away branch
    (the reference's f64::round) in the search and in every force kernel.
    The pair sets are the oracle's in both precisions, FP64 forces and
    energies at the parity bar, mixed forces at the mixed bar.  (The mixed
    energies are not compared: on an exact lattice with the cutoff beyond L/2
    the Coulomb terms, up to ~1e5 kJ/mol in total, cancel to E_CRF = 0.35, below
    what FP32 pair energies resolve.)"""
    s = gen.spc_water_box(64, box=2.0, jitter=0.0, seed=5, rcutp=1.4, rcutl=1.4, name="lattice_64")
    o = oracle.OracleSystem(s)
    ds = rt.DeviceSystem(s, precision=prec)
    ds.pairlist_update()
    assert_same_lists(ds.pairlist(), oracle.pairlist(o))
    g = ds.forcefield_apply()
    r = oracle.single_point(o)
    if prec == rt.FP64:
        assert rel(g["e_lj"], r["e_lj"]) < E_TOL and rel(g["e_crf"], r["e_crf"]) < E_TOL
        assert force_err(g["forces"], r["forces"]) < F_TOL
    else:
        assert force_err(g["forces"], r["forces"]) < MX_F_TOL
    ds.close()


def test_water216_cutoff_beyond_half_box():
    s = gen.water_216()
    o = oracle.OracleSystem(s)
    ds = rt.DeviceSystem(s)
    ds.pairlist_update()
    assert_same_lists(ds.pairlist(), oracle.pairlist(o))
    g = ds.forcefield_apply()
    r = oracle.single_point(o)
    assert rel(g["e_lj"], r["e_lj"]) < E_TOL and rel(g["e_crf"], r["e_crf"]) < E_TOL
    assert force_err(g["forces"], r["forces"]) < F_TOL


def test_water216_md_steps_match_oracle():
    """Per-step parity: from the same state, one GPU step equals one oracle step;
    and a 100-step run stays within a tolerance set by trajectory divergence."""
    s = gen.water_216()
    o = oracle.OracleSystem(s)
    ds = rt.DeviceSystem(s)
    ds.pairlist_update()
    ge = ds.md(100)
    md = oracle.MD(o)
    oe = md.run(100)
    # first steps: identical within 1e-10 (chaos has not amplified rounding yet)
    for k in range(5):
        for c in range(4):
            assert rel(ge[k, c], oe[k, c]) < 1e-10, (k, c, ge[k, c], oe[k, c])
    x_g, v_g = ds.get_state()
    x_o, v_o, _ = md.state()
    # whole run: rounding grows along the trajectory but stays far below physics
    assert np.abs(x_g - x_o).max() < 1e-6
    assert np.all(np.abs(ge[:, 2] - oe[:, 2]) / np.abs(oe[:, 2]) < 1e-6)


def test_settle_fixture_md():
    s, d = golden_io.load("nacl_1water_settle")
    ds = rt.DeviceSystem(s)
    ds.pairlist_update()
    fr = d["expected_energies"]
    e = ds.md(len(fr))
    for k, f in enumerate(fr):
        assert rel(e[k, 2], f["e_pot"]) < 6e-10
        assert rel(e[k, 3], f["e_kin"]) < 6e-10


def test_water32768_full_size():
    """The benchmark system: exact pair set against the oracle's O(N^2)
    standard list, forces against the oracle, Newton's third law."""
    s = gen.water_32768()
    o = oracle.OracleSystem(s)
    ds = rt.DeviceSystem(s)
    ds.pairlist_update()
    gl = ds.pairlist()
    ol = oracle.pairlist(o)
    assert_same_lists(gl, ol)
    g = ds.forcefield_apply()
    r = oracle.single_point(o, parallel=True)
    assert rel(g["e_lj"], r["e_lj"]) < E_TOL and rel(g["e_crf"], r["e_crf"]) < E_TOL
    assert force_err(g["forces"], r["forces"]) < F_TOL
    assert np.abs(g["forces"].sum(0)).max() < 1e-6


def test_hetero_60000_full_size():
    """The heterogeneous single point at its full size: 6,000-atom random-walk
    solute in charge groups of 1-4 atoms, 1-2 / 1-3 exclusions, 1-4 pairs, 18,000
    SPC waters, L = 8.4 nm, rc 1.4 nm (60,000 atoms): exact four-list pair set
    (exclusions included) against the oracle's standard list, FP64 forces and
    energies at the parity tolerance, and the mixed path at its tolerance."""
    s = gen.hetero_system()
    assert s.n_atoms == 60000
    o = oracle.OracleSystem(s)
    ol = oracle.pairlist(o)
    r = oracle.single_point(o, parallel=True)
    for prec in (rt.FP64, rt.MIXED):
        ds = rt.DeviceSystem(s, precision=prec)
        ds.pairlist_update()
        assert_same_lists(ds.pairlist(), ol)
        g = ds.forcefield_apply()
        ds.close()
        if prec == rt.FP64:
            assert rel(g["e_lj"], r["e_lj"]) < E_TOL and rel(g["e_crf"], r["e_crf"]) < E_TOL
            assert force_err(g["forces"], r["forces"]) < F_TOL
        else:
            mixed_energy_ok(g, r)
            assert force_err(g["forces"], r["forces"]) < MX_F_TOL
            assert force_err_median(g["forces"], r["forces"]) < MX_F_MED


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_remd_replica_single_point(prec):
    """One replica of the exchange batch: 4,000 SPC waters (12,000 atoms,
    L = 4.93 nm), rc 1.2 nm: exact pair set and forces against the oracle."""
    s = gen.water_replica(seed=100, temperature=280.0, n_waters=4000, rc=1.2)
    o = oracle.OracleSystem(s)
    ds = rt.DeviceSystem(s, precision=prec)
    ds.pairlist_update()
    assert_same_lists(ds.pairlist(), oracle.pairlist(o))
    g = ds.forcefield_apply()
    ds.close()
    r = oracle.single_point(o, parallel=True)
    if prec == rt.FP64:
        assert rel(g["e_lj"], r["e_lj"]) < E_TOL and rel(g["e_crf"], r["e_crf"]) < E_TOL
        assert force_err(g["forces"], r["forces"]) < F_TOL
    else:
        mixed_energy_ok(g, r)
        assert force_err(g["forces"], r["forces"]) < MX_F_TOL
        assert force_err_median(g["forces"], r["forces"]) < MX_F_MED


# --------------------------------------------------------------------------
# mixed precision (FP32 pair math, int64 fixed-point forces, FP64 sums)
# tolerance (per atom |dF_i| / max(|F_i|, 1 kJ/mol/nm)): median <= 2e-6, max
# <= 1e-4.  FP32 pair forces carry ~2-3 ulp (1.5e-7) each; an atom's net force
# cancels ~1000 terms up to 100x larger than itself, so the worst atoms of a
# 98k-atom water box sit near 3e-5 (a float64 emulation of the same FP32
# operation sequence gives 1.4e-5 even with a correctly rounded rsqrt), and a
# water atom with |F| = 32 next to +-0.5 e ions reaches 7e-5.  Pairs whose
# first atom is perturbed are evaluated in FP64 and held to 1e-5.
# Energies: 1e-5 of the job's potential-energy scale |E_LJ| + |E_CRF| (FP32
# pair energies summed in FP64; a component that is itself a near-cancelling
# sum, as E_CRF of randomly oriented water, is held to the scale of the terms
# it is made of, not to its own small value).
# --------------------------------------------------------------------------
MX_F_TOL = 1e-4
MX_F_TOL_PERT = 1e-5
MX_F_MED = 2e-6
MX_E_TOL = 1e-5


def mixed_energy_ok(g, r, keys=("e_lj", "e_crf")):
    scale = max(abs(r["e_lj"]) + abs(r["e_crf"]), 1.0)
    for k in keys:
        assert abs(g[k] - r[k]) <= MX_E_TOL * scale, (k, g[k], r[k], scale)


def force_err_median(f, ref):
    d = np.linalg.norm(f - ref, axis=1)
    s = np.maximum(np.linalg.norm(ref, axis=1), F_FLOOR)
    return float(np.median(d / s))


@pytest.mark.parametrize("name", ["water_10_box", "nacl_water_box", "aladip_solvated", "water_1000_spc_gridcell",
                                  "water_3_box_twinrange"])
def test_mixed_fixture_forces(name):
    s, _ = golden_io.load(name)
    ds = rt.DeviceSystem(s, precision=rt.MIXED)
    ds.pairlist_update()
    assert_same_lists(ds.pairlist(), oracle.pairlist(oracle.OracleSystem(s)))
    g = ds.forcefield_apply()
    r = oracle.single_point(oracle.OracleSystem(s))
    assert force_err(g["forces"], r["forces"]) < MX_F_TOL
    mixed_energy_ok(g, r)


def test_mixed_fep_mix_jobs_meet_the_force_bar():
    """Jobs of the 64-job FEP mix after 300 relaxation steps (the state the
    mix's force error is quoted on): the mixed path's per-atom force error
    against plain FP64 stays within 1e-5 (near solvent partners in FP64,
    the rest FP32)."""
    srv = importlib.import_module("gromos-rs_amd.server")
    mix = gen.fep_mix()
    pick = [0, 5, 17, 33, 48, 63]
    systems = [gen.fep_job(mix[k]["n_atoms"], mix[k]["n_pert"], mix[k]["lam"], mix[k]["seed"]) for k in pick]
    systems = srv.equilibrate(systems, [300.0] * len(systems))
    worst = 0.0
    for q in systems:
        out = {}
        for prec in (rt.FP64, rt.MIXED):
            ds = rt.DeviceSystem(q, precision=prec)
            ds.pairlist_update()
            out[prec] = ds.forcefield_apply()
            ds.close()
        worst = max(worst, force_err(out[rt.MIXED]["forces"], out[rt.FP64]["forces"]))
        mixed_energy_ok(out[rt.MIXED], out[rt.FP64])
    assert worst <= 1e-5, worst


def test_mixed_water32768_error_and_determinism():
    s = gen.water_32768()
    r = oracle.single_point(oracle.OracleSystem(s), parallel=True)
    out = []
    for _ in range(2):
        ds = rt.DeviceSystem(s, precision=rt.MIXED)
        ds.pairlist_update()
        out.append(ds.forcefield_apply())
        ds.close()
    g = out[0]
    assert force_err(g["forces"], r["forces"]) < MX_F_TOL
    assert force_err_median(g["forces"], r["forces"]) < MX_F_MED
    # the lattice's CRF energy is a small difference of large pair sums: each
    # component is held against the system's nonbonded energy scale, as in the
    # other mixed-path tests
    mixed_energy_ok(g, r)
    # fixed order everywhere: bit-identical run to run
    assert np.array_equal(out[0]["forces"], out[1]["forces"])
    assert out[0]["e_lj"] == out[1]["e_lj"] and out[0]["e_crf"] == out[1]["e_crf"]


def test_mixed_md_tracks_fp64():
    s = gen.spc_water_box(4096, seed=11)
    res = {}
    for prec in (rt.FP64, rt.MIXED):
        ds = rt.DeviceSystem(s, precision=prec)
        ds.pairlist_update()
        res[prec] = ds.md(200)
        ds.close()
    e64, emx = res[rt.FP64], res[rt.MIXED]
    tot64, totmx = e64[:, 2] + e64[:, 3], emx[:, 2] + emx[:, 3]
    # same physics step by step at the start
    assert np.all(np.abs(emx[:10, 2] - e64[:10, 2]) <= 1e-5 * np.abs(e64[:10, 2]))
    # energy drift over the run agrees within 2% of the kinetic energy scale
    d64, dmx = tot64[-1] - tot64[0], totmx[-1] - totmx[0]
    assert abs(dmx - d64) <= 0.02 * abs(e64[-1, 3]), (d64, dmx)


@pytest.mark.parametrize("prec,pairwise_rf", [(rt.FP64, False), (rt.MIXED, False), (rt.MIXED, True)])
def test_resident_virial_and_rf_forms(prec, pairwise_rf, monkeypatch):
    """Virial of the device-resident path against the oracle.  The mixed kernel
    sums the RF polynomial per neutral molecule pair (dipole form); the per-atom-
    pair form (GROMOS_AMD_VV_PAIRWISE_RF=1) must give the same results within
    the mixed tolerance."""
    if pairwise_rf:
        monkeypatch.setenv("GROMOS_AMD_VV_PAIRWISE_RF", "1")
    s = gen.spc_water_box(1500, seed=15)
    ds = rt.DeviceSystem(s, precision=prec, do_virial=True)
    ds.pairlist_update()
    g = ds.forcefield_apply()
    ds.close()
    r = oracle.single_point(oracle.OracleSystem(s))
    vscale = float(np.abs(r["virial"]).max())
    if prec == rt.FP64:
        assert np.abs(g["virial"] - r["virial"]).max() <= 1e-10 * vscale
        assert rel(g["e_lj"], r["e_lj"]) < E_TOL and rel(g["e_crf"], r["e_crf"]) < E_TOL
        assert force_err(g["forces"], r["forces"]) < F_TOL
    else:
        assert np.abs(g["virial"] - r["virial"]).max() <= 1e-5 * vscale, (g["virial"], r["virial"])
        mixed_energy_ok(g, r)
        assert force_err(g["forces"], r["forces"]) < MX_F_TOL
        assert force_err_median(g["forces"], r["forces"]) < MX_F_MED


# --------------------------------------------------------------------------
# free energy: device-resident soft-core correction passes against the oracle
# (soft-core CRF with charges, dH/dlambda, excluded / 1-4 / atom-pair passes are
# parity-unpinned against the reference: this checks GPU == oracle restatement)
# --------------------------------------------------------------------------
def _fep_case(s, prec=None):
    o = oracle.OracleSystem(s)
    ds = rt.DeviceSystem(s, precision=rt.FP64 if prec is None else prec)
    ds.pairlist_update()
    assert_same_lists(ds.pairlist(), oracle.pairlist(o))
    g = ds.forcefield_apply()
    r = oracle.single_point(o)
    if prec == rt.MIXED:
        mixed_energy_ok(g, r, ("e_lj", "e_crf", "dhdl"))
        assert force_err(g["forces"], r["forces"]) < MX_F_TOL
    else:
        for k in ("e_lj", "e_crf", "dhdl"):
            assert abs(g[k] - r[k]) <= E_TOL * max(abs(r[k]), 1.0), (k, g[k], r[k])
        assert force_err(g["forces"], r["forces"]) < F_TOL
    ds.close()
    return g, r


@pytest.mark.parametrize("lam", [0.0, 0.35, 1.0])
def test_resident_fep_job(lam):
    s = gen.fep_job(3000, 12, lam, seed=21)
    _fep_case(s)


@pytest.mark.parametrize("prec", [None, rt.MIXED])
def test_resident_fep_job_cutoff_beyond_half_box(prec):
    """The smallest jobs of the FEP mix: 1,000 atoms, L = 2.15 nm < 2 rc."""
    s = gen.fep_job(1000, 9, 0.7, seed=4)
    assert s.box[0] < 2 * s.rcutl
    g, r = _fep_case(s, prec)
    assert g["dhdl"] != 0.0


def test_resident_fep_bonded_solute_all_passes():
    s = gen.hetero_system(n_chain=300, n_waters=700, box=3.2, seed=8, pert_every=5, lam=0.4)
    # PERTATOMPAIR entries on some 1-4 pairs (a_type 0 / 1, b_type absent or 1-4)
    s.pert.atom_pairs = np.array([[0, 3, 0, -1], [5, 8, 1, 0], [10, 13, 0, 1]], np.int32)
    g, r = _fep_case(s)
    assert g["dhdl"] != 0.0


def test_mixed_fep_job():
    s = gen.fep_job(3000, 12, 0.45, seed=22)
    g, r = _fep_case(s, prec=rt.MIXED)
    # the perturbed atoms' own forces come from FP64 pairs (all their partners
    # are listed after them) plus FP64 excluded / self terms
    pert = np.where(s.pert.flag)[0]
    assert force_err(g["forces"][pert], r["forces"][pert]) < MX_F_TOL_PERT


# --------------------------------------------------------------------------
# Berendsen weak coupling (berendsen_thermostat.rs) on the device vs the oracle
# --------------------------------------------------------------------------
@pytest.mark.parametrize("tau", [0.1, 0.0])
def test_berendsen_md_matches_oracle(tau):
    s = gen.water_216()
    s.thermostat = (350.0, tau, 3 * s.n_atoms - 3 * s.nsm)
    o = oracle.OracleSystem(s)
    ds = rt.DeviceSystem(s)
    ds.pairlist_update()
    ge = ds.md(40)
    oe = oracle.MD(o).run(40)
    for k in range(5):
        for c in range(4):
            assert rel(ge[k, c], oe[k, c]) < 1e-10, (k, c, ge[k, c], oe[k, c])
    assert np.all(np.abs(ge[:, 3] - oe[:, 3]) / oe[:, 3] < 1e-7)
    ds.close()


# --------------------------------------------------------------------------
# every lane count the autotuner may pick gives the oracle's result
# --------------------------------------------------------------------------
@pytest.mark.parametrize("lanes", [4, 8, 16])
@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_vv_lane_variants(lanes, prec, monkeypatch):
    monkeypatch.setenv("GROMOS_AMD_VV_LANES", str(lanes))
    s = gen.spc_water_box(1500, seed=12)
    ds = rt.DeviceSystem(s, precision=prec)
    ds.pairlist_update()
    g = ds.forcefield_apply()
    assert ds.vv_lanes == lanes
    ds.close()
    r = oracle.single_point(oracle.OracleSystem(s))
    if prec == rt.FP64:
        assert rel(g["e_lj"], r["e_lj"]) < E_TOL and rel(g["e_crf"], r["e_crf"]) < E_TOL
        assert force_err(g["forces"], r["forces"]) < F_TOL
    else:
        mixed_energy_ok(g, r)
        assert force_err(g["forces"], r["forces"]) < MX_F_TOL


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_step_graphs_match_separate_launches(prec, monkeypatch):
    """Plain steps replayed from captured graphs (with the eager rebuild steps and
    host synchronisations of every md call between replays) give the same bits
    as every step launched separately."""
    s = gen.spc_water_box(1500, seed=16)
    s.thermostat = (300.0, 0.1, 3 * s.n_atoms - 3 * s.nsm)
    out = {}
    for graphs in ("1", "0"):   # opt-in replay vs the default separate launches
        monkeypatch.setenv("GROMOS_AMD_GRAPHS", graphs)
        ds = rt.DeviceSystem(s, precision=prec)
        ds.pairlist_update()
        e = np.concatenate([ds.md(7), ds.md(1), ds.md(13), ds.md(9)])
        x, v = ds.get_state()
        out[graphs] = (e, x, v, ds.graph_launches)
        ds.close()
    assert out["1"][3] > 0 and out["0"][3] == 0
    for k in range(3):
        assert np.array_equal(out["1"][k], out["0"][k])


def test_side_stream_matches_single_stream(monkeypatch):
    """The mixed md step's solute pass on the side stream (concurrent with the
    solvent pair kernel) gives the same bits as the whole step on one stream,
    for one FEP job and for a session batch of two."""
    a = gen.fep_job(3000, 8, 0.5, seed=7)
    b = gen.fep_job(1500, 5, 0.25, seed=8)
    out = {}
    for side in ("1", "0"):
        monkeypatch.setenv("GROMOS_AMD_SIDE_STREAM", side)
        ds = rt.DeviceSystem(a, precision=rt.MIXED)
        ds.pairlist_update()
        e = np.concatenate([ds.md(11), ds.md(6)])
        x, v = ds.get_state()
        ds.close()
        sess = rt.Session(precision=rt.MIXED, nsnb=5)
        ja, jb = sess.admit(a, 12), sess.admit(b, 9)
        sess.step(12)
        out[side] = (e, x, v, sess.energies(ja), sess.energies(jb), *sess.fetch(ja), *sess.fetch(jb))
        sess.close()
    for k in range(len(out["1"])):
        assert np.array_equal(out["1"][k], out["0"][k]), k


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_refilled_lists_match(prec, monkeypatch):
    """Slots sized at exactly the previous counts make most pair-list builds
    overflow and refill, including builds inside md steps, whose first pair
    kernels were already enqueued on the truncated lists and run again: the
    results equal those of the default slots bit for bit."""
    s = gen.fep_job(2500, 6, 0.4, seed=12)
    out = {}
    for tight in ("1", "0"):
        monkeypatch.setenv("GROMOS_AMD_TIGHT_SLOTS", tight)
        ds = rt.DeviceSystem(s, precision=prec)
        ds.pairlist_update()
        e = np.concatenate([ds.md(12), ds.md(14)])
        x, v = ds.get_state()
        out[tight] = (e, x, v, ds.rebuild_stats)
        ds.close()
    assert out["1"][3][1] >= 3 and out["0"][3][1] == 0, (out["1"][3], out["0"][3])
    for k in range(3):
        assert np.array_equal(out["1"][k], out["0"][k]), k


def test_autotuner_picks_a_variant():
    s = gen.spc_water_box(3000, seed=13)
    ds = rt.DeviceSystem(s, precision=rt.MIXED)
    assert ds.vv_lanes == 0
    ds.pairlist_update()
    ds.forcefield_apply()
    assert ds.vv_lanes in (4, 8, 16)
    ds.close()


def test_mixed_bits_do_not_depend_on_lane_count(monkeypatch):
    s = gen.spc_water_box(1500, seed=14)
    out = []
    for lanes in (4, 8, 16):
        monkeypatch.setenv("GROMOS_AMD_VV_LANES", str(lanes))
        ds = rt.DeviceSystem(s, precision=rt.MIXED)
        ds.pairlist_update()
        out.append(ds.forcefield_apply())
        ds.close()
    for g in out[1:]:
        assert np.array_equal(g["forces"], out[0]["forces"])
        assert g["e_lj"] == out[0]["e_lj"] and g["e_crf"] == out[0]["e_crf"]
