"""GPU parity on triclinic (truncated-octahedron) boxes: Triclinic::nearest_image
(crates/gromos-core/src/math.rs:95-141) in every kernel, the pair list on the
one-cell search that mirrors the reference's O(N^2) fallback for non-rectangular
boxes (pairlist.rs:521-532), against the oracle (itself pinned by the gromosXX
aladip_trunc_oct fixture).  Bars as in test_gpu_parity: exact pair sets, FP64
energies and forces within 1e-10, the mixed path at its tolerance."""
import importlib

import numpy as np
import pytest

import golden_io
import oracle

gamd = importlib.import_module("gromos-rs_amd")
rt = gamd.runtime
gen = importlib.import_module("gromos-rs_amd.generators")

pytestmark = pytest.mark.gpu

E_TOL = 1e-10
F_TOL = 1e-10
MX_F_TOL = 1e-4
MX_E_TOL = 1e-5


def rel(a, b):
    return abs(a - b) / max(abs(b), 1e-300)


def force_err(f, ref):
    d = np.linalg.norm(f - ref, axis=1)
    return float((d / np.maximum(np.linalg.norm(ref, axis=1), 1.0)).max())


def pair_keys(a):
    a = np.asarray(a, np.uint64).reshape(-1, 2)
    return np.sort(a[:, 0] << np.uint64(32) | a[:, 1])


def same_lists(gpu, ref):
    for k in ("solute_short", "solute_long", "solvent_short", "solvent_long"):
        g, r = pair_keys(gpu[k]), pair_keys(ref[k])
        assert len(g) == len(r) and np.array_equal(g, r), k


def test_truncoct_fixture_energies_match_gromosxx():
    """The device path on the gromosXX truncated-octahedron fixture (NTB = -1:
    box converted and coordinates rotated as gromos-rs reads them): nonbonded
    energies equal the fixture's printed 10 significant figures."""
    s, d = golden_io.load("aladip_trunc_oct")
    ds = rt.DeviceSystem(s)
    ds.pairlist_update()
    g = ds.forcefield_apply()
    ds.close()
    e0 = d["expected_energies"][0]
    assert rel(g["e_lj"], e0["e_lj"]) < 6e-10 and rel(g["e_crf"], e0["e_crf"]) < 6e-10


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_truncoct_water(prec):
    s = gen.water_truncoct(14, seed=3, rc=1.2)
    o = oracle.OracleSystem(s)
    ds = rt.DeviceSystem(s, precision=prec)
    ds.pairlist_update()
    same_lists(ds.pairlist(), oracle.pairlist(o))
    g = ds.forcefield_apply()
    ds.close()
    r = oracle.single_point(o, parallel=True)
    if prec == rt.FP64:
        assert rel(g["e_lj"], r["e_lj"]) < E_TOL and rel(g["e_crf"], r["e_crf"]) < E_TOL
        assert force_err(g["forces"], r["forces"]) < F_TOL
    else:
        scale = abs(r["e_lj"]) + abs(r["e_crf"])
        assert abs(g["e_lj"] - r["e_lj"]) < MX_E_TOL * scale and abs(g["e_crf"] - r["e_crf"]) < MX_E_TOL * scale
        assert force_err(g["forces"], r["forces"]) < MX_F_TOL


def test_truncoct_md_matches_oracle():
    """Leap-frog + SETTLE steps on the truncated octahedron: the first steps'
    energies equal the oracle's within 1e-10."""
    s = gen.water_truncoct(8, seed=5, rc=1.0)
    ds = rt.DeviceSystem(s)
    ds.pairlist_update()
    ge = ds.md(12)
    ds.close()
    oe = oracle.MD(oracle.OracleSystem(s)).run(12)
    for k in range(6):
        for c in range(4):
            assert rel(ge[k, c], oe[k, c]) < 1e-10, (k, c, ge[k, c], oe[k, c])


def test_truncoct_explicit_kernels():
    import torch
    nb = importlib.import_module("gromos-rs_amd.nonbonded")
    s = gen.water_truncoct(8, seed=6, rc=1.0)
    o = oracle.OracleSystem(s)
    pl = oracle.pairlist(o)
    crf = rt.crf_new(s.rcutl, 1.0, s.epsrf, s.kappa)
    ocrf = oracle.crf_new(s.rcutl, 1.0, s.epsrf, s.kappa)
    bc = nb.boundary_of(s)
    lj = nb.LJParamMatrix(s.lj, s.n_types)
    st = nb.ForceStorage(s.n_atoms)
    nb.solvent_innerloop(torch.tensor(s.pos), torch.tensor(s.charge), torch.tensor(s.iac),
                         torch.tensor(pl["solvent_short"].astype(np.int32)), lj, crf, bc, 3, s.fpepsi, st)
    torch.cuda.synchronize()
    ref = oracle.innerloop("solvent", o, pl["solvent_short"], ocrf)
    assert rel(st.e_lj, ref.st.e_lj) < E_TOL and rel(st.e_crf, ref.st.e_crf) < E_TOL
    assert force_err(st.forces.cpu().numpy(), ref.forces) < F_TOL
    # all atom pairs as one explicit list through the per-pair nearest image
    i, j = np.triu_indices(s.n_atoms, 1)
    pairs = np.stack([i, j], 1)[(i // 3) != (j // 3)][:50000].astype(np.int32)
    st = nb.ForceStorage(s.n_atoms)
    nb.lj_crf_innerloop(torch.tensor(s.pos), torch.tensor(s.charge), torch.tensor(s.iac), torch.tensor(pairs), lj,
                        crf, bc, s.fpepsi, st)
    torch.cuda.synchronize()
    ref = oracle.innerloop("plain", o, pairs.astype(np.uint32), ocrf)
    assert rel(st.e_lj, ref.st.e_lj) < E_TOL and rel(st.e_crf, ref.st.e_crf) < E_TOL
    assert force_err(st.forces.cpu().numpy(), ref.forces) < F_TOL


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_truncoct_job_in_a_batch_matches_alone(prec):
    """A triclinic job batched with rectangular ones gives the bits it gives alone."""
    specs = [gen.water_truncoct(7, seed=8, rc=1.0), gen.spc_water_box(500, seed=9),
             gen.fep_job(1500, 4, 0.5, seed=10)]
    ref = []
    for s in specs:
        ds = rt.DeviceSystem(s, precision=prec)
        ds.pairlist_update()
        e = ds.md(11)
        ref.append((e, *ds.get_state()))
        ds.close()
    sess = rt.Session(precision=prec, nsnb=5)
    jobs = [sess.admit(s, 11) for s in specs]
    sess.step(11)
    for j, (e, x, v) in zip(jobs, ref):
        assert sess.query(j) == (rt.DONE, 11)
        assert np.array_equal(sess.energies(j), e)
        xx, vv = sess.fetch(j)
        assert np.array_equal(xx, x) and np.array_equal(vv, v)
    sess.close()
