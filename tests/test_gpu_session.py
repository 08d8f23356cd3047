"""GPU tests of the session (continuous batching of independent jobs).

A job's results must not depend on the batch it runs in: each job, admitted at
some clock time into a batch whose other members come and go, is compared with
the same job run alone on a single-job system, and must be bit-identical
(per-job fixed-order reductions, no atomics in the force or energy paths, jobs
on separate 256-atom aligned ranges, pair lists rebuilt on the job's own step
multiples of NSNB).
"""
import importlib

import numpy as np
import pytest

gamd = importlib.import_module("gromos-rs_amd")
rt = gamd.runtime
gen = importlib.import_module("gromos-rs_amd.generators")

pytestmark = pytest.mark.gpu


def alone(s, nsteps, prec, pairs=None):
    ds = rt.DeviceSystem(s, precision=prec)
    ds.pairlist_update()
    p0 = ds.pair_count
    e = ds.md(nsteps)
    if pairs is not None:
        pairs.append(ds.pair_count - p0)
    x, v = ds.get_state()
    ds.close()
    return e, x, v


def job_specs():
    # (system, nsteps, admit clock)
    return [
        (gen.water_216(seed=3), 23, 0),
        (gen.spc_water_box(1200, seed=5), 31, 3),
        (gen.hetero_system(n_chain=240, n_waters=600, box=3.0, seed=4), 17, 10),
        (gen.fep_job(2000, 6, 0.4, seed=3), 26, 12),
        (gen.spc_water_box(300, seed=9), 9, 12),
    ]


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_session_jobs_match_alone(prec):
    specs = job_specs()
    pairs = []
    ref = [alone(s, n, prec, pairs) for s, n, _ in specs]
    sess = rt.Session(precision=prec, nsnb=5)
    ids = [None] * len(specs)
    chunks = [3, 4, 6, 1, 7, 2, 5, 8, 3, 11, 9, 4]
    k = 0
    while True:
        clock = sess.info()["clock"]
        for i, (s, n, at) in enumerate(specs):
            if ids[i] is None and clock >= at:
                ids[i] = sess.admit(s, n)
        if all(j is not None for j in ids) and all(sess.query(j)[0] == rt.DONE for j in ids):
            break
        sess.step(chunks[k % len(chunks)])
        k += 1
        assert k < 200
    info = sess.info()
    assert info["layouts"] >= 3
    # a job that reached its step count waits for the next batch change without
    # being evaluated: the session counts exactly the jobs' own pair interactions
    assert info["pair_count"] == sum(pairs), (info["pair_count"], pairs)
    for i, (s, n, _) in enumerate(specs):
        st, done = sess.query(ids[i])
        assert st == rt.DONE and done == n
        e = sess.energies(ids[i])
        x, v = sess.fetch(ids[i])
        e0, x0, v0 = ref[i]
        assert np.array_equal(e, e0), (i, np.max(np.abs(e - e0)))
        assert np.array_equal(x, x0), (i, np.max(np.abs(x - x0)))
        assert np.array_equal(v, v0), (i, np.max(np.abs(v - v0)))
    sess.close()


def test_session_admission_by_footprint():
    a = gen.spc_water_box(900, seed=1)
    b = gen.spc_water_box(1000, seed=2)
    fa, fb = rt.Session.footprint(a), rt.Session.footprint(b)
    sess = rt.Session(precision=rt.FP64, nsnb=5, hbm_budget=fa + fb // 2)
    ja = sess.admit(a, 10)
    with pytest.raises(rt.Busy):
        sess.admit(b, 10)
    info = sess.info()
    assert info["reserved_bytes"] == fa and info["queued"] == 1
    sess.step(10)
    assert sess.query(ja)[0] == rt.DONE
    # a finished job leaves at the boundary, releasing its reservation
    assert sess.info()["reserved_bytes"] == 0
    jb = sess.admit(b, 5)
    sess.step(5)
    assert sess.query(jb) == (rt.DONE, 5)
    e_alone, x_alone, _ = alone(b, 5, rt.FP64)
    assert np.array_equal(sess.energies(jb), e_alone)
    assert np.array_equal(sess.fetch(jb)[0], x_alone)
    # a job larger than the whole budget is refused outright
    with pytest.raises(rt.GromosAmdError):
        rt.Session(precision=rt.FP64, nsnb=5, hbm_budget=fa // 2).admit(a, 1)
    sess.close()


def test_session_evict_and_idle_clock():
    sess = rt.Session(precision=rt.FP64, nsnb=5)
    sess.step(3)                      # nothing resident: the clock advances
    s = gen.spc_water_box(400, seed=7)
    j = sess.admit(s, 20)
    sess.step(2)                      # joins at clock 5
    assert sess.query(j) == (rt.RUNNING, 0)
    sess.step(6)
    assert sess.query(j) == (rt.RUNNING, 6)
    sess.evict(j)
    sess.step(5)
    assert sess.info()["resident_jobs"] == 0
    with pytest.raises(rt.GromosAmdError):
        sess.query(j)
    sess.close()


def test_replica_exchange_primitives():
    a = gen.water_replica(seed=1, temperature=280.0, n_waters=500)
    b = gen.water_replica(seed=2, temperature=320.0, n_waters=500)
    sess = rt.Session(precision=rt.FP64, nsnb=5)
    ja, jb = sess.admit(a, 100), sess.admit(b, 100)
    sess.step(10)
    xa, va = sess.fetch(ja)
    xb, vb = sess.fetch(jb)
    ke = sess.kinetic_energy([ja, jb])
    m = a.mass[:, None]
    assert np.allclose(ke, [0.5 * np.sum(m * va ** 2), 0.5 * np.sum(m * vb ** 2)], rtol=1e-12)
    sess.swap_configurations(ja, jb)
    xa2, va2 = sess.fetch(ja)
    xb2, vb2 = sess.fetch(jb)
    assert np.array_equal(xa2, xb) and np.array_equal(va2, vb)
    assert np.array_equal(xb2, xa) and np.array_equal(vb2, va)
    sess.scale_velocities([ja, jb], [0.5, 2.0])
    ke2 = sess.kinetic_energy([ja, jb])
    assert np.allclose(ke2, [0.25 * ke[1], 4.0 * ke[0]], rtol=1e-12)
    # continuing after the exchange: job a now runs b's configuration; the same
    # as a single system started from that state
    sess.step(5)
    s2 = gen.water_replica(seed=1, temperature=280.0, n_waters=500)
    s2.pos, s2.vel = xb, vb * 0.5
    e_alone, x_alone, _ = alone(s2, 5, rt.FP64)
    assert np.array_equal(sess.energies(ja, first=10), e_alone)
    assert np.array_equal(sess.fetch(ja)[0], x_alone)
    sess.close()


def test_replica_exchange_run():
    remd = importlib.import_module("gromos-rs_amd.remd")
    temps = [280.0, 300.0, 320.0, 340.0]
    systems = [gen.water_replica(seed=20 + k, temperature=t, n_waters=500) for k, t in enumerate(temps)]
    sess = rt.Session(precision=rt.MIXED, nsnb=5)
    rx = remd.ReplicaExchange(sess, systems, temps, exchange_interval=10, scheme="odd_even", seed=1)
    rx.run(30)
    assert rx.attempts == 2 + 1 + 2
    assert [h[1:3] for h in rx.history[:3]] == [(0, 1), (2, 3), (1, 2)]
    for j in rx.jobs:
        assert sess.query(j)[1] == 30
        e = sess.energies(j)
        assert np.all(np.isfinite(e)) and e.shape == (30, 5)
    sess.close()


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_remd_batch_128_replicas_match_alone(prec):
    """The replica-exchange batch at its full shape: 128 replicas of 4,000 SPC
    waters (12,000 atoms, rc 1.2 nm) at T_k = 280 (450/280)^(k/127) K in one
    session; sampled replicas are bit-identical to the same replica run alone."""
    n, steps = 128, 12
    temps = [280.0 * (450.0 / 280.0) ** (k / (n - 1)) for k in range(n)]
    systems = [gen.water_replica(seed=100 + k, temperature=temps[k], n_waters=4000, rc=1.2) for k in range(n)]
    sess = rt.Session(precision=prec, nsnb=5)
    jobs = [sess.admit(s, steps) for s in systems]
    sess.step(steps)
    assert sess.info()["resident_jobs"] == n or all(sess.query(j)[0] == rt.DONE for j in jobs)
    for k in (0, 61, 127):
        st, done = sess.query(jobs[k])
        assert st == rt.DONE and done == steps
        e0, x0, v0 = alone(systems[k], steps, prec)
        assert np.array_equal(sess.energies(jobs[k]), e0), k
        x, v = sess.fetch(jobs[k])
        assert np.array_equal(x, x0) and np.array_equal(v, v0), k
    sess.close()


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_failed_job_is_isolated(prec):
    """A job whose SETTLE solve fails (one hydrogen moved 1 nm from its oxygen:
    sin(angle) > 1 in the analytical solve, constraints.rs:896-951) is marked
    FAILED with the reason; the jobs batched with it finish and are
    bit-identical to the same jobs run alone."""
    good = [gen.spc_water_box(600, seed=31), gen.fep_job(1500, 5, 0.3, seed=32)]
    bad = gen.spc_water_box(500, seed=33)
    bad.pos[bad.n_solute_atoms + 3 * 40 + 1] += np.array([1.0, 0.0, 0.0])   # the oracle fails at step 1 here
    steps = 15
    sess = rt.Session(precision=prec, nsnb=5)
    jg = [sess.admit(s, steps) for s in good]
    jb = sess.admit(bad, steps)
    sess.step(steps)
    st, done = sess.query(jb)
    assert st == rt.FAILED and done < steps
    assert "SETTLE" in rt.lib().gromos_amd_last_error().decode()
    for j, s in zip(jg, good):
        assert sess.query(j) == (rt.DONE, steps)
        e0, x0, v0 = alone(s, steps, prec)
        assert np.array_equal(sess.energies(j), e0)
        x, v = sess.fetch(j)
        assert np.array_equal(x, x0) and np.array_equal(v, v0)
    # the failed job leaves at the next boundary; the session goes on
    j2 = sess.admit(gen.spc_water_box(300, seed=34), 5)
    sess.step(5)
    assert sess.query(j2) == (rt.DONE, 5)
    sess.close()


@pytest.mark.parametrize("prec", [rt.FP64, rt.MIXED])
def test_diverged_force_fails_only_its_job(prec):
    """A job whose forces blow up (one water moved onto another, first atoms
    1e-4 nm apart: the O-O repulsion is ~1e47 kJ/mol/nm, infinite in FP32) is
    marked FAILED for a diverged force at its first step (the fixed-point
    conversions saturate instead of wrapping); the jobs batched with it finish
    bit-identical to the same jobs run alone."""
    good = [gen.spc_water_box(600, seed=51), gen.fep_job(1500, 5, 0.3, seed=52)]
    bad = gen.spc_water_box(500, seed=53)
    a, b = bad.n_solute_atoms + 3 * 10, bad.n_solute_atoms + 3 * 11
    bad.pos[b:b + 3] += bad.pos[a] + np.array([1e-4, 0.0, 0.0]) - bad.pos[b]
    steps = 10
    sess = rt.Session(precision=prec, nsnb=5)
    jg = [sess.admit(s, steps) for s in good]
    jb = sess.admit(bad, steps)
    sess.step(steps)
    st, done = sess.query(jb)
    assert st == rt.FAILED and done == 0
    assert "force" in rt.lib().gromos_amd_last_error().decode()
    for j, s in zip(jg, good):
        assert sess.query(j) == (rt.DONE, steps)
        e0, x0, v0 = alone(s, steps, prec)
        assert np.array_equal(sess.energies(j), e0)
        x, v = sess.fetch(j)
        assert np.array_equal(x, x0) and np.array_equal(v, v0)
    sess.close()
    # the single-system entry reports the same as an error
    ds = rt.DeviceSystem(bad, precision=prec)
    ds.pairlist_update()
    with pytest.raises(rt.GromosAmdError, match="force"):
        ds.md(3)
    ds.close()


def test_layout_failure_leaves_session_consistent(monkeypatch):
    """A batch layout that runs out of device memory (forced from the host: a
    layout of more atoms than GROMOS_AMD_LAYOUT_ATOM_LIMIT fails before any
    allocation): session_step returns the error, the continuing job stops as
    FAILED with its last state fetchable, the queued job stays queued with its
    reservation, and once memory allows it runs and matches the same job alone."""
    a = gen.spc_water_box(600, seed=41)
    b = gen.spc_water_box(500, seed=42)
    sess = rt.Session(precision=rt.FP64, nsnb=5)
    ja = sess.admit(a, 20)
    sess.step(5)
    x5, v5 = sess.fetch(ja)
    jb = sess.admit(b, 10)
    reserved = sess.info()["reserved_bytes"]
    monkeypatch.setenv("GROMOS_AMD_LAYOUT_ATOM_LIMIT", str(a.n_atoms + b.n_atoms - 1))
    with pytest.raises(rt.GromosAmdError):
        sess.step(5)
    st, done = sess.query(ja)
    assert st == rt.FAILED and done == 5
    assert "layout" in rt.lib().gromos_amd_last_error().decode()
    x, v = sess.fetch(ja)
    assert np.array_equal(x, x5) and np.array_equal(v, v5)
    assert sess.query(jb) == (rt.QUEUED, 0)
    info = sess.info()
    assert info["resident_jobs"] == 0 and info["reserved_bytes"] == reserved - rt.Session.footprint(a)
    monkeypatch.delenv("GROMOS_AMD_LAYOUT_ATOM_LIMIT")
    sess.step(10)
    assert sess.query(jb) == (rt.DONE, 10)
    e0, x0, _ = alone(b, 10, rt.FP64)
    assert np.array_equal(sess.energies(jb), e0)
    assert np.array_equal(sess.fetch(jb)[0], x0)
    sess.close()


def test_admission_refuses_lists_beyond_32bit_offsets():
    """A job whose neighbour lists would not fit the kernels' 32-bit list
    offsets (a vacuum system of 50,000 single-atom charge groups: every pair is
    a candidate, about 3e9 entries) is refused at admission, before anything
    reaches the device."""
    n = 50000
    rng = np.random.default_rng(1)
    s = gamd.system.System(n_atoms=n, n_solute_atoms=n, cg_start=np.arange(n + 1, dtype=np.int32), n_solute_cg=n,
                           charge=np.zeros(n), iac=np.zeros(n, np.int32), mass=np.ones(n), n_types=1,
                           lj=np.array([1e-3, 1e-6, 1e-3, 1e-6]), excl_off=np.zeros(n + 1, np.int32),
                           excl=np.zeros(0, np.int32), o14_off=np.zeros(n + 1, np.int32), o14=np.zeros(0, np.int32),
                           atoms_per_solvent=0, pos=rng.uniform(0, 20, (n, 3)), vel=np.zeros((n, 3)),
                           box=np.zeros(3), boundary=0).validate()
    sess = rt.Session(precision=rt.FP64, nsnb=5)
    with pytest.raises(rt.GromosAmdError, match="32-bit"):
        sess.admit(s, 10)
    assert sess.info()["reserved_bytes"] == 0
    sess.close()
