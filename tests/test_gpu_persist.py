"""GPU tests of the far-pair solvent kernel's persistent form
(GROMOS_AMD_VV_PERSIST=1: resident waves take logical waves from per-XCD
counters): bit-identical to the one-launch-slot-per-wave form for a water box,
a batch of uneven FEP jobs (chunks of every size, a job boundary inside
resident waves' work), over repeated launches and successive systems (the
counters return to zero after every launch)."""
import importlib

import numpy as np
import pytest

gamd = importlib.import_module("gromos-rs_amd")
rt = gamd.runtime
gen = importlib.import_module("gromos-rs_amd.generators")

pytestmark = pytest.mark.gpu


def water_md(s, n, persist, monkeypatch):
    monkeypatch.setenv("GROMOS_AMD_VV_PERSIST", "1" if persist else "0")
    ds = rt.DeviceSystem(s, precision=rt.MIXED)
    ds.pairlist_update()
    e = ds.md(n)
    x, v = ds.get_state()
    sp = ds.forcefield_apply()
    ds.close()
    return e, x, v, sp["forces"]


@pytest.mark.parametrize("n_waters", [3000, 20000])
def test_persistent_form_matches_bitwise(n_waters, monkeypatch):
    s = gen.spc_water_box(n_waters, seed=4)
    ref = water_md(s, 12, False, monkeypatch)
    for _ in range(2):   # a second system after the first: the counters were cleared
        got = water_md(s, 12, True, monkeypatch)
        for a, b in zip(got, ref):
            assert np.array_equal(a, b)


def test_persistent_form_in_a_batch_matches_bitwise(monkeypatch):
    jobs = [gen.fep_job(n, 5, lam, seed=k) for k, (n, lam) in enumerate(((1500, 0.0), (7000, 0.35), (2600, 1.0)))]
    out = {}
    for persist in (False, True):
        monkeypatch.setenv("GROMOS_AMD_VV_PERSIST", "1" if persist else "0")
        sess = rt.Session(precision=rt.MIXED, nsnb=5)
        ids = [sess.admit(j, 11) for j in jobs]
        k = 0
        while not all(sess.query(i)[0] == rt.DONE for i in ids):
            sess.step(5)
            k += 1
            assert k < 10
        out[persist] = [(sess.energies(i), *sess.fetch(i)) for i in ids]
        sess.close()
    for a, b in zip(out[True], out[False]):
        for x, y in zip(a, b):
            assert np.array_equal(x, y)
