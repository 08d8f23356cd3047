"""The allocation-free variant of the pair-list kernel entry point
(gromos_amd_lj_crf_innerloop_noalloc) against the regular one, through the C
ABI: energies and virial bit for bit (both fold the same per-block partial sums
in the same order), forces within 1e-10 per atom (FP64 atomics add in
scheduling order); two calls in a row on one device, the second smaller, so
the second call reads only what it wrote into the shared static array."""
import ctypes as C
import importlib

import numpy as np
import pytest

gamd = importlib.import_module("gromos-rs_amd")
rt = gamd.runtime
gen = importlib.import_module("gromos-rs_amd.generators")

pytestmark = pytest.mark.gpu


def run(fn, s, pairs):
    import torch
    lib = rt.lib()
    pos = torch.tensor(s.pos, device="cuda")
    q = torch.tensor(s.charge, device="cuda")
    iac = torch.tensor(s.iac.astype(np.int32), device="cuda")
    pl = torch.tensor(pairs.astype(np.int32), device="cuda").contiguous()
    lj = torch.tensor(s.lj, device="cuda")
    f = torch.zeros((s.n_atoms, 3), dtype=torch.float64, device="cuda")
    e = torch.zeros(2, dtype=torch.float64, device="cuda")
    v = torch.zeros(9, dtype=torch.float64, device="cuda")
    crf = rt.crf_new(s.rcutl, 1.0, s.epsrf, s.kappa)
    box = (rt.D * 3)(*s.box)
    rc = getattr(lib, fn)(rt.ptr(pos), rt.ptr(q), rt.ptr(iac), rt.ptr(pl), len(pairs), rt.ptr(lj), s.n_types, crf,
                          s.boundary, box, s.fpepsi, 1, rt.ptr(f), rt.ptr(e), rt.ptr(v),
                          torch.cuda.current_stream().cuda_stream)
    rt.check(rc)
    torch.cuda.synchronize()
    return f.cpu().numpy(), e.cpu().numpy(), v.cpu().numpy()


def all_pairs_within(s, rc):
    x = s.pos
    i, j = np.triu_indices(s.n_atoms, 1)
    d = x[i] - x[j]
    d -= s.box * np.round(d / s.box)
    keep = (d * d).sum(1) <= rc * rc
    return np.stack([i[keep], j[keep]], 1).astype(np.uint32)


@pytest.mark.parametrize("sizes", [(900, 300), (1500, 40)])
def test_noalloc_matches_regular_entry_point(sizes):
    for k, n_w in enumerate(sizes):
        s = gen.spc_water_box(n_w, seed=50 + k)
        pairs = all_pairs_within(s, 1.0)
        assert len(pairs) > 256 * 8
        fa, ea, va = run("gromos_amd_lj_crf_innerloop", s, pairs)
        fb, eb, vb = run("gromos_amd_lj_crf_innerloop_noalloc", s, pairs)
        assert np.array_equal(ea, eb), (ea, eb)
        assert np.array_equal(va, vb)
        d = np.linalg.norm(fa - fb, axis=1) / np.maximum(np.linalg.norm(fa, axis=1), 1.0)
        assert d.max() < 1e-10


def test_noalloc_refuses_beyond_its_limit():
    """More pairs than the static array holds: an error before any launch."""
    import torch
    lib = rt.lib()
    s = gen.spc_water_box(10, seed=1)
    crf = rt.crf_new(1.4)
    box = (rt.D * 3)(*s.box)
    dummy = torch.zeros(16, dtype=torch.float64, device="cuda")
    rc = lib.gromos_amd_lj_crf_innerloop_noalloc(rt.ptr(dummy), rt.ptr(dummy), rt.ptr(dummy), rt.ptr(dummy),
                                                 65536 * 256 + 1, rt.ptr(dummy), 1, crf, 1, box, 138.9354, 0,
                                                 rt.ptr(dummy), rt.ptr(dummy), None, None)
    assert rc == -1
    assert b"limit" in lib.gromos_amd_last_error()
