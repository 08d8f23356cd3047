"""CPU-side checks of the HIP library: it loads, exports every entry point the
header declares, and its host-side helpers agree with the oracle."""
import importlib
import os
import re

import pytest

import oracle

gamd = importlib.import_module("gromos-rs_amd")
rt = gamd.runtime
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_symbols():
    names = set()
    for f in os.listdir(os.path.join(ROOT, "include")):
        if f.endswith(".h"):
            text = open(os.path.join(ROOT, "include", f)).read()
            names |= set(re.findall(r"\b(gromos_amd_[a-z0-9_]+)\s*\(", text))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    lib = rt.lib()
    syms = declared_symbols()
    assert len(syms) >= 20
    for s in syms:
        assert hasattr(lib, s), s


def test_crf_new_matches_oracle():
    for rc, eps, kappa in [(1.4, 61.0, 0.0), (0.9, 62.0, 0.0), (1.2, 78.5, 0.3), (1.4, 1.0, 0.0)]:
        a = rt.crf_new(rc, 1.0, eps, kappa)
        b = oracle.crf_new(rc, 1.0, eps, kappa)
        for k in ("crf_cut", "crf_2cut3i", "crf_cut3i", "cutoff_sq"):
            assert getattr(a, k) == getattr(b, k), k


@pytest.mark.parametrize("lam,n", [(0.0, 1), (0.35, 1), (0.5, 2), (1.0, 3)])
def test_lambda_from_matches_oracle(lam, n):
    a = rt.lambda_from(lam, n)
    b = oracle.lambda_params(lam, n)
    for k, _ in oracle.Lambda._fields_:
        assert getattr(a, k) == getattr(b, k), k


def test_errors_are_reported_not_crashing():
    # bad arguments are rejected with a message before any device work
    with pytest.raises(rt.GromosAmdError, match="cutoff"):
        rt.crf_new(0.0)


def test_job_footprint_host_side():
    """admission estimate: computed on the host from the topology alone"""
    gen = importlib.import_module("gromos-rs_amd.generators")
    small, big = gen.spc_water_box(1000, seed=1), gen.spc_water_box(8000, seed=1)
    fs, fb = rt.Session.footprint(small), rt.Session.footprint(big)
    assert 0 < fs < fb
    # per-atom arrays plus neighbour entries at the same density: close to linear
    assert 6.0 < fb / fs < 10.0
    # footprint of the 98k-atom water box (per-atom arrays, lists with their
    # growth allowance) stays below 3 KB per atom
    assert rt.Session.footprint(gen.water_32768()) < 3072 * 98304
    bad = gen.spc_water_box(10, seed=1)
    bad.iac = bad.iac.copy()
    bad.iac[0] = bad.n_types       # type out of range
    with pytest.raises(rt.GromosAmdError, match="type out of range"):
        rt.Session.footprint(bad)


def test_rust_binding_declares_every_entry_point():
    """The Rust extern block a maintainer would add binds every C entry point."""
    text = open(os.path.join(ROOT, "INTEGRATION.md")).read()
    block = text[text.index('#[link(name = "gromos_amd")]'):]
    block = block[:block.index("\n}\n")]
    bound = set(re.findall(r"pub fn (gromos_amd_[a-z0-9_]+)\(", block))
    missing = [s for s in declared_symbols() if s not in bound]
    assert not missing, missing
