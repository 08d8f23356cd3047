"""End to end from GROMOS files on the device: the product's readers, an md run
(gromos_io.run_md) and its trajectory writers, against the gromosXX fixture's
expected output (10 significant figures for energies, 9 decimals for forces)."""
import importlib
import os

import numpy as np
import pytest

import golden_io

gamd = importlib.import_module("gromos-rs_amd")
gio = importlib.import_module("gromos-rs_amd.gromos_io")

pytestmark = pytest.mark.gpu

INPUTS = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden", "gromos")


def files(name):
    d = os.path.join(INPUTS, name)
    fs = os.listdir(d)
    pick = lambda *ext: next((os.path.join(d, f) for f in fs if f.endswith(ext)), None)
    return pick(".topo", ".top"), pick(".conf", ".cnf"), pick(".in"), pick(".ptp")


def test_settle_run_writes_the_fixture_trajectory(tmp_path):
    """nacl_1water_settle: 10 md steps with SETTLE; every written energy frame and
    the step-0 forces equal gromosXX's."""
    top, conf, imd, _ = files("nacl_1water_settle")
    out = str(tmp_path / "run")
    s, e = gio.run_md(top, conf, imd, out)
    _, d = golden_io.load("nacl_1water_settle")
    tre = gio.read_energy_trajectory(out + ".tre")
    assert len(tre) == len(d["expected_energies"])
    for k, f in enumerate(d["expected_energies"]):
        assert abs(tre[k, 2] - f["e_pot"]) <= 6e-10 * abs(f["e_pot"]), (k, tre[k, 2], f["e_pot"])
        assert abs(tre[k, 1] - f["e_kin"]) <= 6e-10 * abs(f["e_kin"]), (k, tre[k, 1], f["e_kin"])
    (step, f0), = gio.read_force_trajectory(out + ".trf")
    ref = np.array(d["expected_forces"][0]["f"]).reshape(-1, 3)
    assert step == 0 and np.abs(f0 - ref).max() <= 2e-9
    c = gio.read_configuration(out + ".cnf")
    assert c.pos.shape == (s.n_atoms, 3) and np.all(np.isfinite(c.pos))


@pytest.mark.parametrize("name", ["water_10_box", "nacl_water_box", "ch4_water_fep"])
def test_single_point_from_files(name, tmp_path):
    """step-0 energies and forces of a run started from the fixture's files"""
    top, conf, imd, ptp = files(name)
    out = str(tmp_path / name)
    s, e = gio.run_md(top, conf, imd, out, ptp=ptp, nsteps=1)
    _, d = golden_io.load(name)
    e0 = d["expected_energies"][0]
    tre = gio.read_energy_trajectory(out + ".tre")
    assert abs(tre[0, 11] - e0["e_lj"]) <= 6e-10 * abs(e0["e_lj"])
    assert abs(tre[0, 12] - e0["e_crf"]) <= 6e-10 * abs(e0["e_crf"])
    (step, f0), = gio.read_force_trajectory(out + ".trf")
    ref = np.array(d["expected_forces"][0]["f"]).reshape(-1, 3)
    assert np.abs(f0 - ref).max() <= 2e-9
