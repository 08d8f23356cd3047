"""Load the golden gromosXX vectors (tests/golden/*.json) into System objects."""
import importlib
import json
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
GOLDEN = os.path.join(HERE, "golden")

gamd = importlib.import_module("gromos-rs_amd")


def names():
    return sorted(f[:-5] for f in os.listdir(GOLDEN) if f.endswith(".json"))


def load_json(name):
    with open(os.path.join(GOLDEN, name + ".json")) as fh:
        return json.load(fh)


def _bonded(d):
    b = d.get("bonded")
    if not b or not (b["bonds"] or b["angles"] or b["impropers"] or b["dihedrals"]):
        return None
    return gamd.system.Bonded.of(b["bonds"], b["bond_types"], b["angles"], b["angle_types"], b["impropers"],
                                 b["improper_types"], b["dihedrals"], b["dihedral_types"],
                                 b.get("ntf", (1, 1, 1, 1)))


def load(name):
    d = load_json(name)
    n = d["n_atoms"]
    excl_off, excl = gamd.system.csr_from_lists(d["excl"], n)
    o14_off, o14 = gamd.system.csr_from_lists(d["o14"], n)
    vel = np.array(d["vel"], float).reshape(n, 3) if d["vel"] else np.zeros((n, 3))
    d_oh, d_hh = 0.1, 0.163299
    for i, j, l in d["solvent_constr"]:
        if {i, j} == {1, 2}:
            d_hh = l
        else:
            d_oh = l
    pert = None
    if d["pert"]:
        p = d["pert"]
        pert = gamd.system.Perturbation.empty(n)
        for a in p["atoms"]:
            k = a["seq"]
            pert.flag[k] = 1
            pert.a_iac[k], pert.b_iac[k] = a["a_iac"], a["b_iac"]
            pert.a_q[k], pert.b_q[k] = a["a_q"], a["b_q"]
            pert.alpha_lj[k] = a["alj"] * p["alphlj"]
            pert.alpha_crf[k] = a["acrf"] * p["alphc"]
        pert.lam, pert.nlam = p["rlam"], p["nlam"]
        pert.atom_pairs = np.array(p["atom_pairs"], np.int32).reshape(-1, 4)
    s = gamd.system.System(
        n_atoms=n, n_solute_atoms=d["n_solute_atoms"], cg_start=np.array(d["cg_start"], np.int32),
        n_solute_cg=d["n_solute_cg"], charge=np.array(d["charge"], float),
        iac=np.array(d["iac"], np.int32), mass=np.array(d["mass"], float), n_types=d["n_types"],
        lj=np.array(d["lj"], float), excl_off=excl_off, excl=excl, o14_off=o14_off, o14=o14,
        atoms_per_solvent=d["atoms_per_solvent"], pos=np.array(d["pos"], float).reshape(n, 3),
        vel=vel, box=np.array(d["box"], float), boundary=d["ntb"], rcutp=d["rcutp"],
        rcutl=d["rcutl"], nsnb=d["nsnb"], algorithm=d["algorithm"], epsrf=d["epsrf"],
        kappa=d["appak"], fpepsi=d["fpepsi"], dt=d["dt"], dist_oh=d_oh, dist_hh=d_hh, pert=pert,
        name=name,
        box_matrix=(np.array(d["box_matrix"], float).reshape(3, 3).T if d.get("box_matrix") else None),
        bonded=_bonded(d))
    return s.validate(), d
