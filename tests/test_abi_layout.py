"""The C ABI's structs as the Python mirror (gromos-rs_amd/runtime.py) lays them
out: every struct of include/gromos_amd.h has the same size and every field the
same offset in ctypes as in a C compiler's layout of the header (gcc on the host;
no GPU)."""
import ctypes
import importlib
import os
import shutil
import subprocess

import pytest

rt = importlib.import_module("gromos-rs_amd.runtime")
ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))

STRUCTS = {
    "gromos_amd_crf": rt.Crf,
    "gromos_amd_lambda": rt.Lambda,
    "gromos_amd_pert": rt.Pert,
    "gromos_amd_bonded": rt.Bonded,
    "gromos_amd_topology": rt.Topology,
    "gromos_amd_params": rt.Params,
    "gromos_amd_session_params": rt.SessionParams,
    "gromos_amd_session_info": rt.SessionInfo,
}
C_NAME = {"lam": "lambda"}   # a Python keyword on the C side


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs a host C compiler")
def test_struct_layouts_match_the_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "gromos_amd.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for f in py._fields_:
            lines.append(f'printf("{cname} {f[0]} %zu\\n", offsetof({cname}, {C_NAME.get(f[0], f[0])}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, py in STRUCTS.items():
        assert got[(cname, "size")] == ctypes.sizeof(py), cname
        for f in py._fields_:
            assert got[(cname, f[0])] == getattr(py, f[0]).offset, (cname, f[0])
