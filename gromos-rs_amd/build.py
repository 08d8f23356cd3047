"""Build recipe of the HIP library (libgromos_amd.so, in-tree, gfx950 only).

    python -m importlib ... or: python gromos-rs_amd/build.py

Flags:
  -ffp-contract=off            no fused multiply-add: the FP64 path restates the
                               reference arithmetic operation by operation;
  -Xclang -target-feature -Xclang -packed-fp32-ops
                               keep the compiler from emitting packed FP32 ops.
"""
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgromos_amd.so")
SOURCES = ["explicit.hip", "bonded.hip", "system_kernels.hip", "system_host.hip", "capi.hip", "session.hip"]
ARCH = os.environ.get("GROMOS_AMD_ARCH", "gfx950")


def hipcc():
    for c in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if c and os.path.exists(c):
            return c
    raise RuntimeError("hipcc not found (ROCm is required to build libgromos_amd.so)")


def flags():
    return ["-O3", "-std=c++17", f"--offload-arch={ARCH}", "-fPIC", "-ffp-contract=off",
            "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
            "-Wno-unused-command-line-argument", "-I", os.path.join(HERE, "..", "include")]


def build(verbose=False, jobs=None):
    """Compile every source to an object in parallel, then link the shared library."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for s in SOURCES:
        o = os.path.join(objdir, s.replace(".hip", ".o"))
        objs.append(o)
        cmd = [cc] + flags() + ["-c", os.path.join(CSRC, s), "-o", o]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((s, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
    failed = []
    for s, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        # the host pass does not know the packed-fp32 target feature: harmless
        lines = [l for l in text.splitlines() if "packed-fp32-ops" not in l and "warning generated" not in l
                 and "warnings generated" not in l and l.strip()]
        if p.returncode != 0:
            failed.append(s)
            print(text, file=sys.stderr)
        elif verbose and lines:
            print("\n".join(lines))
    if failed:
        raise RuntimeError(f"hipcc failed for {failed}")
    tmp = LIB + ".tmp"
    cmd = [cc, "--offload-arch=" + ARCH, "-shared", "-fPIC", "-o", tmp] + objs
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
