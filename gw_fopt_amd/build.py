"""Builds libbubble_gw_c.so (the gfx950 kernels + the C ABI) in-tree.

Run ``python -m gw_fopt_amd.build``.  The library is written next to this
file, so it travels with the source tree.  Every HIP translation unit is
compiled for gfx950 only, with
  -ffp-contract=off   the collision geometry must round like the reference's
                      Rust code (no implicit a*b+c fusion; FMAs are explicit);
  -Xclang -target-feature -Xclang -packed-fp32-ops
                      no packed fp32 instructions are formed.
"""
from __future__ import annotations

import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
INCLUDE = os.path.join(ROOT, "include")
LIB_NAME = "libbubble_gw_c.so"
LIB_PATH = os.path.join(HERE, LIB_NAME)

SOURCES = ["bgw_kernels.hip", "bgw_ensemble.hip", "bgw_nucleation.hip", "bgw_matrix.hip",
           "bgw_host.cpp"]
# per-source flags: the matrix-core kernels keep their accumulators in VGPRs
# (the AGPR form copies every accumulator out and back around each step)
EXTRA_FLAGS = {"bgw_matrix.hip": ["-mllvm", "-amdgpu-mfma-vgpr-form=1"]}
OFFLOAD_ARCH = "gfx950"


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), shutil.which("hipcc"), "/opt/rocm/bin/hipcc"):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the gfx950 kernels need the ROCm toolchain")


def common_flags() -> list[str]:
    return [
        "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
        "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
        f"-I{CSRC}", f"-I{INCLUDE}",
        "-Wall", "-Wno-unused-function",
    ]


def build(verbose: bool = False, jobs: int = 4) -> str:
    """Compile every source for gfx950 and link the shared library."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    procs = []
    objs = []
    for src in SOURCES:
        obj = os.path.join(objdir, src + ".o")
        objs.append(obj)
        lang = ["-x", "hip", f"--offload-arch={OFFLOAD_ARCH}"]
        cmd = [cc, *lang, *common_flags(), *EXTRA_FLAGS.get(src, []), "-c",
               os.path.join(CSRC, src), "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((src, subprocess.Popen(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT)))
        if len(procs) >= jobs:
            _wait(procs)
            procs = []
    _wait(procs)
    tmp = LIB_PATH + ".tmp"
    cmd = [cc, f"--offload-arch={OFFLOAD_ARCH}", "-shared", "-fPIC", "-o", tmp, *objs,
           "-Wl,-soname," + LIB_NAME]
    if verbose:
        print(" ".join(cmd), flush=True)
    out = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    if out.returncode != 0:
        raise RuntimeError(f"link failed:\n{out.stdout}")
    os.replace(tmp, LIB_PATH)
    return LIB_PATH


def _wait(procs):
    for src, p in procs:
        out, _ = p.communicate()
        text = out.decode(errors="replace")
        lines = [ln for ln in text.splitlines() if "packed-fp32-ops" not in ln and ln.strip()]
        if p.returncode != 0:
            raise RuntimeError(f"hipcc failed on {src}:\n" + "\n".join(lines))


if __name__ == "__main__":
    path = build(verbose="-v" in sys.argv)
    print(path)
