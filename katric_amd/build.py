"""Build recipe for the katric_amd device library (gfx950).

Compiles every HIP source under katric_amd/csrc for gfx950 (hipcc -c, in
parallel), the host-only graph input layer with g++, and links them into one
shared library, katric_amd/_lib/libkatric_amd.so.  After linking, the
offload bundles embedded in the library are checked: every one must target
gfx950.  Run as ``python -m katric_amd.build``.
"""
import os
import re
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
OBJDIR = os.path.join(HERE, "_build")
LIBDIR = os.path.join(HERE, "_lib")
LIB = os.path.join(LIBDIR, "libkatric_amd.so")
INCLUDE = os.path.normpath(os.path.join(HERE, "..", "include"))
ARCH = "gfx950"

HIP_FLAGS = [
    f"--offload-arch={ARCH}",
    "-O3",
    "-std=c++17",
    "-fPIC",
    "-Wall",
    "-Wno-unused-function",
    # keep the compiler from forming packed-FP32 instructions
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-I", INCLUDE,
]
CXX_FLAGS = ["-O3", "-std=c++17", "-fPIC", "-Wall", "-I", INCLUDE]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found")


def _run(cmd, verbose):
    if verbose:
        print(" ".join(cmd), flush=True)
    res = subprocess.run(cmd, capture_output=True, text=True)
    if res.returncode != 0:
        sys.stderr.write(" ".join(cmd) + "\n" + res.stdout + res.stderr)
        raise RuntimeError(f"build step failed: {cmd[0]} {cmd[-1]}")
    return res


def bundle_targets(path):
    """Offload-bundle target ids found in a built shared object."""
    with open(path, "rb") as f:
        blob = f.read()
    return sorted(set(m.decode() for m in re.findall(rb"hipv4-amdgcn-amd-amdhsa--[a-z0-9:+-]*", blob)))


def build(verbose: bool = False) -> str:
    os.makedirs(OBJDIR, exist_ok=True)
    os.makedirs(LIBDIR, exist_ok=True)
    hip = sorted(f for f in os.listdir(CSRC) if f.endswith(".hip"))
    cpp = sorted(f for f in os.listdir(CSRC) if f.endswith(".cpp"))
    jobs = []
    for f in hip:
        obj = os.path.join(OBJDIR, f[:-4] + ".o")
        jobs.append([hipcc()] + HIP_FLAGS + ["-c", os.path.join(CSRC, f), "-o", obj])
    for f in cpp:
        obj = os.path.join(OBJDIR, f[:-4] + ".o")
        jobs.append(["g++"] + CXX_FLAGS + ["-c", os.path.join(CSRC, f), "-o", obj])
    with ThreadPoolExecutor(max_workers=min(8, len(jobs))) as ex:
        list(ex.map(lambda c: _run(c, verbose), jobs))
    objs = [j[-1] for j in jobs]
    tmp = LIB + ".tmp"
    _run([hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lpthread"], verbose)
    targets = bundle_targets(tmp)
    if not targets or any(not t.endswith("--" + ARCH) for t in targets):
        raise RuntimeError(f"device code bundles {targets} do not all target {ARCH}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True))
    print("bundles:", bundle_targets(LIB))
