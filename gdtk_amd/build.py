"""Build recipe of the HIP library libgdtkhip.so (in-tree, gfx950).

Run as `python -m gdtk_amd.build` or through __graft_entry__.build().  The library is
written next to this file so that it travels with the repository tree.
"""
from __future__ import annotations

import glob
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
CSRC = os.path.join(HERE, "csrc")
LIB = os.path.join(HERE, "libgdtkhip.so")
ARCH = os.environ.get("GDTK_OFFLOAD_ARCH", "gfx950")

# -ffp-contract=off keeps every expression in the written (reference) order, no FMA
# contraction.  The packed-fp32 feature is turned off for all device code.
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC", "-ffp-contract=off",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wall", "-Wno-unused-function", "-Wno-unused-variable",
]


def hipcc() -> str:
    for cand in (os.environ.get("HIPCC"), "/opt/rocm/bin/hipcc", shutil.which("hipcc")):
        if cand and os.path.exists(cand):
            return cand
    raise RuntimeError("hipcc not found: the HIP library cannot be built")


def _up_to_date(obj: str, dep: str) -> bool:
    """True if obj exists and is newer than every file its dependency list names."""
    if not (os.path.exists(obj) and os.path.exists(dep)):
        return False
    with open(dep) as f:
        text = f.read().replace("\\\n", " ")
    files = text.split(":", 1)[1].split() if ":" in text else []
    t = os.path.getmtime(obj)
    return bool(files) and all(os.path.exists(x) and os.path.getmtime(x) < t for x in files)


def build(verbose: bool = False, incremental: bool = False) -> str:
    """Compile every source and link the library.  incremental=True skips a source whose
    object is newer than every file the compiler's dependency list for it names (a
    developer convenience; the default compiles everything)."""
    cc = hipcc()
    objdir = os.path.join(HERE, "_build")
    os.makedirs(objdir, exist_ok=True)
    objs = []
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip"))) + sorted(glob.glob(os.path.join(CSRC, "*.cpp")))
    procs = []
    for src in srcs:
        obj = os.path.join(objdir, os.path.basename(src) + ".o")
        dep = obj[:-2] + ".d"
        if incremental and _up_to_date(obj, dep):
            objs.append(obj)
            continue
        lang = ["-x", "hip"] if src.endswith(".hip") else []
        cmd = [cc] + HIP_FLAGS + lang + ["-MMD", "-MF", dep, "-c", src, "-o", obj]
        if verbose:
            print(" ".join(cmd), flush=True)
        procs.append((subprocess.Popen(cmd, stderr=subprocess.PIPE, text=True), src))
        objs.append(obj)
    failed = []
    for p, src in procs:
        _, err = p.communicate()
        if p.returncode != 0:
            failed.append((src, err))
        elif verbose and err.strip():
            print(err, file=sys.stderr)
    if failed:
        msg = "\n".join(f"--- {s}\n{e}" for s, e in failed)
        raise RuntimeError(f"HIP build failed:\n{msg}")
    tmp = LIB + ".tmp"
    cmd = [cc, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", tmp] + objs + ["-lz", "-pthread"]
    subprocess.run(cmd, check=True)
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    print(build(verbose=True, incremental="--incremental" in sys.argv))
