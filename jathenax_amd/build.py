"""Build libjathenax_amd.so in-tree: hipcc for gfx950 device code, g++ for the seed path.

Usage: python -m jathenax_amd.build [-j JOBS] [-v]

FP64 translation units are compiled with -ffp-contract=off so that every multiply and
add rounds as in the reference's IEEE float64 expressions.  The FP32 stage and batch units
use -ffp-contract=on: a product is fused only into an add of the same source expression, so
which operations are fused is fixed by the source text.  With "fast" the backend also fuses
across statements, and the result then depends on how the row loop is unrolled and so on
the tile shape.  All HIP code is built with packed FP32 instructions disabled.
"""

import argparse
import os
import shutil
import subprocess
import sys
from concurrent.futures import ThreadPoolExecutor
from pathlib import Path

PKG = Path(__file__).resolve().parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libjathenax_amd.so"
ARCH = os.environ.get("JX_OFFLOAD_ARCH", "gfx950")

HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument", "-Wno-unused-value",
]


def _hipcc():
    h = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
    if not Path(h).exists():
        raise RuntimeError("hipcc not found: the HIP toolchain is required to build")
    return h


def _jobs(n=None):
    if n:
        return n
    return max(1, min(16, os.cpu_count() or 1))


def compile_commands():
    cmds = []
    for src in sorted(CSRC.glob("*.hip")):
        contract = "on" if src.name.startswith(("stage_f32_", "batch_f32")) else "off"
        obj = OBJ / (src.stem + ".o")
        cmds.append((src, obj, [_hipcc(), *HIP_FLAGS, f"-ffp-contract={contract}", "-c",
                                str(src), "-o", str(obj)]))
    for src in sorted(CSRC.glob("*.cpp")):
        obj = OBJ / (src.stem + ".o")
        cmds.append((src, obj, ["g++", "-O2", "-fPIC", "-std=c++17", "-ffp-contract=off",
                                "-c", str(src), "-o", str(obj)]))
    return cmds


def build(jobs=None, verbose=False, only=None):
    """Compile every translation unit (or, with `only`, just the named source stems, reusing
    the other objects of an earlier build) and link the library."""
    OBJ.mkdir(exist_ok=True)
    cmds = compile_commands()
    if only:
        missing = [obj for src, obj, _ in cmds if src.stem not in only and not obj.exists()]
        if missing:
            raise RuntimeError(f"no earlier object for {missing[0].name}: run a full build")
        # Every unit includes the shared headers (kernel argument structs among them), so an
        # object older than any header may disagree with the rebuilt units on a layout.
        headers = [*CSRC.glob("*.hpp"), *(PKG.parent / "include").glob("*.h")]
        newest = max((h.stat().st_mtime for h in headers), default=0.0)
        stale = [obj for src, obj, _ in cmds
                 if src.stem not in only and obj.stat().st_mtime < newest]
        if stale:
            raise RuntimeError(f"{stale[0].name} is older than a shared header: run a full "
                               "build")
        keep = [(src, obj, None) for src, obj, _ in cmds if src.stem not in only]
        cmds = [c for c in cmds if c[0].stem in only]

    def run(item):
        src, obj, cmd = item
        if verbose:
            print(" ".join(cmd), flush=True)
        res = subprocess.run(cmd, capture_output=True, text=True)
        if res.returncode != 0:
            raise RuntimeError(f"compiling {src.name} failed:\n{res.stdout}\n{res.stderr}")
        return obj

    with ThreadPoolExecutor(_jobs(jobs)) as pool:
        objs = list(pool.map(run, cmds))
    if only:
        objs = sorted(objs + [obj for _, obj, _ in keep])
    tmp = LIB.with_suffix(".so.tmp")
    link = [_hipcc(), f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp),
            *map(str, objs)]
    if verbose:
        print(" ".join(link), flush=True)
    res = subprocess.run(link, capture_output=True, text=True)
    if res.returncode != 0:
        raise RuntimeError(f"linking failed:\n{res.stdout}\n{res.stderr}")
    os.replace(tmp, LIB)
    return LIB


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("-j", "--jobs", type=int)
    ap.add_argument("-v", "--verbose", action="store_true")
    ap.add_argument("--only", nargs="+", metavar="STEM",
                    help="recompile only these sources (e.g. jx_capi), keep the other objects")
    args = ap.parse_args()
    print(build(args.jobs, args.verbose, args.only))
    sys.exit(0)
