"""Build libaero_hip.so (gfx950) in-tree with hipcc.

Each translation unit under aero_amd/csrc/ is compiled to an object in
parallel, then linked into aero_amd/libaero_hip.so.  No CUDA, no JIT cache:
the shared library sits next to this file so it travels with the source tree.

    python -m aero_amd.build          # incremental
    python -m aero_amd.build --force  # rebuild everything
"""

from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import pathlib
import shutil
import subprocess
import sys

PKG = pathlib.Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_build"
LIB = PKG / "libaero_hip.so"
ARCH = os.environ.get("AERO_OFFLOAD_ARCH", "gfx950")

HIPCC = shutil.which("hipcc") or "/opt/rocm/bin/hipcc"
ROCM_LIB = "/opt/rocm/lib"
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    # packed FP32 instructions are kept out of the generated code
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    # Do not merge stores from the two sides of a branch into one store to a
    # selected address: on the per-cell population arrays that turns a
    # constant index into a runtime one and moves the array to scratch memory.
    "-mllvm", "-simplifycfg-sink-common=false",
    "-Wno-unused-parameter", "-Wno-unused-command-line-argument",
    f"-I{ROOT / 'include'}", f"-I{CSRC}",
]


def _sources():
    return sorted(CSRC.glob("*.hip"))


def _headers():
    return sorted(CSRC.glob("*.h")) + sorted((ROOT / "include").glob("*.h"))


def _needs(obj: pathlib.Path, deps) -> bool:
    if not obj.exists():
        return True
    t = obj.stat().st_mtime
    return any(d.stat().st_mtime > t for d in deps)


def _compile(src: pathlib.Path, force: bool) -> pathlib.Path:
    obj = OBJ / (src.stem + ".o")
    if force or _needs(obj, [src, *_headers()]):
        tmp = obj.with_suffix(".o.tmp")
        cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(tmp)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"hipcc failed for {src.name}:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, obj)
    return obj


def build(force: bool = False, jobs: int | None = None, verbose: bool = False) -> pathlib.Path:
    OBJ.mkdir(exist_ok=True)
    srcs = _sources()
    jobs = jobs or min(len(srcs), max(1, (os.cpu_count() or 4)))
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        objs = list(ex.map(lambda s: _compile(s, force), srcs))
    if force or _needs(LIB, objs):
        tmp = LIB.with_name(LIB.name + ".tmp")
        # hipFFT: the SEM inlet's divergence-free projection (sem.hip)
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", *map(str, objs), "-o", str(tmp),
               f"-L{ROCM_LIB}", "-lhipfft", f"-Wl,-rpath,{ROCM_LIB}"]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
        os.replace(tmp, LIB)
    if verbose:
        print(f"built {LIB} ({LIB.stat().st_size / 1e6:.1f} MB)")
    return LIB


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("-j", "--jobs", type=int, default=None)
    a = ap.parse_args(argv)
    build(force=a.force, jobs=a.jobs, verbose=True)


if __name__ == "__main__":
    sys.exit(main())
