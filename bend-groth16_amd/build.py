"""Build the gfx950 library libbg16.so (and the bg16_prove CLI) in-tree.

Every ``csrc/*.hip`` and ``csrc/*.cpp`` is compiled with hipcc for gfx950 into
``_obj/`` (in parallel), then linked into ``bend-groth16_amd/libbg16.so``.  Objects are
rebuilt when their source or any ``csrc/*.hpp`` / ``include/*.h`` header is newer.
"""
from __future__ import annotations

import argparse
import concurrent.futures as cf
import os
import shutil
import subprocess
import sys
from pathlib import Path

PKG = Path(__file__).resolve().parent
ROOT = PKG.parent
CSRC = PKG / "csrc"
OBJ = PKG / "_obj"
LIB = PKG / "libbg16.so"
CLI = PKG / "bg16_prove"
INCLUDE = ROOT / "include"

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("BG16_ARCH", "gfx950")
FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-std=c++17", "-fPIC",
    # keep hipcc from forming packed-FP32 instructions (integer code, but stated for every TU)
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-result", "-I", str(INCLUDE),
]
CLI_SOURCES = {"prove_main.cpp"}


def _newest_header() -> float:
    hs = list(CSRC.glob("*.hpp")) + list(INCLUDE.glob("*.h"))
    return max((h.stat().st_mtime for h in hs), default=0.0)


def _compile(src: Path, force: bool) -> tuple[Path, str]:
    obj = OBJ / (src.name + ".o")
    if not force and obj.exists():
        t = obj.stat().st_mtime
        if t >= src.stat().st_mtime and t >= _newest_header():
            return obj, ""
    cmd = [HIPCC, *FLAGS, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC, "-x", "hip", *FLAGS, "-c", str(src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    log = "\n".join(l for l in (p.stdout + p.stderr).splitlines()
                    if "packed-fp32-ops" not in l and l.strip())
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{log}")
    return obj, log


def build(force: bool = False, jobs: int | None = None, verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    lib_srcs = [s for s in srcs if s.name not in CLI_SOURCES]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, force), srcs))
    objs = {o.name: o for o, _ in results}
    if verbose:
        for o, log in results:
            if log:
                print(f"[{o.name}]\n{log}")
    lib_objs = [objs[s.name + ".o"] for s in lib_srcs]
    tmp = LIB.with_suffix(".so.tmp")
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC", "-o", str(tmp), *map(str, lib_objs), "-lpthread"]
    p = subprocess.run(cmd, capture_output=True, text=True)
    if p.returncode != 0:
        raise RuntimeError("link failed\n" + p.stdout + p.stderr)
    os.replace(tmp, LIB)
    cli_objs = [objs[s + ".o"] for s in CLI_SOURCES if (s + ".o") in objs]
    if cli_objs:
        cmd = [HIPCC, f"--offload-arch={ARCH}", "-o", str(CLI), *map(str, cli_objs),
               f"-L{PKG}", "-lbg16", f"-Wl,-rpath,$ORIGIN", "-lpthread"]
        p = subprocess.run(cmd, capture_output=True, text=True)
        if p.returncode != 0:
            raise RuntimeError("CLI link failed\n" + p.stdout + p.stderr)
    return LIB


def clean() -> None:
    shutil.rmtree(OBJ, ignore_errors=True)
    for f in (LIB, CLI):
        if f.exists():
            f.unlink()


if __name__ == "__main__":
    ap = argparse.ArgumentParser()
    ap.add_argument("--force", action="store_true")
    ap.add_argument("--clean", action="store_true")
    ap.add_argument("-j", type=int, default=None)
    ap.add_argument("-v", action="store_true")
    a = ap.parse_args()
    if a.clean:
        clean()
        sys.exit(0)
    print(build(force=a.force, jobs=a.j, verbose=a.v))
