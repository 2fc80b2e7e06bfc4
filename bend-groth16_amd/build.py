"""Build the gfx950 library libbg16.so (and the bg16_prove CLI) in-tree.

Every ``csrc/*.hip`` and ``csrc/*.cpp`` is compiled with hipcc for gfx950 into
``_obj/`` (in parallel), then linked into ``bend-groth16_amd/libbg16.so``.  Objects are
rebuilt when their source or any ``csrc/*.hpp`` / ``include/*.h`` header is newer.

Every kernel's resource usage is read from the compiler's kernel-resource-usage
remarks and checked: no kernel may use a dynamic stack, kernels of the proving path
(every translation unit but the diagnostics in ``bg_micro.hip``) may use no scratch
memory (private segment) at all, and the diagnostic kernels at most
``DIAG_SCRATCH_MAX`` bytes per lane.  A kernel that spills or calls out-of-line
functions shows up here instead of on the device.  The table of every kernel's
VGPRs, scratch and occupancy is written to ``_obj/kernel_resources.txt``.
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
RESOURCE_FLAG = "-Rpass-analysis=kernel-resource-usage"
# translation units whose kernels are diagnostics (known-answer and element-wise checks,
# throughput probes), not part of the proving path
DIAG_SOURCES = {"bg_micro.hip"}
DIAG_SCRATCH_MAX = 512  # bytes per lane


def _newest_header() -> float:
    hs = list(CSRC.glob("*.hpp")) + list(INCLUDE.glob("*.h"))
    return max((h.stat().st_mtime for h in hs), default=0.0)


def parse_resources(text: str) -> list[dict]:
    """Kernel entries of the compiler's kernel-resource-usage remarks: name, VGPRs,
    ScratchSize (bytes per lane), dynamic stack, occupancy (waves per SIMD)."""
    out: list[dict] = []
    cur: dict | None = None
    for line in text.splitlines():
        if "remark:" not in line or "kernel-resource-usage" not in line:
            continue
        body = line.split("remark:", 1)[1].rsplit("[-Rpass", 1)[0].strip()
        key, _, val = body.partition(":")
        key, val = key.strip(), val.strip()
        if key == "Function Name":
            cur = {"name": val}
            out.append(cur)
        elif cur is not None:
            cur[key] = val
    return out


def check_resources(src_name: str, kernels: list[dict]) -> list[str]:
    """Violations of the resource rules for the kernels of one translation unit."""
    bad = []
    limit = DIAG_SCRATCH_MAX if src_name in DIAG_SOURCES else 0
    for k in kernels:
        scratch = int(k.get("ScratchSize [bytes/lane]", "0"))
        if k.get("Dynamic Stack", "False") != "False":
            bad.append(f"{src_name}: {k['name']} uses a dynamic stack")
        if scratch > limit:
            bad.append(f"{src_name}: {k['name']} uses {scratch} B/lane of scratch (limit {limit})")
    return bad


def _compile(src: Path, force: bool) -> tuple[Path, str, list[dict]]:
    obj = OBJ / (src.name + ".o")
    res = OBJ / (src.name + ".res")
    if not force and obj.exists() and res.exists():
        t = obj.stat().st_mtime
        if t >= src.stat().st_mtime and t >= _newest_header():
            return obj, "", parse_resources(res.read_text())
    cmd = [HIPCC, *FLAGS, RESOURCE_FLAG, "-c", str(src), "-o", str(obj)]
    if src.suffix == ".cpp":
        cmd = [HIPCC, "-x", "hip", *FLAGS, RESOURCE_FLAG, "-c", str(src), "-o", str(obj)]
    p = subprocess.run(cmd, capture_output=True, text=True)
    text = p.stdout + p.stderr
    log = "\n".join(l for l in text.splitlines()
                    if "packed-fp32-ops" not in l and l.strip() and "kernel-resource-usage" not in l
                    and not l.lstrip().startswith(("|", "^")) and not l.strip().isdigit())
    if p.returncode != 0:
        raise RuntimeError(f"compile failed: {src.name}\n{log}")
    res.write_text(text)
    return obj, log, parse_resources(text)


def build(force: bool = False, jobs: int | None = None, verbose: bool = False) -> Path:
    OBJ.mkdir(exist_ok=True)
    srcs = sorted(list(CSRC.glob("*.hip")) + list(CSRC.glob("*.cpp")))
    lib_srcs = [s for s in srcs if s.name not in CLI_SOURCES]
    jobs = jobs or min(8, os.cpu_count() or 4)
    with cf.ThreadPoolExecutor(max_workers=jobs) as ex:
        results = list(ex.map(lambda s: _compile(s, force), srcs))
    objs = {o.name: o for o, _, _ in results}
    if verbose:
        for o, log, _ in results:
            if log:
                print(f"[{o.name}]\n{log}")
    violations, table = [], []
    for src, (_, _, kernels) in zip(srcs, results):
        violations += check_resources(src.name, kernels)
        for k in kernels:
            table.append(f"{src.name}\t{k['name']}\tVGPRs {k.get('VGPRs', '?')}\t"
                         f"scratch {k.get('ScratchSize [bytes/lane]', '?')}\t"
                         f"dynamic stack {k.get('Dynamic Stack', '?')}\t"
                         f"waves/SIMD {k.get('Occupancy [waves/SIMD]', '?')}")
    (OBJ / "kernel_resources.txt").write_text("\n".join(table) + "\n")
    if violations:
        raise RuntimeError("kernel resource check failed:\n" + "\n".join(violations))
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
