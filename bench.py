#!/usr/bin/env python3
"""Benchmark: one Groth16 proof of SquareChain(2^18) on one MI355X per step.

Workload (the reference's headline configuration, the single 2^18 proof):
SquareChain with N = 2^K constraints (x_{i+1} = x_i^2), domain d = 2^(K+1), proving key
from this library's seeded trapdoor setup (synthetic, seed 42), key resident in HBM and
the packed witness already resident in HBM when the timed region starts.  A step is a
full prove: evaluation vectors, qap_h (7 NTTs of 2^(K+1)), the five MSMs (A, B1, H, L in
G1; B2 in G2) and the proof combination; the output is the affine proof (A, B, C).

Prints one JSON line (rank 0).  With --gpus N > 1 it exits: this build measures one GPU.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import sys
import time
from pathlib import Path

import numpy as np

ROOT = Path(__file__).resolve().parent
sys.path[:0] = [str(ROOT), str(ROOT / "bend-groth16_amd")]

import bendg  # noqa: E402

METRIC = "ms/proof SquareChain 2^18 (bit-exact)"


def log(msg: str) -> None:
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


class Hip:
    """Minimal HIP runtime plumbing (device buffer for the resident witness)."""

    def __init__(self):
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.L.hipDeviceSynchronize.argtypes = []

    def upload(self, arr: np.ndarray) -> int:
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), arr.nbytes) == 0
        assert self.L.hipMemcpy(p, arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0
        return p.value

    def sync(self) -> None:
        assert self.L.hipDeviceSynchronize() == 0

    def free(self, p: int) -> None:
        self.L.hipFree(ctypes.c_void_p(p))


def ensure_key(k: int, seed: int, base: Path) -> Path:
    d = base / f"square_chain_K{k}_s{seed}"
    if (d / "header.bin").exists() and (d / "pk_l.bin").exists():
        return d
    d.mkdir(parents=True, exist_ok=True)
    log(f"generating SquareChain 2^{k} key in {d}")
    t = time.perf_counter()
    bendg.setup_square_chain(d, k, seed=seed)
    log(f"setup done in {time.perf_counter() - t:.1f} s")
    return d


# field products per group operation, and 32x32->64 multiplies per Fq product (64 a*b,
# 64 m*p, 8 reduction digits)
MULS_PER_FQ = 136
FQ_PER_MADD = {1: 10, 2: 28}   # madd-2008-s: 8M + 2S (G2: Karatsuba 3 / complex square 2)


def msm_sizes(hdr):
    nc, ni, nw, lg, na, nb, nh, nl = hdr
    return {1: [na, nb, nh, nl], 2: [nb]}


def standalone_msm_pass(key: Path, hdr, reps: int) -> dict:
    """Each MSM kernel alone on the device (no overlap): the H query (the largest G1 MSM)
    and the B2 query (the G2 MSM) with seeded uniform scalars, `reps` launches each;
    returns per-kernel average ms per launch from the library's HIP events."""
    rng = np.random.default_rng(7)
    out = {}
    for g, fname, n in ((1, "pk_h.bin", hdr[6]), (2, "pk_b2.bin", hdr[5])):
        pts = bendg.pack16(np.fromfile(key / fname, dtype=np.uint32).reshape(-1, 16)).reshape(n, -1)
        sc = rng.integers(0, 2**32, size=(n, 8), dtype=np.uint32)
        sc[:, 7] &= np.uint32(0x0FFFFFFF)
        f = bendg.msm_g1 if g == 1 else bendg.msm_g2
        f(pts, sc)  # warm-up (allocations)
        bendg.timing_reset()
        bendg.set_timing(True)
        for _ in range(reps):
            f(pts, sc)
        bendg.set_timing(False)
        for name, (avg, cnt) in bendg.timing_report().items():
            out[name] = {"avg_ms": round(avg, 4), "n": n, "launches": cnt}
    return out


def roofline(standalone: dict, peak_mul: float) -> dict:
    """Roofline of the G1 accumulation kernel (the MSMs' dominant kernel) from its
    standalone average launch duration: one mixed add (10 Fq products x 136 multiplies)
    per sorted (bucket, point) entry."""
    k = standalone["msm_g1_accumulate"]
    entries = bendg.msm_plan_info(k["n"], 1)["entries"]
    work = entries * FQ_PER_MADD[1] * MULS_PER_FQ
    achieved = work / (k["avg_ms"] * 1e-3) / 1e12
    return {"bound": "valu", "kernel": "msm_g1_accumulate", "achieved": round(achieved, 3),
            "peak": round(peak_mul / 1e12, 3), "unit": "T int32-mul/s", "frac": round(achieved / (peak_mul / 1e12), 4),
            "traffic": None, "work_per_launch": work, "points_per_launch": k["n"], "avg_launch_ms": k["avg_ms"]}


def cpu_baseline(key: Path) -> dict:
    """The C oracle's full prove of the same instance on the host cores (one proof)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle import oracle as O  # cpu_baseline leg only
    from instance import load
    inst = load(key)
    t = time.perf_counter()
    O.prove(inst)
    ms = (time.perf_counter() - t) * 1e3
    return {"value": round(ms, 1), "unit": "ms/proof", "cores": O.threads(), "kind": "port",
            "sample": "1 full SquareChain proof of the same key by the C oracle (OpenMP, reference-shaped "
                      "Pippenger c=8/10 unsigned windows, radix-2 NTT)"}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--k", type=int, default=18, help="log2 of the constraint count")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--keys", default=os.environ.get("BG16_KEYS", "/tmp/bg16_keys"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2

    key = ensure_key(a.k, a.seed, Path(a.keys))
    bendg.init(0)
    t = time.perf_counter()
    pk = bendg.ProvingKey(key)
    load_ms = (time.perf_counter() - t) * 1e3
    hdr = pk.header
    log(f"key loaded in {load_ms:.0f} ms ({pk.device_bytes() / 2**20:.0f} MiB in HBM), header {hdr}")
    z16 = np.fromfile(key / "witness.bin", dtype=np.uint32)
    rs16 = np.fromfile(key / "rs.bin", dtype=np.uint32)
    z = bendg.pack16(z16.reshape(-1, 16)).reshape(-1)
    rs = bendg.pack16(rs16.reshape(-1, 16)).reshape(-1)
    hip = Hip()
    dz = hip.upload(z)

    ref = None
    for i in range(a.warmup):
        p, ph = pk.prove_device(dz, rs)
        ref = p if ref is None else ref
        log(f"warmup {i}: " + " ".join(f"{k}={v:.2f}" for k, v in ph.items()))
    # host-buffer (PCIe-inclusive) proof for reference
    p_host, ph_host = pk.prove(z16, rs16, fmt="words16")
    pcie_ms = ph_host["total"]
    bendg.set_timing(True)
    bendg.timing_reset()
    hip.sync()
    t0 = time.perf_counter()
    walls = []
    for _ in range(a.steps):
        s = time.perf_counter()
        p, _ = pk.prove_device(dz, rs)
        walls.append((time.perf_counter() - s) * 1e3)
        if ref is None:
            ref = p
        if not np.array_equal(p, ref):
            log("proof differs between steps")
            return 1
    hip.sync()
    elapsed = time.perf_counter() - t0
    timing = bendg.timing_report()
    bendg.set_timing(False)
    if not np.array_equal(bendg.pack16(p_host.reshape(-1, 16)).reshape(-1), ref):
        log("host-witness proof differs from the device-witness proof")
        return 1
    ms = elapsed * 1e3 / a.steps
    per_proof = {k: round(v[0] * v[1] / a.steps, 4) for k, v in sorted(timing.items(), key=lambda kv: -kv[1][0] * kv[1][1])}
    peak, _ = bendg.probe_rate(0, 4096)
    log("standalone MSM pass")
    alone = standalone_msm_pass(key, hdr, 5)
    for k, v in sorted(alone.items()):
        log(f"  {k:22s} {v['avg_ms']:8.3f} ms/launch (n={v['n']})")
    roof = roofline(alone, peak)
    line = {
        "metric": METRIC, "value": round(ms, 3), "unit": "ms/proof", "n_gpus": 1, "steps": a.steps,
        "warmup": a.warmup, "ms_per_step": round(ms, 3), "higher_is_better": False, "scaling": "weak",
        "vs_baseline": None, "dtype": "u32 (exact BN254 integer arithmetic)", "data": "synthetic",
        "config": {"workload": f"SquareChain 2^{a.k} single proof, key and witness resident in HBM",
                   "constraints": 1 << a.k, "domain": 1 << hdr[3], "msm_g1_points": msm_sizes(hdr)[1],
                   "msm_g2_points": msm_sizes(hdr)[2], "parallelism": "1 GPU"},
        "proofs_per_s": round(1e3 / ms, 2),
        "latency_ms": {"p50": round(float(np.percentile(walls, 50)), 3),
                       "p99": round(float(np.percentile(walls, 99)), 3)},
        "pcie_inclusive_ms": round(pcie_ms, 3),
        "key_load_ms": round(load_ms, 1),
        "kernel_ms_per_proof_overlapped": per_proof,
        "kernel_ms_standalone": {k: v["avg_ms"] for k, v in alone.items()},
        "roofline": roof,
    }
    if not a.no_cpu_baseline:
        log("cpu baseline (C oracle, one proof)")
        line["cpu_baseline"] = cpu_baseline(key)
    print(json.dumps(line), flush=True)
    hip.free(dz)
    pk.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
