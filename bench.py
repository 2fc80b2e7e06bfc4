#!/usr/bin/env python3
"""Benchmark: one Groth16 proof of SquareChain(2^18) on one MI355X per step.

Workload (the reference's headline configuration, the single 2^18 proof):
SquareChain with N = 2^K constraints (x_{i+1} = x_i^2), domain d = 2^(K+1), proving key
from this library's seeded trapdoor setup (synthetic, seed 42), key resident in HBM and
the packed witness already resident in HBM when the timed region starts.  A step is a
full prove: evaluation vectors, qap_h (7 NTTs of 2^(K+1)), the five MSMs (A, B1, H, L in
G1; B2 in G2) and the proof combination; the output is the affine proof (A, B, C).  After
the timed region the same number of proofs is made from the host witness (FORMAT.md words,
the PCIe-inclusive time, reported beside `value`), the proof is checked with the pairing
equation against vk.bin (host verifier, "proof_verified"), and the line carries the
batched mix (64 proofs, K uniform in 10..18, own keys: engine vs one at a time, every
proof bit-equal) and the CPU baseline (the C oracle's proof of the same key on the host
cores, bit-equal to the GPU proof, plus a one-thread run).

Prints one JSON line (rank 0).  With --gpus N > 1 it exits: this build measures one GPU.

Other modes (one JSON line each):
  --mode batch       64 SquareChain proofs, K uniform in 10..18 by seed, each with its own
                     key and witness: proofs/s one proof at a time vs the proving engine
                     (concurrent slots); plus 256 independent G1 MSMs with sizes
                     log-uniform in 2^12..2^20: one call each vs one batched call.
  --mode serve       Poisson stream of proof requests, K in {12,14,16,18,20} with weights
                     0.3/0.3/0.2/0.15/0.05, all five keys resident, swept arrival rates:
                     proofs/s and p50/p99 latency; every proof equals the same proof made
                     alone.
  --mode primitives  G1 MSM 2^20 and 2^22, G2 MSM 2^20 (bases k_i G, uniform scalars,
                     checked against the closed form) and Fr NTT 2^20 / 2^24 forward and
                     inverse (round trip checked).
  --mode tune        sweeps window width, slice target and reduction shape of the MSMs and
                     tile and radix of the NTT on the device and writes the parameter table
                     the planners read (tune_gfx950.txt).
  --mode roofline    the standalone MSM pass and the transforms alone (profiled by rocprofv3).
  --mode sizes       single proofs at K = 10..18, idle gaps between sizes (for a kernel trace of
                     the single path's device-busy fraction by size).
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

    def alloc(self, nbytes: int) -> int:
        p = ctypes.c_void_p()
        assert self.L.hipMalloc(ctypes.byref(p), max(nbytes, 4)) == 0
        return p.value

    def upload(self, arr: np.ndarray) -> int:
        p = self.alloc(arr.nbytes)
        assert self.L.hipMemcpy(ctypes.c_void_p(p), arr.ctypes.data_as(ctypes.c_void_p), arr.nbytes, 1) == 0
        return p

    def download(self, p: int, arr: np.ndarray) -> np.ndarray:
        assert self.L.hipMemcpy(arr.ctypes.data_as(ctypes.c_void_p), ctypes.c_void_p(p), arr.nbytes, 2) == 0
        return arr

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


# field products per group operation, and the algorithmic 32x32->64 multiplies per Fq
# product: 64 a*b, 64 m*p and 8 reduction digits of the 32-bit-digit Montgomery product (the
# work unit; the kernels' 29-bit-digit product issues 171 multiplies for the same product)
MULS_PER_FQ = 136
# Fq products per mixed addition in the reference's formula (madd-2007-bl, 7M + 4S; G2 with
# Karatsuba Fq2 products: 29), the algorithmic work unit of a sorted (bucket, point) entry;
# the kernel's own XYZZ mixed addition (madd-2008-s) takes 10 (G1) / 28 (G2)
FQ_PER_MADD = {1: 11, 2: 29}


R_MOD = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def uniform_fr(rng: np.random.Generator, n: int) -> np.ndarray:
    """(n, 8) u32 canonical scalars uniform in [0, r): uniform 256-bit words with the top
    word drawn below r's, redrawn where the whole value lands at or above r."""
    w = rng.integers(0, 2**64, size=(n, 4), dtype=np.uint64)
    top = R_MOD >> 192
    w[:, 3] = rng.integers(0, top + 1, size=n, dtype=np.uint64)
    for i in np.nonzero(w[:, 3] == top)[0]:
        while sum(int(w[i, j]) << (64 * j) for j in range(4)) >= R_MOD:
            w[i] = rng.integers(0, 2**64, size=4, dtype=np.uint64)
            w[i, 3] = rng.integers(0, top + 1, dtype=np.uint64)
    return w.view(np.uint32).reshape(n, 8)


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
        sc = uniform_fr(rng, n)
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


def probe_peak() -> float:
    """Measured issue rate of v_mad_u64_u32 (16 independent chains per lane, four waves
    per SIMD), best of three ~20 ms launches after a warm-up: the binding resource of the
    MSM kernels."""
    bendg.probe_rate(0, 65536)
    return max(bendg.probe_rate(0, 262144)[0] for _ in range(3))


def traffic_from_profile() -> dict | None:
    """HBM bytes per launch of the G1 accumulation kernel and its VALU issue fraction, from
    the committed rocprofv3 --pmc summary (FETCH_SIZE + WRITE_SIZE scaled by a device copy
    of known size; SQ_INSTS_VALU over the launch's duration)."""
    f = ROOT / "profiles" / "msm_g1_accumulate_2^19_traffic.md"
    if not f.exists():
        return None
    out = {"source": f"profiles/{f.name}"}
    for line in f.read_text().splitlines():
        key, _, val = line.partition(":")
        if key in ("traffic_bytes_per_launch", "valu_issue_frac_published_clock", "valu_issue_frac_measured_clock"):
            out[key] = float(val)
    return out if "traffic_bytes_per_launch" in out else None


# MI355X's published VALU issue rate: 256 CUs x 64 lanes x 2.4 GHz peak engine clock, one
# 32-bit multiply(-add) per lane per clock
SPEC_MUL_RATE = 256 * 64 * 2.4e9


def roofline(standalone: dict, peak_mul: float) -> dict:
    """Roofline of the G1 accumulation kernel (the proof's dominant kernel) from its
    standalone average launch duration (HIP events on the library stream): one mixed add
    (11 Fq products x 136 multiplies, the reference's madd-2007-bl) per sorted (bucket,
    point) entry.  `peak` is the
    published issue rate; `peak_measured` is the v_mad_u64_u32 probe on this machine."""
    k = standalone["msm_g1_accumulate"]
    entries = bendg.msm_plan_info(k["n"], 1)["entries"]
    work = entries * FQ_PER_MADD[1] * MULS_PER_FQ
    achieved = work / (k["avg_ms"] * 1e-3) / 1e12
    spec = SPEC_MUL_RATE / 1e12
    tr = traffic_from_profile()
    return {"bound": "valu", "kernel": "k_accumulate<G1> (msm_g1_accumulate)", "achieved": round(achieved, 3),
            "peak": round(spec, 3), "unit": "T int32-mul/s", "frac": round(achieved / spec, 4),
            "peak_measured": round(peak_mul / 1e12, 3), "frac_of_measured": round(achieved / (peak_mul / 1e12), 4),
            "traffic": tr["traffic_bytes_per_launch"] if tr else None, "traffic_source": tr["source"] if tr else None,
            "valu_issue_frac": tr.get("valu_issue_frac_published_clock") if tr else None,
            "valu_issue_frac_measured_clock": tr.get("valu_issue_frac_measured_clock") if tr else None,
            "algorithmic_bytes": entries * (8 + 64), "work_per_launch": work, "points_per_launch": k["n"],
            "entries_per_launch": entries, "avg_launch_ms": k["avg_ms"],
            "peak_note": "published VALU issue rate 256 CU x 64 lanes x 2.4 GHz; peak_measured: "
                         "v_mad_u64_u32 issue-rate probe (16 independent chains per lane, 4 waves per SIMD)"}


def cpu_baseline(key: Path, gpu_proof: np.ndarray, keys_dir: Path) -> dict:
    """The C oracle's full prove of the same instance on all host cores (one proof),
    checked bit-equal to the GPU proof; then the same on one thread.  The one-thread run
    is bounded: the full proof when the all-core time times the core count stays under a
    minute, else one proof of the 2^16 SquareChain (named in its sample)."""
    sys.path.insert(0, str(ROOT / "tests"))
    from oracle import oracle as O  # cpu_baseline leg only
    from instance import load
    inst = load(key)
    cores = O.threads()
    t = time.perf_counter()
    A, B, C, _ = O.prove(inst)
    ms = (time.perf_counter() - t) * 1e3
    got = np.ascontiguousarray(gpu_proof, dtype=np.uint32).reshape(-1).view(np.uint64)
    equal = bool(np.array_equal(got[:8], A) and np.array_equal(got[8:24], B) and np.array_equal(got[24:32], C))
    if not equal:
        raise SystemExit("the CPU baseline's proof differs from the GPU proof")
    one = {}
    O.set_threads(1)
    try:
        if ms * cores <= 60e3:
            t = time.perf_counter()
            O.prove(inst)
            one = {"value": round((time.perf_counter() - t) * 1e3, 1), "unit": "ms/proof", "cores": 1,
                   "sample": "1 full SquareChain proof of the same key, one thread"}
        else:
            k16 = load(ensure_key(16, 42, keys_dir))
            t = time.perf_counter()
            O.prove(k16)
            one = {"value": round((time.perf_counter() - t) * 1e3, 1), "unit": "ms/proof (SquareChain 2^16)",
                   "cores": 1, "sample": "1 full SquareChain 2^16 proof, one thread (the 2^18 proof on one "
                                         "thread would exceed the bench's time bound)"}
    finally:
        O.set_threads(cores)
    return {"value": round(ms, 1), "unit": "ms/proof", "cores": cores, "kind": "port",
            "equal_to_gpu_proof": equal,
            "sample": "1 full SquareChain proof of the same key by the C oracle (OpenMP over all host cores, "
                      "reference-shaped Pippenger c=8/10 unsigned windows, radix-2 NTT), bit-equal to the GPU proof",
            "one_thread": one}


# ------------------------------------------------------------------ engine / batch modes

def engine_rate(pk, dz, rs, count: int, slots: int, ref) -> dict:
    """count proofs of one key through the proving engine; proofs/s and bit-equality."""
    eng = bendg.Engine(slots=slots)
    for t in [eng.submit(pk, dz, rs) for _ in range(slots)]:  # slot work arrays
        eng.wait(t)
    t0 = time.perf_counter()
    tickets = [eng.submit(pk, dz, rs) for _ in range(count)]
    lat = []
    for t in tickets:
        p, tm = eng.wait(t)
        lat.append(tm["latency"])
        if not np.array_equal(p, ref):
            raise SystemExit("engine proof differs from the single proof")
    dt = time.perf_counter() - t0
    eng.close()
    return {"proofs": count, "slots": slots, "proofs_per_s": round(count / dt, 2)}


def witness_files(d: Path):
    z = bendg.pack16(np.fromfile(d / "witness.bin", dtype=np.uint32).reshape(-1, 16)).reshape(-1)
    rs = bendg.pack16(np.fromfile(d / "rs.bin", dtype=np.uint32).reshape(-1, 16)).reshape(-1)
    return z, rs


def batched_mix(a, hip, slots_list) -> dict:
    """64 SquareChain proofs with K uniform in 10..18 (seeded), each with its own key and
    witness (witnesses resident in HBM): one proof at a time on the single-proof path vs
    all 64 submitted at once to the proving engine; every engine proof must equal the
    single proof.  Also reports the single path's time by K and the share of it spent in
    the K >= 17 proofs: one such proof already keeps the device busy (the same-key burst
    at 2^18 gains nothing from batching), so 1 / share bounds what any batching of this
    mix can gain."""
    rng = np.random.default_rng(a.seed)
    ks = [int(k) for k in rng.integers(10, 19, size=64)]
    base = Path(a.keys) / "batch"
    t = time.perf_counter()
    dirs = [ensure_key(k, 1000 + i, base) for i, k in enumerate(ks)]
    pks = [bendg.ProvingKey(d) for d in dirs]
    wz = [witness_files(d) for d in dirs]
    dzs = [hip.upload(z) for z, _ in wz]
    log(f"64 keys ready in {time.perf_counter() - t:.1f} s, K = {ks}")
    # one proof at a time (the single-proof path), after two warm proofs per key (the
    # second records the recorded launch sequences of the sizes that replay them, so the
    # timed proofs replay them as a serving loop would)
    ref = [pk.prove_device(dz, rs)[0] for pk, dz, (_, rs) in zip(pks, dzs, wz)]
    for pk, dz, (_, rs) in zip(pks, dzs, wz):
        pk.prove_device(dz, rs)
    hip.sync()
    passes = []
    for _ in range(3):  # the median of three passes over the 64 proofs
        per = []
        t0 = time.perf_counter()
        for pk, dz, (_, rs), r in zip(pks, dzs, wz, ref):
            s0 = time.perf_counter()
            p, _ = pk.prove_device(dz, rs)
            per.append(time.perf_counter() - s0)
            if not np.array_equal(p, r):
                raise SystemExit("single proof not reproducible")
        passes.append((time.perf_counter() - t0, per))
    single_s, per = sorted(passes, key=lambda x: x[0])[1]
    res = {}
    for slots in slots_list:
        eng = bendg.Engine(slots=slots)
        for _ in range(2):  # warm: every slot's work arrays at their largest
            for t in [eng.submit(pk, dz, rs) for pk, dz, (_, rs) in zip(pks, dzs, wz)]:
                eng.wait(t)
        runs = []
        for _ in range(3):  # the median of three passes of all 64 (one pass is ~0.1 s)
            t0 = time.perf_counter()
            tickets = [eng.submit(pk, dz, rs) for pk, dz, (_, rs) in zip(pks, dzs, wz)]
            lat = []
            for t, r in zip(tickets, ref):
                p, tm = eng.wait(t)
                lat.append(tm["latency"])
                if not np.array_equal(p, r):
                    raise SystemExit("batched proof differs from the single proof")
            runs.append((time.perf_counter() - t0, lat))
        eng.close()
        dt, lat = sorted(runs, key=lambda x: x[0])[1]
        res[slots] = {"proofs_per_s": round(64 / dt, 2), "wall_ms": round(dt * 1e3, 1),
                      "p50_ms": round(float(np.percentile(lat, 50)), 2),
                      "p99_ms": round(float(np.percentile(lat, 99)), 2),
                      "passes_wall_ms": [round(x[0] * 1e3, 1) for x in runs]}
        log(f"engine slots={slots}: {res[slots]}")
    for pk in pks:
        pk.close()
    for dz in dzs:
        hip.free(dz)
    best = max(res.values(), key=lambda v: v["proofs_per_s"])
    single_rate = 64 / single_s
    by_k = {}
    for k, dt in zip(ks, per):
        by_k.setdefault(k, []).append(dt * 1e3)
    big = sum(dt for k, dt in zip(ks, per) if k >= 17)
    share = big / sum(per)
    return {"proofs": 64, "engine_proofs_per_s": best["proofs_per_s"], "single_proofs_per_s": round(single_rate, 2),
            "vs_single": round(best["proofs_per_s"] / single_rate, 3),
            "engine": {str(k): v for k, v in res.items()}, "ks": ks,
            "single_ms_by_k": {str(k): round(float(np.median(v)), 3) for k, v in sorted(by_k.items())},
            "single_time_share_k17_18": round(share, 3), "ceiling_vs_single": round(1.0 / share, 3),
            "all_equal_single": True}


def batch_mode(a) -> dict:
    bendg.init(0)
    hip = Hip()
    mix = batched_mix(a, hip, a.slots)
    msm = batch_msm_256(a, hip)
    return {"metric": "batched proofs/s, 64 SquareChain proofs K~U{10..18} (bit-exact)",
            "value": mix["engine_proofs_per_s"], "unit": "proofs/s", "n_gpus": 1, "higher_is_better": True,
            "single_proofs_per_s": mix["single_proofs_per_s"], "batched_vs_single": mix["vs_single"],
            "mix": mix, "dtype": "u32 (exact BN254 integer arithmetic)",
            "data": "synthetic (own trapdoor setup per proof, seeds 1000+i)", "msm_256": msm,
            "config": {"workload": "64 proofs, K uniform in 10..18, own key and witness each, witnesses resident in HBM",
                       "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")}}


def batch_msm_256(a, hip) -> dict:
    """256 independent G1 MSMs, sizes 2^U(12,20): one bg_msm_g1 call each vs one
    bg_msm_batch call (device-resident inputs); results must agree bit for bit."""
    rng = np.random.default_rng(a.seed + 1)
    ns = [int(round(2 ** x)) for x in rng.uniform(12, 20, size=256)]
    pool_n = 1 << 20
    pts = bendg.gen_bases(1, 77, pool_n)
    dpts = hip.upload(pts)
    tot = sum(ns)
    sc = uniform_fr(rng, tot)
    dsc = hip.upload(sc)
    offs = np.concatenate([[0], np.cumsum(ns)[:-1]]).astype(np.uint64)
    L = bendg.lib()
    flags = bendg.FMT_PACKED | bendg.MEM_DEVICE
    dout = hip.alloc(256 * 64)
    one = []

    def seq():
        for i, n in enumerate(ns):
            bendg._check(L.bg_msm_g1(ctypes.c_void_p(dpts), ctypes.c_void_p(dsc + 32 * int(offs[i])), n,
                                     ctypes.c_void_p(dout + 64 * i), flags, 0))

    def bat():
        pp = (ctypes.c_void_p * 256)(*([dpts] * 256))
        sp = (ctypes.c_void_p * 256)(*[dsc + 32 * int(o) for o in offs])
        nn = np.array(ns, dtype=np.uint64)
        bendg._check(L.bg_msm_batch(1, 256, ctypes.cast(pp, ctypes.c_void_p), ctypes.cast(sp, ctypes.c_void_p),
                                    bendg._p(nn), ctypes.c_void_p(dout), flags, 0))

    seq()
    hip.sync()
    t0 = time.perf_counter()
    seq()
    hip.sync()
    t_seq = time.perf_counter() - t0
    r_seq = hip.download(dout, np.zeros((256, 16), np.uint32)).copy()
    bat()
    hip.sync()
    t0 = time.perf_counter()
    bat()
    hip.sync()
    t_bat = time.perf_counter() - t0
    r_bat = hip.download(dout, np.zeros((256, 16), np.uint32))
    if not np.array_equal(r_seq, r_bat):
        raise SystemExit("batched MSM results differ from the one-at-a-time results")
    for p in (dpts, dsc, dout):
        hip.free(p)
    log(f"256 MSMs ({tot} points): one at a time {t_seq * 1e3:.1f} ms, batched {t_bat * 1e3:.1f} ms")
    return {"msms": 256, "points": tot, "sequential_ms": round(t_seq * 1e3, 2), "batched_ms": round(t_bat * 1e3, 2),
            "batched_vs_sequential": round(t_seq / t_bat, 3), "points_per_s_batched": round(tot / t_bat, 0)}


def serve_mode(a) -> dict:
    import threading
    kset = [12, 14, 16, 18, 20]
    wts = [0.3, 0.3, 0.2, 0.15, 0.05]
    dirs = {k: ensure_key(k, 42 + k, Path(a.keys)) for k in kset}
    bendg.init(0)
    hip = Hip()
    pks = {k: bendg.ProvingKey(d) for k, d in dirs.items()}
    wz = {k: witness_files(d) for k, d in dirs.items()}
    dz = {k: hip.upload(wz[k][0]) for k in kset}
    solo, alone = {}, {}
    for k in kset:
        # two warm proofs (the second records the recorded launch sequences where the size
        # replays them), then the median of ten
        for _ in range(2):
            solo[k], _ = pks[k].prove_device(dz[k], wz[k][1])
        walls = []
        for _ in range(10):
            t0 = time.perf_counter()
            p, _ = pks[k].prove_device(dz[k], wz[k][1])
            walls.append(1e3 * (time.perf_counter() - t0))
            if not np.array_equal(p, solo[k]):
                raise SystemExit("single proof not reproducible")
        alone[str(k)] = round(float(np.median(walls)), 2)
        log(f"K={k}: alone {alone[str(k)]} ms")
    rates = []
    for rate in a.rates:
        eng = bendg.Engine(slots=a.slots[0])
        for k in sorted(kset, reverse=True):  # warm every slot, largest size first
            for t in [eng.submit(pks[k], dz[k], wz[k][1]) for _ in range(a.slots[0])]:
                eng.wait(t)
        rng = np.random.default_rng(a.seed + int(rate))
        n = int(rate * a.duration)
        gaps = rng.exponential(1.0 / rate, size=n)
        kinds = rng.choice(kset, size=n, p=wts)
        done = []
        lock = threading.Lock()
        q = []
        cv = threading.Condition(lock)
        finished = [False]

        def collector():
            while True:
                with cv:
                    while not q and not finished[0]:
                        cv.wait()
                    if not q:
                        return
                    t, k, t_arr = q.pop(0)
                p, tm = eng.wait(t)
                # latency = arrival to the engine's completion time (not to this wait,
                # which runs in submission order)
                done.append((k, tm["latency"] * 1e-3, tm, np.array_equal(p, solo[k])))

        th = threading.Thread(target=collector)
        th.start()
        t_start = time.perf_counter()
        t_next = t_start
        t_log = t_start
        for g, k in zip(gaps, kinds):
            t_next += g
            dt = t_next - time.perf_counter()
            if dt > 0:
                time.sleep(dt)
            t_arr = time.perf_counter()
            if t_arr - t_log >= 30.0:  # progress of a long run
                t_log = t_arr
                log(f"rate {rate}/s: {t_arr - t_start:.0f} s, {len(done)} proofs done")
            t = eng.submit(pks[int(k)], dz[int(k)], wz[int(k)][1])  # submit time ~ arrival time
            with cv:
                q.append((t, int(k), t_arr))
                cv.notify()
        with cv:
            finished[0] = True
            cv.notify()
        th.join()
        wall = time.perf_counter() - t_start
        eng.close()
        lat = np.array([d[1] for d in done]) * 1e3
        ok = all(d[3] for d in done)
        if not ok:
            raise SystemExit("a served proof differs from the same proof made alone")
        per_k = {str(k): round(float(np.percentile([d[1] * 1e3 for d in done if d[0] == k], 99)), 2)
                 for k in kset if any(d[0] == k for d in done)}
        r = {"offered_per_s": rate, "proofs": len(done), "proofs_per_s": round(len(done) / wall, 2),
             "p50_ms": round(float(np.percentile(lat, 50)), 2), "p99_ms": round(float(np.percentile(lat, 99)), 2),
             "p99_ms_by_k": per_k, "all_equal_alone": ok}
        log(f"rate {rate}/s: {r}")
        rates.append(r)
    for pk in pks.values():
        pk.close()
    best = max(rates, key=lambda v: v["proofs_per_s"])
    return {"metric": "served proofs/s and p99 latency, Poisson arrivals, K in {12..20} (bit-exact)",
            "value": best["proofs_per_s"], "unit": "proofs/s", "n_gpus": 1, "higher_is_better": True,
            "p99_ms_at_best": best["p99_ms"], "rates": rates, "alone_ms": alone, "slots": a.slots[0],
            "duration_s_per_rate": a.duration, "dtype": "u32 (exact BN254 integer arithmetic)",
            "data": "synthetic (own trapdoor setup per K, seed 42+K)",
            "config": {"workload": "Poisson requests, K 12/14/16/18/20 w 0.3/0.3/0.2/0.15/0.05, keys resident",
                       "hw_queues": os.environ.get("GPU_MAX_HW_QUEUES")}}


def primitives_mode(a) -> dict:
    bendg.init(0)
    hip = Hip()
    L = bendg.lib()
    out = {}
    peak = probe_peak() / 1e12  # T mul/s, measured
    spec = SPEC_MUL_RATE / 1e12
    out["mul_peak_T_per_s"] = round(spec, 3)
    out["mul_peak_measured_T_per_s"] = round(peak, 3)
    rng = np.random.default_rng(a.seed)
    R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
    flags = bendg.FMT_PACKED | bendg.MEM_DEVICE
    for group, lg in ((1, 20), (1, 22), (2, 20)):
        n = 1 << lg
        t = time.perf_counter()
        pts, ks = bendg.gen_bases(group, 100 + lg, n, with_k=True)
        log(f"G{group} 2^{lg}: bases in {time.perf_counter() - t:.1f} s")
        sc = uniform_fr(rng, n)
        dp, ds = hip.upload(pts), hip.upload(sc)
        pw = 16 if group == 1 else 32
        dout = hip.alloc(pw * 4)
        f = L.bg_msm_g1 if group == 1 else L.bg_msm_g2
        call = lambda: bendg._check(f(ctypes.c_void_p(dp), ctypes.c_void_p(ds), n, ctypes.c_void_p(dout), flags, 0))  # noqa: E731
        call()
        hip.sync()
        reps = 5
        t0 = time.perf_counter()
        for _ in range(reps):
            call()
        hip.sync()
        ms = (time.perf_counter() - t0) * 1e3 / reps
        got = hip.download(dout, np.zeros(pw, np.uint32))
        # closed form: (sum a_i k_i mod r) G
        a64 = sc.view(np.uint64).reshape(-1, 4)
        k64 = ks.view(np.uint64).reshape(-1, 4)
        tot = 0
        for i in range(0, n, 1 << 16):
            av = [int(x[0]) | int(x[1]) << 64 | int(x[2]) << 128 | int(x[3]) << 192 for x in a64[i:i + (1 << 16)]]
            kv = [int(x[0]) | int(x[1]) << 64 | int(x[2]) << 128 | int(x[3]) << 192 for x in k64[i:i + (1 << 16)]]
            tot = (tot + sum(x * y for x, y in zip(av, kv))) % R
        kt = np.array([(tot >> (32 * j)) & 0xFFFFFFFF for j in range(8)], dtype=np.uint32)
        want = bendg.fixed_base_mul(group, kt).reshape(-1)
        ok = bool(np.array_equal(got, want))
        info = bendg.msm_plan_info(n, group)
        # algorithmic work at the planned c in Fq products (Pippenger with the reference's
        # per-operation costs: G1 madd 11, add 16, dbl 7; G2 29 / 43 / 16), 136 multiplies each
        c = info["c"]
        w = -(-254 // c)
        madd, add, dbl = (11, 16, 7) if group == 1 else (29, 43, 16)
        fqm = n * w * madd + w * (1 << (c - 1)) * 2 * add + 254 * dbl + w * add
        tmul = fqm * MULS_PER_FQ / (ms * 1e-3) / 1e12
        out[f"msm_g{group}_2^{lg}"] = {"ms": round(ms, 3), "points_per_s": round(n / ms * 1e3, 0), "c": c,
                                       "algorithmic_T_mul_per_s": round(tmul, 2), "valu_frac": round(tmul / spec, 3),
                                       "valu_frac_of_measured": round(tmul / peak, 3), "closed_form_ok": ok}
        log(f"  MSM G{group} 2^{lg}: {ms:.3f} ms, closed form {'ok' if ok else 'MISMATCH'}")
        for p in (dp, ds, dout):
            hip.free(p)
        if not ok:
            raise SystemExit("primitive MSM differs from its closed form")
    for lg in (20, 24):
        n = 1 << lg
        x = uniform_fr(rng, n)
        d = hip.upload(x)
        res = {}
        for name, dirn in (("fft", bendg.NTT_FFT), ("ifft", bendg.NTT_IFFT)):
            bendg._check(L.bg_ntt(ctypes.c_void_p(d), lg, dirn, flags))
            hip.sync()
            reps = 5
            t0 = time.perf_counter()
            for _ in range(reps):
                bendg._check(L.bg_ntt(ctypes.c_void_p(d), lg, dirn, flags))
            hip.sync()
            res[name] = (time.perf_counter() - t0) * 1e3 / reps
        # round trip: 6 ffts and 6 iffts above cancel exactly
        back = hip.download(d, np.zeros_like(x))
        ok = bool(np.array_equal(back, x))
        hip.free(d)
        byts = 64 * n
        # both roofs for the transform: algorithmic HBM bytes (one 32 B read and write per
        # element) against 8 TB/s, and (n/2) log2 n Fr products x 136 multiplies against the
        # published multiply issue rate (and the measured one)
        gbps = byts / res["fft"] / 1e6
        tmul = (n // 2) * lg * MULS_PER_FQ / (res["fft"] * 1e-3) / 1e12
        out[f"ntt_2^{lg}"] = {"fft_ms": round(res["fft"], 3), "ifft_ms": round(res["ifft"], 3),
                              "fft_GBps_algorithmic": round(gbps, 1), "hbm_frac": round(gbps / 8000.0, 4),
                              "fft_T_mul_per_s": round(tmul, 2), "valu_frac": round(tmul / spec, 3),
                              "valu_frac_of_measured": round(tmul / peak, 3), "round_trip_ok": ok}
        log(f"  NTT 2^{lg}: fft {res['fft']:.3f} ms, ifft {res['ifft']:.3f} ms, round trip {'ok' if ok else 'BAD'}")
        if not ok:
            raise SystemExit("NTT round trip failed")
    return {"metric": "primitive MSM / NTT times (bit-exact)", "value": out["msm_g1_2^20"]["ms"], "unit": "ms",
            "n_gpus": 1, "higher_is_better": False, "primitives": out, "dtype": "u32 (exact BN254 integer arithmetic)",
            "data": "synthetic (bases k_i G seeded, uniform scalars)",
            "config": {"workload": "G1 MSM 2^20/2^22, G2 MSM 2^20, Fr NTT 2^20/2^24"}}


def roofline_mode(a) -> dict:
    """The default line's standalone pass alone (profiled by rocprofv3 for the committed
    roofline and counter summaries), plus a 256 MiB device copy whose counters calibrate
    FETCH_SIZE / WRITE_SIZE, then five 2^20 forward NTTs and five qap_h of the proof's
    domain."""
    key = ensure_key(a.k, a.seed, Path(a.keys))
    bendg.init(0)
    pk = bendg.ProvingKey(key)
    hdr = pk.header
    pk.close()
    hip = Hip()
    nb = 256 << 20
    src, dst = hip.alloc(nb), hip.alloc(nb)
    hip.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    assert hip.L.hipMemcpy(ctypes.c_void_p(dst), ctypes.c_void_p(src), nb, 3) == 0
    hip.sync()
    hip.free(src)
    hip.free(dst)
    peak = probe_peak()
    alone = standalone_msm_pass(key, hdr, 5)
    # the transforms alone: a 2^20 forward NTT and the proof's qap_h at 2^19 (device
    # buffers, uniform Montgomery inputs), five launches each
    rng = np.random.default_rng(11)
    L = bendg.lib()
    flags = bendg.FMT_PACKED | bendg.MEM_DEVICE
    d = hip.upload(uniform_fr(rng, 1 << 20))
    lg = hdr[3]
    abc = [hip.upload(uniform_fr(rng, 1 << lg)) for _ in range(3)]
    dh = hip.alloc(32 << lg)
    # one untimed call of each first: the domains' tables are built on first use
    bendg._check(L.bg_ntt(ctypes.c_void_p(d), 20, bendg.NTT_FFT, flags))
    bendg._check(L.bg_qap_h(*(ctypes.c_void_p(x) for x in abc), lg, ctypes.c_void_p(dh), flags))
    bendg.timing_reset()
    bendg.set_timing(True)
    for _ in range(5):
        bendg._check(L.bg_ntt(ctypes.c_void_p(d), 20, bendg.NTT_FFT, flags))
        bendg._check(L.bg_qap_h(*(ctypes.c_void_p(x) for x in abc), lg, ctypes.c_void_p(dh), flags))
    bendg.set_timing(False)
    transforms = {k: round(v[0], 4) for k, v in bendg.timing_report().items()}
    for x in [d, dh] + abc:
        hip.free(x)
    return {"metric": "roofline of the G1 accumulation kernel", "value": roofline(alone, peak)["frac"],
            "unit": "fraction of the published VALU issue rate", "roofline": roofline(alone, peak),
            "kernel_ms_standalone": {k: v["avg_ms"] for k, v in alone.items()}, "copy_bytes": nb,
            "transform_ms": transforms}


def sizes_mode(a) -> dict:
    """One proof at a time at K = 10, 12, 14, 16, 18 (`steps` timed proofs each, witnesses
    resident), 100 ms of idle device between sizes so that a kernel trace of this run splits
    by size: with rocprofv3 --kernel-trace, the union of the kernels' intervals over a size's
    proofs against their wall time is the single path's device-busy fraction."""
    bendg.init(0)
    hip = Hip()
    out = {}
    for k in (10, 12, 14, 16, 18):
        kdir = ensure_key(k, 42 + k, Path(a.keys))
        pk = bendg.ProvingKey(kdir)
        z, rs = witness_files(kdir)
        dz = hip.upload(z)
        for _ in range(3):
            pk.prove_device(dz, rs)
        hip.sync()
        time.sleep(0.1)
        walls = []
        for _ in range(a.steps):
            t0 = time.perf_counter()
            pk.prove_device(dz, rs)
            walls.append((time.perf_counter() - t0) * 1e3)
        hip.sync()
        time.sleep(0.1)
        out[str(k)] = {"median_ms": round(float(np.median(walls)), 3), "proofs": a.steps}
        log(f"K={k}: {out[str(k)]}")
        hip.free(dz)
        pk.close()
    return {"metric": "single-proof wall time by size", "value": out["18"]["median_ms"], "unit": "ms/proof",
            "n_gpus": 1, "sizes": out}


def launch_mode(a) -> dict:
    """Launch shapes the prover's recorded sequences choose between (bg_probe_launch): 96
    short kernels eagerly or from HIP graphs, on one stream or forked over four, at two
    kernel lengths; and single proofs by size with the recorded sequences off and on."""
    bendg.init(0)
    out = {}
    for spin in (0, 2000):
        out[f"spin_{spin}"] = bendg.probe_launch(96, spin, reps=40)
        log(f"96 kernels, spin {spin}: {out[f'spin_{spin}']} us")
    proofs = {}
    hip = Hip()
    for k in (10, 12, 14, 16, 18):
        kdir = ensure_key(k, 42 + k, Path(a.keys))
        pk = bendg.ProvingKey(kdir)
        z, rs = witness_files(kdir)
        dz = hip.upload(z)
        row = {}
        for g in ("0", "1", "0", "1"):  # alternating, twice
            os.environ["BG16_GRAPH"] = g
            for _ in range(3):
                pk.prove_device(dz, rs)
            walls = []
            for _ in range(a.steps):
                t0 = time.perf_counter()
                pk.prove_device(dz, rs)
                walls.append((time.perf_counter() - t0) * 1e3)
            row.setdefault("eager" if g == "0" else "graphs", []).append(round(float(np.median(walls)), 3))
        os.environ.pop("BG16_GRAPH", None)
        proofs[str(k)] = row
        log(f"K={k}: {row}")
        hip.free(dz)
        pk.close()
    out["proof_ms_median"] = proofs
    return {"metric": "launch probe (us per repetition) and single proofs eager vs recorded graphs", "value": None,
            "unit": "us", "n_gpus": 1, **out}


def tune_mode(a) -> dict:
    """Device-measured parameter table (replaces Msm.pick_c / pick_fd, bend2/src/msm.bend:37-44,
    and Ntt.fd, bend2/src/ntt.bend:382-385).  Each candidate runs with a one-entry table (plus
    the entries already chosen for the same proof) loaded by bg_tune_load, timed on device
    buffers; the best of each sweep is kept:
      - NTT tile (2^9 / 2^10 / 2^11 elements) x radix (2 / 4) per log2 n: qap_h for 2^11..2^21
        (the proofs' domains), the forward transform for 2^22..2^24;
      - windowed MSMs (arbitrary bases) per log2 n, G1 2^10..2^22 and G2 2^10..2^20: window
        width c around the cost model's, then the accumulation slice target L, the level-1
        group m and the level-2 block K;
      - fixed-base MSMs of SquareChain proofs 2^10..2^20: c of the G1 batch (A, B1, L), of H
        and of B2, one after the other, by whole-proof time (the key is reloaded per candidate:
        its tables are built for the width chosen at load).
    Writes the table to <keys>/tune_gfx950.txt (copied next to the library to take effect)
    and prints it in the JSON line."""
    import math
    bendg.init(0)
    hip = Hip()
    L = bendg.lib()
    flags = bendg.FMT_PACKED | bendg.MEM_DEVICE
    base = Path(a.keys)
    base.mkdir(parents=True, exist_ok=True)
    cand = base / "tune_candidate.txt"
    reps = 5

    def load(lines):
        cand.write_text("\n".join(lines) + "\n")
        bendg.tune_load(str(cand))

    def timed(fn, reps=reps):
        fn()
        hip.sync()
        ts = []
        for _ in range(reps):
            t0 = time.perf_counter()
            fn()
            hip.sync()
            ts.append(time.perf_counter() - t0)
        return float(np.median(ts)) * 1e3

    def keep_model(res, model_key, margin=0.015):
        """The fastest candidate, unless it beats the model's choice by less than the margin
        (measurement noise): then the model's choice."""
        best = min(res, key=res.get)
        return best if res[best] < (1 - margin) * res[model_key] else model_key

    table, results = [], {"ntt": {}, "msm": {}, "fixed": {}, "validation": {}}
    t_start = time.perf_counter()
    rng = np.random.default_rng(a.seed)
    # ---- NTT tile and radix
    for lg in range(11, 25):
        n = 1 << lg
        if lg <= 21:
            bufs = [hip.upload(uniform_fr(rng, n)) for _ in range(3)]
            dh = hip.alloc(32 * n)
            fn = lambda: bendg._check(L.bg_qap_h(*(ctypes.c_void_p(x) for x in bufs), lg, ctypes.c_void_p(dh), flags))  # noqa: E731
        else:
            bufs = [hip.upload(uniform_fr(rng, n))]
            dh = None
            fn = lambda: bendg._check(L.bg_ntt(ctypes.c_void_p(bufs[0]), lg, bendg.NTT_FFT, flags))  # noqa: E731
        res = {}
        for tile in (9, 10, 11):
            for radix in (2, 4):
                load([f"ntt {lg} tile={tile} radix={radix}"])
                res[(tile, radix)] = timed(fn)
        best = keep_model(res, (10, 4))
        table.append(f"ntt {lg} tile={best[0]} radix={best[1]}")
        results["ntt"][lg] = {f"{t}/{r}": round(v, 4) for (t, r), v in res.items()}
        log(f"ntt 2^{lg}: {table[-1]} ({res[best]:.3f} ms; default {res[(10, 4)]:.3f} ms)")
        for x in bufs + ([dh] if dh else []):
            hip.free(x)
    ntt_lines = list(table)
    # ---- windowed MSMs
    for group, top in ((1, 22), (2, 20)):
        pts = bendg.gen_bases(group, 99 + group, 1 << top)
        dp = hip.upload(pts)
        ds = hip.upload(uniform_fr(rng, 1 << top))
        dout = hip.alloc(128)
        f = L.bg_msm_g1 if group == 1 else L.bg_msm_g2
        for lg in range(10, top + 1):
            n = 1 << lg
            fn = lambda: bendg._check(f(ctypes.c_void_p(dp), ctypes.c_void_p(ds), n, ctypes.c_void_p(dout), flags, 0))  # noqa: E731
            load([])
            c0 = bendg.msm_plan_info(n, group)["c"]
            key = f"msm {group} windowed {lg}"
            res = {}
            for c in range(max(4, c0 - 3), min(20, c0 + 2) + 1):
                load([f"{key} c={c}"])
                res[f"c={c}"] = timed(fn)
            best = keep_model(res, f"c={c0}")
            sweeps = [("L", (20, 30, 40, 60)), ("m", (2, 4, 8)),
                      ("K", (64, 128, 256) if group == 1 else (32, 64, 128))]
            for name, vals in sweeps:
                base_entry = best
                for v in vals:
                    entry = f"{base_entry} {name}={v}"
                    load([f"{key} {entry}"])
                    res[entry] = timed(fn)
                best = keep_model({k: v for k, v in res.items() if k == base_entry or k.startswith(base_entry + " ")},
                                  base_entry)
            table.append(f"{key} {best}")
            results["msm"][f"g{group}_{lg}"] = {k: round(v, 4) for k, v in res.items()}
            log(f"msm G{group} 2^{lg}: {best} ({res[best]:.3f} ms; model c={c0} {res[f'c={c0}']:.3f} ms)")
        for x in (dp, ds, dout):
            hip.free(x)
    # ---- fixed-base MSMs inside proofs
    for k in range(10, 21):
        kdir = ensure_key(k, 42 + k, base)
        z, rs = witness_files(kdir)
        dz = hip.upload(z)
        hdr = [int(x) for x in np.fromfile(kdir / "header.bin", dtype=np.uint32)]
        roles = [("g1", 1, hdr[4] + hdr[5] + hdr[7]), ("h", 1, hdr[6]), ("b2", 2, hdr[5])]
        load(ntt_lines)
        pk = bendg.ProvingKey(kdir)
        c_model = pk.windows()
        pk.close()
        chosen = {}
        res_k = {}
        for role, group, total in roles:
            half = round(2 * math.log2(total)) / 2
            key = f"msm {group} fixed {half:g}"
            res = {}
            c0 = c_model[role]
            for c in range(max(8, c0 - 2), min(20, c0 + 2) + 1):
                load(ntt_lines + list(chosen.values()) + [f"{key} c={c}"])
                pk = bendg.ProvingKey(kdir)
                assert pk.windows()[role] == c, (role, c, pk.windows())
                res[c] = timed(lambda: pk.prove_device(dz, rs), reps=9)
                pk.close()
            bc = keep_model(res, c0)
            chosen[role] = f"{key} c={bc}"
            res_k[role] = {str(c): round(v, 4) for c, v in res.items()}
            log(f"proof 2^{k} {role}: c={bc} ({res[bc]:.3f} ms; model c={c0} {res[c0]:.3f} ms)")
        table += list(chosen.values())
        results["fixed"][k] = res_k
        # validation: the proof under the cost model alone vs under the tuned entries,
        # alternated three times
        val = {"model": [], "tuned": []}
        for _ in range(3):
            for name, lines in (("model", []), ("tuned", ntt_lines + list(chosen.values()))):
                load(lines)
                pk = bendg.ProvingKey(kdir)
                val[name].append(timed(lambda: pk.prove_device(dz, rs), reps=9))
                pk.close()
        results["validation"][k] = {n: round(float(np.median(v)), 4) for n, v in val.items()}
        log(f"proof 2^{k}: model {results['validation'][k]['model']:.3f} ms, tuned "
            f"{results['validation'][k]['tuned']:.3f} ms")
        hip.free(dz)
    text = "\n".join(["# Device-tuned parameters for gfx950 (MI355X), written by bench.py --mode tune.",
                      "# msm <group> <windowed|fixed> <log2 total points, half steps> c= L= m= K=;  ntt <log2 n> tile= radix="]
                     + table) + "\n"
    (base / "tune_gfx950.txt").write_text(text)
    cand.unlink()
    bendg.tune_load(None)
    return {"metric": "device-tuned parameter table", "value": len(table), "unit": "entries", "n_gpus": 1,
            "table": text, "results": results, "tune_s": round(time.perf_counter() - t_start, 1)}


def main() -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    # the first proofs after the key load run ~3 % slower (fresh tables and buffers): ten
    # untimed proofs put the timed ones in steady state
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--k", type=int, default=18, help="log2 of the constraint count")
    ap.add_argument("--seed", type=int, default=42)
    ap.add_argument("--keys", default=os.environ.get("BG16_KEYS", "/tmp/bg16_keys"))
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-mix", action="store_true", help="single mode: skip the 64-proof batched mix")
    ap.add_argument("--mode", default="single", choices=["single", "batch", "serve", "primitives", "roofline", "tune", "sizes",
                                                               "launch"])
    ap.add_argument("--slots", type=lambda v: [int(x) for x in v.split(",")], default=[4, 8],
                    help="engine slots (batch: comma list to sweep)")
    ap.add_argument("--rates", type=lambda v: [float(x) for x in v.split(",")], default=[100.0, 200.0, 300.0],
                    help="serve: arrival rates in proofs/s")
    ap.add_argument("--duration", type=float, default=20.0, help="serve: seconds per arrival rate")
    ap.add_argument("--hw-queues", type=int, default=16,
                    help="HIP hardware queues for this process (each proof in flight uses four streams, "
                         "so the engine's slots need more than the default 4); 0 keeps the environment's value")
    a = ap.parse_args()
    if a.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) > 1:
        print("bench.py measures one GPU only: this build has no multi-GPU path (run with --gpus 1)",
              file=sys.stderr)
        return 2
    # must precede the first HIP call of the process
    hwq = a.hw_queues
    if hwq:
        os.environ["GPU_MAX_HW_QUEUES"] = str(min(32, max(1, hwq)))
    if a.mode != "single":
        line = {"batch": batch_mode, "serve": serve_mode, "primitives": primitives_mode,
                "roofline": roofline_mode, "tune": tune_mode, "sizes": sizes_mode, "launch": launch_mode}[a.mode](a)
        print(json.dumps(line), flush=True)
        return 0

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
    # timed region: `steps` proofs back to back, key and packed witness resident in HBM,
    # no per-kernel event timing
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
    ms = elapsed * 1e3 / a.steps
    # the same `steps` proofs with the witness handed over in host memory as FORMAT.md
    # words (the reference's boundary: packed into pinned staging and copied in by the
    # call); one untimed call first sizes the staging buffer.  Reported beside `value`.
    p_host, _ = pk.prove(z16, rs16, fmt="words16")
    hip.sync()
    t0 = time.perf_counter()
    pcie_walls = []
    for _ in range(a.steps):
        s = time.perf_counter()
        p_h, _ = pk.prove(z16, rs16, fmt="words16")
        pcie_walls.append((time.perf_counter() - s) * 1e3)
        if not np.array_equal(p_h, p_host):
            log("host-witness proofs differ between calls")
            return 1
    pcie_ms = (time.perf_counter() - t0) * 1e3 / a.steps
    if not np.array_equal(bendg.pack16(p_host.reshape(-1, 16)).reshape(-1), ref):
        log("host-witness proof differs from the device-witness proof")
        return 1
    # the same proofs with the combination and affine normalisation on the device
    # (BG_COMBINE_DEVICE): reported beside `value`, every proof equal to the host-combined one
    pk.prove_device(dz, rs, device_combine=True)
    hip.sync()
    t0 = time.perf_counter()
    for _ in range(a.steps):
        p_d, _ = pk.prove_device(dz, rs, device_combine=True)
        if not np.array_equal(p_d, ref):
            log("device-combined proof differs from the host-combined proof")
            return 1
    device_combine_ms = (time.perf_counter() - t0) * 1e3 / a.steps
    # per-kernel HIP-event times (overlapped, per proof) from separate proofs
    bendg.set_timing(True)
    bendg.timing_reset()
    nt = max(3, min(a.steps, 10))
    for _ in range(nt):
        pk.prove_device(dz, rs)
    timing = bendg.timing_report()
    bendg.set_timing(False)
    per_proof = {k: round(v[0] * v[1] / nt, 4) for k, v in sorted(timing.items(), key=lambda kv: -kv[1][0] * kv[1][1])}
    peak = probe_peak()
    log("standalone MSM pass")
    alone = standalone_msm_pass(key, hdr, 5)
    for k, v in sorted(alone.items()):
        log(f"  {k:22s} {v['avg_ms']:8.3f} ms/launch (n={v['n']})")
    roof = roofline(alone, peak)
    # the benchmarked proof checked with the pairing equation (host verifier, vk.bin)
    verified = None
    if (key / "vk.bin").exists():
        vk = np.fromfile(key / "vk.bin", dtype=np.uint32)
        verified = bool(bendg.verify(vk, p_host, z16[16:32]))
        log(f"proof verifies against vk.bin: {verified}")
        if not verified:
            return 1
    log("engine (concurrent proofs of this key)")
    batched = engine_rate(pk, dz, rs, 32, 4, ref)
    batched["vs_single"] = round(batched["proofs_per_s"] / (1e3 / ms), 3)
    mix = None
    if not a.no_mix:
        log("batched mix: 64 proofs, K uniform in 10..18, own keys")
        mix = batched_mix(a, hip, [2, 3, 4])
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
        "device_combine_ms": round(device_combine_ms, 3),
        "pcie_inclusive_latency_ms": {"p50": round(float(np.percentile(pcie_walls, 50)), 3),
                                      "p99": round(float(np.percentile(pcie_walls, 99)), 3)},
        "proof_verified": verified,
        "key_load_ms": round(load_ms, 1),
        "kernel_ms_per_proof_overlapped": per_proof,
        "kernel_ms_standalone": {k: v["avg_ms"] for k, v in alone.items()},
        "roofline": roof,
        "batched_same_key": batched,
        "batched_mix": mix,
    }
    if not a.no_cpu_baseline:
        log("cpu baseline (C oracle, one proof)")
        line["cpu_baseline"] = cpu_baseline(key, ref, Path(a.keys))
    print(json.dumps(line), flush=True)
    hip.free(dz)
    pk.close()
    return 0


if __name__ == "__main__":
    sys.exit(main())
