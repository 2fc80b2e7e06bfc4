"""Python host binding of libbg16.so (the gfx950 Groth16/BN254 hot path).

Mirrors the reference's operator surface for the device path -- ``msm_g1`` / ``msm_g2``
(bend2/src/msm.bend:293-298, :525-530), ``ntt`` (bend2/src/ntt.bend:431-442),
``qap_h`` (:459) and the prover entry (bend2/prove.bend:129-168) -- over the C ABI in
``include/bg16.h``.  Arrays are numpy ``uint32`` in one of the two element encodings:

* ``"words16"``: the reference's ``Array<U32>`` layout (bend2/FORMAT.md: 16 words of
  16-bit limbs per field element);
* ``"packed"``: 8 x u32 little-endian words per element (the same integers).

There is no CPU fallback: every call runs the HIP kernels and raises if the library is
missing or the device call fails.
"""
from __future__ import annotations

import ctypes
import os
from pathlib import Path

import numpy as np

PKG_DIR = Path(__file__).resolve().parent.parent
# BG16_LIB selects another build of the library (A/B timing of two builds)
LIB_PATH = Path(os.environ["BG16_LIB"]) if os.environ.get("BG16_LIB") else PKG_DIR / "libbg16.so"

FMT_WORDS16 = 0
FMT_PACKED = 1
MEM_HOST = 0
MEM_DEVICE = 2
COMBINE_DEVICE = 4  # bg_prove: combination and affine proof points on the device

NTT_FFT, NTT_IFFT, NTT_COSET_FFT, NTT_COSET_IFFT = 0, 1, 2, 3

_ERRORS = {-1: "bad argument", -2: "device error", -3: "out of memory", -4: "I/O error", -5: "internal error"}


class BgError(RuntimeError):
    def __init__(self, code: int, msg: str):
        super().__init__(f"{_ERRORS.get(code, code)}: {msg}")
        self.code = code


_lib = None


def lib() -> ctypes.CDLL:
    """Load libbg16.so; fails loudly when the native library is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not LIB_PATH.exists():
        raise RuntimeError(f"{LIB_PATH} is not built (run `python bend-groth16_amd/build.py`)")
    L = ctypes.CDLL(str(LIB_PATH))
    P, U64, I = ctypes.c_void_p, ctypes.c_uint64, ctypes.c_int
    sigs = {
        "bg_init": [I],
        "bg_last_error": [],
        "bg_set_timing": [I],
        "bg_timing_report": [P, P, P, I],
        "bg_timing_reset": [],
        "bg_msm_g1": [P, P, U64, P, I, I],
        "bg_msm_g2": [P, P, U64, P, I, I],
        "bg_ntt": [P, I, I, I],
        "bg_qap_h": [P, P, P, I, P, I],
        "bg_pk_load": [ctypes.c_char_p, ctypes.POINTER(ctypes.c_void_p)],
        "bg_pk_free": [P],
        "bg_pk_header": [P, P],
        "bg_pk_device_bytes": [P, P],
        "bg_prove": [P, P, P, P, I, P],
        "bg_graph_info": [P],
        "bg_proof_format": [P, ctypes.c_char_p, I],
        "bg_setup_square_chain": [ctypes.c_char_p, I, U64, I],
        "bg_kat_field_mul_xor": [I, P],
        "bg_kat_g1_add_chain": [U64, P],
        "bg_field_op": [I, I, P, P, P, U64],
        "bg_curve_op": [I, I, P, P, P, U64],
        "bg_probe_rate": [I, U64, P],
        "bg_probe_launch": [I, I, I, P],
        "bg_msm_plan_info": [U64, I, I, P],
        "bg_msm_batch": [I, ctypes.c_uint32, P, P, P, P, I, I],
        "bg_engine_create": [I, U64, ctypes.POINTER(ctypes.c_void_p)],
        "bg_engine_submit": [P, P, P, P, I, ctypes.POINTER(ctypes.c_uint64)],
        "bg_engine_wait": [P, U64, P, P],
        "bg_engine_info": [P, P],
        "bg_engine_destroy": [P],
        "bg_prove_batch": [I, P, P, P, P, I, I],
        "bg_fixed_base_mul": [I, P, U64, P, I],
        "bg_gen_bases": [I, U64, U64, P, P, I],
        "bg_verify": [P, U64, P, P, ctypes.c_uint32, I],
        "bg_pairing_check": [P, P, ctypes.c_uint32, I],
        "bg_tune_load": [ctypes.c_char_p],
        "bg_pk_windows": [P, P],
        "bg_tune_lookup": [I, I, I, U64, P],
        "bg_qap_h_noalloc": [P, P, P, I, P],
    }
    for name, args in sigs.items():
        fn = getattr(L, name, None)
        if fn is None:
            continue
        fn.argtypes = args
        fn.restype = I
    L.bg_last_error.restype = ctypes.c_char_p
    _lib = L
    return L


def exported_symbols() -> list[str]:
    return [n for n in ("bg_init",) if hasattr(lib(), n)]


def _check(rc: int) -> int:
    if rc < 0:
        raise BgError(rc, lib().bg_last_error().decode(errors="replace"))
    return rc


def _p(a: np.ndarray | None):
    if a is None:
        return None
    assert a.flags["C_CONTIGUOUS"], "arrays must be C-contiguous"
    return a.ctypes.data_as(ctypes.c_void_p)


def _u32(a) -> np.ndarray:
    return np.ascontiguousarray(np.asarray(a).view(np.uint32) if np.asarray(a).dtype != np.uint32 else a,
                                dtype=np.uint32)


def init(device: int = 0) -> None:
    _check(lib().bg_init(device))


def graph_info() -> dict:
    """Process-wide counts of the single-proof HIP graphs (``bg_graph_info``)."""
    v = np.zeros(3, dtype=np.uint64)
    _check(lib().bg_graph_info(_p(v)))
    return {"recorded": int(v[0]), "launches": int(v[1]), "failed": int(v[2])}


# ---------------------------------------------------------------- encodings

def pack16(words16: np.ndarray) -> np.ndarray:
    """FORMAT.md 16-bit-limb words (..., 16k) -> packed u32 words (..., 8k)."""
    w = np.ascontiguousarray(words16, dtype=np.uint32)
    lo = w[..., 0::2]
    hi = w[..., 1::2]
    return np.ascontiguousarray(lo | (hi << np.uint32(16)))


def unpack16(packed: np.ndarray) -> np.ndarray:
    p = np.ascontiguousarray(packed, dtype=np.uint32)
    out = np.empty(p.shape[:-1] + (p.shape[-1] * 2,), dtype=np.uint32)
    out[..., 0::2] = p & np.uint32(0xFFFF)
    out[..., 1::2] = p >> np.uint32(16)
    return out


# ---------------------------------------------------------------- diagnostics

def kat_field_mul_xor(log2_count: int) -> int:
    out = np.zeros(1, dtype=np.uint32)
    _check(lib().bg_kat_field_mul_xor(log2_count, _p(out)))
    return int(out[0])


def kat_g1_add_chain(k: int) -> int:
    out = np.zeros(1, dtype=np.uint32)
    _check(lib().bg_kat_g1_add_chain(k, _p(out)))
    return int(out[0])


def field_op(field: int, op: int, a: np.ndarray, b: np.ndarray | None = None) -> np.ndarray:
    """field 0 Fr / 1 Fq / 2 Fq2; op 0 add 1 sub 2 mul 3 neg 4 inv 5 sqr 6 to_mont 7 from_mont."""
    w = 16 if field == 2 else 8
    a = _u32(a).reshape(-1, w)
    b = a if b is None else _u32(b).reshape(-1, w)
    assert a.shape == b.shape
    out = np.zeros_like(a)
    _check(lib().bg_field_op(field, op, _p(a), _p(b), _p(out), a.shape[0]))
    return out


def curve_op(group: int, op: int, p: np.ndarray, q: np.ndarray | None = None) -> np.ndarray:
    """group 1 G1 / 2 G2; op 0 add (mixed), 1 double, 2 neg, 3 add (full XYZZ)."""
    w = 16 if group == 1 else 32
    p = _u32(p).reshape(-1, w)
    q = p if q is None else _u32(q).reshape(-1, w)
    out = np.zeros_like(p)
    _check(lib().bg_curve_op(group, op, _p(p), _p(q), _p(out), p.shape[0]))
    return out


def probe_rate(kind: int, iters: int) -> tuple[float, float]:
    r = np.zeros(2, dtype=np.float64)
    _check(lib().bg_probe_rate(kind, iters, _p(r)))
    return float(r[0]), float(r[1])


LAUNCH_VARIANTS = ("eager_linear", "graph_linear", "eager_fork", "graph_fork", "graph_fork_event_nodes",
                   "graph_per_stream")


def probe_launch(kernels: int, spin: int, reps: int = 40) -> dict:
    """``bg_probe_launch``: microseconds per repetition of each launch shape."""
    r = np.zeros(6, dtype=np.float64)
    _check(lib().bg_probe_launch(kernels, spin, reps, _p(r)))
    return {k: round(float(v), 1) for k, v in zip(LAUNCH_VARIANTS, r)}


# ---------------------------------------------------------------- hot path

def _fmt_flag(fmt: str) -> int:
    if fmt == "packed":
        return FMT_PACKED
    if fmt == "words16":
        return FMT_WORDS16
    raise ValueError(fmt)


def msm_g1(points: np.ndarray, scalars: np.ndarray, fmt: str = "packed", window_bits: int = 0) -> np.ndarray:
    """sum_i scalars[i] * points[i] over G1 (Msm.g1, bend2/src/msm.bend:297)."""
    ew = 8 if fmt == "packed" else 16
    points = _u32(points).reshape(-1, 2 * ew)
    scalars = _u32(scalars).reshape(-1, ew)
    n = scalars.shape[0]
    if points.shape[0] < n:
        raise ValueError("fewer points than scalars")
    out = np.zeros(2 * ew, dtype=np.uint32)
    _check(lib().bg_msm_g1(_p(points), _p(scalars), n, _p(out), _fmt_flag(fmt), window_bits))
    return out


def msm_g2(points: np.ndarray, scalars: np.ndarray, fmt: str = "packed", window_bits: int = 0) -> np.ndarray:
    """sum_i scalars[i] * points[i] over G2 (Msm.g2, bend2/src/msm.bend:529)."""
    ew = 8 if fmt == "packed" else 16
    points = _u32(points).reshape(-1, 4 * ew)
    scalars = _u32(scalars).reshape(-1, ew)
    n = scalars.shape[0]
    if points.shape[0] < n:
        raise ValueError("fewer points than scalars")
    out = np.zeros(4 * ew, dtype=np.uint32)
    _check(lib().bg_msm_g2(_p(points), _p(scalars), n, _p(out), _fmt_flag(fmt), window_bits))
    return out


def msm_batch(group: int, points: list, scalars: list, window_bits: int = 0) -> np.ndarray:
    """Batch of independent MSMs of one group (packed host arrays) in one launch sequence;
    returns (count, 16 | 32) packed affine results."""
    count = len(points)
    pw = 16 if group == 1 else 32
    ps = [_u32(p).reshape(-1, pw) for p in points]
    ss = [_u32(s).reshape(-1, 8) for s in scalars]
    ns = np.array([s.shape[0] for s in ss], dtype=np.uint64)
    for p, s in zip(ps, ss):
        if p.shape[0] < s.shape[0]:
            raise ValueError("fewer points than scalars")
    pp = (ctypes.c_void_p * count)(*[p.ctypes.data for p in ps])
    sp = (ctypes.c_void_p * count)(*[s.ctypes.data for s in ss])
    out = np.zeros((count, pw), dtype=np.uint32)
    _check(lib().bg_msm_batch(group, count, ctypes.cast(pp, ctypes.c_void_p), ctypes.cast(sp, ctypes.c_void_p),
                              _p(ns), _p(out), FMT_PACKED, window_bits))
    return out


def ntt(data: np.ndarray, log2n: int, direction: int, fmt: str = "packed") -> np.ndarray:
    """fft / ifft / coset_fft / coset_ifft (bend2/src/ntt.bend:431-442); returns a new array."""
    ew = 8 if fmt == "packed" else 16
    x = _u32(data).reshape(-1, ew).copy()
    if x.shape[0] != 1 << log2n:
        raise ValueError("data length must be 2^log2n elements")
    _check(lib().bg_ntt(_p(x), log2n, direction, _fmt_flag(fmt)))
    return x


def qap_h(a: np.ndarray, b: np.ndarray, c: np.ndarray, log2n: int, fmt: str = "packed") -> np.ndarray:
    """Ntt.qap_h (bend2/src/ntt.bend:459)."""
    ew = 8 if fmt == "packed" else 16
    a, b, c = (_u32(v).reshape(-1, ew) for v in (a, b, c))
    for v in (a, b, c):
        if v.shape[0] != 1 << log2n:
            raise ValueError("inputs must hold 2^log2n elements")
    h = np.zeros_like(a)
    _check(lib().bg_qap_h(_p(a), _p(b), _p(c), log2n, _p(h), _fmt_flag(fmt)))
    return h


def fixed_base_mul(group: int, scalars: np.ndarray, threads: int = 0) -> np.ndarray:
    """k_i * G for canonical packed scalars (n, 8) u32 -> (n, 16 | 32) packed affine."""
    s = _u32(scalars).reshape(-1, 8)
    out = np.zeros((s.shape[0], 16 if group == 1 else 32), dtype=np.uint32)
    _check(lib().bg_fixed_base_mul(group, _p(s), s.shape[0], _p(out), threads))
    return out


def gen_bases(group: int, seed: int, n: int, with_k: bool = False, threads: int = 0):
    """n seeded synthetic bases k_i G (packed affine); with_k also returns the k_i."""
    out = np.zeros((n, 16 if group == 1 else 32), dtype=np.uint32)
    k = np.zeros((n, 8), dtype=np.uint32) if with_k else None
    _check(lib().bg_gen_bases(group, seed, n, _p(out), _p(k), threads))
    return (out, k) if with_k else out


def verify(vk: np.ndarray, proof: np.ndarray, inputs: np.ndarray, fmt: str = "words16") -> bool:
    """Groth16 check e(A, B) = e(alpha, beta) e(IC(x), gamma) e(C, delta) (ref/src/verify.rs:8-27).

    vk: the words of a vk.bin; proof: A, B, C (the prover's output encoding); inputs: the
    public inputs as Montgomery Fr.  False for a proof that fails the equation; BgError for
    a point off its curve or outside G2 (arkworks refuses those too)."""
    v, pr, x = _u32(vk).reshape(-1), _u32(proof).reshape(-1), _u32(inputs).reshape(-1)
    ew = 16 if fmt == "words16" else 8
    return _check(lib().bg_verify(_p(v), v.size, _p(pr), _p(x), x.size // ew, _fmt_flag(fmt))) == 1


def pairing_check(g1s: np.ndarray, g2s: np.ndarray, fmt: str = "packed") -> bool:
    """prod_i e(P_i, Q_i) == 1 for affine G1 points P_i and G2 points Q_i (Montgomery)."""
    a, b = _u32(g1s).reshape(-1), _u32(g2s).reshape(-1)
    ew = 16 if fmt == "words16" else 8
    n = a.size // (2 * ew)
    if b.size != 4 * ew * n:
        raise ValueError("g1s and g2s hold different point counts")
    return _check(lib().bg_pairing_check(_p(a), _p(b), n, _fmt_flag(fmt))) == 1


def msm_plan_info(n: int, group: int = 1, window_bits: int = 0) -> dict:
    info = np.zeros(8, dtype=np.uint64)
    _check(lib().bg_msm_plan_info(n, group, window_bits, _p(info)))
    keys = ("c", "windows", "buckets", "entries", "L", "acc_threads", "groups1", "outputs2")
    return {k: int(v) for k, v in zip(keys, info)}


def tune_load(path: str | None = None) -> None:
    """(Re)load the device-tuned parameter table (None: the default next to the library)."""
    _check(lib().bg_tune_load(path.encode() if path else None))


def tune_lookup(kind: str, size: int, group: int = 1, fixed: bool = False) -> dict | None:
    """The table's entry: kind "msm" (size = total points) -> {c, L, m, K}; kind "ntt"
    (size = log2 n) -> {tile, radix}; None when the table has no entry."""
    v = np.zeros(4, dtype=np.int32)
    rc = lib().bg_tune_lookup(0 if kind == "msm" else 1, group, 1 if fixed else 0, size, _p(v))
    if rc == 1:
        return None
    _check(rc)
    if kind == "msm":
        return dict(zip(("c", "L", "m", "K"), (int(x) for x in v)))
    return {"tile": int(v[0]), "radix": int(v[1])}


# ---------------------------------------------------------------- timing

def set_timing(on: bool) -> None:
    _check(lib().bg_set_timing(1 if on else 0))


def timing_reset() -> None:
    _check(lib().bg_timing_reset())


def timing_report() -> dict[str, tuple[float, int]]:
    cap = 64
    names = ctypes.create_string_buffer(48 * cap)
    avg = np.zeros(cap, dtype=np.float64)
    cnt = np.zeros(cap, dtype=np.uint64)
    n = _check(lib().bg_timing_report(names, _p(avg), _p(cnt), cap))
    raw = names.raw
    out = {}
    for i in range(n):
        nm = raw[48 * i:48 * i + 48].split(b"\0", 1)[0].decode()
        out[nm] = (float(avg[i]), int(cnt[i]))
    return out


from .prover import Engine, ProvingKey, prove_batch, setup_square_chain, format_proof  # noqa: E402,F401
