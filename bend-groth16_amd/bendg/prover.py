"""Resident-key prover binding: the device-side counterpart of bend2/prove.bend's Pr.run
(:129-162) behind ``bg_pk_load`` / ``bg_prove``, plus the SquareChain key generator and the
reference's proof text format (bend2/FORMAT.md:33)."""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import COMBINE_DEVICE, FMT_PACKED, FMT_WORDS16, _check, _p, lib

PHASES = ("upload", "evals", "qap", "msm_a", "msm_b1", "msm_b2", "msm_h", "msm_l", "total")


class ProvingKey:
    """A FORMAT.md key directory loaded into HBM (``bg_pk_load``)."""

    def __init__(self, directory: str | Path):
        self.dir = Path(directory)
        h = ctypes.c_void_p()
        _check(lib().bg_pk_load(str(self.dir).encode(), ctypes.byref(h)))
        self._h = h
        hdr = np.zeros(8, dtype=np.uint32)
        _check(lib().bg_pk_header(self._h, _p(hdr)))
        self.header = [int(x) for x in hdr]

    @property
    def handle(self):
        return self._h

    def windows(self) -> dict:
        """Window widths of the resident fixed-base tables: G1 batch (A, B1, L), H, B2."""
        c = np.zeros(3, dtype=np.int32)
        _check(lib().bg_pk_windows(self._h, _p(c)))
        return {"g1": int(c[0]), "h": int(c[1]), "b2": int(c[2])}

    def device_bytes(self) -> int:
        b = np.zeros(1, dtype=np.uint64)
        _check(lib().bg_pk_device_bytes(self._h, _p(b)))
        return int(b[0])

    def prove_device(self, witness_ptr: int, rs: np.ndarray, device_combine: bool = False):
        """Proof with the packed witness already resident in HBM (device pointer);
        rs is host (packed).  device_combine: the combination and affine points on the
        device (BG_COMBINE_DEVICE).  Returns (packed proof, phase times)."""
        r = np.ascontiguousarray(rs, dtype=np.uint32).reshape(-1)
        if r.size != 16:
            raise ValueError("rs must hold two packed Fr elements")
        out = np.zeros(64, dtype=np.uint32)
        t = np.zeros(9, dtype=np.float64)
        flags = FMT_PACKED | 2 | (COMBINE_DEVICE if device_combine else 0)
        _check(lib().bg_prove(self._h, ctypes.c_void_p(witness_ptr), _p(r), _p(out), flags, _p(t)))
        return out, dict(zip(PHASES, (float(x) for x in t)))

    def prove(self, witness: np.ndarray, rs: np.ndarray, fmt: str = "words16", device_combine: bool = False):
        """witness / rs in FORMAT.md words (``words16``) or packed u32; returns
        (proof uint32 array [A | B | C] in the same encoding, phase times dict).
        device_combine: the combination and affine points on the device."""
        ew = 16 if fmt == "words16" else 8
        w = np.ascontiguousarray(witness, dtype=np.uint32).reshape(-1)
        r = np.ascontiguousarray(rs, dtype=np.uint32).reshape(-1)
        if r.size != 2 * ew:
            raise ValueError("rs must hold two Fr elements")
        n_z = self.header[1] + self.header[2]
        if w.size != n_z * ew:
            raise ValueError(f"witness must hold {n_z} Fr elements")
        out = np.zeros(8 * ew, dtype=np.uint32)
        t = np.zeros(9, dtype=np.float64)
        flag = (FMT_WORDS16 if fmt == "words16" else FMT_PACKED) | (COMBINE_DEVICE if device_combine else 0)
        _check(lib().bg_prove(self._h, _p(w), _p(r), _p(out), flag, _p(t)))
        return out, dict(zip(PHASES, (float(x) for x in t)))

    def close(self):
        """Frees the key.  Raises (and keeps the key loaded) while a proof of it is queued
        or running in an engine."""
        if self._h:
            _check(lib().bg_pk_free(self._h))
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class Engine:
    """Continuous-batching prover (``bg_engine_*``): ``submit`` queues a proof and returns
    a ticket, ``wait`` returns (proof, {queued, proving, latency} ms).  Each proof equals
    what ``ProvingKey.prove`` returns for the same inputs."""

    def __init__(self, slots: int = 4, hbm_budget: int = 0):
        h = ctypes.c_void_p()
        _check(lib().bg_engine_create(slots, hbm_budget, ctypes.byref(h)))
        self._h = h
        self._keep = {}

    def submit(self, pk: ProvingKey, witness, rs, fmt: str = "words16") -> int:
        """witness / rs as ``ProvingKey.prove``; or witness = device pointer (int,
        packed, kept valid by the caller until ``wait``)."""
        t = ctypes.c_uint64()
        r = np.ascontiguousarray(rs, dtype=np.uint32).reshape(-1)
        if isinstance(witness, int):
            if r.size != 16:
                raise ValueError("rs must hold two packed Fr elements")
            _check(lib().bg_engine_submit(self._h, pk.handle, ctypes.c_void_p(witness), _p(r), FMT_PACKED | 2,
                                          ctypes.byref(t)))
        else:
            ew = 16 if fmt == "words16" else 8
            w = np.ascontiguousarray(witness, dtype=np.uint32).reshape(-1)
            if r.size != 2 * ew or w.size != (pk.header[1] + pk.header[2]) * ew:
                raise ValueError("witness / rs sizes do not match the key")
            flag = FMT_WORDS16 if fmt == "words16" else FMT_PACKED
            _check(lib().bg_engine_submit(self._h, pk.handle, _p(w), _p(r), flag, ctypes.byref(t)))
        # the key object stays referenced until the proof is collected, so it cannot be
        # garbage-collected (and freed) while the engine still uses it
        self._keep[t.value] = (16 if (fmt == "words16" and not isinstance(witness, int)) else 8, pk)
        return t.value

    def wait(self, ticket: int):
        ew, _pk = self._keep.pop(ticket, (8, None))
        out = np.zeros(8 * ew, dtype=np.uint32)
        tm = np.zeros(3, dtype=np.float64)
        _check(lib().bg_engine_wait(self._h, ticket, _p(out), _p(tm)))
        return out, {"queued": float(tm[0]), "proving": float(tm[1]), "latency": float(tm[2])}

    def info(self) -> dict:
        v = np.zeros(4, dtype=np.uint64)
        _check(lib().bg_engine_info(self._h, _p(v)))
        return {"slots": int(v[0]), "reserved_bytes": int(v[1]), "budget_bytes": int(v[2]), "completed": int(v[3])}

    def close(self):
        if self._h:
            lib().bg_engine_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def prove_batch(pks, witnesses, rss, fmt: str = "words16", slots: int = 4) -> list:
    """``bg_prove_batch``: one proof per (key, witness, rs), returned in order."""
    n = len(pks)
    if len(witnesses) != n or len(rss) != n:
        raise ValueError("prove_batch: one witness and one rs per key")
    ew = 16 if fmt == "words16" else 8
    ws = [np.ascontiguousarray(w, dtype=np.uint32).reshape(-1) for w in witnesses]
    rr = [np.ascontiguousarray(r, dtype=np.uint32).reshape(-1) for r in rss]
    for pk, w, r in zip(pks, ws, rr):
        if w.size != (pk.header[1] + pk.header[2]) * ew or r.size != 2 * ew:
            raise ValueError("prove_batch: witness / rs sizes do not match the key")
    outs = [np.zeros(8 * ew, dtype=np.uint32) for _ in range(n)]
    arr = lambda xs: ctypes.cast((ctypes.c_void_p * n)(*[x.ctypes.data for x in xs]), ctypes.c_void_p)  # noqa: E731
    kp = ctypes.cast((ctypes.c_void_p * n)(*[pk.handle.value for pk in pks]), ctypes.c_void_p)
    flag = FMT_WORDS16 if fmt == "words16" else FMT_PACKED
    _check(lib().bg_prove_batch(n, kp, arr(ws), arr(rr), arr(outs), flag, slots))
    return outs


def setup_square_chain(directory: str | Path, log2: int, seed: int = 42, threads: int = 0) -> Path:
    """Write a SquareChain(2^log2) key directory (FORMAT.md files + trapdoor.bin)."""
    d = Path(directory)
    d.mkdir(parents=True, exist_ok=True)
    _check(lib().bg_setup_square_chain(str(d).encode(), log2, seed, threads))
    return d


def format_proof(proof_packed: np.ndarray) -> str:
    """A/B/C lines as the reference prover prints them (canonical big-endian hex)."""
    p = np.ascontiguousarray(proof_packed, dtype=np.uint32).reshape(-1)
    buf = ctypes.create_string_buffer(1024)
    n = _check(lib().bg_proof_format(_p(p), buf, 1024))
    return buf.raw[:n].decode()
