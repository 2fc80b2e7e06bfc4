"""Resident-key prover binding: the device-side counterpart of bend2/prove.bend's Pr.run
(:129-162) behind ``bg_pk_load`` / ``bg_prove``, plus the SquareChain key generator and the
reference's proof text format (bend2/FORMAT.md:33)."""
from __future__ import annotations

import ctypes
from pathlib import Path

import numpy as np

from . import FMT_PACKED, FMT_WORDS16, _check, _p, lib

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

    def device_bytes(self) -> int:
        b = np.zeros(1, dtype=np.uint64)
        _check(lib().bg_pk_device_bytes(self._h, _p(b)))
        return int(b[0])

    def prove_device(self, witness_ptr: int, rs: np.ndarray):
        """Proof with the packed witness already resident in HBM (device pointer);
        rs is host (packed).  Returns (packed proof, phase times)."""
        r = np.ascontiguousarray(rs, dtype=np.uint32).reshape(-1)
        if r.size != 16:
            raise ValueError("rs must hold two packed Fr elements")
        out = np.zeros(64, dtype=np.uint32)
        t = np.zeros(9, dtype=np.float64)
        _check(lib().bg_prove(self._h, ctypes.c_void_p(witness_ptr), _p(r), _p(out), FMT_PACKED | 2, _p(t)))
        return out, dict(zip(PHASES, (float(x) for x in t)))

    def prove(self, witness: np.ndarray, rs: np.ndarray, fmt: str = "words16"):
        """witness / rs in FORMAT.md words (``words16``) or packed u32; returns
        (proof uint32 array [A | B | C] in the same encoding, phase times dict)."""
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
        flag = FMT_WORDS16 if fmt == "words16" else FMT_PACKED
        _check(lib().bg_prove(self._h, _p(w), _p(r), _p(out), flag, _p(t)))
        return out, dict(zip(PHASES, (float(x) for x in t)))

    def close(self):
        if self._h:
            lib().bg_pk_free(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


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
