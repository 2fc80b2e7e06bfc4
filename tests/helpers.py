"""Test helpers: seeded random field elements and curve points produced by the CPU
oracle, in the packed (8 x u32 per element) encoding the library takes."""
from __future__ import annotations

import random

import numpy as np

from oracle import oracle as O


def rand_fe(rnd: random.Random, p: int, n: int) -> list[int]:
    return [rnd.randrange(p) for _ in range(n)]


def ints_u32(xs, mont_p=None) -> np.ndarray:
    """ints -> (n, 8) uint32 packed (optionally Montgomery)."""
    return O.ints_to_array(xs, mont_p).view(np.uint32).reshape(len(xs), 8)


def u32_ints(a, unmont_p=None) -> list[int]:
    a = np.ascontiguousarray(a, dtype=np.uint32).reshape(-1, 8)
    return O.array_to_ints(a.view(np.uint64).reshape(-1, 4), unmont_p)


def g1_points(rnd: random.Random, n: int, scalars=None):
    """n affine G1 points k_i * G as (n, 16) uint32 Montgomery, and the k_i."""
    ks = scalars if scalars is not None else [rnd.randrange(1, O.P_R) for _ in range(n)]
    gen = O.g1_pack([O.G1_GEN])[0]
    out = np.zeros((n, 8), dtype=np.uint64)
    for i, k in enumerate(ks):
        kk = np.array(O.to_words(k), dtype=np.uint64)
        r = np.zeros(8, dtype=np.uint64)
        O.lib().or_g1_mul_aff(O._ptr(r), O._ptr(gen), O._ptr(kk))
        out[i] = r
    return out.view(np.uint32).reshape(n, 16), ks


def g2_points(rnd: random.Random, n: int, scalars=None):
    ks = scalars if scalars is not None else [rnd.randrange(1, O.P_R) for _ in range(n)]
    gen = O.g2_pack([O.G2_GEN])[0]
    out = np.zeros((n, 16), dtype=np.uint64)
    for i, k in enumerate(ks):
        kk = np.array(O.to_words(k), dtype=np.uint64)
        r = np.zeros(16, dtype=np.uint64)
        O.lib().or_g2_mul_aff(O._ptr(r), O._ptr(gen), O._ptr(kk))
        out[i] = r
    return out.view(np.uint32).reshape(n, 32), ks


class DeviceBuffers:
    """Device memory for tests that call the C ABI with device pointers (HIP runtime
    through ctypes); every buffer is freed by free_all()."""

    def __init__(self):
        import ctypes
        self.ct = ctypes
        self.L = ctypes.CDLL("libamdhip64.so")
        self.L.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.L.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.L.hipFree.argtypes = [ctypes.c_void_p]
        self.ptrs = []

    def alloc(self, nbytes: int) -> int:
        p = self.ct.c_void_p()
        assert self.L.hipMalloc(self.ct.byref(p), max(nbytes, 4)) == 0
        self.ptrs.append(p.value)
        return p.value

    def upload(self, arr: np.ndarray) -> int:
        a = np.ascontiguousarray(arr)
        p = self.alloc(a.nbytes)
        assert self.L.hipMemcpy(self.ct.c_void_p(p), a.ctypes.data_as(self.ct.c_void_p), a.nbytes, 1) == 0
        return p

    def download(self, p: int, arr: np.ndarray) -> np.ndarray:
        assert self.L.hipMemcpy(arr.ctypes.data_as(self.ct.c_void_p), self.ct.c_void_p(p), arr.nbytes, 2) == 0
        return arr

    def free_all(self) -> None:
        for p in self.ptrs:
            self.L.hipFree(self.ct.c_void_p(p))
        self.ptrs = []
