"""bg_qap_h_noalloc (qap_h with all working memory in a static device array, for callers
that cannot allocate) against bg_qap_h, the entry point it mirrors (itself pinned to the
oracle's qap_h and the QAP identity in test_gpu_ntt.py), bit for bit, through the C ABI
with device buffers: consecutive calls of different sizes (a smaller domain after a
larger one, then the larger again, so the tables held in the array are rebuilt each
time), and a domain beyond the array's bound rejected."""
import ctypes

import numpy as np
import pytest

import bendg
from oracle import oracle as O
from helpers import DeviceBuffers

pytestmark = pytest.mark.gpu

MAX_LOG2 = 17


def _call(dev, a, b, c, log2n):
    da, db, dc = (dev.upload(v) for v in (a, b, c))
    dh = dev.alloc(a.nbytes)
    rc = bendg.lib().bg_qap_h_noalloc(ctypes.c_void_p(da), ctypes.c_void_p(db), ctypes.c_void_p(dc), log2n,
                                      ctypes.c_void_p(dh))
    return rc, dh


def test_qap_h_noalloc_equals_qap_h(bg):
    L = bendg.lib()
    L.bg_qap_h_noalloc.argtypes = [ctypes.c_void_p] * 3 + [ctypes.c_int, ctypes.c_void_p]
    dev = DeviceBuffers()
    try:
        for step, log2n in enumerate((16, 11, MAX_LOG2, 5)):
            rng = np.random.default_rng(60 + step)
            n = 1 << log2n
            # uniform values below r: each is the Montgomery form of a uniform element
            a, b, c = (O.rand_scalars(rng, n).view(np.uint32).reshape(n, 8) for _ in range(3))
            want = bendg.qap_h(a, b, c, log2n)
            print(f"noalloc qap_h call {step}: 2^{log2n}", flush=True)
            rc, dh = _call(dev, a, b, c, log2n)
            assert rc == 0, bendg.lib().bg_last_error()
            got = dev.download(dh, np.zeros_like(a))
            assert np.array_equal(got, want), f"call {step} (2^{log2n})"
            dev.free_all()
        # beyond the static array: an argument error, nothing launched
        rc = L.bg_qap_h_noalloc(ctypes.c_void_p(8), ctypes.c_void_p(8), ctypes.c_void_p(8), MAX_LOG2 + 1,
                                ctypes.c_void_p(8))
        assert rc == -1
    finally:
        dev.free_all()
