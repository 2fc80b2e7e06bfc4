"""The device Montgomery product's 29-bit digit arithmetic (csrc/bg_mul29.hpp), compiled
for the host with g++ and checked against Python big integers (CPU only).

The header is plain C++, so the same source the gfx950 kernels inline is exercised here:
a * b * 2^-256 mod p for Fr and Fq, results below 2p for operands below 2p (the lazy
range the kernels keep), including the extremes 0, 1, p - 1, p, 2p - 1 and operands whose
top digit is maximal."""
import ctypes
import random
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CSRC = ROOT / "bend-groth16_amd" / "csrc"

P_R = 21888242871839275222246405745257275088548364400416034343698204186575808495617
P_Q = 21888242871839275222246405745257275088696311157297823662689037894645226208583

SRC = """
#include "bg_mul29.hpp"
extern "C" void mulw(const uint32_t* a, const uint32_t* b, uint32_t* o, const uint32_t* p29, uint32_t n0) {
  bg::m29::mul_words(a, b, o, p29, n0);
}
extern "C" void mul4w(const uint32_t* a, const uint32_t* b, uint32_t* o, const uint32_t* p29, uint32_t n0) {
  const uint32_t* as[4] = {a, a + 8, a + 16, a + 24};
  const uint32_t* bs[4] = {b, b + 8, b + 16, b + 24};
  bg::m29::mul4_words(as, bs, o, p29, n0);
}
extern "C" void mul2w(const uint32_t* a1, const uint32_t* b1, const uint32_t* a2, const uint32_t* b2, uint32_t* o,
                      const uint32_t* p29, uint32_t n0) {
  bg::m29::mul2_words(a1, b1, a2, b2, o, p29, n0);
}
"""


@pytest.fixture(scope="module")
def lib(tmp_path_factory):
    if not shutil.which("g++"):
        pytest.skip("g++ not available")
    d = tmp_path_factory.mktemp("m29")
    (d / "m29.cpp").write_text(SRC)
    so = d / "libm29.so"
    subprocess.run(["g++", "-O2", "-shared", "-fPIC", "-I", str(CSRC), str(d / "m29.cpp"), "-o", str(so)], check=True)
    return ctypes.CDLL(str(so))


W8 = ctypes.c_uint32 * 8
W9 = ctypes.c_uint32 * 9


def words(x):
    return W8(*[(x >> (32 * i)) & 0xFFFFFFFF for i in range(8)])


def value(w):
    return sum(int(v) << (32 * i) for i, v in enumerate(w))


def digits29(p):
    return W9(*[(p >> (29 * i)) & ((1 << 29) - 1) for i in range(9)])


@pytest.mark.parametrize("p", [P_R, P_Q], ids=["Fr", "Fq"])
def test_montgomery_product_29bit_digits(lib, p):
    n0 = (-pow(p, -1, 1 << 29)) % (1 << 29)
    p29 = digits29(p)
    rinv = pow(1 << 256, -1, p)
    rng = random.Random(p & 0xFFFF)
    edge = [0, 1, 2, p - 1, p, p + 1, 2 * p - 1, 2 * p - 2, (1 << 253), 2 * p - (1 << 100)]
    cases = [(a, b) for a in edge for b in edge]
    cases += [(rng.randrange(2 * p), rng.randrange(2 * p)) for _ in range(4000)]
    out = W8()
    for a, b in cases:
        lib.mulw(words(a), words(b), out, p29, n0)
        r = value(out)
        assert r % p == a * b * rinv % p, (hex(a), hex(b))
        assert r < 2 * p, (hex(a), hex(b), r / p)


@pytest.mark.parametrize("p", [P_R, P_Q], ids=["Fr", "Fq"])
def test_sum_of_two_products(lib, p):
    """(a1 b1 + a2 b2) 2^-256 mod p with one reduction, below 2.6 p for operands below 2p
    (the Fq2 product's two components), at the extremes and at random."""
    n0 = (-pow(p, -1, 1 << 29)) % (1 << 29)
    p29 = digits29(p)
    rinv = pow(1 << 256, -1, p)
    rng = random.Random(p & 0xFFF)
    edge = [0, 1, p - 1, p, 2 * p - 1, (1 << 253)]
    cases = [(a, b, c, d) for a in edge for b in edge for c in (0, 2 * p - 1) for d in (1, 2 * p - 1)]
    cases += [tuple(rng.randrange(2 * p) for _ in range(4)) for _ in range(3000)]
    out = W8()
    worst = 0.0
    for a1, b1, a2, b2 in cases:
        lib.mul2w(words(a1), words(b1), words(a2), words(b2), out, p29, n0)
        r = value(out)
        assert r % p == (a1 * b1 + a2 * b2) * rinv % p
        assert r < 2.6 * p
        worst = max(worst, r / p)
    assert worst > 1.0  # the bound is exercised: results above p occur


@pytest.mark.parametrize("p", [P_R, P_Q], ids=["Fr", "Fq"])
def test_sum_of_four_products(lib, p):
    """(sum of four a_k b_k) 2^-256 mod p with one reduction, below 4.1 p for operands
    below 2p (the G2 y coordinate's components), at the extremes and at random."""
    n0 = (-pow(p, -1, 1 << 29)) % (1 << 29)
    p29 = digits29(p)
    rinv = pow(1 << 256, -1, p)
    rng = random.Random(p & 0xFFFF)
    W32 = ctypes.c_uint32 * 32
    top = 2 * p - 1
    cases = [[top] * 8, [0] * 8, [1] * 8, [p] * 8]
    cases += [[rng.randrange(2 * p) for _ in range(8)] for _ in range(2000)]
    cases += [[rng.choice([0, 1, p - 1, p, top]) for _ in range(8)] for _ in range(300)]
    out = W8()
    worst = 0.0
    for v in cases:
        a = W32(*[w for x in v[:4] for w in words(x)])
        b = W32(*[w for x in v[4:] for w in words(x)])
        lib.mul4w(a, b, out, p29, n0)
        r = value(out)
        assert r % p == sum(x * y for x, y in zip(v[:4], v[4:])) * rinv % p
        assert r < 4.1 * p
        worst = max(worst, r / p)
    assert worst > 2.0  # results above 2p occur: two conditional subtractions are needed
