"""Digit-domain bucket accumulation (bend-groth16_amd/csrc/bg_fd.hpp, bg_curve_fd.hpp).

1. Interval analysis of the lazy mixed addition with the constants the C++ uses: every
   subtraction's bias multiple covers its subtrahend's value and digit width, every
   product column stays below 2^64 for the digit widths that reach it, and the
   accumulator's bounds reach a fixed point (so a run of any length stays inside them).
   The worst case is what this checks; random tests cannot reach it.
2. The host build of the same code (tests/fd/fd_host_test.cpp) against the word-domain
   host arithmetic, with a product accumulator that flags any column above 2^64, plain
   and under AddressSanitizer + UndefinedBehaviorSanitizer.
"""
import math
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
CSRC = ROOT / "bend-groth16_amd" / "csrc"
P = 0x30644e72e131a029b85045b68181585d97816a916871ca8d3c208c16d87cfd47
EPS = P / 2**261
N = 2**29 - 1                      # normalised digit maximum
RAW = 64.0                         # load_raw: 32 W with W < 2p (else reduced below 2.1 p)
RAW_ANY = 2**261 / P               # any 8-word input before the reducing load


class V:
    """A value bound in units of p and the largest digit."""
    def __init__(self, val, dig=N):
        self.val, self.dig = val, dig


def _cols(pairs):
    s = sum(9 * a * b for a, b in pairs) + 9 * N * N + 2**35
    assert s < 2**64, f"product column reaches 2^{math.log2(s):.3f}"


def mul(a, b):
    _cols([(a.dig, b.dig)])
    return V(1 + a.val * b.val * EPS)


def sqr(a):
    assert a.dig <= N, "squaring needs normalised digits"
    # the doubled digit against the digit: at most 5 cross terms, 1 square, 9 reductions
    assert 5 * (2 * N) * N + N * N + 9 * N * N + 2**35 < 2**64
    return V(1 + a.val * a.val * EPS)


def mul2(a, b, c, d):
    _cols([(a.dig, b.dig), (c.dig, d.dig)])
    return V(1 + (a.val * b.val + c.val * d.val) * EPS)


def mul4(xs, ys):
    _cols([(x.dig, y.dig) for x, y in zip(xs, ys)])
    return V(1 + sum(x.val * y.val for x, y in zip(xs, ys)) * EPS)


def sub(a, b, k, t):
    """a + (k p raised by t) - b (bias_digits): digits 1..7 of the bias are at least
    t 2^29 - t, its top digit is (k p)'s lowered by t."""
    assert b.dig <= t * 2**29 - t, f"subtrahend digits {b.dig:#x} above the bias (t={t})"
    assert b.val <= k - 1, f"subtrahend below {b.val:.3f} p needs a bias above {k} p"
    return V(a.val + k, a.dig + (t + 1) * 2**29)


def neg(b, k, t):
    return sub(V(0.0, 0), b, k, t)


def add(a, b):
    return V(a.val + b.val, a.dig + b.dig)


def norm(a):
    assert a.val * P < 2**261
    return V(a.val, N)


def top(a, b):
    return V(max(a.val, b.val), max(a.dig, b.dig))


# constants of bg_curve_fd.hpp
G1K = dict(KX=8, KY=3, KYN=4, KX3=5, KT=8, KU=3, KRAW=65)
G2K = dict(KX=9, KY=3, KYN=6, KX3=6, KT=9, KNT=12, KU=3, KRAW=65, KZ=3, KPP=6, KSP=13, KSR=8)


def reduce_fast(a):
    """bg_fd.hpp reduce_fast: one subtraction of q p, q = floor(d8 / (p8 + 1)); the value
    must stay below 2^262 and the digits below 2^31; the remainder is below p + 2^240."""
    assert a.val * P < 2**262 and a.dig < 2**31
    return V(1 + 2**240 / P)


def start(k):
    # reduce_fast(x), reduce_fast(y or KRAW p - y), ZZ = ZZZ = 1
    x = reduce_fast(V(RAW))
    y = reduce_fast(neg(V(RAW), k["KRAW"], 1))
    return x, y, V(1.0), V(1.0)


def g1_madd(X, Y, ZZ, ZZZ, k):
    ax = ay = V(RAW)
    u2, s2 = mul(ax, ZZ), mul(ay, ZZZ)
    pp_ = norm(sub(u2, X, k["KX"], 1))
    r = norm(top(neg(add(s2, Y), k["KYN"], 2), sub(s2, Y, k["KY"], 1)))
    pp = sqr(pp_)
    zz3, ppp = mul(ZZ, pp), mul(pp_, pp)
    zzz3, q, r2 = mul(ZZZ, ppp), mul(X, pp), sqr(r)
    x3 = norm(sub(r2, add(ppp, add(q, q)), k["KX3"], 3))
    y3 = mul2(r, sub(q, x3, k["KT"], 1), Y, neg(ppp, k["KU"], 1))
    assert zz3.val < 2, "the flag test reads ZZ3 below 2p"
    return x3, y3, zz3, zzz3


def mul2x(a, b, kb):
    assert a.dig <= N and b.dig <= N
    nb1 = neg(b, kb, 1)
    c0 = mul2(a, b, a, nb1)
    c1 = mul2(a, b, a, b)
    return top(c0, c1)


def sqr2x(a, ka):
    assert a.dig <= N
    c0 = mul(norm(add(a, a)), sub(a, a, ka, 1))
    c1 = mul(add(a, a), a)
    return top(c0, c1)


def g2_madd(X, Y, ZZ, ZZZ, k):
    ax = ay = V(RAW)
    u2, s2 = mul2x(ax, ZZ, k["KZ"]), mul2x(ay, ZZZ, k["KZ"])
    pp_ = norm(sub(u2, X, k["KX"], 1))
    r = norm(top(neg(add(s2, Y), k["KYN"], 2), sub(s2, Y, k["KY"], 1)))
    pp = sqr2x(pp_, k["KSP"])
    zz3, ppp = mul2x(ZZ, pp, k["KPP"]), mul2x(pp_, pp, k["KPP"])
    zzz3, q, r2 = mul2x(ZZZ, ppp, k["KZ"]), mul2x(X, pp, k["KPP"]), sqr2x(r, k["KSR"])
    x3 = norm(sub(r2, add(ppp, add(q, q)), k["KX3"], 3))
    t = norm(sub(q, x3, k["KT"], 1))
    nt = norm(neg(t, k["KNT"], 1))
    nppp = norm(neg(ppp, k["KU"], 1))
    y3 = top(mul4([r, r, Y, Y], [t, nt, nppp, ppp]), mul4([r, r, Y, Y], [t, t, nppp, nppp]))
    assert zz3.val < 2
    return x3, y3, zz3, zzz3


def fixed_point(madd, k):
    st = start(k)
    for _ in range(200):
        nxt = madd(*st, k)
        grown = tuple(V(max(a.val, b.val)) for a, b in zip(st, nxt))
        if all(a.val == b.val for a, b in zip(st, grown)):
            return st
        st = grown
    raise AssertionError("the accumulator bounds do not converge")


def test_g1_bounds_fixed_point():
    X, Y, ZZ, ZZZ = fixed_point(g1_madd, G1K)
    # the bounds bg_curve_fd.hpp states (and the host test's stress levels stay inside)
    assert X.val < 6.22 and Y.val < 1.39 and ZZ.val < 1.01 and ZZZ.val < 1.01


def test_g2_bounds_fixed_point():
    X, Y, ZZ, ZZZ = fixed_point(g2_madd, G2K)
    assert X.val < 8.0 and Y.val < 1.85 and ZZ.val < 1.07 and ZZZ.val < 1.035


def test_raw_load_bound():
    # a word above 2p's top word is reduced first: below p (1 + 2^261/p eps) < 2.1 p
    assert mul(V(RAW_ANY), V(1.0)).val < 2.1
    # the store converts (D + k p) / 32, k < 32: below 2p for every accumulator coordinate
    for madd, k in ((g1_madd, G1K), (g2_madd, G2K)):
        assert all((v.val + 31) / 32 < 2 for v in fixed_point(madd, k))
    # 2p's top word 0x60c89ce5 bounds every word-domain value that skips the reduction
    assert 0x60c89ce5 * 2**224 < 2 * P


def test_smaller_biases_fail():
    """The constants are tight where the analysis says so: one step smaller breaks a
    bias condition (a dropped term or a wrong bound in the formula shows up here)."""
    for name in ("KX", "KX3", "KT"):
        k = dict(G1K, **{name: G1K[name] - 2})
        with pytest.raises(AssertionError):
            fixed_point(g1_madd, k)


def test_bias_digits_match_the_header():
    # bias_digits(p29, k, t): k p in digits, low digits raised by t 2^29, top lowered by t
    p29 = [(P >> (29 * i)) & N for i in range(8)] + [P >> 232]
    for k, t in ((8, 1), (5, 3), (65, 1), (13, 1)):
        s = [((k * P) >> (29 * i)) & N for i in range(8)] + [(k * P) >> 232]
        b = [s[0] + (t << 29)] + [s[i] + (t << 29) - t for i in range(1, 8)] + [s[8] - t]
        assert sum(d << (29 * i) for i, d in enumerate(b)) == k * P
        assert all(0 <= d < 2**32 for d in b)
        assert all(d >= t * 2**29 - t for d in b[1:8]) and b[0] >= t * 2**29
    assert p29[0] == 0x187cfd47 and p29[8] == 0x30644e


def _build(tmp: Path, sanitize: bool) -> Path:
    gxx = shutil.which("g++")
    if not gxx:
        pytest.skip("g++ not found")
    exe = tmp / ("fd_host_test_san" if sanitize else "fd_host_test")
    cmd = [gxx, "-O1" if sanitize else "-O2", "-std=c++17", "-Wno-unknown-pragmas", "-Wno-subobject-linkage",
           "-I", str(CSRC), str(ROOT / "tests" / "fd" / "fd_host_test.cpp"), "-o", str(exe)]
    if sanitize:
        cmd[1:1] = ["-fsanitize=address,undefined", "-fno-sanitize-recover=undefined", "-g"]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


@pytest.mark.parametrize("sanitize", [False, True])
def test_host_accumulation(tmp_path, sanitize):
    exe = _build(tmp_path, sanitize)
    r = subprocess.run([str(exe)], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert "checks OK" in r.stdout
