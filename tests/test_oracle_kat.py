"""Pins the CPU oracle against the known-answer values committed in the reference
(tests/golden/reference_kats.json, each with its file:line) -- CPU only."""
import json
from pathlib import Path

import pytest

from oracle import oracle as O

KATS = json.loads((Path(__file__).parent / "golden" / "reference_kats.json").read_text())


@pytest.mark.parametrize("n", ["18", "20", "22"])
def test_field_mul_xor_64bit(n):
    assert O.field_mul_xor(int(n)) == KATS["field_mul_xor"][n]["xor"]


@pytest.mark.parametrize("n", ["18", "20"])
def test_field_mul_xor_16bit_restatement(n):
    # the 16 x 16-bit CIOS restatement of gen_field.py must give the same checksum
    assert O.field_mul_xor(int(n), use16=True) == KATS["field_mul_xor"][n]["xor"]


def test_fe16_matches_f_mul_and_bigint():
    import random
    rnd = random.Random(7)
    for fid, p in ((0, O.P_R), (1, O.P_Q)):
        for _ in range(200):
            a, b = rnd.randrange(p), rnd.randrange(p)
            r16 = O.fe16_mul(fid, O.words16(a), O.words16(b))
            r64 = O.f_binop("or_f_mul", fid, a, b)
            expect = a * b * pow(1 << 256, -1, p) % p
            assert sum(v << (16 * j) for j, v in enumerate(r16)) == expect == r64


def test_kat_operand_generator_below_modulus():
    for s in (0, 1, 12345, 2**31 - 7):
        limbs = O.kat_rnd(s)
        assert sum(v << (16 * j) for j, v in enumerate(limbs)) < O.P_R


@pytest.mark.parametrize("case", KATS["g1_add_sum"]["cases"], ids=lambda c: f"N{c['N']}FD{c['FD']}")
def test_g1_add_chain(case):
    # every leaf runs the same chain, so sum = 2^FD * lo16(X after 2^(N-FD) adds) mod 2^32
    k = 1 << (case["N"] - case["FD"])
    if k > (1 << 14):
        pytest.skip("chain longer than the CPU budget")
    x16 = O.g1_add_chain_x16(k)
    assert (x16 << case["FD"]) % (1 << 32) == case["sum"]


def _limbs_value(limbs):
    return sum(v << (16 * j) for j, v in enumerate(limbs))


def test_fr_constants():
    c = KATS["fr_constants_16bit_limbs"]
    p = O.P_R
    assert _limbs_value(c["modulus"]["limbs"]) == p
    assert _limbs_value(c["one"]["limbs"]) == O.mont(1, p)
    assert _limbs_value(c["two"]["limbs"]) == O.mont(2, p)
    assert _limbs_value(c["neg_one"]["limbs"]) == O.mont(p - 1, p)
    assert _limbs_value(c["r2"]["limbs"]) == (1 << 512) % p
    # the oracle's own constants agree
    import numpy as np
    P = np.zeros(4, dtype=np.uint64); one = np.zeros(4, dtype=np.uint64)
    r2 = np.zeros(4, dtype=np.uint64); n0 = np.zeros(1, dtype=np.uint64)
    O.lib().or_f_consts(0, O._ptr(P), O._ptr(one), O._ptr(r2), O._ptr(n0))
    assert O.from_words(one) == _limbs_value(c["one"]["limbs"])
    assert O.from_words(r2) == _limbs_value(c["r2"]["limbs"])
    assert (int(n0[0]) * p) % (1 << 64) == (1 << 64) - 1


def test_ntt_constants():
    c = KATS["ntt_constants_16bit_limbs"]
    p = O.P_R
    import numpy as np
    root = np.zeros(4, dtype=np.uint64)
    O.lib().or_root28(O._ptr(root))
    assert O.from_words(root) == _limbs_value(c["root"]["limbs"])
    assert _limbs_value(c["root_inv"]["limbs"]) == O.mont(pow(O.ROOT28, -1, p), p)
    assert _limbs_value(c["gen"]["limbs"]) == O.mont(5, p)
    assert _limbs_value(c["gen_inv"]["limbs"]) == O.mont(pow(5, -1, p), p)
    assert _limbs_value(c["two_inv"]["limbs"]) == O.mont(pow(2, -1, p), p)
    # omega(k) has exact order 2^k
    for k in (1, 5, 11, 19):
        w = np.zeros(4, dtype=np.uint64)
        O.lib().or_omega(O._ptr(w), k, 0)
        wv = O.unmont(O.from_words(w), p)
        assert pow(wv, 1 << k, p) == 1 and pow(wv, 1 << (k - 1), p) == p - 1
