"""Pins the CPU oracle against textbook mathematics on small inputs (CPU only):
Python big-integer field arithmetic, the affine group law, brute-force MSMs, the naive
O(n^2) DFT and the QAP identity."""
import random

import numpy as np
import pytest

from oracle import oracle as O

R, Q = O.P_R, O.P_Q


@pytest.mark.parametrize("fid,p", [(0, R), (1, Q)])
def test_field_ops_bigint(fid, p):
    rnd = random.Random(fid)
    Rm = 1 << 256
    for _ in range(300):
        a, b = rnd.randrange(p), rnd.randrange(p)
        assert O.f_binop("or_f_add", fid, a, b) == (a + b) % p
        assert O.f_binop("or_f_sub", fid, a, b) == (a - b) % p
        assert O.f_binop("or_f_mul", fid, a, b) == a * b * pow(Rm, -1, p) % p
        assert O.f_unop("or_f_neg", fid, a) == (-a) % p
        assert O.f_unop("or_f_to_mont", fid, a) == a * Rm % p
        assert O.f_unop("or_f_from_mont", fid, a) == a * pow(Rm, -1, p) % p
        am = a * Rm % p
        assert O.f_unop("or_f_inv", fid, am) == (pow(a, -1, p) * Rm % p if a else 0)
    assert O.f_unop("or_f_inv", fid, 0) == 0


def test_fq2_closed_form():
    rnd = random.Random(2)
    for _ in range(100):
        a = (rnd.randrange(Q), rnd.randrange(Q))
        b = (rnd.randrange(Q), rnd.randrange(Q))
        enc = lambda x: np.ascontiguousarray(O.ints_to_array(list(x), Q).reshape(-1))
        dec = lambda arr: tuple(O.array_to_ints(arr.reshape(2, 4), Q))
        r = np.zeros(8, np.uint64)
        O.lib().or_fq2_mul(O._ptr(r), O._ptr(enc(a)), O._ptr(enc(b)))
        assert dec(r) == O.fq2_mul(a, b)
        O.lib().or_fq2_sqr(O._ptr(r), O._ptr(enc(a)))
        assert dec(r) == O.fq2_mul(a, a)
        O.lib().or_fq2_inv(O._ptr(r), O._ptr(enc(a)))
        assert dec(r) == O.fq2_inv(a)
        assert O.fq2_mul(dec(r), a) == (1, 0)


def test_generators_on_curve_and_order_r():
    assert O.g1_on_curve(O.G1_GEN) and O.g2_on_curve(O.G2_GEN)
    assert O.g1_mul_aff(O.G1_GEN, R) is None
    assert O.g2_mul_aff(O.G2_GEN, R) is None
    assert O.g1_mul_aff(O.G1_GEN, R + 1) == O.G1_GEN


def _jac_affine(group, jac):
    r = np.zeros(8 if group == 1 else 16, np.uint64)
    getattr(O.lib(), f"or_g{group}_to_affine")(O._ptr(r), O._ptr(np.ascontiguousarray(jac)))
    return (O.g1_unpack if group == 1 else O.g2_unpack)(r)[0]


@pytest.mark.parametrize("group", [1, 2])
def test_curve_ops_special_cases(group):
    rnd = random.Random(3 + group)
    mul = O.g1_mul_aff if group == 1 else O.g2_mul_aff
    add = O.g1_add_aff if group == 1 else O.g2_add_aff
    pack = O.g1_pack if group == 1 else O.g2_pack
    gen = O.G1_GEN if group == 1 else O.G2_GEN
    P = mul(gen, rnd.randrange(1, R))
    Qp = mul(gen, rnd.randrange(1, R))
    negP = (P[0], (Q - P[1]) % Q) if group == 1 else (P[0], ((Q - P[1][0]) % Q, (Q - P[1][1]) % Q))
    w = 12 if group == 1 else 24
    for a, b in ((P, Qp), (P, P), (P, negP), (None, P), (P, None), (None, None)):
        ja = np.zeros(w, np.uint64)
        getattr(O.lib(), f"or_g{group}_from_affine")(O._ptr(ja), O._ptr(np.ascontiguousarray(pack([a])[0])))
        jb = np.zeros(w, np.uint64)
        getattr(O.lib(), f"or_g{group}_from_affine")(O._ptr(jb), O._ptr(np.ascontiguousarray(pack([b])[0])))
        out = np.zeros(w, np.uint64)
        getattr(O.lib(), f"or_g{group}_add")(O._ptr(out), O._ptr(ja), O._ptr(jb))
        assert _jac_affine(group, out) == add(a, b)
        getattr(O.lib(), f"or_g{group}_add_mixed")(O._ptr(out), O._ptr(ja), O._ptr(np.ascontiguousarray(pack([b])[0])))
        assert _jac_affine(group, out) == add(a, b)
    ja = np.zeros(w, np.uint64)
    getattr(O.lib(), f"or_g{group}_from_affine")(O._ptr(ja), O._ptr(np.ascontiguousarray(pack([P])[0])))
    out = np.zeros(w, np.uint64)
    getattr(O.lib(), f"or_g{group}_double")(O._ptr(out), O._ptr(ja))
    assert _jac_affine(group, out) == add(P, P)


@pytest.mark.parametrize("group", [1, 2])
@pytest.mark.parametrize("n", [1, 5, 17])
def test_msm_brute_force(group, n):
    rnd = random.Random(10 * n + group)
    mul = O.g1_mul_aff if group == 1 else O.g2_mul_aff
    add = O.g1_add_aff if group == 1 else O.g2_add_aff
    gen = O.G1_GEN if group == 1 else O.G2_GEN
    pts = [mul(gen, rnd.randrange(1, R)) for _ in range(n)]
    ks = [rnd.randrange(R) for _ in range(n)]
    ks[0] = 0
    want = None
    for p_, k in zip(pts, ks):
        want = add(want, mul(p_, k))
    pk = (O.g1_pack if group == 1 else O.g2_pack)(pts)
    sc = O.ints_to_array(ks)
    for c in (0, 3, 8):
        got = (O.msm_g1 if group == 1 else O.msm_g2)(pk, sc, c)
        assert (O.g1_unpack if group == 1 else O.g2_unpack)(got)[0] == want


def test_msm_closed_form_cpu():
    rng = np.random.default_rng(4)
    n = 3000
    ks = O.rand_scalars(rng, n)
    pts = O.g1_gen_mul_batch(ks)
    sc = O.rand_scalars(rng, n)
    got = O.g1_unpack(O.msm_g1(pts, sc))[0]
    assert got == O.g1_mul_aff(O.G1_GEN, O.fr_dot(sc, ks))


@pytest.mark.parametrize("lg", [1, 3, 6])
def test_ntt_naive_dft(lg):
    rnd = random.Random(lg)
    n = 1 << lg
    a = [rnd.randrange(R) for _ in range(n)]
    arr = O.ints_to_array(a, R)
    w = O.omega(lg)
    assert O.array_to_ints(O.ntt(arr, lg, 0), R) == O.naive_dft(a, w)
    wi = O.omega(lg, inverse=True)
    ninv = pow(n, -1, R)
    assert O.array_to_ints(O.ntt(arr, lg, 1), R) == [x * ninv % R for x in O.naive_dft(a, wi)]
    g = 5
    assert O.array_to_ints(O.ntt(arr, lg, 2), R) == O.naive_dft([x * pow(g, j, R) % R for j, x in enumerate(a)], w)
    gi = pow(g, -1, R)
    inv = [x * ninv % R for x in O.naive_dft(a, wi)]
    assert O.array_to_ints(O.ntt(arr, lg, 3), R) == [x * pow(gi, j, R) % R for j, x in enumerate(inv)]


@pytest.mark.parametrize("lg", [8, 12])
def test_ntt_round_trips(lg):
    rng = np.random.default_rng(lg)
    arr = O.to_mont_array(rng, 1 << lg) if hasattr(O, "to_mont_array") else O.rand_scalars(rng, 1 << lg)
    assert np.array_equal(O.ntt(O.ntt(arr, lg, 0), lg, 1), arr)
    assert np.array_equal(O.ntt(O.ntt(arr, lg, 2), lg, 3), arr)


@pytest.mark.parametrize("lg", [3, 7, 11])
def test_qap_identity(lg):
    """A B - C = h (x^d - 1) at a random x and h[d-1] = 0 (restating ref_qap.py:110-114)
    on a satisfied random R1CS-shaped instance: c_i = a_i b_i on the first d-2 points."""
    rnd = random.Random(lg)
    d = 1 << lg
    n = d - 2
    a = [rnd.randrange(R) for _ in range(n)] + [rnd.randrange(R), rnd.randrange(R)]
    b = [rnd.randrange(R) for _ in range(n)] + [0, 0]
    c = [x * y % R for x, y in zip(a[:n], b[:n])] + [0, 0]
    h = O.array_to_ints(O.qap_h(O.ints_to_array(a, R), O.ints_to_array(b, R), O.ints_to_array(c, R), lg), R)
    assert h[d - 1] == 0
    x = rnd.randrange(R)
    ev = lambda vals: O.unmont(O.poly_eval(O.ntt(O.ints_to_array(vals, R), lg, 1), O.mont(x, R)), R)
    hx = O.unmont(O.poly_eval(O.ints_to_array(h, R), O.mont(x, R)), R)
    assert (ev(a) * ev(b) - ev(c)) % R == hx * (pow(x, d, R) - 1) % R
