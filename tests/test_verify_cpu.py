"""Groth16 verification (bg_verify) and the BN254 pairing check (bg_pairing_check), host
code, CPU only.  Mirrors ref/src/verify.rs:29-51: the verifier accepts an honest proof,
rejects a wrong public input ("0x2") and C = infinity, and errors (rather than rejecting)
on an off-curve A.

The pairing is pinned by what fixes it mathematically: bilinearity on random scalars
(e(aP, bQ) e(-abP, Q) = 1, e(aP, Q) e(bP, Q) e(-(a+b)P, Q) = 1), non-degeneracy
(e(G1, G2) != 1) and the subgroup check on a twist point outside G2.  A wrong Miller loop
or final exponentiation leaves a map that is not bilinear, or degenerate; the Groth16
checks then use proofs pinned to the trapdoor closed form (test_setup_cpu.py)."""
import random

import numpy as np
import pytest

import bendg
from oracle import oracle as O
from instance import closed_form, load

R, Q = O.P_R, O.P_Q


def _sc(xs):
    """canonical packed scalars (n, 8) u32"""
    a = np.zeros((len(xs), 8), dtype=np.uint32)
    for i, x in enumerate(xs):
        x %= R
        for j in range(8):
            a[i, j] = (x >> (32 * j)) & 0xFFFFFFFF
    return a


def g1(*ks):
    return bendg.fixed_base_mul(1, _sc(ks))


def g2(*ks):
    return bendg.fixed_base_mul(2, _sc(ks))


def test_pairing_bilinear_and_nondegenerate():
    rng = random.Random(11)
    for _ in range(3):
        a, b = rng.randrange(1, R), rng.randrange(1, R)
        assert bendg.pairing_check(g1(a, -a * b), g2(b, 1))
        assert bendg.pairing_check(g1(a, b, -(a + b)), g2(1, 1, 1))
        assert bendg.pairing_check(g1(a, -1), g2(b, a * b))
        assert not bendg.pairing_check(g1(a, -1), g2(b, a * b + 1))
        assert not bendg.pairing_check(g1(a, b), g2(1, 1))
    assert not bendg.pairing_check(g1(1), g2(1))  # e(G1, G2) != 1
    assert bendg.pairing_check(g1(0), g2(5)) and bendg.pairing_check(g1(3), g2(0))  # infinity pairs to 1
    assert bendg.pairing_check(np.zeros((0, 16), np.uint32), np.zeros((0, 32), np.uint32))  # empty product


def _fq2_sqrt(a):
    """square root in Fq2 = Fq[u]/(u^2 + 1), q = 3 mod 4 (Adj, Rodriguez-Henriquez 2012,
    Algorithm 9), None if a is not a square"""
    e = (Q - 3) // 4

    def pw(x, k):
        r = (1, 0)
        while k:
            if k & 1:
                r = O.fq2_mul(r, x)
            x = O.fq2_mul(x, x)
            k >>= 1
        return r

    a1 = pw(a, e)
    alpha = O.fq2_mul(O.fq2_mul(a1, a1), a)
    x0 = O.fq2_mul(a1, a)
    if alpha == (Q - 1, 0):
        x = O.fq2_mul((0, 1), x0)
    else:
        x = O.fq2_mul(pw(O.fq2_add((1, 0), alpha), (Q - 1) // 2), x0)
    return x if O.fq2_mul(x, x) == a else None


def test_pairing_point_checks():
    bad_g1 = np.zeros((1, 16), np.uint32)
    bad_g1[0, :8] = np.array(O.to_words(O.mont(1, Q)), np.uint64).view(np.uint32)
    bad_g1[0, 8:] = bad_g1[0, :8]  # (1, 1) is not on y^2 = x^3 + 3
    with pytest.raises(bendg.BgError) as ei:
        bendg.pairing_check(bad_g1, g2(1))
    assert ei.value.code == -1
    # a twist point outside the order-r subgroup (the twist's cofactor is 2q - r)
    x = (5, 1)
    while True:
        rhs = O.fq2_add(O.fq2_mul(O.fq2_mul(x, x), x), O.B_G2)
        y = _fq2_sqrt(rhs)
        if y is not None:
            break
        x = (x[0] + 1, x[1])
    assert O.g2_on_curve((x, y))
    pt = O.g2_pack([(x, y)])
    with pytest.raises(bendg.BgError) as ei:
        bendg.pairing_check(g1(1), pt.view(np.uint32))
    assert ei.value.code == -1
    assert "G2" in str(ei.value)
    # the same check accepts a genuine G2 point given the same way
    assert not bendg.pairing_check(g1(1), O.g2_pack([O.g2_mul_aff(O.G2_GEN, 7)]).view(np.uint32))


@pytest.fixture(scope="module", params=[4, 6])
def instance(request, tmp_path_factory):
    d = tmp_path_factory.mktemp(f"v{request.param}")
    bendg.setup_square_chain(d, request.param, seed=300 + request.param)
    inst = load(d)
    A, B, C, _ = O.prove(inst)
    proof = bendg.unpack16(np.concatenate([A, B, C]).view(np.uint32).reshape(-1, 8)).reshape(-1)
    vk = np.fromfile(d / "vk.bin", dtype=np.uint32)
    inputs = np.fromfile(d / "witness.bin", dtype=np.uint32)[16:32]  # z_1 = x_N, the public input
    return inst, vk, proof, inputs


def test_vk_matches_trapdoor(instance):
    """alpha G1, beta/gamma/delta G2, and IC(z) = sum_i z_i gamma_abc_i = ic / gamma G1."""
    inst, vk, _, _ = instance
    td = inst["trapdoor"]
    pk = bendg.pack16(vk[:224].reshape(-1, 16)).reshape(-1).view(np.uint64)
    assert O.g1_unpack(pk[:8].reshape(1, 8))[0] == O.g1_mul_aff(O.G1_GEN, td["alpha"])
    for i, k in enumerate(("beta", "gamma", "delta")):
        got = O.g2_unpack(pk[8 + 16 * i:24 + 16 * i].reshape(1, 16))[0]
        assert got == O.g2_mul_aff(O.G2_GEN, td[k]), k
    assert int(vk[224]) == 2
    ic = O.g1_unpack(bendg.pack16(vk[225:].reshape(-1, 16)).view(np.uint64).reshape(-1, 8))
    z1 = O.unmont(O.from_words(inst["z"][1]), R)
    icz = O.g1_add_aff(ic[0], O.g1_mul_aff(ic[1], z1))
    cf = closed_form(inst)
    assert icz == O.g1_mul_aff(O.G1_GEN, cf["ic"] * pow(td["gamma"], -1, R) % R)


def test_verify_accepts_and_rejects(instance):
    inst, vk, proof, inputs = instance
    assert bendg.verify(vk, proof, inputs)
    two = np.array(O.words16(O.mont(2, R)), dtype=np.uint32)  # the reference's "0x2"
    assert not bendg.verify(vk, proof, two)
    bad_c = proof.copy()
    bad_c[96:] = 0  # C = infinity: a valid point, a wrong proof
    assert not bendg.verify(vk, bad_c, inputs)
    swapped = proof.copy()
    swapped[:32], swapped[96:] = proof[96:], proof[:32]
    assert not bendg.verify(vk, swapped, inputs)
    # packed encoding of the same objects
    pk16 = lambda a: bendg.pack16(a.reshape(-1, 16)).reshape(-1)
    vk_p = np.concatenate([pk16(vk[:224]), vk[224:225], pk16(vk[225:])])
    assert bendg.verify(vk_p, pk16(proof), pk16(inputs), fmt="packed")


def test_verify_errors_on_bad_points(instance):
    _, vk, proof, inputs = instance
    bad_a = proof.copy()
    one16 = np.array(O.words16(O.mont(1, Q)), dtype=np.uint32)
    bad_a[:16], bad_a[16:32] = one16, one16  # A = (1, 1): off the curve
    with pytest.raises(bendg.BgError) as ei:
        bendg.verify(vk, bad_a, inputs)
    assert ei.value.code == -1
    big = proof.copy()
    big[:16] = np.array(O.words16(Q), dtype=np.uint32)  # a coordinate not below q
    with pytest.raises(bendg.BgError):
        bendg.verify(vk, big, inputs)
    with pytest.raises(bendg.BgError):  # two inputs for a one-input key
        bendg.verify(vk, proof, np.concatenate([inputs, inputs]))
    with pytest.raises(bendg.BgError):  # truncated key
        bendg.verify(vk[:-1], proof, inputs)
