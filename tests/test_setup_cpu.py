"""Key generator and CPU oracle prover (CPU only): the trapdoor setup writes valid
FORMAT.md files; the oracle's full prove equals the closed-form Groth16 proof and the
closed form satisfies the verification equation in the exponent."""
import numpy as np
import pytest

import bendg
from oracle import oracle as O
from instance import closed_form, load, points_of

R = O.P_R


@pytest.fixture(scope="module")
def key4(tmp_path_factory):
    d = tmp_path_factory.mktemp("k4")
    bendg.setup_square_chain(d, 4, seed=42)
    return d


def test_files_and_header(key4):
    hdr = np.fromfile(key4 / "header.bin", dtype=np.uint32)
    n = 16
    assert list(hdr) == [n, 2, n, 5, n + 2, n + 2, 31, n]
    sizes = {"witness.bin": (n + 2) * 64, "pk_g1.bin": 3 * 128, "pk_g2.bin": 2 * 256, "pk_a.bin": (n + 2) * 128,
             "pk_b1.bin": (n + 2) * 128, "pk_b2.bin": (n + 2) * 256, "pk_h.bin": 31 * 128, "pk_l.bin": n * 128,
             "rs.bin": 128, "r1cs_a.bin": 4 * (2 + n + 17 * n)}
    for f, sz in sizes.items():
        assert (key4 / f).stat().st_size == sz, f
    for f in key4.glob("*.bin"):
        if f.name not in ("header.bin",) and not f.name.startswith("r1cs"):
            assert (np.fromfile(f, dtype=np.uint32) < (1 << 16)).all(), f.name


def test_witness_satisfies_square_chain(key4):
    inst = load(key4)
    z = [O.unmont(O.from_words(v), R) for v in inst["z"]]
    n = 16
    assert z[0] == 1
    xs = z[2:] + [z[1]]
    assert all(xs[i + 1] == xs[i] * xs[i] % R for i in range(n))


def test_b_query_infinity_for_instance_columns(key4):
    inst = load(key4)
    assert not inst["pk_b1"][:2].any() and not inst["pk_b2"][:2].any()


def test_oracle_prove_equals_closed_form(key4):
    inst = load(key4)
    A, B, C, _ = O.prove(inst)
    cf = closed_form(inst)
    eA, eB, eC = points_of(cf)
    assert np.array_equal(A, eA) and np.array_equal(B, eB) and np.array_equal(C, eC)


def test_verification_equation(key4):
    cf = closed_form(load(key4))
    assert cf["a"] * cf["b"] % R == (cf["alpha"] * cf["beta"] + cf["ic"] + cf["c"] * cf["delta"]) % R
    # the QAP holds at tau: A(t) B(t) - C(t) is divisible by t^d - 1 exactly when h exists;
    # closed_form uses A B - C = h t, so a wrong evaluation vector breaks the equation above
    assert cf["At"] != 0


def test_pk_points_match_trapdoor(key4):
    """a_query[i] = u_i(tau) G1 for the instance columns (input-consistency rows) and
    h_query[k] = tau^k (tau^d - 1) / delta G1."""
    inst = load(key4)
    from instance import lagrange_at
    td = inst["trapdoor"]
    t, de = td["tau"], td["delta"]
    lg, nc = 5, 16
    u0 = lagrange_at(nc, lg, t)
    assert O.g1_unpack(inst["pk_a"][:1])[0] == O.g1_mul_aff(O.G1_GEN, u0)
    zt = (pow(t, 32, R) - 1) % R
    for k in (0, 1, 30):
        hk = pow(t, k, R) * zt * pow(de, -1, R) % R
        assert O.g1_unpack(inst["pk_h"][k:k + 1])[0] == O.g1_mul_aff(O.G1_GEN, hk)


@pytest.mark.parametrize("lg", [8, 10])
def test_oracle_prove_closed_form_larger(tmp_path, lg):
    bendg.setup_square_chain(tmp_path, lg, seed=7)
    inst = load(tmp_path)
    A, B, C, _ = O.prove(inst)
    eA, eB, eC = points_of(closed_form(inst))
    assert np.array_equal(A, eA) and np.array_equal(B, eB) and np.array_equal(C, eC)


def test_setup_deterministic(tmp_path):
    a, b = tmp_path / "a", tmp_path / "b"
    bendg.setup_square_chain(a, 3, seed=5)
    bendg.setup_square_chain(b, 3, seed=5)
    for f in a.glob("*.bin"):
        assert f.read_bytes() == (b / f.name).read_bytes(), f.name
    c = tmp_path / "c"
    bendg.setup_square_chain(c, 3, seed=6)
    assert (a / "pk_a.bin").read_bytes() != (c / "pk_a.bin").read_bytes()


@pytest.mark.parametrize("group", [1, 2])
def test_fixed_base_and_synthetic_bases(group):
    """bg_fixed_base_mul equals the oracle's k*G (affine group law in Python pins the
    oracle); bg_gen_bases returns k_i G for the k_i it reports, the k_i below r."""
    rng = np.random.default_rng(31 + group)
    ks = O.rand_scalars(rng, 40)
    ks[0] = 0
    ks[1] = O.to_words(1)
    ks[2] = O.to_words(R - 1)
    got = bendg.fixed_base_mul(group, ks.view(np.uint32)).view(np.uint64)
    gen = O.G1_GEN if group == 1 else O.G2_GEN
    for i in (0, 1, 2, 7, 39):
        k = O.from_words(ks[i])
        want = O.g1_pack([O.g1_mul_aff(gen, k)])[0] if group == 1 else O.g2_pack([O.g2_mul_aff(gen, k)])[0]
        assert np.array_equal(got[i], want), i
    pts, k = bendg.gen_bases(group, 5, 64, with_k=True)
    assert all(O.from_words(v) < R for v in k.view(np.uint64).reshape(-1, 4))
    assert np.array_equal(pts, bendg.fixed_base_mul(group, k))
    assert np.array_equal(pts, bendg.gen_bases(group, 5, 64))
    assert not np.array_equal(pts, bendg.gen_bases(group, 6, 64))


def test_c_eval_rows_equal_python_rows(key4, tmp_path):
    """The closed form's evaluation vectors come from the oracle's C rows; they equal the
    big-integer rows (CSR x z, plus the instance rows of A) at K=4 and K=7."""
    from instance import eval_vectors, eval_vectors_mont
    from oracle import oracle as O
    d7 = tmp_path / "k7"
    bendg.setup_square_chain(d7, 7, seed=5)
    for d in (key4, d7):
        inst = load(d)
        py = eval_vectors(inst)
        cm = eval_vectors_mont(inst)
        for a, b in zip(py, cm):
            assert a == O.array_to_ints(b, O.P_R)
