"""The benchmark configurations at their full sizes on the device, each checked bit for bit:

- primitives: G1 MSM at 2^20 and 2^22 points and G2 MSM at 2^20 (bases k_i G, scalars
  uniform in [0, r)) against the closed form (sum a_i k_i mod r) G; the Fr NTT at 2^20 and
  2^24 against the oracle's radix-2 transform (itself pinned to the naive DFT) plus the
  inverse round trip;
- batched: uneven MSMs up to 2^20 points in one bg_msm_batch call, each against its own
  closed form; an engine mix of SquareChain proofs with K from 13 to 18 where every proof
  equals the proof made alone and the K=17 / K=18 proofs equal the trapdoor closed form;
- serving: a K=20 proof against the closed form, and a served mix over K in
  {12, 14, 16, 18, 20} (arrivals spread over time, concurrent slots) equal to the proofs
  made alone.

The reference's sized checks are bend2/tests/test_msm.bend:60-67 and
bend2/tests/test_ntt.bend:171-184; each proof is the unit of bend2/prove.bend:129-162."""
import time

import numpy as np
import pytest

import bendg
from oracle import oracle as O
from instance import closed_form, load, points_of

pytestmark = pytest.mark.gpu

R = O.P_R


def _gen_point(group, k: int) -> np.ndarray:
    kk = np.array(O.to_words(k % R), dtype=np.uint64)
    if group == 1:
        out = np.zeros(8, dtype=np.uint64)
        O.lib().or_g1_mul_aff(O._ptr(out), O._ptr(O.g1_pack([O.G1_GEN])[0]), O._ptr(kk))
    else:
        out = np.zeros(16, dtype=np.uint64)
        O.lib().or_g2_mul_aff(O._ptr(out), O._ptr(O.g2_pack([O.G2_GEN])[0]), O._ptr(kk))
    return out


@pytest.fixture(scope="module")
def g1_pool(bg):
    """2^22 G1 bases k_i G (host fixed-base multiplication, pinned to the oracle's affine
    law in test_setup_cpu.py) with their k_i."""
    t = time.perf_counter()
    pts, ks = bendg.gen_bases(1, 2022, 1 << 22, with_k=True)
    print(f"G1 pool of 2^22 bases in {time.perf_counter() - t:.1f} s", flush=True)
    return pts, ks.view(np.uint64).reshape(-1, 4)


@pytest.mark.parametrize("log2n", [20, 22])
def test_msm_g1_primitive_closed_form(g1_pool, log2n):
    pts, ks = g1_pool
    n = 1 << log2n
    rng = np.random.default_rng(300 + log2n)
    sc = O.rand_scalars(rng, n)
    print(f"G1 MSM 2^{log2n}, c = {bendg.msm_plan_info(n, 1)['c']}", flush=True)
    got = bendg.msm_g1(pts[:n], sc.view(np.uint32)).view(np.uint64)
    assert np.array_equal(got, _gen_point(1, O.fr_dot(sc, ks[:n])))


def test_msm_g2_primitive_closed_form(bg):
    n = 1 << 20
    t = time.perf_counter()
    pts, ks = bendg.gen_bases(2, 2020, n, with_k=True)
    print(f"G2 2^20 bases in {time.perf_counter() - t:.1f} s; MSM c = {bendg.msm_plan_info(n, 2)['c']}",
          flush=True)
    rng = np.random.default_rng(320)
    sc = O.rand_scalars(rng, n)
    got = bendg.msm_g2(pts, sc.view(np.uint32)).view(np.uint64)
    assert np.array_equal(got, _gen_point(2, O.fr_dot(sc, ks.view(np.uint64).reshape(-1, 4))))


@pytest.mark.parametrize("log2n", [20, 24])
def test_ntt_primitive_vs_oracle(bg, log2n):
    n = 1 << log2n
    rng = np.random.default_rng(400 + log2n)
    x = O.rand_scalars(rng, n)
    print(f"NTT 2^{log2n} forward", flush=True)
    y = bendg.ntt(x.view(np.uint32), log2n, bendg.NTT_FFT).view(np.uint64).reshape(-1, 4)
    assert np.array_equal(y, O.ntt(x, log2n, 0))
    print(f"NTT 2^{log2n} inverse", flush=True)
    back = bendg.ntt(y.view(np.uint32), log2n, bendg.NTT_IFFT).view(np.uint64).reshape(-1, 4)
    assert np.array_equal(back, x)
    # a few outputs straight from the definition y_k = x(w^k)
    w = pow(O.ROOT28, 1 << (28 - log2n), R)
    for k in (0, 1, n // 2, n - 1):
        assert O.from_words(y[k]) == O.poly_eval(x, O.mont(pow(w, k, R), R))


def test_msm_batch_uneven_to_2_20(g1_pool):
    """Uneven G1 MSMs from 1 point to 2^20 in one batched call (several window-width
    classes), each equal to its own closed form."""
    pts, ks = g1_pool
    sizes = [1, 4096, 3000, 16391, 50000, 1 << 16, (1 << 17) + 1, 1 << 18, 300000, 1 << 19, 1 << 20]
    rng = np.random.default_rng(500)
    offs = rng.integers(0, (1 << 22) - (1 << 20), size=len(sizes))
    scs = [O.rand_scalars(rng, n) for n in sizes]
    print(f"msm_batch sizes {sizes}", flush=True)
    outs = bendg.msm_batch(1, [pts[o:o + n] for o, n in zip(offs, sizes)], [s.view(np.uint32) for s in scs])
    for i, (o, n, s) in enumerate(zip(offs, sizes, scs)):
        want = _gen_point(1, O.fr_dot(s, ks[o:o + n]))
        assert np.array_equal(outs[i].view(np.uint64), want), f"MSM {i} of {n} points"


# ---------------------------------------------------------------- proofs at the served sizes

PROOF_SIZES = (12, 13, 14, 15, 16, 17, 18, 20)


def _files(d):
    z = np.fromfile(d / "witness.bin", dtype=np.uint32)
    rs = np.fromfile(d / "rs.bin", dtype=np.uint32)
    return z, rs


@pytest.fixture(scope="module")
def big_keys(bg, tmp_path_factory):
    out = {}
    for lg in PROOF_SIZES:
        d = tmp_path_factory.mktemp(f"c{lg}")
        t = time.perf_counter()
        bendg.setup_square_chain(d, lg, seed=600 + lg)
        print(f"setup K={lg} in {time.perf_counter() - t:.1f} s", flush=True)
        out[lg] = d
    return out


@pytest.fixture(scope="module")
def big_loaded(big_keys):
    pks = {lg: bendg.ProvingKey(d) for lg, d in big_keys.items()}
    solo = {}
    for lg, pk in pks.items():
        z, rs = _files(big_keys[lg])
        print(f"solo proof K={lg}", flush=True)
        solo[lg], _ = pk.prove(z, rs, fmt="words16")
    yield pks, solo
    for pk in pks.values():
        pk.close()


@pytest.mark.parametrize("lg", [17, 18, 20])
def test_large_proof_closed_form(big_keys, big_loaded, lg):
    _, solo = big_loaded
    inst = load(big_keys[lg])
    eA, eB, eC = points_of(closed_form(inst))
    p = bendg.pack16(solo[lg].reshape(-1, 16)).reshape(-1).view(np.uint64)
    assert np.array_equal(p[:8], eA) and np.array_equal(p[8:24], eB) and np.array_equal(p[24:32], eC)
    vk = np.fromfile(big_keys[lg] / "vk.bin", dtype=np.uint32)
    z, _ = _files(big_keys[lg])
    assert bendg.verify(vk, solo[lg], z[16:32])


@pytest.mark.parametrize("slots", [2, 4])
def test_engine_mix_k13_to_k18(big_keys, big_loaded, slots):
    """A burst of 24 proofs with K drawn from 13..18 (every size at least twice) on
    concurrent slots: every proof equals the proof made alone."""
    pks, solo = big_loaded
    rng = np.random.default_rng(700 + slots)
    order = list(range(13, 19)) * 2 + [int(k) for k in rng.integers(13, 19, size=12)]
    rng.shuffle(order)
    files = {lg: _files(big_keys[lg]) for lg in set(order)}
    eng = bendg.Engine(slots=slots)
    print(f"engine slots={slots} order={order}", flush=True)
    tickets = [(lg, eng.submit(pks[lg], *files[lg])) for lg in order]
    for lg, t in tickets:
        proof, _ = eng.wait(t)
        assert np.array_equal(proof, solo[lg]), f"K={lg} ticket {t}"
    eng.close()


def test_served_mix_k12_to_k20(big_keys, big_loaded):
    """Requests over K in {12, 14, 16, 18, 20} (weights 0.3 / 0.3 / 0.2 / 0.15 / 0.05,
    every size present) arriving spaced in time while earlier ones are on the device:
    each proof equals the proof made alone."""
    pks, solo = big_loaded
    rng = np.random.default_rng(800)
    ks = [12, 14, 16, 18, 20]
    order = ks + [int(k) for k in rng.choice(ks, size=35, p=[0.3, 0.3, 0.2, 0.15, 0.05])]
    rng.shuffle(order)
    gaps = rng.exponential(1.0 / 300.0, size=len(order))
    files = {lg: _files(big_keys[lg]) for lg in ks}
    eng = bendg.Engine(slots=4)
    print(f"served mix order={order}", flush=True)
    tickets = []
    for lg, g in zip(order, gaps):
        time.sleep(float(g))
        tickets.append((lg, eng.submit(pks[lg], *files[lg])))
    for lg, t in tickets:
        proof, tm = eng.wait(t)
        assert np.array_equal(proof, solo[lg]), f"K={lg} ticket {t}"
        assert tm["latency"] >= tm["proving"] >= 0
    eng.close()
