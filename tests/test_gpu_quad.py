"""The quad-cooperative bucket reductions and merge levels (four lanes per point addition)
against the one-lane kernels they replace (BG16_QUAD=0): MSMs and a proof are bit-equal
both ways and equal the oracle or the closed form.  The MSM unit is Msm.g1 / Msm.g2
(bend2/src/msm.bend:293-298, :525-530), the proof one Pr.run (bend2/prove.bend:129-162).

The planner takes the quad path when the level-1 group count is at most 2^15, which holds
for every size used here (asserted through bg_msm_plan_info)."""
import numpy as np
import pytest

import bendg
from oracle import oracle as O
from instance import closed_form, load, points_of

pytestmark = pytest.mark.gpu

R = O.P_R


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _msm(group, pts, sc):
    f = bendg.msm_g1 if group == 1 else bendg.msm_g2
    return f(_u32(pts), _u32(sc)).view(np.uint64)


def _mul_gen(group, k: int):
    kk = np.array(O.to_words(k % R), dtype=np.uint64)
    if group == 1:
        out = np.zeros(8, dtype=np.uint64)
        O.lib().or_g1_mul_aff(O._ptr(out), O._ptr(O.g1_pack([O.G1_GEN])[0]), O._ptr(kk))
    else:
        out = np.zeros(16, dtype=np.uint64)
        O.lib().or_g2_mul_aff(O._ptr(out), O._ptr(O.g2_pack([O.G2_GEN])[0]), O._ptr(kk))
    return out


def _scalars(rng, n, kind):
    if kind == "random":
        return O.rand_scalars(rng, n)
    if kind == "ones":  # witness-like 0/1 values: few buckets, many merge slices
        sc = np.zeros((n, 4), dtype=np.uint64)
        sc[:, 0] = 1
        sc[::5, 0] = 0
        return sc
    vals = O.rand_scalars(rng, 3)  # three distinct values
    return vals[rng.integers(0, 3, size=n)]


@pytest.mark.parametrize("group,n,kind", [
    (1, 1000, "random"), (1, 1 << 14, "random"), (1, 1 << 16, "ones"), (1, 1 << 15, "three"),
    (2, 700, "random"), (2, 1 << 12, "random"), (2, 1 << 13, "ones")])
def test_quad_equals_one_lane(bg, monkeypatch, group, n, kind):
    assert bendg.msm_plan_info(n, group)["groups1"] <= 1 << 15
    rng = np.random.default_rng(1300 + group * 7 + n % 97)
    ks = O.rand_scalars(rng, n)
    pts = O.g1_gen_mul_batch(ks) if group == 1 else O.g2_gen_mul_batch(ks)
    sc = _scalars(rng, n, kind)
    print(f"quad vs one-lane G{group} n={n} {kind}", flush=True)
    monkeypatch.setenv("BG16_QUAD", "0")
    one_lane = _msm(group, pts, sc)
    monkeypatch.delenv("BG16_QUAD")
    quad = _msm(group, pts, sc)
    assert np.array_equal(quad, one_lane)
    assert np.array_equal(quad, _mul_gen(group, O.fr_dot(sc, ks)))


@pytest.mark.parametrize("group", [1, 2], ids=["G1", "G2"])
def test_quad_doubling_cases(bg, monkeypatch, group):
    """Repeated and opposite bases under equal scalars: bucket sums that hit P + P (the
    doubling branch of the exact quad addition) and P + (-P) (the point at infinity)."""
    rng = np.random.default_rng(1400 + group)
    n = 600
    ks = O.rand_scalars(rng, n)
    pts = O.g1_gen_mul_batch(ks) if group == 1 else O.g2_gen_mul_batch(ks)
    sc = O.rand_scalars(rng, n)
    q = O.P_Q
    half = pts.shape[1] // 2
    for i in range(0, 80, 2):  # P_i = P_{i+100}: doubling in the same bucket
        pts[i] = pts[i + 100]
        sc[i] = sc[i + 100]
    for i in range(1, 80, 2):  # P_i = -P_{i+100}: cancellation in the same bucket
        y = O.array_to_ints(pts[i + 100, half:].reshape(-1, 4), None)
        pts[i, :half] = pts[i + 100, :half]
        pts[i, half:] = O.ints_to_array([(q - O.unmont(v, q)) % q for v in y], q).reshape(-1)
        sc[i] = sc[i + 100]
    monkeypatch.setenv("BG16_QUAD", "0")
    one_lane = _msm(group, pts, sc)
    monkeypatch.delenv("BG16_QUAD")
    quad = _msm(group, pts, sc)
    want = O.msm_g1(pts, sc) if group == 1 else O.msm_g2(pts, sc)
    assert np.array_equal(one_lane, want)
    assert np.array_equal(quad, want)


def test_quad_proof(bg, monkeypatch, tmp_path):
    """A 2^10 proof (every MSM on the quad path) equals the one-lane proof and the
    trapdoor closed form."""
    bendg.setup_square_chain(tmp_path, 10, seed=1500)
    inst = load(tmp_path)
    pk = bendg.ProvingKey(tmp_path)
    z = np.fromfile(tmp_path / "witness.bin", dtype=np.uint32)
    rs = np.fromfile(tmp_path / "rs.bin", dtype=np.uint32)
    try:
        monkeypatch.setenv("BG16_QUAD", "0")
        one_lane, _ = pk.prove(z, rs)
        monkeypatch.delenv("BG16_QUAD")
        for i in range(3):  # eager, then recorded and replayed where graphs are on
            quad, _ = pk.prove(z, rs)
            assert np.array_equal(quad, one_lane), f"proof {i}"
    finally:
        pk.close()
    packed = bendg.pack16(one_lane.reshape(-1, 16)).reshape(-1)
    assert np.array_equal(packed.view(np.uint64), np.concatenate(points_of(closed_form(inst))))
