"""GPU parity of the Pippenger MSM (Msm.g1 / Msm.g2, bend2/src/msm.bend:293-298, :525-530)
against the CPU oracle (brute-force-pinned Pippenger restatement) and, at large n, the
closed form MSM = (sum a_i k_i mod r) G for bases P_i = k_i G.  Affine output is unique,
so the comparison is bit-exact.

Every test runs with each bucket sort: the size rule (auto), the single-pass counting sort
(BG16_SORT=1) and the two-pass sort (BG16_SORT=2) forced at every size."""
import numpy as np
import pytest

import bendg
from oracle import oracle as O

pytestmark = pytest.mark.gpu

R = O.P_R


@pytest.fixture(autouse=True, params=["auto", "1", "2"], ids=["sort-auto", "sort-single", "sort-two-pass"])
def sort_mode(request, monkeypatch):
    if request.param != "auto":
        monkeypatch.setenv("BG16_SORT", request.param)
    return request.param


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


def _mul_gen(group, k: int):
    kk = np.array(O.to_words(k % R), dtype=np.uint64)
    if group == 1:
        out = np.zeros(8, dtype=np.uint64)
        O.lib().or_g1_mul_aff(O._ptr(out), O._ptr(O.g1_pack([O.G1_GEN])[0]), O._ptr(kk))
    else:
        out = np.zeros(16, dtype=np.uint64)
        O.lib().or_g2_mul_aff(O._ptr(out), O._ptr(O.g2_pack([O.G2_GEN])[0]), O._ptr(kk))
    return out


def _bases(group, rng, n):
    ks = O.rand_scalars(rng, n)
    pts = O.g1_gen_mul_batch(ks) if group == 1 else O.g2_gen_mul_batch(ks)
    return pts, ks


def _msm(group, pts, sc, **kw):
    f = bendg.msm_g1 if group == 1 else bendg.msm_g2
    return f(_u32(pts), _u32(sc), **kw).view(np.uint64)


def _oracle(group, pts, sc):
    return O.msm_g1(pts, sc) if group == 1 else O.msm_g2(pts, sc)


@pytest.mark.parametrize("group", [1, 2], ids=["G1", "G2"])
@pytest.mark.parametrize("n", [1, 2, 3, 8, 64, 1024])
def test_msm_small_vs_oracle(bg, group, n):
    rng = np.random.default_rng(100 + n + 7 * group)
    pts, _ = _bases(group, rng, n)
    sc = O.rand_scalars(rng, n)
    print(f"msm G{group} n={n}", flush=True)
    assert np.array_equal(_msm(group, pts, sc), _oracle(group, pts, sc))


@pytest.mark.parametrize("group", [1, 2], ids=["G1", "G2"])
def test_msm_edge_cases(bg, group):
    """zero / one / r-1 / >= r scalars, equal scalars (one heavy bucket), infinity bases,
    repeated and opposite bases (bucket doubling and cancellation)."""
    rng = np.random.default_rng(5 + group)
    n = 600
    pts, _ = _bases(group, rng, n)
    sc = O.rand_scalars(rng, n)
    sc[0] = 0
    sc[1] = O.to_words(1)
    sc[2] = O.to_words(R - 1)
    sc[3] = O.to_words((1 << 256) - 1)          # reduced mod r by the MSM
    sc[4] = O.to_words(R)                        # == 0 mod r
    sc[10:300] = sc[9]                           # 290 equal scalars -> heavy buckets
    pts[20] = 0                                  # infinity bases
    pts[21] = 0
    pts[30] = pts[31]                            # same base twice
    sc[30] = sc[31]
    pts[40] = pts[41]                            # P and -P with equal scalars cancel
    half = pts.shape[1] // 2
    negy = O.ints_to_array([(O.P_Q - O.unmont(v, O.P_Q)) % O.P_Q for v in
                            O.array_to_ints(pts[41, half:].reshape(-1, 4), None)], O.P_Q)
    pts[40, half:] = negy.reshape(-1)
    sc[40] = sc[41]
    want_sc = sc.copy()
    want_sc[3] = O.to_words(((1 << 256) - 1) % R)
    want_sc[4] = 0
    got = _msm(group, pts, sc)
    assert np.array_equal(got, _oracle(group, pts, want_sc))


@pytest.mark.parametrize("c", [2, 3, 5, 8, 11, 13, 16])
def test_msm_window_widths(bg, c):
    rng = np.random.default_rng(c)
    n = 300
    pts, _ = _bases(1, rng, n)
    sc = O.rand_scalars(rng, n)
    assert np.array_equal(_msm(1, pts, sc, window_bits=c), _oracle(1, pts, sc))


def test_msm_all_zero_and_empty(bg):
    rng = np.random.default_rng(9)
    pts, _ = _bases(1, rng, 16)
    sc = np.zeros((16, 4), dtype=np.uint64)
    assert not _msm(1, pts, sc).any()
    out = bendg.msm_g1(np.zeros((0, 16), np.uint32), np.zeros((0, 8), np.uint32))
    assert not out.any()


def test_msm_words16_encoding(bg):
    rng = np.random.default_rng(12)
    n = 100
    pts, _ = _bases(1, rng, n)
    sc = O.rand_scalars(rng, n)
    p16 = bendg.unpack16(_u32(pts).reshape(n, 16))
    s16 = bendg.unpack16(_u32(sc).reshape(n, 8))
    got16 = bendg.msm_g1(p16, s16, fmt="words16")
    assert np.array_equal(bendg.pack16(got16).view(np.uint64), _oracle(1, pts, sc))


@pytest.mark.parametrize("group,log2n", [(1, 14), (1, 16), (1, 18), (2, 12), (2, 14)])
def test_msm_closed_form(bg, group, log2n):
    rng = np.random.default_rng(1000 + log2n + group)
    n = 1 << log2n
    pts, ks = _bases(group, rng, n)
    sc = O.rand_scalars(rng, n)
    want = _mul_gen(group, O.fr_dot(sc, ks))
    print(f"msm closed form G{group} 2^{log2n}", flush=True)
    assert np.array_equal(_msm(group, pts, sc), want)


@pytest.mark.parametrize("group", [1, 2], ids=["G1", "G2"])
def test_msm_repeated_and_opposite_bases(bg, group):
    """Buckets whose generic-formula accumulation meets x(acc) = x(P): the same base many
    times (doublings, spanning many accumulation slices) and P / -P pairs (cancellation)
    mixed with ordinary entries; exact recomputation of those buckets must give the sum."""
    rng = np.random.default_rng(77 + group)
    n = 2500
    pts, ks = _bases(group, rng, n)
    sc = O.rand_scalars(rng, n)
    pts[100:1300] = pts[99]          # 1201 copies of one base ...
    sc[100:1300] = sc[99]            # ... with one scalar: heavy buckets of equal points
    half = pts.shape[1] // 2
    for i in range(1400, 1420):      # opposite bases with equal scalars
        y = O.array_to_ints(pts[i + 40, half:].reshape(-1, 4), None)
        pts[i, :half] = pts[i + 40, :half]
        pts[i, half:] = O.ints_to_array([(O.P_Q - O.unmont(v, O.P_Q)) % O.P_Q for v in y], O.P_Q).reshape(-1)
        sc[i] = sc[i + 40]
    assert np.array_equal(_msm(group, pts, sc), _oracle(group, pts, sc))


@pytest.mark.parametrize("group", [1, 2], ids=["G1", "G2"])
def test_msm_batch_uneven(bg, group):
    """Independent MSMs of uneven sizes (including empty and single-point ones) in one
    batched launch sequence equal each MSM computed alone."""
    rng = np.random.default_rng(300 + group)
    sizes = [1, 0, 37, 1000, 5, 4096, 2] if group == 1 else [3, 0, 500, 1, 64]
    pts_l, sc_l, want = [], [], []
    for n in sizes:
        pts, _ = _bases(group, rng, max(n, 1))
        pts = pts[:n]
        sc = O.rand_scalars(rng, n)
        pts_l.append(_u32(pts))
        sc_l.append(_u32(sc))
        want.append(_oracle(group, pts, sc) if n else np.zeros(8 if group == 1 else 16, np.uint64))
    got = bendg.msm_batch(group, pts_l, sc_l).view(np.uint64)
    for i, n in enumerate(sizes):
        assert np.array_equal(got[i], want[i]), f"MSM {i} (n={n})"


def test_msm_batch_closed_form_large(bg):
    rng = np.random.default_rng(400)
    sizes = [1 << 12, 3 << 14, 1 << 16, 70001]
    pts_l, sc_l, want = [], [], []
    for n in sizes:
        pts, ks = _bases(1, rng, n)
        sc = O.rand_scalars(rng, n)
        pts_l.append(_u32(pts))
        sc_l.append(_u32(sc))
        want.append(_mul_gen(1, O.fr_dot(sc, ks)))
    got = bendg.msm_batch(1, pts_l, sc_l).view(np.uint64)
    for i in range(len(sizes)):
        assert np.array_equal(got[i], want[i])


def _closed(group, pts, ks, sc, **kw):
    want = _mul_gen(group, O.fr_dot(sc, ks))
    return np.array_equal(_msm(group, pts, sc, **kw), want)


@pytest.mark.parametrize("group,log2n,kind", [
    (1, 16, "ones"), (1, 16, "equal"), (1, 16, "small"), (1, 17, "three"),
    (2, 13, "ones"), (2, 13, "three")])
def test_msm_skewed_scalars(bg, group, log2n, kind):
    """Scalar distributions that concentrate digits in few buckets (witness values 0/1,
    one repeated value, small integers, three distinct values): buckets spread over
    thousands of accumulation slices, merged by the segmented tree levels."""
    rng = np.random.default_rng(500 + log2n + group)
    n = 1 << log2n
    pts, ks = _bases(group, rng, n)
    if kind == "ones":
        sc = np.zeros((n, 4), dtype=np.uint64)
        sc[:, 0] = 1
        sc[::7, 0] = 0
    elif kind == "equal":
        sc = np.repeat(O.rand_scalars(rng, 1), n, axis=0)
    elif kind == "small":
        sc = np.zeros((n, 4), dtype=np.uint64)
        sc[:, 0] = rng.integers(0, 16, size=n)
    else:
        vals = O.rand_scalars(rng, 3)
        sc = vals[rng.integers(0, 3, size=n)]
    print(f"msm skewed G{group} 2^{log2n} {kind}", flush=True)
    assert _closed(group, pts, ks, sc)


@pytest.mark.parametrize("group,log2n,c", [(1, 17, 12), (1, 17, 14), (2, 13, 12), (2, 13, 13)])
def test_msm_short_top_window(bg, group, log2n, c):
    """Window widths whose top window holds only a few digit values (254-bit scalars):
    every scalar lands in one of a handful of top-window buckets."""
    rng = np.random.default_rng(600 + c + group)
    n = 1 << log2n
    pts, ks = _bases(group, rng, n)
    sc = O.rand_scalars(rng, n)
    print(f"msm top window G{group} 2^{log2n} c={c}", flush=True)
    assert _closed(group, pts, ks, sc, window_bits=c)


@pytest.mark.parametrize("group", [1, 2], ids=["G1", "G2"])
@pytest.mark.parametrize("kval", ["1", "3", "100", "4096"])
def test_msm_reduction_block_override(bg, monkeypatch, group, kval):
    """BG16_MSM_K (level-2 block width) is rounded down to a power of two and capped by
    the group count, so any value keeps the result exact."""
    rng = np.random.default_rng(300 + group)
    n = 3000
    pts, _ = _bases(group, rng, n)
    sc = O.rand_scalars(rng, n)
    monkeypatch.setenv("BG16_MSM_K", kval)
    got = _msm(group, pts, sc)
    monkeypatch.delenv("BG16_MSM_K")
    assert np.array_equal(got, _oracle(group, pts, sc))


@pytest.mark.parametrize("group", [1, 2], ids=["G1", "G2"])
def test_msm_plan_slices(bg, group):
    """Accumulation slicing of the windowed planner: every sorted entry belongs to exactly
    one slice of 2..64 entries; G2 MSMs of 2^16 entries or more take slices of at least 8;
    MSMs of 2^24 entries or more take slices of at least 20 (one device round holds far
    fewer threads than entries / 20)."""
    for log2n in (4, 10, 14, 16, 18, 20, 22):
        n = 1 << log2n
        info = bg.msm_plan_info(n, group)
        E, L, T = info["entries"], info["L"], info["acc_threads"]
        assert E == info["windows"] * n, (log2n, info)
        assert 2 <= L <= 64, (log2n, info)
        assert T == -(-E // L), (log2n, info)
        if group == 2 and E >= 1 << 16:
            assert L >= 8, (log2n, info)
        if E >= 1 << 24:
            assert L >= 20, (log2n, info)
