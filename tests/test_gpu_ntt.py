"""GPU parity of the NTT family and qap_h (bend2/src/ntt.bend:431-459) against the CPU
oracle (pinned by the naive DFT and the QAP identity), bit-exact."""
import numpy as np
import pytest

import bendg
from oracle import oracle as O

pytestmark = pytest.mark.gpu


def _u32(a):
    return np.ascontiguousarray(a).view(np.uint32)


@pytest.mark.parametrize("lg", [1, 2, 5, 10, 11, 13, 16])
@pytest.mark.parametrize("direction", [0, 1, 2, 3], ids=["fft", "ifft", "coset_fft", "coset_ifft"])
def test_ntt_vs_oracle(bg, lg, direction):
    rng = np.random.default_rng(lg * 4 + direction)
    a = O.rand_scalars(rng, 1 << lg)
    got = bendg.ntt(_u32(a), lg, direction).view(np.uint64)
    assert np.array_equal(got, O.ntt(a, lg, direction))


@pytest.mark.parametrize("lg", [20, 22])
def test_ntt_round_trip_large(bg, lg):
    rng = np.random.default_rng(lg)
    a = _u32(O.rand_scalars(rng, 1 << lg))
    f = bendg.ntt(a, lg, 0)
    assert np.array_equal(bendg.ntt(f, lg, 1), a.reshape(-1, 8))
    cf = bendg.ntt(a, lg, 2)
    assert np.array_equal(bendg.ntt(cf, lg, 3), a.reshape(-1, 8))


def test_ntt_large_vs_oracle(bg):
    lg = 19
    rng = np.random.default_rng(19)
    a = O.rand_scalars(rng, 1 << lg)
    assert np.array_equal(bendg.ntt(_u32(a), lg, 0).view(np.uint64), O.ntt(a, lg, 0))


def test_ntt_words16(bg):
    rng = np.random.default_rng(3)
    lg = 9
    a = O.rand_scalars(rng, 1 << lg)
    w16 = bendg.unpack16(_u32(a).reshape(-1, 8))
    got = bendg.pack16(bendg.ntt(w16, lg, 0, fmt="words16"))
    assert np.array_equal(got.view(np.uint64), O.ntt(a, lg, 0))


def _satisfied_abc(rng, lg):
    d = 1 << lg
    n = d - 2
    a = O.rand_scalars(rng, d)
    b = O.rand_scalars(rng, d)
    b[n:] = 0
    am = O.array_to_ints(a[:n])
    bm = O.array_to_ints(b[:n])
    # c = a * b in Montgomery form: (aR)(bR)/R = abR
    c = np.zeros_like(a)
    cm = np.zeros((n, 4), np.uint64)
    O.lib().or_f_mul_vec(0, O._ptr(cm), O._ptr(np.ascontiguousarray(a[:n])), O._ptr(np.ascontiguousarray(b[:n])), n)
    c[:n] = cm
    return a, b, c


@pytest.mark.parametrize("lg", [2, 5, 11, 15, 19])
def test_qap_h_vs_oracle(bg, lg):
    rng = np.random.default_rng(100 + lg)
    a, b, c = _satisfied_abc(rng, lg)
    got = bendg.qap_h(_u32(a), _u32(b), _u32(c), lg).view(np.uint64)
    want = O.qap_h(a, b, c, lg)
    assert np.array_equal(got, want)
    assert not want[-1].any()  # h[d-1] = 0 for a satisfied instance


@pytest.mark.parametrize("tile", [9, 10, 11])
@pytest.mark.parametrize("radix", [2, 4])
def test_ntt_tile_radix_variants(bg, tmp_path, tile, radix):
    """Every pass variant the tuning table can select (tile 2^9 / 2^10 / 2^11 elements,
    radix 2 / 4) gives the same transforms and qap_h, at a one-pass and a multi-pass size."""
    t = tmp_path / "tune.txt"
    sizes = (8, 12, 19)
    t.write_text("".join(f"ntt {lg} tile={tile} radix={radix}\n" for lg in sizes))
    bendg.tune_load(str(t))
    try:
        for lg in sizes:
            rng = np.random.default_rng(1000 + 10 * tile + radix + lg)
            a = O.rand_scalars(rng, 1 << lg)
            print(f"ntt tile={tile} radix={radix} 2^{lg}", flush=True)
            for direction in (0, 3):
                got = bendg.ntt(_u32(a), lg, direction).view(np.uint64)
                assert np.array_equal(got, O.ntt(a, lg, direction)), (lg, direction)
            if lg <= 12:
                b, c = O.rand_scalars(rng, 1 << lg), O.rand_scalars(rng, 1 << lg)
                h = bendg.qap_h(_u32(a), _u32(b), _u32(c), lg).view(np.uint64)
                assert np.array_equal(h, O.qap_h(a, b, c, lg))
    finally:
        bendg.tune_load(None)
