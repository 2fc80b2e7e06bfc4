"""End-to-end GPU prove (bend2/prove.bend's Pr.run on the device) against the CPU oracle
prover and the trapdoor closed form, bit-exact; repeated proofs are identical."""
import numpy as np
import pytest

import bendg
from oracle import oracle as O
from instance import closed_form, load, points_of

pytestmark = pytest.mark.gpu


def _files(d):
    z = np.fromfile(d / "witness.bin", dtype=np.uint32)
    rs = np.fromfile(d / "rs.bin", dtype=np.uint32)
    return z, rs


def _split(proof_packed):
    p = proof_packed.view(np.uint64)
    return p[:8], p[8:24], p[24:32]


@pytest.fixture(scope="module")
def keys(tmp_path_factory):
    out = {}
    for lg in (3, 4, 10, 12):
        d = tmp_path_factory.mktemp(f"k{lg}")
        bendg.setup_square_chain(d, lg, seed=42 + lg)
        out[lg] = d
    return out


@pytest.mark.parametrize("lg", [3, 4, 10, 12])
def test_prove_vs_oracle_and_closed_form(bg, keys, lg):
    d = keys[lg]
    print(f"prove K={lg}", flush=True)
    pk = bendg.ProvingKey(d)
    z, rs = _files(d)
    proof16, t = pk.prove(z, rs, fmt="words16")
    A, B, C = _split(bendg.pack16(proof16.reshape(-1, 16)).reshape(-1))
    inst = load(d)
    oA, oB, oC, _ = O.prove(inst)
    assert np.array_equal(A, oA) and np.array_equal(B, oB) and np.array_equal(C, oC)
    eA, eB, eC = points_of(closed_form(inst))
    assert np.array_equal(A, eA) and np.array_equal(B, eB) and np.array_equal(C, eC)
    # the device proof passes the pairing equation (bg_verify, host)
    vk = np.fromfile(d / "vk.bin", dtype=np.uint32)
    assert bendg.verify(vk, proof16, z[16:32])
    assert not bendg.verify(vk, proof16, z[32:48])
    pk.close()


def test_prove_packed_and_repeat(bg, keys):
    d = keys[10]
    pk = bendg.ProvingKey(d)
    z, rs = _files(d)
    zp = bendg.pack16(z.reshape(-1, 16)).reshape(-1)
    rp = bendg.pack16(rs.reshape(-1, 16)).reshape(-1)
    first, _ = pk.prove(zp, rp, fmt="packed")
    for _ in range(3):
        again, _ = pk.prove(zp, rp, fmt="packed")
        assert np.array_equal(first, again)
    p16, _ = pk.prove(z, rs, fmt="words16")
    assert np.array_equal(bendg.pack16(p16.reshape(-1, 16)).reshape(-1), first)
    text = bendg.format_proof(first)
    lines = text.strip().split("\n")
    assert [l.split()[0] for l in lines] == ["A", "B", "C"] and len(lines[1].split()) == 5
    ax = int(lines[0].split()[1], 16)
    assert ax == O.unmont(O.from_words(first.view(np.uint64)[:4]), O.P_Q)
    pk.close()


def test_two_keys_resident(bg, keys):
    """two proving keys resident at once give the same proofs as alone"""
    pa, pb = bendg.ProvingKey(keys[4]), bendg.ProvingKey(keys[10])
    za, ra = _files(keys[4])
    zb, rb = _files(keys[10])
    a1, _ = pa.prove(za, ra)
    b1, _ = pb.prove(zb, rb)
    a2, _ = pa.prove(za, ra)
    assert np.array_equal(a1, a2)
    inst = load(keys[10])
    oA, oB, oC, _ = O.prove(inst)
    A, B, C = _split(bendg.pack16(b1.reshape(-1, 16)).reshape(-1))
    assert np.array_equal(A, oA) and np.array_equal(C, oC)
    pa.close()
    pb.close()


@pytest.mark.parametrize("lg", [16, 18])
def test_prove_large_closed_form(bg, tmp_path, lg):
    print(f"setup K={lg}", flush=True)
    bendg.setup_square_chain(tmp_path, lg, seed=42)
    print(f"prove K={lg}", flush=True)
    pk = bendg.ProvingKey(tmp_path)
    z, rs = _files(tmp_path)
    proof16, t = pk.prove(z, rs, fmt="words16")
    proof16, t = pk.prove(z, rs, fmt="words16")
    print(f"K={lg} phases: " + " ".join(f"{k}={v:.2f}" for k, v in t.items()), flush=True)
    A, B, C = _split(bendg.pack16(proof16.reshape(-1, 16)).reshape(-1))
    print("closed form", flush=True)
    eA, eB, eC = points_of(closed_form(load(tmp_path)))
    assert np.array_equal(A, eA) and np.array_equal(B, eB) and np.array_equal(C, eC)
    assert bendg.verify(np.fromfile(tmp_path / "vk.bin", dtype=np.uint32), proof16, z[16:32])
    pk.close()


def test_prove_windowed_and_fixed_base_agree(bg, keys, monkeypatch):
    """Keys load with fixed-base tables by default (2^(wc) P_i per window, one set of
    buckets for all windows); BG16_FIXED_BASE=0 keeps the plain windowed MSMs.  Both give
    the same proof, equal to the oracle's."""
    d = keys[12]
    z, rs = _files(d)
    pk_fixed = bendg.ProvingKey(d)
    monkeypatch.setenv("BG16_FIXED_BASE", "0")
    pk_plain = bendg.ProvingKey(d)
    monkeypatch.delenv("BG16_FIXED_BASE")
    a, _ = pk_fixed.prove(z, rs)
    b, _ = pk_plain.prove(z, rs)
    assert np.array_equal(a, b)
    assert pk_fixed.device_bytes() > pk_plain.device_bytes()
    oA, oB, oC, _ = O.prove(load(d))
    A, B, C = _split(bendg.pack16(a.reshape(-1, 16)).reshape(-1))
    assert np.array_equal(A, oA) and np.array_equal(B, oB) and np.array_equal(C, oC)
    pk_fixed.close()
    pk_plain.close()


def test_prove_tuned_widths_used_and_exact(bg, keys, tmp_path):
    """A key loaded under a tuning table takes the table's fixed-base widths for its G1
    batch, H and B2 (the planner reads the table first), and its proof is unchanged and
    equal to the closed form."""
    d = keys[10]
    z, rs = _files(d)
    pk0 = bendg.ProvingKey(d)
    ref, _ = pk0.prove(z, rs)
    w0 = pk0.windows()
    pk0.close()
    hdr = [int(x) for x in np.fromfile(d / "header.bin", dtype=np.uint32)]
    import math
    half = lambda n: round(2 * math.log2(n)) / 2  # noqa: E731
    want = {"g1": 9 if w0["g1"] != 9 else 10, "h": 11 if w0["h"] != 11 else 12, "b2": 8 if w0["b2"] != 8 else 9}
    t = tmp_path / "tune.txt"
    t.write_text(f"msm 1 fixed {half(hdr[4] + hdr[5] + hdr[7]):g} c={want['g1']}\n"
                 f"msm 1 fixed {half(hdr[6]):g} c={want['h']}\n"
                 f"msm 2 fixed {half(hdr[5]):g} c={want['b2']}\n")
    bendg.tune_load(str(t))
    try:
        pk = bendg.ProvingKey(d)
        assert pk.windows() == want
        got, _ = pk.prove(z, rs)
        pk.close()
    finally:
        bendg.tune_load(None)
    assert np.array_equal(got, ref)
    eA, eB, eC = points_of(closed_form(load(d)))
    A, B, C = _split(bendg.pack16(got.reshape(-1, 16)).reshape(-1))
    assert np.array_equal(A, eA) and np.array_equal(B, eB) and np.array_equal(C, eC)


@pytest.mark.parametrize("m", ["2", "4", "8"])
def test_prove_g2_group_widths_agree(bg, keys, monkeypatch, m):
    """The G2 bucket reduction's level-1 group width (pairs below 2^16 fixed-base buckets,
    fours above, eights with a bucket set per window) does not change the proof."""
    d = keys[12]
    z, rs = _files(d)
    pk = bendg.ProvingKey(d)
    ref, _ = pk.prove(z, rs)
    monkeypatch.setenv("BG16_MSM_M_G2", m)
    got, _ = pk.prove(z, rs)
    monkeypatch.delenv("BG16_MSM_M_G2")
    assert np.array_equal(got, ref)
    pk.close()


def test_cli_prove_and_verify(bg, keys):
    """bg16_prove (the bend2/prove.bend command-line shape) prints the library's proof and,
    with --verify, checks it against vk.bin."""
    import subprocess
    from pathlib import Path
    cli = Path(bendg.__file__).resolve().parent.parent / "bg16_prove"
    d = keys[10]
    out = subprocess.run([str(cli), "--verify", str(d)], capture_output=True, text=True, timeout=120)
    assert out.returncode == 0, out.stderr
    lines = out.stdout.strip().split("\n")
    pk = bendg.ProvingKey(d)
    z, rs = _files(d)
    proof16, _ = pk.prove(z, rs, fmt="words16")
    pk.close()
    want = bendg.format_proof(bendg.pack16(proof16.reshape(-1, 16)).reshape(-1)).strip().split("\n")
    assert lines[:3] == want
    assert lines[-1] == "V ok"
    assert any(l.startswith("T total ") for l in lines)


@pytest.mark.parametrize("sched", ["5", "9"])
def test_cli_schedules(bg, keys, sched):
    """Every proof schedule gives the same proof (5: one stream); a value outside 0..5
    keeps the default."""
    import os
    import subprocess
    from pathlib import Path
    cli = Path(bendg.__file__).resolve().parent.parent / "bg16_prove"
    d = keys[10]
    ref = subprocess.run([str(cli), str(d)], capture_output=True, text=True, timeout=120)
    env = dict(os.environ, BG16_PROVE_SCHEDULE=sched)
    out = subprocess.run([str(cli), "--verify", str(d)], capture_output=True, text=True, timeout=120, env=env)
    assert ref.returncode == 0 and out.returncode == 0, out.stderr
    assert out.stdout.split("\n")[:3] == ref.stdout.split("\n")[:3]
    assert out.stdout.strip().split("\n")[-1] == "V ok"


@pytest.mark.parametrize("lg", [4, 8, 10, 12, 16])
def test_device_combination_equals_host(bg, tmp_path, lg):
    """BG_COMBINE_DEVICE (MSM finishing, groth16.bend's combination and the affine
    normalisation on the device) gives the proof the host combination gives, bit for bit,
    in both encodings and on repeated calls; the K=12 proof also equals the closed form."""
    d = tmp_path / f"k{lg}"
    bendg.setup_square_chain(d, lg, seed=500 + lg)
    z, rs = _files(d)
    pk = bendg.ProvingKey(d)
    host16, _ = pk.prove(z, rs, fmt="words16")
    print(f"device combination K={lg}", flush=True)
    dev16, _ = pk.prove(z, rs, fmt="words16", device_combine=True)
    assert np.array_equal(dev16, host16)
    zp = bendg.pack16(z.reshape(-1, 16)).reshape(-1)
    rp = bendg.pack16(rs.reshape(-1, 16)).reshape(-1)
    for _ in range(2):
        devp, _ = pk.prove(zp, rp, fmt="packed", device_combine=True)
        assert np.array_equal(devp, bendg.pack16(host16.reshape(-1, 16)).reshape(-1))
    if lg == 12:
        eA, eB, eC = points_of(closed_form(load(d)))
        A, B, C = _split(devp)
        assert np.array_equal(A, eA) and np.array_equal(B, eB) and np.array_equal(C, eC)
    pk.close()


def test_device_combination_needs_fixed_base(bg, keys, monkeypatch):
    d = keys[4]
    monkeypatch.setenv("BG16_FIXED_BASE", "0")
    pk = bendg.ProvingKey(d)
    monkeypatch.delenv("BG16_FIXED_BASE")
    z, rs = _files(d)
    with pytest.raises(bendg.BgError) as ei:
        pk.prove(z, rs, device_combine=True)
    assert ei.value.code == -1
    pk.close()
