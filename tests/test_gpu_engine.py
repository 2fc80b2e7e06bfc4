"""Proving engine (continuous batching, bg_engine_* / bg_prove_batch): every proof made in a
mix of concurrent requests of different sizes equals the proof of the same inputs made
alone by bg_prove (itself pinned to the oracle prover and the trapdoor closed form in
test_gpu_prove.py), bit for bit; a device-memory budget that admits one proof at a time
changes nothing but the schedule.  Queued requests of one key shape are proved together
(one launch sequence for up to 8 proofs, keys may differ); those batches are checked the
same way, including batches mixing two keys and both witness encodings."""
import numpy as np
import pytest

import bendg
from instance import closed_form, load, points_of

pytestmark = pytest.mark.gpu

SIZES = (4, 8, 10, 12)


def _files(d):
    z = np.fromfile(d / "witness.bin", dtype=np.uint32)
    rs = np.fromfile(d / "rs.bin", dtype=np.uint32)
    return z, rs


@pytest.fixture(scope="module")
def keys(tmp_path_factory):
    out = {}
    for lg in SIZES:
        d = tmp_path_factory.mktemp(f"e{lg}")
        bendg.setup_square_chain(d, lg, seed=900 + lg)
        out[lg] = d
    # a second key of the K=10 shape (own trapdoor and witness): batches mix keys
    d = tmp_path_factory.mktemp("e10b")
    bendg.setup_square_chain(d, 10, seed=77)
    out["10b"] = d
    return out


@pytest.fixture(scope="module")
def loaded(bg, keys):
    pks = {lg: bendg.ProvingKey(d) for lg, d in keys.items()}
    solo = {}
    for lg, pk in pks.items():
        z, rs = _files(keys[lg])
        solo[lg], _ = pk.prove(z, rs, fmt="words16")
    yield pks, solo
    for pk in pks.values():
        pk.close()


def test_solo_proofs_are_closed_form(loaded, keys):
    _, solo = loaded
    lg = 10
    inst = load(keys[lg])
    p = bendg.pack16(solo[lg].reshape(-1, 16)).reshape(-1).view(np.uint64)
    eA, eB, eC = points_of(closed_form(inst))
    assert np.array_equal(p[:8], eA) and np.array_equal(p[8:24], eB) and np.array_equal(p[24:32], eC)


@pytest.mark.parametrize("slots", [1, 3, 6])
def test_engine_mix_equals_solo(loaded, keys, slots):
    pks, solo = loaded
    rng = np.random.default_rng(slots)
    order = [SIZES[i] for i in rng.integers(0, len(SIZES), size=24)]
    eng = bendg.Engine(slots=slots)
    files = {lg: _files(keys[lg]) for lg in SIZES}
    print(f"engine slots={slots} order={order}", flush=True)
    tickets = [(lg, eng.submit(pks[lg], *files[lg], fmt="words16")) for lg in order]
    for lg, t in tickets:
        proof, tm = eng.wait(t)
        assert np.array_equal(proof, solo[lg]), f"K={lg} ticket {t}"
        assert tm["latency"] >= tm["proving"] >= 0
    info = eng.info()
    assert info["completed"] == len(order) and info["slots"] == slots
    eng.close()


def test_engine_budget_admits_one_at_a_time(loaded, keys):
    """A budget just above one K=12 proof's work arrays: requests wait for memory, idle
    slots give their arrays back, and results are unchanged."""
    pks, solo = loaded
    eng0 = bendg.Engine(slots=1)
    z, rs = _files(keys[12])
    eng0.wait(eng0.submit(pks[12], z, rs))
    need = eng0.info()["reserved_bytes"]
    eng0.close()
    eng = bendg.Engine(slots=4, hbm_budget=int(need * 1.3))
    order = [12, 4, 12, 10, 12, 8, 4, 12]
    tickets = [(lg, eng.submit(pks[lg], *_files(keys[lg]))) for lg in order]
    for lg, t in tickets:
        proof, _ = eng.wait(t)
        assert np.array_equal(proof, solo[lg])
    assert eng.info()["reserved_bytes"] <= int(need * 1.3) + need
    eng.close()


def test_engine_rejects_key_over_budget(loaded, keys):
    pks, _ = loaded
    eng = bendg.Engine(slots=2, hbm_budget=1 << 20)
    z, rs = _files(keys[12])
    with pytest.raises(bendg.BgError) as ei:
        eng.submit(pks[12], z, rs)
    assert ei.value.code == -3
    eng.close()


def test_prove_batch_packed(loaded, keys):
    pks, solo = loaded
    order = [10, 4, 12, 8, 10, 12]
    ws, rr = [], []
    for lg in order:
        z, rs = _files(keys[lg])
        ws.append(bendg.pack16(z.reshape(-1, 16)).reshape(-1))
        rr.append(bendg.pack16(rs.reshape(-1, 16)).reshape(-1))
    outs = bendg.prove_batch([pks[lg] for lg in order], ws, rr, fmt="packed", slots=3)
    for lg, o in zip(order, outs):
        assert np.array_equal(o, bendg.pack16(solo[lg].reshape(-1, 16)).reshape(-1))


@pytest.mark.parametrize("slots", [1, 2])
def test_engine_batches_two_keys_of_one_shape(loaded, keys, slots):
    """16 queued K=10 requests over two keys (words16 and packed witnesses interleaved) and
    a K=12 request between them: the slots take them in batches of one shape; every
    proof equals its solo proof."""
    pks, solo = loaded
    eng = bendg.Engine(slots=slots)
    order = ["10b", 10] * 4 + [12] + [10, "10b"] * 4
    tickets = []
    for i, k in enumerate(order):
        z, rs = _files(keys[k])
        if i % 3 == 0:
            zp = bendg.pack16(z.reshape(-1, 16)).reshape(-1)
            rp = bendg.pack16(rs.reshape(-1, 16)).reshape(-1)
            tickets.append((k, "packed", eng.submit(pks[k], zp, rp, fmt="packed")))
        else:
            tickets.append((k, "words16", eng.submit(pks[k], z, rs, fmt="words16")))
    print(f"engine batch slots={slots} order={order}", flush=True)
    for k, fmt, t in tickets:
        proof, _ = eng.wait(t)
        want = solo[k] if fmt == "words16" else bendg.pack16(solo[k].reshape(-1, 16)).reshape(-1)
        assert np.array_equal(proof, want), f"key {k} {fmt} ticket {t}"
    eng.close()


def test_key_free_refused_while_queued(loaded, keys):
    """bg_pk_free refuses a key whose proof is still queued in an engine (the key stays
    loaded and the proof completes unchanged); once the proof is collected it frees."""
    _, solo = loaded
    pk = bendg.ProvingKey(keys[4])
    eng = bendg.Engine(slots=1)
    pks, _ = loaded
    z12, rs12 = _files(keys[12])
    ahead = [eng.submit(pks[12], z12, rs12) for _ in range(8)]  # keep the one slot busy
    z, rs = _files(keys[4])
    t = eng.submit(pk, z, rs)
    with pytest.raises(bendg.BgError) as ei:
        pk.close()
    assert ei.value.code == -1
    for a in ahead:
        eng.wait(a)
    proof, _ = eng.wait(t)
    assert np.array_equal(proof, solo[4])
    pk.close()  # no proof of the key left: frees
    eng.close()


def test_unreduced_inputs_rejected(loaded, keys):
    """A witness element or r / s at or above r is an argument error (FORMAT.md stores
    reduced Montgomery values), from bg_prove and from bg_engine_submit alike."""
    pks, solo = loaded
    z, rs = _files(keys[8])
    r_mod = 21888242871839275222246405745257275088548364400416034343698204186575808495617
    limbs = np.array([(r_mod >> (16 * i)) & 0xFFFF for i in range(16)], dtype=np.uint32)
    zb = z.copy()
    zb[16 * 3:16 * 4] = limbs
    rb = rs.copy()
    rb[16:32] = limbs
    eng = bendg.Engine(slots=1)
    for w, r in ((zb, rs), (z, rb)):
        with pytest.raises(bendg.BgError) as ei:
            pks[8].prove(w, r, fmt="words16")
        assert ei.value.code == -1
        with pytest.raises(bendg.BgError) as ei:
            eng.submit(pks[8], w, r)
        assert ei.value.code == -1
    proof, _ = pks[8].prove(z, rs, fmt="words16")  # the key still proves
    assert np.array_equal(proof, solo[8])
    eng.close()
