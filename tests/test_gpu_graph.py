"""Single proofs replayed from a recorded HIP graph (bg_prove after the first proof of a
key on a slot): every replay reads the new witness, equals the eagerly launched proof and
the trapdoor closed form, and keys alternating on one engine slot each keep their own
graph.  The unit is bend2/prove.bend:129-162 (one Pr.run); the graph changes only how the
device work is launched."""
import random

import numpy as np
import pytest

import bendg
from oracle import oracle as O
from instance import closed_form, load, points_of

pytestmark = pytest.mark.gpu

R = O.P_R


def chain_witness(n: int, x0: int) -> np.ndarray:
    """SquareChain witness z = (1, x_n, x_0 .. x_{n-1}) with x_{i+1} = x_i^2, as packed
    Montgomery words (n + 2 rows of 4 u64)."""
    xs = [x0 % R]
    for _ in range(n):
        xs.append(xs[-1] * xs[-1] % R)
    return O.ints_to_array([1, xs[n]] + xs[:n], mont_p=R)


def _expect(inst: dict, z: np.ndarray) -> np.ndarray:
    inst = dict(inst)
    inst["z"] = z
    return np.concatenate(points_of(closed_form(inst)))


def _packed(a16: np.ndarray) -> np.ndarray:
    return bendg.pack16(a16.reshape(-1, 16)).reshape(-1)


@pytest.fixture(scope="module")
def gkeys(tmp_path_factory):
    out = {}
    for lg in (4, 10, 12):
        d = tmp_path_factory.mktemp(f"g{lg}")
        bendg.setup_square_chain(d, lg, seed=900 + lg)
        out[lg] = d
    return out


def test_graph_replays_new_witnesses(bg, gkeys, monkeypatch):
    monkeypatch.setenv("BG16_GRAPH", "1")
    d = gkeys[10]
    inst = load(d)
    n = int(inst["header"][0])
    pk = bendg.ProvingKey(d)
    rp = _packed(np.fromfile(d / "rs.bin", dtype=np.uint32))
    before = bendg.graph_info()
    rnd = random.Random(11)
    for i in range(5):
        z = chain_witness(n, rnd.randrange(1, R))
        print(f"K=10 proof {i}", flush=True)
        proof, t = pk.prove(z.view(np.uint32).reshape(-1), rp, fmt="packed")
        assert np.array_equal(proof.view(np.uint64), _expect(inst, z)), f"proof {i}"
        assert list(t.values())[8] > 0
    after = bendg.graph_info()
    # the first proof runs eagerly, the second records the graph and launches it, the
    # rest launch it
    assert after["failed"] == before["failed"]
    assert after["recorded"] - before["recorded"] == 1
    assert after["launches"] - before["launches"] == 4
    pk.close()


def test_graph_equals_eager_with_phase_times(bg, gkeys, monkeypatch):
    d = gkeys[12]
    pk = bendg.ProvingKey(d)
    z = np.fromfile(d / "witness.bin", dtype=np.uint32)
    rs = np.fromfile(d / "rs.bin", dtype=np.uint32)
    monkeypatch.setenv("BG16_GRAPH", "0")
    eager, _ = pk.prove(z, rs)
    monkeypatch.setenv("BG16_GRAPH", "1")
    before = bendg.graph_info()
    for i in range(4):
        proof, t = pk.prove(z, rs)
        assert np.array_equal(proof, eager), f"proof {i}"
        # phases: device times (or -1 where the runtime cannot time a graph's event node)
        tv = list(t.values())
        assert all(x >= 0 or x == -1 for x in tv[:8]) and tv[8] > 0
    after = bendg.graph_info()
    assert after["recorded"] - before["recorded"] == 1 and after["launches"] - before["launches"] == 3
    # a changed BG16_ setting is a different launch sequence: recorded anew, same proof
    monkeypatch.setenv("BG16_MSM_L", "24")
    for _ in range(3):
        proof, _ = pk.prove(z, rs)
        assert np.array_equal(proof, eager)
    assert bendg.graph_info()["recorded"] - after["recorded"] == 1
    pk.close()


def test_graph_device_witness(bg, gkeys, monkeypatch):
    """a packed witness already in HBM (BG_MEM_DEVICE): replays read each new device witness"""
    from helpers import DeviceBuffers
    monkeypatch.setenv("BG16_GRAPH", "1")
    d = gkeys[12]
    inst = load(d)
    n = int(inst["header"][0])
    pk = bendg.ProvingKey(d)
    rp = _packed(np.fromfile(d / "rs.bin", dtype=np.uint32))
    before = bendg.graph_info()
    rnd = random.Random(23)
    dev = DeviceBuffers()
    try:
        for i in range(4):  # a new device buffer each time
            z = chain_witness(n, rnd.randrange(1, R))
            ptr = dev.upload(z.view(np.uint32).reshape(-1))
            proof, _ = pk.prove_device(ptr, rp)
            assert np.array_equal(proof.view(np.uint64), _expect(inst, z)), f"device witness proof {i}"
    finally:
        dev.free_all()
    after = bendg.graph_info()
    assert after["recorded"] - before["recorded"] == 1 and after["launches"] - before["launches"] == 3
    pk.close()


@pytest.mark.parametrize("order", [(12, 4, 10), (4, 10, 12)], ids=["large-first", "growing"])
def test_engine_slot_alternating_keys(bg, gkeys, monkeypatch, order):
    """keys of three sizes take turns on one engine slot: each keeps its own graph and
    every proof equals the eager proof made alone (growing: the slot's work buffers are
    reallocated after smaller keys ran, so sequences recorded before must not be replayed)"""
    pks = {lg: bendg.ProvingKey(d) for lg, d in gkeys.items()}
    files = {lg: (np.fromfile(d / "witness.bin", dtype=np.uint32), np.fromfile(d / "rs.bin", dtype=np.uint32))
             for lg, d in gkeys.items()}
    monkeypatch.setenv("BG16_GRAPH", "0")
    solo = {lg: pks[lg].prove(*files[lg])[0] for lg in pks}
    monkeypatch.setenv("BG16_GRAPH", "1")
    before = bendg.graph_info()
    eng = bendg.Engine(slots=1)
    try:
        for rnd in range(4):
            for lg in order:
                print(f"engine round {rnd} K={lg}", flush=True)
                t = eng.submit(pks[lg], *files[lg])
                proof, _ = eng.wait(t)
                assert np.array_equal(proof, solo[lg]), f"round {rnd} K={lg}"
    finally:
        eng.close()
    after = bendg.graph_info()
    assert after["failed"] == before["failed"]
    assert after["recorded"] - before["recorded"] >= 3
    assert after["launches"] - before["launches"] >= 6
    for pk in pks.values():
        pk.close()
