"""The device-tuned parameter table (bg_tune_load / bg_tune_lookup, replacing the
reference's Msm.pick_c / pick_fd, bend2/src/msm.bend:37-44, and Ntt.fd, bend2/src/ntt.bend:
382-385): parsing, half-octave matching, the planner taking the table's entries over its
cost model, malformed tables rejected, and the committed table valid.  No device needed."""
from pathlib import Path

import pytest

import bendg

ROOT = Path(__file__).resolve().parent.parent
TABLE = ROOT / "bend-groth16_amd" / "tune_gfx950.txt"


@pytest.fixture
def restore_table():
    yield
    bendg.tune_load(None)


def test_table_entries_drive_the_planner(tmp_path, restore_table):
    t = tmp_path / "t.txt"
    bendg.tune_load(str(tmp_path / "missing.txt"))  # no file: an empty table
    model = bendg.msm_plan_info(1 << 14, 1)
    assert bendg.tune_lookup("msm", 1 << 14) is None
    c = 9 if model["c"] != 9 else 10
    t.write_text("# comment\n"
                 f"msm 1 windowed 14 c={c} L=20 m=4 K=64   # trailing comment\n"
                 "msm 2 fixed 15.5 c=13\n"
                 "ntt 19 tile=11 radix=2\n")
    bendg.tune_load(str(t))
    assert bendg.tune_lookup("msm", 1 << 14) == {"c": c, "L": 20, "m": 4, "K": 64}
    # nearest half octave: 2^14 * 1.1 still maps to 14, 3 * 2^14 (log2 15.58) to 15.5
    assert bendg.tune_lookup("msm", int((1 << 14) * 1.1))["c"] == c
    assert bendg.tune_lookup("msm", 3 << 14, group=2, fixed=True) == {"c": 13, "L": 0, "m": 0, "K": 0}
    assert bendg.tune_lookup("msm", 3 << 14, group=2, fixed=False) is None
    assert bendg.tune_lookup("ntt", 19) == {"tile": 11, "radix": 2}
    assert bendg.tune_lookup("ntt", 18) is None
    info = bendg.msm_plan_info(1 << 14, 1)
    assert info["c"] == c and info["c"] != model["c"]
    # an explicit width still wins over the table
    assert bendg.msm_plan_info(1 << 14, 1, window_bits=7)["c"] == 7
    # sizes the table does not hold keep the cost model
    bendg.tune_load(str(tmp_path / "missing.txt"))
    other = bendg.msm_plan_info(1 << 16, 1)
    bendg.tune_load(str(t))
    assert bendg.msm_plan_info(1 << 16, 1) == other


@pytest.mark.parametrize("line", ["msm 3 windowed 14 c=9", "msm 1 sideways 14 c=9", "msm 1 windowed 14 c=40",
                                  "msm 1 windowed 14 m=3", "msm 1 windowed 14 K=100", "ntt 19 tile=12",
                                  "ntt 19 radix=8", "fft 19 tile=10", "msm 1 windowed 14 q=1"])
def test_malformed_table_rejected(tmp_path, line, restore_table):
    t = tmp_path / "bad.txt"
    t.write_text(line + "\n")
    with pytest.raises(bendg.BgError) as e:
        bendg.tune_load(str(t))
    assert e.value.code == -4


def test_committed_table_is_valid(restore_table):
    if not TABLE.exists():
        pytest.skip("no tuned table committed")
    bendg.tune_load(str(TABLE))
    lines = [l.split("#")[0].split() for l in TABLE.read_text().splitlines()]
    lines = [l for l in lines if l]
    assert any(l[0] == "ntt" for l in lines) and any(l[0] == "msm" for l in lines)
    for l in lines:
        if l[0] == "ntt":
            assert bendg.tune_lookup("ntt", int(l[1])) is not None
        else:
            n = int(round(2 ** float(l[3])))
            e = bendg.tune_lookup("msm", n, group=int(l[1]), fixed=l[2] == "fixed")
            want = {kv.split("=")[0]: int(kv.split("=")[1]) for kv in l[4:]}
            assert all(e[k] == v for k, v in want.items()), (l, e)
    # the library's default location is this file
    bendg.tune_load(None)
    l0 = next(l for l in lines if l[0] == "ntt")
    assert bendg.tune_lookup("ntt", int(l0[1])) is not None
