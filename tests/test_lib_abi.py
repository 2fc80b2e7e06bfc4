"""The C-ABI library loads and exports every entry point include/bg16.h declares, and
host-side argument / file validation rejects bad input before any device work (CPU)."""
import re
import shutil
from pathlib import Path

import numpy as np
import pytest

import bendg

ROOT = Path(__file__).resolve().parent.parent


def declared():
    text = (ROOT / "include" / "bg16.h").read_text()
    return sorted(set(re.findall(r"^\s*(?:int|const char\*)\s+(bg_\w+)\s*\(", text, re.M)))


def test_header_declares_entry_points():
    names = declared()
    for n in ("bg_msm_g1", "bg_msm_g2", "bg_ntt", "bg_qap_h", "bg_pk_load", "bg_prove", "bg_setup_square_chain"):
        assert n in names


def test_library_exports_every_declared_symbol():
    L = bendg.lib()
    missing = [n for n in declared() if not hasattr(L, n)]
    assert not missing, missing


def test_bad_arguments_rejected_without_device():
    with pytest.raises(bendg.BgError) as e:
        bendg.setup_square_chain("/nonexistent-dir/x", 0)
    assert e.value.code == -1


def _copy_key(src: Path, dst: Path) -> Path:
    shutil.copytree(src, dst)
    return dst


@pytest.fixture(scope="module")
def key3(tmp_path_factory):
    d = tmp_path_factory.mktemp("k3")
    bendg.setup_square_chain(d, 3, seed=1)
    return d


@pytest.mark.parametrize("case", ["missing", "truncated", "limb", "header", "column"])
def test_pk_load_rejects_malformed_keys(key3, tmp_path, case):
    d = _copy_key(key3, tmp_path / case)
    if case == "missing":
        (d / "pk_h.bin").unlink()
    elif case == "truncated":
        data = (d / "pk_l.bin").read_bytes()
        (d / "pk_l.bin").write_bytes(data[:-4])
    elif case == "limb":
        w = np.fromfile(d / "pk_a.bin", dtype=np.uint32)
        w[5] = 0x10000
        w.tofile(d / "pk_a.bin")
    elif case == "header":
        h = np.fromfile(d / "header.bin", dtype=np.uint32)
        h[6] = 1 << 20  # n_h beyond the domain
        h.tofile(d / "header.bin")
    elif case == "column":
        w = np.fromfile(d / "r1cs_b.bin", dtype=np.uint32)
        nc = 8
        w[2 + nc] = 999  # first entry's column
        w.tofile(d / "r1cs_b.bin")
    with pytest.raises(bendg.BgError) as e:
        bendg.ProvingKey(d)
    assert e.value.code == -4  # BG_ERR_IO, raised before any device work
