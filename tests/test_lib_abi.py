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


P_Q = 21888242871839275222246405745257275088696311157297823662689037894645226208583
P_R = 21888242871839275222246405745257275088548364400416034343698204186575808495617


def _limbs16(x: int) -> np.ndarray:
    return np.array([(x >> (16 * i)) & 0xFFFF for i in range(16)], dtype=np.uint32)


def _int16(w) -> int:
    return sum(int(v) << (16 * i) for i, v in enumerate(w))


@pytest.mark.parametrize("case", ["missing", "truncated", "limb", "header", "column", "coef_unreduced",
                                  "point_unreduced", "off_curve_g1", "off_curve_g2", "off_curve_pk_g1"])
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
    elif case == "coef_unreduced":
        w = np.fromfile(d / "r1cs_a.bin", dtype=np.uint32)
        nc = 8
        w[2 + nc + 1:2 + nc + 17] = _limbs16(P_R)  # first coefficient = r (stored values are below r)
        w.tofile(d / "r1cs_a.bin")
    elif case == "point_unreduced":
        w = np.fromfile(d / "pk_b1.bin", dtype=np.uint32)
        x = _int16(w[:16])
        w[:16] = _limbs16(x + P_Q)  # same residue, not below q
        w.tofile(d / "pk_b1.bin")
    elif case in ("off_curve_g1", "off_curve_pk_g1"):
        name = "pk_l.bin" if case == "off_curve_g1" else "pk_g1.bin"
        w = np.fromfile(d / name, dtype=np.uint32)
        y = _int16(w[16:32])
        w[16:32] = _limbs16((y + 1) % P_Q)  # y + 1: off the curve
        w.tofile(d / name)
    elif case == "off_curve_g2":
        w = np.fromfile(d / "pk_b2.bin", dtype=np.uint32)
        x0 = _int16(w[:16])
        w[:16] = _limbs16((x0 + 1) % P_Q)
        w.tofile(d / "pk_b2.bin")
    with pytest.raises(bendg.BgError) as e:
        bendg.ProvingKey(d)
    assert e.value.code == -4  # BG_ERR_IO, raised before any device work


def test_kernel_resource_check():
    """The build's resource check flags dynamic stacks and scratch in proving-path
    kernels (parsed from the compiler's kernel-resource-usage remarks), and the last
    build's proving-path kernels use neither."""
    import build  # bend-groth16_amd/build.py
    text = "\n".join([
        "x.hip:1:1: remark: Function Name: k_a [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     VGPRs: 130 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     ScratchSize [bytes/lane]: 0 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     Dynamic Stack: False [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark: Function Name: k_b [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     ScratchSize [bytes/lane]: 16 [-Rpass-analysis=kernel-resource-usage]",
        "x.hip:1:1: remark:     Dynamic Stack: True [-Rpass-analysis=kernel-resource-usage]",
    ])
    ks = build.parse_resources(text)
    assert [k["name"] for k in ks] == ["k_a", "k_b"]
    bad = build.check_resources("bg_msm.hip", ks)
    assert len(bad) == 2 and all("k_b" in b for b in bad)
    assert len(build.check_resources("bg_micro.hip", ks)) == 1  # scratch within the diagnostic bound
    obj = ROOT / "bend-groth16_amd" / "_obj"
    reports = sorted(obj.glob("*.hip.res"))
    if not reports:
        pytest.skip("library not built in this tree")
    names = set()
    for r in reports:
        kernels = build.parse_resources(r.read_text())
        assert not build.check_resources(r.name[:-4], kernels), r.name
        names |= {k["name"] for k in kernels}
    assert any("k_accumulate" in n for n in names) and any("k_ntt_pass" in n for n in names)
