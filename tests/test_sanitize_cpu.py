"""Host code under sanitizers (CPU only): the library's host half (key generation, the key
reader and its validation, the pairing verifier, host scalar multiplication, the tuning
table, every entry point's no-device error path) under AddressSanitizer + UBSan; the proving
engine's scheduler on a simulated prover under ThreadSanitizer and ASan + UBSan; the CPU
oracle under ASan + UBSan.  Builds tests/sanitize/ (make) and runs each program, which
checks its own results and exits non-zero on a check failure or a sanitizer report."""
import os
import subprocess
from pathlib import Path

import numpy as np
import pytest

import bendg
from oracle import oracle as O
from instance import load

ROOT = Path(__file__).resolve().parent.parent
SAN = ROOT / "tests" / "sanitize"


@pytest.fixture(scope="module")
def built():
    p = subprocess.run(["make", "-C", str(SAN), "-j8", "all"], capture_output=True, text=True)
    if p.returncode != 0:
        pytest.fail("sanitizer build failed:\n" + p.stdout[-3000:] + p.stderr[-3000:])
    return SAN / "_build"


@pytest.fixture(scope="module")
def work(tmp_path_factory):
    w = tmp_path_factory.mktemp("san")
    bendg.setup_square_chain(w / "key5", 5, seed=5)
    A, B, C, _ = O.prove(load(w / "key5"))
    np.concatenate([A, B, C]).astype(np.uint64).tofile(w / "proof.bin")
    return w


def _run(args, **env):
    e = dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1",
             TSAN_OPTIONS="halt_on_error=1", **env)
    p = subprocess.run([str(a) for a in args], capture_output=True, text=True, env=e, timeout=600)
    assert p.returncode == 0, p.stdout[-2000:] + p.stderr[-4000:]
    return p.stdout


def test_scheduler_under_tsan(built):
    assert "scheduler simulation OK" in _run([built / "sched_tsan"])


def test_scheduler_under_asan_ubsan(built):
    assert "scheduler simulation OK" in _run([built / "sched_asan"])


def test_oracle_under_asan_ubsan(built, work):
    assert "oracle checks OK" in _run([built / "oracle_asan", work / "key5", work / "proof.bin"])


def test_library_host_code_under_asan_ubsan(built, work):
    out = _run([built / "host_asan", work, work / "proof.bin", ROOT / "bend-groth16_amd" / "tune_gfx950.txt"],
               HIP_VISIBLE_DEVICES="")
    assert "host checks OK" in out
