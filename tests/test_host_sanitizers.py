"""Host-side checks built with the compilers' sanitizers (CPU only): the
device-memory pool, the classify tile map and the emit's chunk tickets under
AddressSanitizer and UndefinedBehaviorSanitizer, the oracle under both, and a stress emulation of
the scan's decoupled look-back protocol (plain and under ThreadSanitizer)."""
from __future__ import annotations

import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
GOLDEN = os.path.join(ROOT, "tests", "golden")
CSRC = os.path.join(ROOT, "framework_amd", "csrc")
ASAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
SAN_ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="halt_on_error=1",
               TSAN_OPTIONS="halt_on_error=1")


def _run(cmd, exe, args=(), timeout=600):
    subprocess.run(cmd, check=True)
    res = subprocess.run([exe, *map(str, args)], capture_output=True, text=True, env=SAN_ENV, timeout=timeout)
    assert res.returncode == 0 and res.stdout.strip().endswith("OK"), res.stdout[-2000:] + res.stderr[-4000:]


@pytest.mark.parametrize("src", ["pool_check.cpp", "tile_map_check.cpp", "emit_ticket_check.cpp"])
def test_host_checks_under_asan_ubsan(tmp_path, src):
    exe = str(tmp_path / src.replace(".cpp", ""))
    _run(["g++", "-std=c++17", *ASAN, "-I", CSRC, os.path.join(GOLDEN, src), "-o", exe], exe)


def test_oracle_under_asan_ubsan(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    _run(["gcc", "-std=c99", *ASAN, "-ffp-contract=off", "-fopenmp", "-I", os.path.join(ROOT, "oracle"),
          "-I", CSRC, os.path.join(GOLDEN, "oracle_sanitize.c"), os.path.join(ROOT, "oracle", "dmc_oracle.c"),
          "-lm", "-o", exe], exe)


@pytest.mark.parametrize("tiles,threads,seed", [(20000, 8, 1), (3000, 16, 2), (64 * 65 + 1, 4, 3)])
def test_lookback_protocol_emulation(tmp_path, tiles, threads, seed):
    exe = str(tmp_path / "lookback")
    _run(["g++", "-std=c++17", "-O2", "-pthread", os.path.join(GOLDEN, "lookback_emulation.cpp"), "-o", exe], exe,
         (tiles, threads, seed))


def test_lookback_protocol_under_tsan(tmp_path):
    if shutil.which("g++") is None:
        pytest.skip("no g++")
    exe = str(tmp_path / "lookback_tsan")
    _run(["g++", "-std=c++17", "-O1", "-g", "-fsanitize=thread", "-pthread",
          os.path.join(GOLDEN, "lookback_emulation.cpp"), "-o", exe], exe, (2000, 6, 4))
