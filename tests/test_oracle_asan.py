"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer: a
standalone driver (oracle/asan_check.cpp) calls every mesh entry point of the
restatement on small seeded inputs (5 x 6 mesh, 3 bands, dense and packed
exchange, Hartree, reference density, contact term; SCF linear / DIIS / ODA,
DM CG / Newton / deflation) and must finish without a sanitizer report."""

from __future__ import annotations

import shutil
import subprocess
from pathlib import Path

import pytest

ORACLE = Path(__file__).resolve().parents[1] / "oracle"


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_entry_points_clean_under_asan_and_ubsan():
    build = subprocess.run(["make", "-s", "-C", str(ORACLE), "asan"], capture_output=True,
                           text=True, timeout=600)
    if build.returncode != 0 and "asan" in (build.stderr + build.stdout).lower():
        pytest.skip("the compiler has no AddressSanitizer runtime: " + build.stderr[-300:])
    assert build.returncode == 0, build.stderr[-2000:]
    run = subprocess.run([str(ORACLE / "_build" / "hfcpu_asan_check")], capture_output=True,
                         text=True, timeout=300)
    assert run.returncode == 0, (run.stdout[-2000:], run.stderr[-4000:])
    assert "asan check ok" in run.stdout
    assert "runtime error" not in run.stderr  # UBSan reports
