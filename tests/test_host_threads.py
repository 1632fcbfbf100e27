"""Host threading of the library: loading libcfdgpu.so gives its (LLVM) OpenMP workers a 1 ms
block time unless the process sets KMP_BLOCKTIME or OMP_WAIT_POLICY itself (spinning workers
after the host-side setup pushed a CPU-quota-limited process into throttling that stalled queued
solves). Each case runs in a fresh process: the runtime reads its settings once."""
from __future__ import annotations

import os
import subprocess
import sys
from pathlib import Path

import pytest

from cfd_amd import _lib as L

ROOT = Path(__file__).resolve().parents[1]
LIBOMP = Path("/opt/rocm/llvm/lib/libomp.so")

PROBE = """
import ctypes, sys
lib = ctypes.CDLL(sys.argv[1])      # the library's load-time setting runs here
omp = ctypes.CDLL(sys.argv[2])      # the same OpenMP runtime the library linked
print(omp.kmp_get_blocktime())
"""


def _blocktime(extra_env: dict) -> int:
    env = {k: v for k, v in os.environ.items() if k not in ("KMP_BLOCKTIME", "OMP_WAIT_POLICY")}
    env.update(extra_env)
    out = subprocess.run([sys.executable, "-c", PROBE, str(L.LIB_PATH), str(LIBOMP)], env=env,
                         capture_output=True, text=True, timeout=120, check=True)
    return int(out.stdout.strip().splitlines()[-1])


@pytest.mark.skipif(not L.LIB_PATH.exists() or not LIBOMP.exists(), reason="library or libomp not present")
def test_load_sets_a_short_block_time():
    assert _blocktime({}) == 1


@pytest.mark.skipif(not L.LIB_PATH.exists() or not LIBOMP.exists(), reason="library or libomp not present")
def test_process_settings_take_precedence():
    assert _blocktime({"KMP_BLOCKTIME": "50"}) == 50
    assert _blocktime({"OMP_WAIT_POLICY": "PASSIVE"}) == 0
