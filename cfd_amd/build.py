"""Builds libcfdgpu.so for gfx950 in-tree (cfd_amd/lib/) and, for tests and the CPU baseline,
the oracle under oracle/ (plus oracle/_ref from the reference checkout when it is present).

    python cfd_amd/build.py              # library + oracle, every source compiled again
    python cfd_amd/build.py --lib-only
    python cfd_amd/build.py --incremental   # only what changed
"""
from __future__ import annotations

import argparse
import os
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent


def _make(directory: Path, *targets: str, force: bool = True) -> None:
    jobs = str(min(16, os.cpu_count() or 4))
    cmd = ["make", "-C", str(directory), f"-j{jobs}", *(["-B"] if force else []), *targets]
    subprocess.run(cmd, check=True)


def build_library(force: bool = True) -> Path:
    _make(ROOT / "cfd_amd" / "csrc", force=force)
    out = ROOT / "cfd_amd" / "lib" / "libcfdgpu.so"
    if not out.exists():
        raise RuntimeError("libcfdgpu.so was not produced")
    return out


def build_oracle(force: bool = True) -> None:
    _make(ROOT / "oracle", force=force)
    # the integration guide's binding over libcfdgpu, for its tests (needs the reference header)
    subprocess.run([sys.executable, str(ROOT / "oracle" / "build_shim.py")], check=True)


def main(argv=None) -> int:
    ap = argparse.ArgumentParser()
    ap.add_argument("--lib-only", action="store_true")
    ap.add_argument("--incremental", action="store_true")
    args = ap.parse_args(argv)
    build_library(force=not args.incremental)
    if not args.lib_only:
        build_oracle(force=not args.incremental)
    return 0


if __name__ == "__main__":
    sys.exit(main())
