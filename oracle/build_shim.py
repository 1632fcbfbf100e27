"""Builds oracle/_ref/libshimtest.so: the reference-side binding of the integration guide
(extracted verbatim to oracle/_ref/krylov_hip.cpp) compiled with
oracle/shim_driver.cpp against the reference's own cfd::gpu header and linked to libcfdgpu.
Test infrastructure: it needs the reference checkout (its header only) and runs in the build
container; the built library travels with oracle/_ref."""
from __future__ import annotations

import re
import shutil
import subprocess
import sys
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]
REF_INCLUDE = Path("/root/reference/src/gpu/include")
OUT = ROOT / "oracle" / "_ref"


def build() -> bool:
    """Returns False (and builds nothing) where the reference header or g++ is missing."""
    if not (REF_INCLUDE / "cfd" / "gpu" / "krylov.hpp").exists() or shutil.which("g++") is None:
        return False
    text = (ROOT / "INTEGRATION.md").read_text()
    section = text[text.index("## 2. The binding"):]
    code = re.search(r"```cpp\n(.*?)```", section, re.S).group(1)
    OUT.mkdir(parents=True, exist_ok=True)
    (OUT / "krylov_hip.cpp").write_text(code)
    lib = ROOT / "cfd_amd" / "lib"
    cmd = ["g++", "-std=c++20", "-O2", "-Wall", "-Werror", "-fPIC", "-shared",
           str(ROOT / "oracle" / "shim_driver.cpp"), f"-I{REF_INCLUDE}", f"-I{ROOT / 'include'}", f"-I{OUT}",
           f"-L{lib}", "-lcfdgpu", "-Wl,--no-undefined", "-Wl,-rpath,$ORIGIN/../../cfd_amd/lib",
           "-o", str(OUT / "libshimtest.so")]
    subprocess.run(cmd, check=True)
    return True


if __name__ == "__main__":
    if not build():
        print("build_shim: reference header or g++ not available; libshimtest.so not built")
    sys.exit(0)
