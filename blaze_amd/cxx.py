"""Builds the C++ front-end programs (host C++ linked against libblaze_gpu.so; no HIP headers)."""
from __future__ import annotations

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parent.parent
BIN = ROOT / "_build" / "bin"
PROGRAMS = {"test_frontend": ROOT / "tests" / "cpp" / "test_frontend.cpp"}


def build_examples() -> dict[str, Path]:
    BIN.mkdir(parents=True, exist_ok=True)
    lib = ROOT / "blaze_amd" / "lib"
    out = {}
    for name, src in PROGRAMS.items():
        exe = BIN / name
        cmd = ["g++", "-std=c++17", "-O2", "-Wall", f"-I{ROOT / 'include'}", str(src), f"-L{lib}", "-lblaze_gpu",
               f"-Wl,-rpath,{lib}", "-Wl,-rpath-link,/opt/rocm/lib", "-o", str(exe)]
        r = subprocess.run(cmd, capture_output=True, text=True)
        if r.returncode != 0:
            raise RuntimeError("C++ front-end build failed:\n" + r.stderr)
        out[name] = exe
    return out


if __name__ == "__main__":
    print(build_examples())
