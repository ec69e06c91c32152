"""Builds tests/cpp/api_check.cpp against include/giao_hip.hpp and the HIP library."""

import subprocess
from pathlib import Path

ROOT = Path(__file__).resolve().parents[1]


def build(tmp: Path) -> Path:
    exe = tmp / "api_check"
    lib = ROOT / "giaoint_amd"
    subprocess.run(
        ["g++", "-std=c++20", "-O1", "-Wall", "-Wextra", "-Werror", f"-I{ROOT / 'include'}",
         str(ROOT / "tests" / "cpp" / "api_check.cpp"), f"-L{lib}", "-lgiaohip", f"-Wl,-rpath,{lib}", "-o", str(exe)],
        check=True,
    )
    return exe


def run(exe: Path, mode: str, tmp: Path) -> dict:
    out = subprocess.run([str(exe), mode, str(tmp)], check=True, capture_output=True, text=True, timeout=300).stdout
    result = {}
    for line in out.splitlines():
        name, _, rest = line.partition(" ")
        result[name] = rest
    return result
