"""The device top-gamma selection (tgtp_amd/csrc/select.hpp) must permute
the frontier exactly as libstdc++'s std::nth_element does (aesc.h:417):
compiled for the host and compared on tie-heavy and adversarial inputs.
CPU only."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_select_matches_std_nth_element(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "select_vs_std"
    subprocess.run([cxx, "-std=c++20", "-O2", f"-I{ROOT / 'tgtp_amd' / 'csrc'}",
                    str(ROOT / "tests" / "cpp" / "select_vs_std.cpp"), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe), "24000"], capture_output=True, text=True)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok 24000")
