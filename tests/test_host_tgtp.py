"""Host helpers of the tgtp pipeline (tgtp_amd/csrc/host_tgtp.cpp), compiled
for the host with AddressSanitizer and UBSan: the symmetrization map of
aesc.h:538-545 and its refusal of unsorted rows and missing reverse edges
(tests/cpp/symmetrize_check.cpp).  CPU only."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]


def test_symmetrize_sources_map_and_refusals(tmp_path):
    cxx = shutil.which("g++")
    if cxx is None:
        pytest.skip("g++ not available")
    exe = tmp_path / "symmetrize_check"
    build = subprocess.run(
        [cxx, "-std=c++20", "-O1", "-g", "-fopenmp", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=all", f"-I{ROOT / 'tgtp_amd' / 'csrc'}",
         str(ROOT / "tests" / "cpp" / "symmetrize_check.cpp"),
         str(ROOT / "tgtp_amd" / "csrc" / "host_tgtp.cpp"), "-o", str(exe)],
        capture_output=True, text=True)
    if build.returncode != 0 and "asan" in build.stderr.lower():
        pytest.skip("sanitizer runtime not available")
    assert build.returncode == 0, build.stderr[-2000:]
    out = subprocess.run([str(exe), "2000"], capture_output=True, text=True, timeout=300)
    assert out.returncode == 0, out.stdout + out.stderr
    assert out.stdout.startswith("ok 2000")
