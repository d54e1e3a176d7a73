"""The host-only C and C++ code under AddressSanitizer, LeakSanitizer and
UndefinedBehaviorSanitizer: the seeded initial conditions (nb_ic.cpp), the
particle-file reader and writer (nb_io.cpp) and the CPU oracle (oracle.c), driven
by tests/sanitize_driver.cpp over small and edge-case inputs (n = 1, coincident
bodies, capped, short, torn and headerless files, rejected arguments, a node-pool
overflow). Built with gcc/g++ only: no device code is involved."""
import shutil
import subprocess
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parents[1]
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
       "-fno-sanitize-recover=all"]


def run(cmd, cwd):
    return subprocess.run(cmd, cwd=cwd, capture_output=True, text=True, timeout=300)


def test_host_code_under_sanitizers(tmp_path):
    if not shutil.which("gcc") or not shutil.which("g++"):
        pytest.skip("gcc/g++ not found")
    probe = tmp_path / "probe.c"
    probe.write_text("int main(void) { return 0; }\n")
    if run(["gcc", *SAN, str(probe), "-o", str(tmp_path / "probe")], tmp_path).returncode != 0:
        pytest.skip("this gcc cannot link the sanitizer runtimes")
    inc = ["-I", str(ROOT / "include"), "-I", str(ROOT / "oracle")]
    steps = [
        ["gcc", "-std=c11", "-fopenmp", *SAN, "-c", str(ROOT / "oracle" / "oracle.c"), "-o",
         "oracle.o"],
        ["g++", "-std=c++17", *SAN, *inc, "-c", str(ROOT / "n-body_amd" / "csrc" / "nb_ic.cpp"),
         "-o", "nb_ic.o"],
        ["g++", "-std=c++17", *SAN, *inc, "-c", str(ROOT / "n-body_amd" / "csrc" / "nb_io.cpp"),
         "-o", "nb_io.o"],
        ["g++", "-std=c++17", "-Wall", *SAN, *inc, "-c", str(ROOT / "tests" / "sanitize_driver.cpp"),
         "-o", "driver.o"],
        ["g++", "-fopenmp", *SAN, "driver.o", "nb_ic.o", "nb_io.o", "oracle.o", "-lm", "-o",
         "driver"],
    ]
    for cmd in steps:
        r = run(cmd, tmp_path)
        assert r.returncode == 0, " ".join(cmd) + "\n" + r.stderr
    scratch = tmp_path / "scratch"
    scratch.mkdir()
    r = run([str(tmp_path / "driver"), str(scratch)], tmp_path)
    print(r.stdout + r.stderr)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "0 failures" in r.stdout
    assert "Sanitizer" not in r.stderr and "runtime error" not in r.stderr
