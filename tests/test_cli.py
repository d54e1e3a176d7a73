"""Headless CLI argument handling (host side; no GPU needed)."""
import subprocess
from pathlib import Path

import pytest

CLI = Path(__file__).resolve().parents[1] / "n-body_amd" / "bin" / "nbody_cli"


def run(*args):
    return subprocess.run([str(CLI), *map(str, args)], capture_output=True, text=True, timeout=60)


@pytest.mark.parametrize("v", [1, 2, 5, 6])
def test_cpu_versions_are_rejected(v):
    r = run("-v", v)
    assert r.returncode != 0 and "not part of this engine" in r.stderr


@pytest.mark.parametrize("args,msg", [(("-v", 9), "Invalid version"), (("-n", 0), "Invalid number"),
                                      (("--bogus",), "unknown argument"), (("-t",), "missing value"),
                                      (("--frames", 2), "go together"),
                                      (("-i", 7), "Invalid initialization"),
                                      (("-i", 0), "Invalid initialization")])
def test_bad_arguments(args, msg):
    r = run(*args)
    assert r.returncode != 0 and msg in r.stderr


def test_missing_file(tmp_path):
    r = run("-f", tmp_path / "none.txt")
    assert r.returncode != 0 and "Failed to open" in r.stderr
