"""The C-ABI from a plain C program (tests/c_abi/fs_c_smoke.c): compiled with
gcc against include/, linked with the library, run as a separate process."""
import os
import subprocess

import pytest

from conftest import ROOT

LIBDIR = os.path.join(ROOT, "felspa_amd")


def _build(tmp_path):
    exe = str(tmp_path / "fs_c_smoke")
    cmd = ["gcc", "-std=c99", "-O1", "-Wall", "-Werror", "-I", os.path.join(ROOT, "include"),
           os.path.join(ROOT, "tests", "c_abi", "fs_c_smoke.c"), "-L", LIBDIR, "-lfelspa_amd",
           "-Wl,-rpath," + LIBDIR, "-lm", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def test_c_program_host_part(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fs_c_smoke ok" in r.stdout


@pytest.mark.gpu
def test_c_program_solve(tmp_path):
    exe = _build(tmp_path)
    r = subprocess.run([exe, "--gpu"], capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "fs_c_smoke ok (solve" in r.stdout
