"""CPU test of the LDS bond scheduler (peridigm_amd/csrc/pdg_sched.hpp),
compiled with the host compiler: row orders stay permutations of the rows
and the bank-conflict cost falls to near one per 16-lane group iteration."""
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SRC = os.path.join(ROOT, "tests", "cpp", "test_sched.cpp")


@pytest.mark.skipif(shutil.which("g++") is None, reason="no host C++ compiler")
def test_lds_bond_schedule(tmp_path):
    exe = str(tmp_path / "test_sched")
    subprocess.run(["g++", "-O2", "-std=c++17", "-o", exe, SRC], check=True, timeout=300)
    p = subprocess.run([exe], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASSED" in p.stdout
