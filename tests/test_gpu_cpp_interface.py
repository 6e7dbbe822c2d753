"""Runs the C++ restatement of the reference's ElasticMaterial unit tests
(tests/cpp/test_elastic_material.cpp) through the pdgpu::ElasticMaterial /
CriticalStretchDamageModel interface mirror on the GPU."""
import os
import subprocess

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "test_elastic_material")


def test_reference_unit_tests_through_cpp_interface():
    assert os.path.exists(EXE), "built by __graft_entry__.build()"
    p = subprocess.run([EXE], capture_output=True, text=True, timeout=120)
    print(p.stdout)
    assert p.returncode == 0, p.stdout + p.stderr
    assert "PASSED" in p.stdout
