"""Runs the C++ operator-level tests (tests/cpp/operator_test.cpp): the GPU drop-in operators driven
through the reference's Operator / OperatorFactory surface with the reference's join and aggregation
fixtures."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "omnioperator_amd", "lib", "omni_operator_test")


@pytest.mark.gpu
def test_operator_surface_cpp():
    assert os.path.exists(EXE), "build the native library first (python -m omnioperator_amd.build)"
    r = subprocess.run([EXE], stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=120)
    print(r.stdout)
    assert r.returncode == 0, r.stdout
    assert "0 failure(s)" in r.stdout
