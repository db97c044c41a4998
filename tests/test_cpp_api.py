"""The C++ host layer (include/katric/cetric.hpp) compiles against the
library (CPU) and passes its checks on the GPU."""
import os
import subprocess

import pytest

from tests.conftest import ROOT
from tests.cpp import build as cppbuild


def test_cpp_layer_compiles():
    out = cppbuild.build()
    assert os.path.exists(out)


@pytest.mark.gpu
def test_cpp_layer_runs():
    if not os.path.exists(cppbuild.OUT):
        pytest.fail("C++ test program not built (run __graft_entry__.build())")
    res = subprocess.run([cppbuild.OUT, os.path.join(ROOT, "tests", "golden")], capture_output=True, text=True,
                         timeout=120)
    assert res.returncode == 0, res.stdout + res.stderr
    assert "ALL OK" in res.stdout
