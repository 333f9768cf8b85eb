"""The C++ facade (include/symla/solver.hpp) driven by the reference's own
smoke / multifrontal / static-pivoting assertions, on the device path."""
import subprocess
from pathlib import Path

import pytest

pytestmark = pytest.mark.gpu

BIN = Path(__file__).resolve().parent / "cpp" / "facade_test"


def test_cpp_facade_program():
    assert BIN.exists(), "tests/cpp/facade_test is not built; run `python -m ma57_amd.build`"
    res = subprocess.run([str(BIN)], capture_output=True, text=True, timeout=120)
    assert res.returncode == 0, res.stderr + res.stdout
    assert "all checks passed" in res.stdout
