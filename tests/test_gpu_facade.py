"""The C++ facade (include/GSHTrans) on the GPU: tests/cpp/test_facade.cpp,
built by __graft_entry__.build(), runs the reference's accessor checks and
coefficient round trips (tests/CheckCoeff2Coeff.h:74-83) through
GaussLegendreGrid / CanonicalComponentField / CanonicalComponentExpansion."""
import os
import subprocess

import pytest

from tests.conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_facade_round_trips():
    binary = os.path.join(ROOT, "tests", "cpp", "test_facade")
    assert os.path.exists(binary), "build() compiles tests/cpp/test_facade"
    r = subprocess.run([binary], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
