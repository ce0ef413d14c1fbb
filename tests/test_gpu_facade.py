"""The C++ facade (include/GSHTrans) on the GPU: tests/cpp/test_facade.cpp,
built by __graft_entry__.build() with a C++23 compiler, runs the reference's
accessor checks, coefficient round trips (tests/CheckCoeff2Coeff.h:74-83) and
Wigner view checks (tests/CheckLegendre.h, tests/CheckAdditionTheorem.h call
shapes) through GaussLegendreGrid / CanonicalComponentField /
CanonicalComponentExpansion / Wigner, and writes the inputs and results of
its device-resident containers, which are held here to the oracle's
synthesis and analysis of the same coefficients (<= 1e-12 relative L2)."""
import os
import subprocess

import numpy as np
import pytest

import oracle
from tests.conftest import ROOT

pytestmark = pytest.mark.gpu


def test_cpp_facade_round_trips(tmp_path):
    binary = os.path.join(ROOT, "tests", "cpp", "test_facade")
    assert os.path.exists(binary), "build() compiles tests/cpp/test_facade"
    dump = tmp_path / "device_containers.bin"
    r = subprocess.run([binary, str(dump)], capture_output=True, text=True, timeout=240)
    print(r.stdout)
    assert r.returncode == 0, r.stdout + r.stderr
    assert r.stdout.strip().endswith("OK")
    # The device containers against the oracle: complex n = 1 and real n = 0
    # fields at L = 32 (the sizes test_facade.cpp uses).
    L = 32
    nt, nphi = L + 1, 2 * L
    C1 = (L + 1) ** 2 - 1
    R0 = (L + 1) * (L + 2) // 2
    raw = np.fromfile(dump, dtype=np.float64)
    sizes = [2 * C1, 2 * nt * nphi, 2 * C1, 2 * R0, nt * nphi, 2 * R0]
    assert raw.size == sum(sizes)
    parts = np.split(raw, np.cumsum(sizes)[:-1])
    u, f, v = parts[0].view(np.complex128), parts[1].view(np.complex128), parts[2].view(
        np.complex128)
    ru, rf, rv = parts[3].view(np.complex128), parts[4], parts[5].view(np.complex128)
    og = oracle.OracleGrid(L, 2, table=True)
    f_ref = og.inverse(L, 1, u).reshape(-1)
    v_ref = og.forward(L, 1, f_ref.reshape(nt, nphi))
    rf_ref = og.inverse(L, 0, ru, real=True).reshape(-1)
    rv_ref = og.forward(L, 0, rf_ref.reshape(nt, nphi), real=True)

    def rel(a, b):
        return float(np.linalg.norm(a - b) / np.linalg.norm(b))

    errs = {"complex synthesis": rel(f, f_ref), "complex analysis": rel(v, v_ref),
            "real synthesis": rel(rf, rf_ref), "real analysis": rel(rv, rv_ref)}
    print("device containers vs oracle:", {k: f"{e:.2e}" for k, e in errs.items()})
    assert max(errs.values()) <= 1e-12
