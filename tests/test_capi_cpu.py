"""CPU-side checks of the C ABI library (no GPU compute): it loads, exports
every symbol include/gshtrans_amd.h declares, and its host-only helpers
(Gauss-Legendre nodes, coefficient sizes, Wigner seeds, status strings,
argument validation) behave like the reference's."""
import ctypes
import math
import os
import re

import numpy as np
import pytest

import gshtrans_amd as G
import oracle
from tests.conftest import ROOT, gpu_available


def header_symbols():
    txt = open(os.path.join(ROOT, "include", "gshtrans_amd.h")).read()
    return sorted(set(re.findall(r"\b(gsh_[a-z_0-9]+)\s*\(", txt)))


def test_exports_every_declared_symbol():
    lib = G.library()
    declared = header_symbols()
    assert set(declared) == set(G.EXPORTS)
    for name in declared:
        assert hasattr(lib, name), name


def test_nm_lists_symbols():
    import subprocess
    out = subprocess.run(["nm", "-D", "--defined-only", G.LIB_PATH], capture_output=True,
                         text=True).stdout
    for name in header_symbols():
        assert re.search(r"\bT " + name + r"\b", out), name


@pytest.mark.parametrize("L", [0, 1, 8, 64, 200, 1024])
def test_gl_nodes_match_oracle_and_numpy(L):
    th, w = G.gl_nodes(L)
    tho, wo = oracle.gl_nodes(L)
    assert np.array_equal(th, tho) and np.array_equal(w, wo)
    if L <= 300:
        x, wn = np.polynomial.legendre.leggauss(L + 1)
        assert np.max(np.abs(th - np.arccos(-x))) < 1e-13
        assert np.max(np.abs(w - wn)) < 1e-13


def test_coefficient_sizes():
    for (lmax, n) in [(0, 0), (5, 0), (5, 2), (5, -5), (64, 1), (1024, 2)]:
        assert G.coefficient_size(lmax, n) == oracle.gsh_size(lmax, lmax, n, "All")
        assert G.coefficient_size(lmax, n, real=True) == oracle.gsh_size(lmax, lmax, n,
                                                                          "NonNegative")
        for (l, m) in [(abs(n), 0), (lmax, lmax), (lmax, -lmax)]:
            assert G.coefficient_index(lmax, n, l, m) == oracle.gsh_index(lmax, lmax, n, l, m)
        assert G.coefficient_index(lmax, n, lmax, lmax, True) == \
            oracle.gsh_index(lmax, lmax, n, lmax, lmax, "NonNegative")
    with pytest.raises(ValueError):
        G.coefficient_size(3, 4)


def test_wigner_seed_bit_identical_to_oracle():
    for (l, n, th) in [(0, 0, 0.3), (1024, 2, 0.0023), (1024, -2, 1.5707963267948966),
                       (513, 1, 3.1), (7, -7, 0.0), (7, 7, math.pi)]:
        assert G.wigner_seed(l, n, th) == oracle.wigner_min_order(l, n, th)


def test_status_strings_and_errors():
    lib = G.library()
    assert lib.gsh_status_string(0) == b"ok"
    for s in range(1, 9):
        assert len(lib.gsh_status_string(s)) > 0
    th = np.zeros(3)
    assert lib.gsh_gl_nodes(-1, G._dp(th), G._dp(th)) == 1
    v = ctypes.c_double()
    assert lib.gsh_wigner_seed(2, 3, 0.1, ctypes.byref(v)) == 1
    h = ctypes.c_void_p()
    assert lib.gsh_grid_create(4, 5, 0, 0, ctypes.byref(h)) == 1  # nMax > lMax
    assert lib.gsh_grid_create(4, 1, 7, 0, ctypes.byref(h)) == 1  # bad MRange


@pytest.mark.skipif(gpu_available(), reason="checks the no-device path")
def test_grid_create_without_device_fails_loudly():
    with pytest.raises(G.GSHError) as e:
        G.GaussLegendreGrid(8, 2)
    assert e.value.status == 7


def test_random_coefficients_masks_and_reproducible():
    a = G.random_coefficients(16, 2, seed=5)
    b = G.random_coefficients(16, 2, seed=5)
    c = G.random_coefficients(16, 2, seed=6)
    assert np.array_equal(a, b) and not np.array_equal(a, c)
    assert a[-1] == 0  # (L, L) zeroed for complex fields at lmax == L
    r = G.random_coefficients(16, 0, real=True, seed=1)
    for l in range(0, 17):
        assert r[G.coefficient_index(16, 0, l, 0, True)].imag == 0
    assert r[-1].imag == 0
    t = G.random_coefficients(16, 1, seed=3, grid_lmax=32)
    assert t[-1] != 0
    bt = G.random_coefficients(8, 0, seed=2, batch=3)
    assert bt.shape == (3, G.coefficient_size(8, 0))
    # roughly standard normal before the (l+1)^-decay scaling
    big = G.random_coefficients(200, 0, seed=9, decay=0.0)
    assert abs(np.std(big.real) - 1) < 0.02 and abs(np.mean(big.imag)) < 0.02


def test_cpp_facade_compiles_and_reports_errors():
    """include/GSHTrans (the C++ facade mirroring the reference's
    GaussLegendreGrid / CanonicalComponentField / Wigner API) compiles as
    C++23 with the reference's call shapes (tests/cpp/test_facade.cpp, built
    with the compiler gshtrans_amd.build uses), and its umbrella header also
    compiles as C++20, where the two-argument subscripts are operator();
    without a device the grid constructor raises the library's NO_DEVICE
    status."""
    import subprocess
    from gshtrans_amd.build import cxx23
    inc = os.path.join(ROOT, "include")
    src = os.path.join(ROOT, "tests", "cpp", "test_facade.cpp")
    r = subprocess.run([cxx23(), "-std=c++23", "-fsyntax-only", "-Wall", "-I", inc, src],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    snippet = (
        "#include <GSHTrans/All>\n"
        "using namespace GSHTrans;\n"
        "double f(const Wigner<double, Ortho, All, All, Multiple>& d,\n"
        "         ComplexCanonicalComponentExpansion<0, GaussLegendreGrid<double, All, All>>& u) {\n"
        "  u(1, 0) = 1.0;\n"
        "  return d(0, 0)(1, 0) + d(0, 0)[1][0] + u(1, 0).real();\n"
        "}\n")
    r = subprocess.run(["g++", "-std=c++20", "-fsyntax-only", "-Wall", "-I", inc, "-x", "c++",
                        "-"], input=snippet, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    binary = os.path.join(ROOT, "tests", "cpp", "test_facade")
    if gpu_available() or not os.path.exists(binary):
        return
    r = subprocess.run([binary], capture_output=True, text=True, timeout=60)
    assert r.returncode != 0 and "no gfx950 device" in r.stderr


def test_wrapper_argument_validation():
    """The Python wrappers check dtype, row count and per-row size of every
    array before they pass a raw pointer to the C ABI, so a wrong tensor
    raises ValueError instead of letting a kernel read or write past the
    allocation.  Host tensors pass the dtype and shape checks and then fail
    the device check."""
    import torch
    from gshtrans_amd import _check_coeffs, _check_field
    c = torch.zeros((3, 10), dtype=torch.complex128)
    _bad = [
        (lambda: _check_coeffs(c.to(torch.complex64), 3, 10, torch.complex128, "t"), "complex128"),
        (lambda: _check_coeffs(c, 4, 10, torch.complex128, "t"), "(4, >= 10)"),
        (lambda: _check_coeffs(c, 3, 11, torch.complex128, "t"), "(3, >= 11)"),
        (lambda: _check_coeffs(c[:, ::2], 3, 5, torch.complex128, "t"), "contiguous"),
        (lambda: _check_coeffs(c, 3, 10, torch.complex128, "t"), "cuda"),
        (lambda: _check_coeffs(np.zeros((3, 10), complex), 3, 10, torch.complex128, "t"),
         "torch tensor"),
    ]
    f = torch.zeros((2, 5, 8), dtype=torch.float64)
    _bad += [
        (lambda: _check_field(f, 2, 5, 8, torch.complex128, "t"), "complex128"),
        (lambda: _check_field(f.to(torch.float32), 2, 5, 8, torch.float64, "t"), "float64"),
        (lambda: _check_field(f, 3, 5, 8, torch.float64, "t"), "(3, 5, 8)"),
        (lambda: _check_field(f[:, :4], 2, 4, 8, torch.float64, "t"), "contiguous"),
        (lambda: _check_field(f, 2, 5, 8, torch.float64, "t"), "cuda"),
    ]
    for fn, msg in _bad:
        with pytest.raises(ValueError) as e:
            fn()
        assert msg in str(e.value), (msg, str(e.value))
