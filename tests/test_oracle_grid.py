"""Pins the oracle's grid restatement (oracle/gsh_oracle.cpp, following
GSHTrans/src/GaussLegendreGrid.h:38-321) with checks independent of it:

* Gauss-Legendre nodes/weights against numpy's leggauss, exact polynomial
  quadrature and node symmetry (GaussQuad itself is not vendored);
* the FFT against numpy.fft with FFTW's conventions;
* synthesis->analysis round trips (tests/CheckCoeff2Coeff.h:40-84), with the
  reference's 5e4*eps absolute tolerance;
* brute-force synthesis/analysis at tiny L built from mpmath Wigner values
  and explicit exponentials (no FFT, no recurrence);
* single-mode synthesis, Parseval, the lMax == 0 branches and argument errors.
"""
import math

import mpmath as mp
import numpy as np
import pytest

import oracle
from oracle import OracleGrid, gsh_index
from tests.golden.make_wigner_golden import wigner_d

EPS = np.finfo(np.float64).eps


def random_coeffs(rng, L, lmax, n, real):
    """GridBase::Random{Complex,Real}Coefficient masks (GridBase.h:115-172)."""
    mr = "NonNegative" if real else "All"
    size = oracle.gsh_size(lmax, lmax, n, mr)
    c = rng.normal(size=size) + 1j * rng.normal(size=size)
    if real:
        for l in range(abs(n), lmax + 1):
            i = gsh_index(lmax, lmax, n, l, 0, mr)
            c[i] = c[i].real
        if lmax == L:
            c[-1] = c[-1].real
    elif lmax == L:
        c[-1] = 0
    return c


# --------------------------------------------------------------------------
# Quadrature
# --------------------------------------------------------------------------
@pytest.mark.parametrize("L", [0, 1, 2, 7, 64, 200])
def test_gl_nodes_vs_numpy(L):
    th, w = oracle.gl_nodes(L)
    x, wn = np.polynomial.legendre.leggauss(L + 1)
    assert np.all(np.diff(th) > 0)
    assert np.max(np.abs(th - np.arccos(-x))) < 1e-13
    assert np.max(np.abs(w - wn)) < 1e-13


@pytest.mark.parametrize("L", [10, 513, 1024])
def test_gl_exactness_and_symmetry(L):
    x, w = oracle.gl_x_ld(L + 1)
    x = x.astype(np.longdouble)
    assert abs(float(np.sum(w)) - 2.0) < 1e-15
    for k in (2, 10, 2 * L):
        exact = 2.0 / (k + 1)
        assert abs(float(np.sum(w * x ** k)) - exact) < 1e-14
    assert abs(float(np.sum(w * x ** (2 * L + 1)))) < 1e-15
    th, wd = oracle.gl_nodes(L)
    assert np.max(np.abs(th + th[::-1] - np.pi)) < 4 * EPS * np.pi
    assert np.max(np.abs(wd - wd[::-1])) < 4 * EPS


# --------------------------------------------------------------------------
# FFT with FFTW semantics
# --------------------------------------------------------------------------
@pytest.mark.parametrize("N", [2, 8, 128, 2048, 14, 30, 100, 194, 66])
def test_fft_matches_numpy(N):
    rng = np.random.default_rng(N)
    x = rng.normal(size=N) + 1j * rng.normal(size=N)
    fwd = oracle.dft(x)
    bwd = oracle.dft(x, backward=True)
    scale = np.sqrt(N) * np.max(np.abs(x))
    assert np.max(np.abs(fwd - np.fft.fft(x))) < 1e-14 * scale * np.log2(N + 1)
    assert np.max(np.abs(bwd - N * np.fft.ifft(x))) < 1e-14 * scale * np.log2(N + 1)
    xr = rng.normal(size=N)
    assert np.max(np.abs(oracle.r2c(xr) - np.fft.rfft(xr))) < 1e-13 * np.sqrt(N)
    X = rng.normal(size=N // 2 + 1) + 1j * rng.normal(size=N // 2 + 1)
    # c2r ignores Im of bins 0 and N/2 (Hermitian input); numpy irfft does too.
    assert np.max(np.abs(oracle.c2r(X, N) - N * np.fft.irfft(X, N))) < 1e-13 * N


# --------------------------------------------------------------------------
# Round trips (tests/CheckCoeff2Coeff.h)
# --------------------------------------------------------------------------
@pytest.mark.parametrize("seed", range(8))
@pytest.mark.parametrize("real", [False, True])
def test_coeff2coeff(seed, real):
    rng = np.random.default_rng(100 + seed)
    L = int(rng.integers(4, 65))
    lmax = int(rng.integers(4, L + 1))
    nmax = min(lmax, 4)
    n = int(rng.integers(-nmax, nmax + 1))
    g = OracleGrid(L, nmax, table=bool(seed % 2))
    c = random_coeffs(rng, L, lmax, n, real)
    f = g.inverse(lmax, n, c, real)
    c2 = g.forward(lmax, n, f, real)
    assert np.max(np.abs(c2 - c)) < 5e4 * EPS


def test_round_trip_field_space():
    """analysis->synthesis reproduces a band-limited field."""
    rng = np.random.default_rng(7)
    L, n = 24, -2
    g = OracleGrid(L, 2, table=False)
    c = random_coeffs(rng, L, L, n, False)
    f = g.inverse(L, n, c)
    f2 = g.inverse(L, n, g.forward(L, n, f))
    assert np.max(np.abs(f2 - f)) < 1e-12 * np.max(np.abs(f))


# --------------------------------------------------------------------------
# Brute force at tiny L (mpmath Wigner values, explicit sums)
# --------------------------------------------------------------------------
def _D_table(L, n, thetas):
    D = {}
    for i, th in enumerate(thetas):
        for l in range(abs(n), L + 1):
            for m in range(-l, l + 1):
                D[(i, l, m)] = float(wigner_d(l, m, n, th) *
                                     mp.sqrt(2 * l + 1) / (2 * mp.sqrt(mp.pi)))
    return D


@pytest.mark.parametrize("L,lmax,n", [(5, 5, 0), (6, 4, 1), (4, 4, -2), (6, 6, 2)])
def test_brute_force_complex(L, lmax, n):
    rng = np.random.default_rng(L * 10 + lmax)
    g = OracleGrid(L, 2)
    th, w = g.nodes()
    N = 2 * L
    phi = np.arange(N) * (2 * np.pi / N)
    D = _D_table(L, n, th)
    c = random_coeffs(rng, L, lmax, n, False)
    f = np.zeros((L + 1, N), dtype=complex)
    for i in range(L + 1):
        for l in range(abs(n), lmax + 1):
            for m in range(-l, l + 1):
                f[i] += c[gsh_index(lmax, lmax, n, l, m)] * D[(i, l, m)] * np.exp(1j * m * phi)
    fo = g.inverse(lmax, n, c)
    assert np.max(np.abs(fo - f)) < 1e-13 * np.max(np.abs(f))
    # analysis of an arbitrary (not band-limited) field
    h = rng.normal(size=(L + 1, N)) + 1j * rng.normal(size=(L + 1, N))
    co = g.forward(lmax, n, h)
    cb = np.zeros_like(co)
    for l in range(abs(n), lmax + 1):
        for m in range(-l, l + 1):
            acc = 0j
            for i in range(L + 1):
                acc += w[i] * (2 * np.pi / N) * D[(i, l, m)] * np.sum(h[i] * np.exp(-1j * m * phi))
            cb[gsh_index(lmax, lmax, n, l, m)] = acc
    if lmax == L:
        cb[-1] = 0  # GaussLegendreGrid.h:209-215
    assert np.max(np.abs(co - cb)) < 1e-13 * np.max(np.abs(cb))


@pytest.mark.parametrize("L,lmax,n", [(5, 5, 0), (6, 3, 2), (4, 4, -1)])
def test_brute_force_real(L, lmax, n):
    rng = np.random.default_rng(L * 7 + lmax)
    g = OracleGrid(L, 2)
    th, w = g.nodes()
    N = 2 * L
    phi = np.arange(N) * (2 * np.pi / N)
    D = _D_table(L, n, th)
    c = random_coeffs(rng, L, lmax, n, True)
    f = np.zeros((L + 1, N))
    for i in range(L + 1):
        F = np.zeros(lmax + 1, dtype=complex)
        for l in range(abs(n), lmax + 1):
            for m in range(0, l + 1):
                F[m] += c[gsh_index(lmax, lmax, n, l, m, "NonNegative")] * D[(i, l, m)]
        # c2r: Re(F_0) + 2 Re sum_{0<m<N/2} F_m e^{im phi} + Re(F_{N/2}) (-1)^j
        f[i] = F[0].real
        for m in range(1, lmax + 1):
            if m == N // 2:
                f[i] += F[m].real * np.cos(m * phi)
            else:
                f[i] += 2 * np.real(F[m] * np.exp(1j * m * phi))
    fo = g.inverse(lmax, n, c, real=True)
    assert np.max(np.abs(fo - f)) < 1e-13 * np.max(np.abs(f))
    h = rng.normal(size=(L + 1, N))
    co = g.forward(lmax, n, h, real=True)
    cb = np.zeros_like(co)
    for l in range(abs(n), lmax + 1):
        for m in range(0, l + 1):
            acc = 0j
            for i in range(L + 1):
                acc += w[i] * (2 * np.pi / N) * D[(i, l, m)] * np.sum(h[i] * np.exp(-1j * m * phi))
            cb[gsh_index(lmax, lmax, n, l, m, "NonNegative")] = acc
    assert np.max(np.abs(co - cb)) < 1e-13 * np.max(np.abs(cb))


# --------------------------------------------------------------------------
# Single mode, Parseval, quirks, errors
# --------------------------------------------------------------------------
def test_single_mode():
    L, n = 32, 1
    g = OracleGrid(L, 2)
    th, _ = g.nodes()
    phi = np.arange(2 * L) * np.pi / L
    for (l0, m0) in [(1, 0), (5, -3), (20, 17), (32, -32)]:
        c = np.zeros(oracle.gsh_size(L, L, n), dtype=complex)
        c[gsh_index(L, L, n, l0, m0)] = 1.0
        f = g.inverse(L, n, c)
        for i in (0, 9, 16, 32):
            dv = oracle.wigner_block(L, L, n, th[i])[gsh_index(L, L, n, l0, m0)]
            assert np.max(np.abs(f[i] - dv * np.exp(1j * m0 * phi))) < 1e-13


def test_parseval():
    rng = np.random.default_rng(11)
    L, n = 40, 2
    g = OracleGrid(L, 2, table=False)
    _, w = g.nodes()
    c = random_coeffs(rng, L, L, n, False)
    f = g.inverse(L, n, c)
    e_field = np.sum(w[:, None] * (2 * np.pi / (2 * L)) * np.abs(f) ** 2)
    assert abs(e_field - np.sum(np.abs(c) ** 2)) < 1e-12 * np.sum(np.abs(c) ** 2)


def test_lmax_zero_branches():
    """GaussLegendreGrid.h:137-140 and :251-260 act on element 0 only."""
    g = OracleGrid(8, 0)
    c = np.array([1.5 - 0.5j])
    f = g.inverse(0, 0, c)
    assert f[0, 0] == c[0] * 2 / (1 / math.sqrt(math.pi))
    assert np.all(f.ravel()[1:] == 0)
    fr = g.inverse(0, 0, c, real=True)
    assert fr[0, 0] == 1.5 * 2 / (1 / math.sqrt(math.pi))
    h = np.full((9, 16), 3.0 + 1j)
    out = g.forward(0, 0, h)
    assert out[0] == h[0, 0] * (1 / math.sqrt(math.pi)) / 2


def test_argument_errors():
    g = OracleGrid(8, 2, nrange="NonNegative")
    with pytest.raises(ValueError):
        g.inverse(8, -1, np.zeros(oracle.gsh_size(8, 8, -1), dtype=complex))
    with pytest.raises(ValueError):
        g.inverse(9, 0, np.zeros(oracle.gsh_size(9, 9, 0), dtype=complex))
    g2 = OracleGrid(8, 2, mrange="NonNegative")
    with pytest.raises(ValueError):
        g2.inverse(8, 0, np.zeros(oracle.gsh_size(8, 8, 0), dtype=complex), real=False)
    with pytest.raises(ValueError):
        OracleGrid(4, 5)


def test_nonnegative_mrange_grid_real():
    """GaussLegendreGrid<double, NonNegative, ...> gives the same real-field
    transforms as the All grid (the iterator skip at :200-202, :305-307)."""
    rng = np.random.default_rng(5)
    L, n = 20, 1
    ga = OracleGrid(L, 2)
    gn = OracleGrid(L, 2, mrange="NonNegative")
    c = random_coeffs(rng, L, L, n, True)
    fa = ga.inverse(L, n, c, real=True)
    fn = gn.inverse(L, n, c, real=True)
    assert np.array_equal(fa, fn)
    assert np.array_equal(ga.forward(L, n, fa, real=True), gn.forward(L, n, fa, real=True))


def test_long_double_grid_agrees():
    rng = np.random.default_rng(9)
    L, n = 48, -1
    g = OracleGrid(L, 2, table=False)
    gl = OracleGrid(L, 2, table=False, long_double=True)
    c = random_coeffs(rng, L, L, n, False)
    f = g.inverse(L, n, c)
    fl = gl.inverse(L, n, c).astype(np.complex128)
    assert np.linalg.norm(f - fl) / np.linalg.norm(fl) < 1e-14
    c2 = g.forward(L, n, f)
    c2l = gl.forward(L, n, fl).astype(np.complex128)
    assert np.linalg.norm(c2 - c2l) / np.linalg.norm(c2l) < 1e-13
