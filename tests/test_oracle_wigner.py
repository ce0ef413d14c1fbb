"""Pins the oracle's Wigner restatement (oracle/gsh_oracle.cpp, following
GSHTrans/src/Wigner.h:24-486) against checks that do not retype its formulas:

* the reference's own tests, re-expressed with seeded inputs:
  tests/CheckLegendre.h:11-44 (n = 0 against std::sph_legendre) and
  tests/CheckAdditionTheorem.h:13-48 (sum_m d_mn d_mn' = delta_nn');
* mpmath values of Wigner's explicit sum up to l = 1024
  (tests/golden/wigner_mpmath.json, written by make_wigner_golden.py);
* symmetries of d^l_{mn} and a long double run of the same restatement.
A sign, index or transposition slip in the recurrence breaks one of these.
"""
import json
import math
import os

import numpy as np
import pytest

import oracle
from oracle import gsh_index, wigner_block, wigner_block_ld

EPS = np.finfo(np.float64).eps


def amp(l):
    return math.sqrt(2 * l + 1) / (2 * math.sqrt(math.pi))


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_check_legendre(seed):
    """tests/CheckLegendre.h: L=300, n=0, Ortho vs std::sph_legendre."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0, np.pi)
    L = 300
    d = wigner_block(L, L, 0, theta)
    tiny = 1000 * np.finfo(np.float64).tiny
    worst = 0.0
    for l in range(L + 1):
        for m in range(0, l + 1):
            ref = oracle.std_sph_legendre(l, m, theta)
            if abs(ref) > tiny:
                worst = max(worst, abs(d[gsh_index(L, L, 0, l, m)] - ref))
    assert worst < 1e5 * EPS


def test_legendre_scipy_closed_form():
    """D^l_{m0} = Y_l^m(theta, 0) (scipy, Condon-Shortley phase); also m < 0."""
    from scipy.special import sph_harm_y
    theta = 0.7314
    L = 120
    d = wigner_block(L, L, 0, theta)
    for l in range(0, L + 1, 7):
        for m in range(-l, l + 1, 3):
            ref = float(np.real(sph_harm_y(l, m, theta, 0.0)))
            assert abs(d[gsh_index(L, L, 0, l, m)] - ref) < 1e-12


@pytest.mark.parametrize("seed", [3, 4])
def test_addition_theorem(seed):
    """tests/CheckAdditionTheorem.h: FourPi, L=40, all n, n'."""
    rng = np.random.default_rng(seed)
    theta = rng.uniform(0, np.pi)
    L = 40
    blocks = {n: wigner_block(L, L, n, theta, norm="FourPi") for n in range(-L, L + 1)}
    for n in range(-L, L + 1):
        for npr in range(-L, L + 1):
            for l in range(max(abs(n), abs(npr)), L + 1):
                a = blocks[n][gsh_index(L, L, n, l, -l):gsh_index(L, L, n, l, l) + 1]
                b = blocks[npr][gsh_index(L, L, npr, l, -l):gsh_index(L, L, npr, l, l) + 1]
                s = float(np.dot(a, b)) - (1.0 if n == npr else 0.0)
                assert abs(s) < 1e3 * EPS


def test_mpmath_golden(golden_dir):
    with open(os.path.join(golden_dir, "wigner_mpmath.json")) as fh:
        samples = json.load(fh)["samples"]
    assert len(samples) > 100
    for s in samples:
        l, m, n, th = s["l"], s["m"], s["n"], s["theta"]
        idx = gsh_index(l, l, n, l, m)
        v = wigner_block(l, l, n, th)[idx]
        vl = float(wigner_block_ld(l, l, n, th)[idx])
        e = s["D_ortho"]
        # The double recurrence carries ~l*eps of the Ortho amplitude.
        assert abs(v - e) <= 2e-12 * amp(l), (l, m, n, th, v, e)
        assert abs(vl - e) <= 2e-14 * amp(l), (l, m, n, th, vl, e)
        if abs(e) > 0.1 * amp(l):
            assert abs(v - e) <= 1e-10 * abs(e)
        # FourPi differs from Ortho only by the scale.
    s = samples[5]
    raw = wigner_block(s["l"], s["l"], s["n"], s["theta"], norm="FourPi")
    assert abs(raw[gsh_index(s["l"], s["l"], s["n"], s["l"], s["m"])] - s["d"]) < 1e-12


def test_seed_closed_form():
    """WignerMinOrder = sqrt(C(2l, l+n)) sin^{l+n}(t/2) cos^{l-n}(t/2)."""
    for (l, n, th) in [(0, 0, 0.3), (5, 2, 1.0), (40, -7, 2.5), (300, 3, 0.9)]:
        ref = math.sqrt(math.comb(2 * l, l + n)) * math.sin(th / 2) ** (l + n) * \
            math.cos(th / 2) ** (l - n)
        assert abs(oracle.wigner_min_order(l, n, th) - ref) <= 1e-12 * max(abs(ref), 1e-300)
    # Pole flags (Wigner.h:33-36, 61-69).
    assert oracle.wigner_min_order(3, -3, 0.0) == 1.0
    assert oracle.wigner_min_order(3, 2, 0.0) == 0.0
    assert oracle.wigner_min_order(3, 3, math.pi) == 1.0


@pytest.mark.parametrize("L,n,theta", [(30, 0, 0.4), (30, 3, 1.3), (25, -2, 2.9), (64, 4, 0.05)])
def test_symmetries(L, n, theta):
    """d_mn = (-1)^(m-n) d_{-m,-n} = (-1)^(m-n) d_nm; d_mn(pi-t) = (-1)^(l+m) d_{m,-n}(t)."""
    dn = wigner_block(L, L, n, theta, norm="FourPi")
    dmn = wigner_block(L, L, -n, theta, norm="FourPi")
    dflip = wigner_block(L, L, -n, math.pi - theta, norm="FourPi")
    for l in range(abs(n), L + 1):
        for m in range(-l, l + 1):
            v = dn[gsh_index(L, L, n, l, m)]
            sgn = -1.0 if (m - n) % 2 else 1.0
            assert abs(v - sgn * dmn[gsh_index(L, L, -n, l, -m)]) < 1e-13
            sgn2 = -1.0 if (l + m) % 2 else 1.0
            assert abs(v - sgn2 * dflip[gsh_index(L, L, -n, l, m)]) < 1e-12
    # d_mn = (-1)^(m-n) d_nm for a few m
    for m in (-3, 0, 2):
        if abs(m) > L:
            continue
        dm = wigner_block(L, L, m, theta, norm="FourPi")
        for l in range(max(abs(n), abs(m)), L + 1):
            sgn = -1.0 if (m - n) % 2 else 1.0
            assert abs(dn[gsh_index(L, L, n, l, m)] - sgn * dm[gsh_index(L, L, m, l, n)]) < 1e-13


def test_layout_variants_consistent():
    """NonNegative and mMax < lMax blocks equal the slices of the full block
    (Indexing.h:104-131; Wigner.h:398-471 boundary branches)."""
    L, n, th = 50, 2, 1.1
    full = wigner_block(L, L, n, th)
    nonneg = wigner_block(L, L, n, th, mrange="NonNegative")
    for l in range(abs(n), L + 1):
        for m in range(0, l + 1):
            assert nonneg[gsh_index(L, L, n, l, m, "NonNegative")] == \
                full[gsh_index(L, L, n, l, m)]
    for mmax in (0, 1, 5, 17):
        part = wigner_block(L, mmax, n, th)
        partnn = wigner_block(L, mmax, n, th, mrange="NonNegative")
        for l in range(abs(n), L + 1):
            for m in range(-min(l, mmax), min(l, mmax) + 1):
                assert part[gsh_index(L, mmax, n, l, m)] == full[gsh_index(L, L, n, l, m)]
                if m >= 0:
                    assert partnn[gsh_index(L, mmax, n, l, m, "NonNegative")] == \
                        full[gsh_index(L, L, n, l, m)]


def test_long_double_agrees():
    for (L, n, th) in [(200, 0, 0.77), (300, -2, 2.2), (100, 1, 0.01)]:
        a = wigner_block(L, L, n, th)
        b = wigner_block_ld(L, L, n, th).astype(np.float64)
        assert np.max(np.abs(a - b)) < 1e-12 * amp(L)


def test_index_layout_matches_library():
    for (lmax, mmax, n, mr) in [(10, 10, 0, "All"), (10, 4, -3, "All"), (9, 9, 2, "NonNegative"),
                                (9, 2, 4, "NonNegative"), (6, 6, 6, "All")]:
        size = oracle.gsh_size(lmax, mmax, n, mr)
        seen = []
        for l in range(abs(n), lmax + 1):
            mm = min(l, mmax)
            for m in (range(-mm, mm + 1) if mr == "All" else range(0, mm + 1)):
                seen.append(gsh_index(lmax, mmax, n, l, m, mr))
        assert seen == list(range(size))
