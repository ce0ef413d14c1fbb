"""CPU oracle for the GSH transform hot path -- TEST INFRASTRUCTURE ONLY.

ctypes wrapper around ``oracle/libgsh_oracle.so`` (built from
``oracle/gsh_oracle.cpp`` by ``oracle/Makefile``), a step-for-step CPU
restatement of GSHTrans's ``Wigner::Compute`` (GSHTrans/src/Wigner.h:317-486)
and ``GaussLegendreGrid::{Forward,Inverse}Transformation``
(GSHTrans/src/GaussLegendreGrid.h:120-321).

Only ``tests/``, ``__graft_entry__.smoke()`` and ``bench.py``'s cpu_baseline
leg may import this package, and only as the checker or the timed CPU
baseline.  The product package ``gshtrans_amd`` never imports it.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "libgsh_oracle.so")
_lib = None

MRANGE = {"All": 0, "NonNegative": 1}
NRANGE = {"All": 0, "NonNegative": 1, "Single": 2}


def build(force: bool = False) -> str:
    """Compile the oracle with its Makefile (g++); returns the library path."""
    if force or not os.path.exists(_LIB_PATH):
        subprocess.run(["make", "-C", _HERE, "-s", "-B" if force else "all"],
                       check=True)
    return _LIB_PATH


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        L = ctypes.CDLL(_LIB_PATH)
        dp = ctypes.POINTER(ctypes.c_double)
        ldp = ctypes.POINTER(ctypes.c_longdouble)
        c_long, c_int, vp = ctypes.c_long, ctypes.c_int, ctypes.c_void_p
        L.gsho_wigner_size.restype = c_long
        L.gsho_wigner_size.argtypes = [c_long, c_long, c_long, c_int]
        L.gsho_wigner_block.argtypes = [c_long, c_long, c_long, c_int, c_int,
                                        ctypes.c_double, dp]
        L.gsho_wigner_block_ld.argtypes = [c_long, c_long, c_long, c_int, c_int,
                                           ctypes.c_longdouble, ldp]
        L.gsho_wigner_min_order.restype = ctypes.c_double
        L.gsho_wigner_min_order.argtypes = [c_long, c_long, ctypes.c_double]
        L.gsho_std_sph_legendre.restype = ctypes.c_double
        L.gsho_std_sph_legendre.argtypes = [ctypes.c_uint, ctypes.c_uint,
                                            ctypes.c_double]
        L.gsho_gl_nodes.argtypes = [c_long, dp, dp]
        L.gsho_gl_nodes_x_ld.argtypes = [c_long, ldp, ldp]
        L.gsho_dft.argtypes = [c_long, c_int, dp, dp]
        L.gsho_r2c.argtypes = [c_long, dp, dp]
        L.gsho_c2r.argtypes = [c_long, dp, dp]
        L.gsho_grid_create.restype = vp
        L.gsho_grid_create.argtypes = [c_long, c_long, c_int, c_int, c_int]
        L.gsho_grid_destroy.argtypes = [vp]
        L.gsho_grid_nodes.argtypes = [vp, dp, dp]
        L.gsho_inverse.argtypes = [vp, c_long, c_long, c_int, dp, dp]
        L.gsho_forward.argtypes = [vp, c_long, c_long, c_int, dp, dp]
        L.gsho_grid_create_ld.restype = vp
        L.gsho_grid_create_ld.argtypes = [c_long, c_long, c_int, c_int, c_int]
        L.gsho_grid_destroy_ld.argtypes = [vp]
        L.gsho_inverse_ld.argtypes = [vp, c_long, c_long, c_int, ldp, ldp]
        L.gsho_forward_ld.argtypes = [vp, c_long, c_long, c_int, ldp, ldp]
        L.gsho_num_threads.restype = c_int
        _lib = L
    return _lib


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


def _ldp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_longdouble))


# ---------------------------------------------------------------------------
# Layouts (GSHTrans/src/Indexing.h:104-150)
# ---------------------------------------------------------------------------
def gsh_size(lmax: int, mmax: int, n: int, mrange: str = "All") -> int:
    return int(lib().gsho_wigner_size(lmax, mmax, n, MRANGE[mrange]))


def gsh_index(lmax: int, mmax: int, n: int, l: int, m: int, mrange: str = "All") -> int:
    """Index(l, m) of GSHIndices<MRange>(lMax, mMax, n) (Indexing.h:104-146)."""
    mmax = min(lmax, mmax)
    nabs = abs(n)
    if mrange == "All":
        if mmax >= nabs:
            off = (l * l - nabs * nabs if l <= mmax else
                   (mmax + 1) ** 2 - nabs * nabs + (l - 1 - mmax) * (2 * mmax + 1))
        else:
            off = (l - nabs) * (2 * mmax + 1)
        return off + m + min(l, mmax)
    if mmax >= nabs:
        off = ((l * (l + 1)) // 2 - (nabs * (nabs + 1)) // 2 if l <= mmax else
               ((mmax + 1) * (mmax + 2)) // 2 - (nabs * (nabs + 1)) // 2
               + (l - 1 - mmax) * (mmax + 1))
    else:
        off = (l - nabs) * (mmax + 1)
    return off + m


# ---------------------------------------------------------------------------
# Wigner (Wigner.h:317-486)
# ---------------------------------------------------------------------------
def wigner_block(lmax, mmax, n, theta, mrange="All", norm="Ortho"):
    size = gsh_size(lmax, mmax, n, mrange)
    out = np.zeros(size, dtype=np.float64)
    rc = lib().gsho_wigner_block(lmax, mmax, n, MRANGE[mrange],
                                 0 if norm == "Ortho" else 1, float(theta), _dp(out))
    if rc:
        raise ValueError("invalid Wigner arguments")
    return out


def wigner_block_ld(lmax, mmax, n, theta, mrange="All", norm="Ortho"):
    size = gsh_size(lmax, mmax, n, mrange)
    out = np.zeros(size, dtype=np.longdouble)
    rc = lib().gsho_wigner_block_ld(lmax, mmax, n, MRANGE[mrange],
                                    0 if norm == "Ortho" else 1,
                                    np.longdouble(theta), _ldp(out))
    if rc:
        raise ValueError("invalid Wigner arguments")
    return out


def wigner_min_order(l, n, theta):
    return lib().gsho_wigner_min_order(l, n, float(theta))


def std_sph_legendre(l, m, theta):
    """libstdc++ std::sph_legendre (the routine tests/CheckLegendre.h:30 uses)."""
    return lib().gsho_std_sph_legendre(l, m, float(theta))


def gl_nodes(L):
    th = np.zeros(L + 1)
    w = np.zeros(L + 1)
    lib().gsho_gl_nodes(L, _dp(th), _dp(w))
    return th, w


def gl_x_ld(npts):
    x = np.zeros(npts, dtype=np.longdouble)
    w = np.zeros(npts, dtype=np.longdouble)
    lib().gsho_gl_nodes_x_ld(npts, _ldp(x), _ldp(w))
    return x, w


# ---------------------------------------------------------------------------
# FFT with FFTW semantics
# ---------------------------------------------------------------------------
def dft(x, backward=False):
    x = np.ascontiguousarray(x, dtype=np.complex128)
    out = np.zeros_like(x)
    lib().gsho_dft(x.size, 1 if backward else 0, _dp(x.view(np.float64)),
                   _dp(out.view(np.float64)))
    return out


def r2c(x):
    x = np.ascontiguousarray(x, dtype=np.float64)
    out = np.zeros(x.size // 2 + 1, dtype=np.complex128)
    lib().gsho_r2c(x.size, _dp(x), _dp(out.view(np.float64)))
    return out


def c2r(X, N):
    X = np.ascontiguousarray(X, dtype=np.complex128)
    assert X.size == N // 2 + 1
    out = np.zeros(N)
    lib().gsho_c2r(N, _dp(X.view(np.float64)), _dp(out))
    return out


# ---------------------------------------------------------------------------
# GaussLegendreGrid (GaussLegendreGrid.h:24-333)
# ---------------------------------------------------------------------------
class OracleGrid:
    """CPU restatement of GaussLegendreGrid<double, MRange, NRange>.

    ``table=True`` precomputes the reference's Wigner table with OpenMP
    (Wigner.h:305-314); ``table=False`` recomputes each (n, ring) block with
    identical arithmetic, which needs only O(L^2) memory.
    """

    def __init__(self, L, nmax, mrange="All", nrange="All", table=True,
                 long_double=False):
        self.L, self.nmax = L, nmax
        self.mrange, self.nrange = mrange, nrange
        self.ld = long_double
        create = lib().gsho_grid_create_ld if long_double else lib().gsho_grid_create
        self._h = create(L, nmax, MRANGE[mrange], NRANGE[nrange], 1 if table else 0)
        if not self._h:
            raise ValueError("invalid grid arguments")

    def close(self):
        if self._h:
            (lib().gsho_grid_destroy_ld if self.ld else lib().gsho_grid_destroy)(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def ntheta(self):
        return self.L + 1

    @property
    def nphi(self):
        return 2 * self.L

    def nodes(self):
        th = np.zeros(self.L + 1)
        w = np.zeros(self.L + 1)
        lib().gsho_grid_nodes(self._h, _dp(th), _dp(w))
        return th, w

    def coeff_size(self, lmax, n, real=False):
        return gsh_size(lmax, lmax, n, "NonNegative" if real else "All")

    def inverse(self, lmax, n, coeffs, real=False):
        """Synthesis: coefficients -> field (ntheta x nphi)."""
        if self.ld:
            c = np.ascontiguousarray(coeffs, dtype=np.clongdouble)
            out = np.zeros((self.ntheta, self.nphi),
                           dtype=np.longdouble if real else np.clongdouble)
            rc = lib().gsho_inverse_ld(self._h, lmax, n, 1 if real else 0,
                                       _ldp(c.view(np.longdouble)),
                                       _ldp(out.view(np.longdouble)))
        else:
            c = np.ascontiguousarray(coeffs, dtype=np.complex128)
            out = np.zeros((self.ntheta, self.nphi),
                           dtype=np.float64 if real else np.complex128)
            rc = lib().gsho_inverse(self._h, lmax, n, 1 if real else 0,
                                    _dp(c.view(np.float64)), _dp(out.view(np.float64)))
        if rc:
            raise ValueError("invalid transform arguments")
        return out

    def forward(self, lmax, n, field, real=False):
        """Analysis: field (ntheta x nphi) -> coefficients."""
        size = self.coeff_size(lmax, n, real)
        if self.ld:
            f = np.ascontiguousarray(field, dtype=np.longdouble if real else np.clongdouble)
            out = np.zeros(size, dtype=np.clongdouble)
            rc = lib().gsho_forward_ld(self._h, lmax, n, 1 if real else 0,
                                       _ldp(f.view(np.longdouble)),
                                       _ldp(out.view(np.longdouble)))
        else:
            f = np.ascontiguousarray(field, dtype=np.float64 if real else np.complex128)
            out = np.zeros(size, dtype=np.complex128)
            rc = lib().gsho_forward(self._h, lmax, n, 1 if real else 0,
                                    _dp(f.view(np.float64)), _dp(out.view(np.float64)))
        if rc:
            raise ValueError("invalid transform arguments")
        return out
