"""gshtrans_amd -- MI355X (gfx950) GSH transform hot path.

Python host layer over the C ABI in ``include/gshtrans_amd.h``
(``gshtrans_amd/libgshtrans_hip.so``).  It mirrors the reference's
``GaussLegendreGrid`` interface for the transform path
(GSHTrans/src/GaussLegendreGrid.h:24-333, GridBase.h:12-180) with the same
method names and argument meaning; device arrays are torch tensors, used only
for memory and streams.  There is no CPU fallback: every transform runs the
HIP kernels and fails loudly when the library or a gfx950 device is missing.
"""
from __future__ import annotations

import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libgshtrans_hip.so")

GSH_OK = 0
GSH_ERR_SCRATCH_LIMIT = 9
STATUS = {
    0: "GSH_OK", 1: "GSH_ERR_INVALID_ARGUMENT", 2: "GSH_ERR_UPPER_INDEX",
    3: "GSH_ERR_DEGREE", 4: "GSH_ERR_VALUE_KIND", 5: "GSH_ERR_UNSUPPORTED",
    6: "GSH_ERR_DEVICE", 7: "GSH_ERR_NO_DEVICE", 8: "GSH_ERR_OUT_OF_MEMORY",
    9: "GSH_ERR_SCRATCH_LIMIT",
}
RANGE = {"All": 0, "NonNegative": 1, "Single": 2}
# Pointwise field expressions of gsh_forward_expr (gsh_field_expr).
EXPR = {"product": 1, "sum": 2, "difference": 3, "quotient": 4, "negate": 5, "conj": 6,
        "real": 7, "imag": 8, "scale": 9, "shift": 10}
PRECISION = {"fp64": 0, "fp32": 1, "bf16x3": 2, "bf16x6": 3}
STORAGE = {"fp64": 0, "fp32": 1}
REAL, COMPLEX = 0, 1

# Every symbol declared in include/gshtrans_amd.h.
EXPORTS = [
    "gsh_status_string", "gsh_version", "gsh_gl_nodes", "gsh_coefficient_size",
    "gsh_wigner_seed", "gsh_grid_create", "gsh_grid_destroy", "gsh_grid_info",
    "gsh_grid_colatitudes", "gsh_grid_prepare", "gsh_inverse", "gsh_forward",
    "gsh_adjoint_inverse", "gsh_inverse_host", "gsh_forward_host",
    "gsh_timing_enable", "gsh_timing_read", "gsh_plan_create", "gsh_plan_destroy",
    "gsh_plan_inverse", "gsh_plan_forward", "gsh_plan_info", "gsh_device_alloc",
    "gsh_device_free", "gsh_copy", "gsh_synchronize", "gsh_wigner_size", "gsh_wigner_offset",
    "gsh_wigner_table", "gsh_wigner_table_host", "gsh_forward_product", "gsh_integrate",
    "gsh_random_coefficients", "gsh_integrate_noalloc", "gsh_forward_expr",
    "gsh_plan_analysis_pass",
]
KERNELS = ["legendre_synthesis", "fft_synthesis", "fft_analysis", "legendre_analysis",
           "coef_split", "bin_split"]


class GSHError(RuntimeError):
    def __init__(self, status, what=""):
        self.status = status
        msg = _lib().gsh_status_string(status).decode() if _LIB is not None else STATUS.get(status)
        super().__init__(f"{what}: {STATUS.get(status, status)} ({msg})")


_LIB = None


def _lib():
    global _LIB
    if _LIB is None:
        if not os.path.exists(LIB_PATH):
            raise ImportError(
                f"{LIB_PATH} is missing: build it with `python -m gshtrans_amd.build`")
        # torch's wheel carries its own HIP runtime (SONAME libamdhip64.so.7,
        # requested by name as libamdhip64.so).  Loading torch first lets this
        # library bind to that runtime by SONAME, so the process holds one
        # HIP runtime and torch's streams and allocations are this library's.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        i, l, d, vp = ctypes.c_int, ctypes.c_long, ctypes.c_double, ctypes.c_void_p
        dp = ctypes.POINTER(ctypes.c_double)
        ip = ctypes.POINTER(ctypes.c_int)
        L.gsh_status_string.restype = ctypes.c_char_p
        L.gsh_status_string.argtypes = [i]
        L.gsh_version.restype = i
        L.gsh_gl_nodes.argtypes = [i, dp, dp]
        L.gsh_coefficient_size.restype = l
        L.gsh_coefficient_size.argtypes = [i, i, i]
        L.gsh_wigner_seed.argtypes = [i, i, d, dp]
        L.gsh_grid_create.argtypes = [i, i, i, i, ctypes.POINTER(vp)]
        L.gsh_grid_destroy.argtypes = [vp]
        L.gsh_grid_info.argtypes = [vp, ip, ip, ip, ip, ip]
        L.gsh_grid_colatitudes.argtypes = [vp, dp, dp]
        L.gsh_grid_prepare.argtypes = [vp, i]
        L.gsh_inverse.argtypes = [vp, i, i, i, i, i, i, vp, l, vp, l, vp]
        L.gsh_forward.argtypes = [vp, i, i, i, i, i, i, vp, l, vp, l, i, vp]
        L.gsh_adjoint_inverse.argtypes = [vp, i, i, i, i, i, i, vp, l, vp, l, vp]
        L.gsh_inverse_host.argtypes = [vp, i, i, i, dp, dp]
        L.gsh_forward_host.argtypes = [vp, i, i, i, dp, dp, i]
        L.gsh_timing_enable.argtypes = [vp, i]
        L.gsh_timing_read.argtypes = [vp, dp, ctypes.POINTER(l), i]
        L.gsh_forward_product.argtypes = [vp, i, i, i, i, i, vp, vp, l, vp, l, i, vp]
        L.gsh_forward_expr.argtypes = [vp, i, i, i, i, i, i, vp, vp, l, d, d, vp, l, i, vp]
        L.gsh_integrate.argtypes = [vp, i, i, vp, vp, l, vp, vp]
        L.gsh_integrate_noalloc.argtypes = [vp, i, i, vp, vp, l, vp, vp]
        L.gsh_random_coefficients.argtypes = [vp, i, i, i, ctypes.c_ulonglong, d, i, vp, l, vp]
        L.gsh_wigner_size.restype = l
        L.gsh_wigner_size.argtypes = [i, i, i, i, i, i]
        L.gsh_wigner_offset.restype = l
        L.gsh_wigner_offset.argtypes = [i, i, i, i, i, i, i, i, i]
        L.gsh_wigner_table.argtypes = [i, i, i, i, i, dp, i, i, i, vp, vp]
        L.gsh_wigner_table_host.argtypes = [i, i, i, i, i, dp, i, i, i, dp]
        for name in ("gsh_gl_nodes", "gsh_wigner_seed", "gsh_grid_create", "gsh_grid_destroy",
                     "gsh_grid_info", "gsh_grid_colatitudes", "gsh_grid_prepare",
                     "gsh_inverse", "gsh_forward", "gsh_adjoint_inverse",
                     "gsh_inverse_host", "gsh_forward_host", "gsh_timing_enable",
                     "gsh_timing_read", "gsh_wigner_table", "gsh_wigner_table_host",
                     "gsh_forward_product", "gsh_forward_expr", "gsh_integrate",
                     "gsh_integrate_noalloc",
                     "gsh_random_coefficients"):
            getattr(L, name).restype = i
        _LIB = L
    return _LIB


def library():
    """The loaded ctypes library (raises if the HIP library is not built)."""
    return _lib()


def _check(status, what):
    if status != GSH_OK:
        raise GSHError(status, what)


def _dp(a):
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


# ---------------------------------------------------------------------------
# Host-side helpers (no device)
# ---------------------------------------------------------------------------
def gl_nodes(lmax):
    """Colatitudes theta_i = acos(-x_i) (ascending) and Gauss-Legendre weights
    of the (lmax+1)-point grid (GaussLegendreGrid.h:47-49)."""
    th = np.zeros(lmax + 1)
    w = np.zeros(lmax + 1)
    _check(_lib().gsh_gl_nodes(lmax, _dp(th), _dp(w)), "gsh_gl_nodes")
    return th, w


def coefficient_size(lmax, n, real=False):
    """GSHIndices<All|NonNegative>(lmax, lmax, n).Size() (Indexing.h:148-150)."""
    s = _lib().gsh_coefficient_size(lmax, n, REAL if real else COMPLEX)
    if s < 0:
        raise ValueError("invalid (lmax, n)")
    return int(s)


def coefficient_index(lmax, n, l, m, real=False):
    """GSHIndices<All|NonNegative>(lmax, lmax, n).Index(l, m) (Indexing.h:143-146)."""
    na = abs(n)
    if real:
        return l * (l + 1) // 2 - na * (na + 1) // 2 + m
    return l * l - na * na + m + l


def wigner_seed(l, n, theta):
    """WignerMinOrder(l, n, theta) (Wigner.h:52-80)."""
    v = ctypes.c_double()
    _check(_lib().gsh_wigner_seed(l, n, float(theta), ctypes.byref(v)), "gsh_wigner_seed")
    return v.value


def random_coefficients(lmax, n, real=False, seed=0, decay=1.0, grid_lmax=None,
                        batch=None):
    """Seeded synthetic coefficients with the reference generator's masks
    (GridBase.h:115-172): complex normal c_lm scaled by (l+1)^-decay;
    (L, L) zeroed for complex fields when lmax equals the grid's L; Im = 0 at
    m = 0 and at (L, L) for real fields.  Counter-based splitmix64 ->
    Box-Muller, so every host produces the same values for a seed."""
    if grid_lmax is None:
        grid_lmax = lmax
    size = coefficient_size(lmax, n, real)
    shape = (size,) if batch is None else (batch, size)
    count = int(np.prod(shape))
    golden = np.uint64(0x9E3779B97F4A7C15)

    def mix(z):
        z = (z ^ (z >> np.uint64(30))) * np.uint64(0xBF58476D1CE4E5B9)
        z = (z ^ (z >> np.uint64(27))) * np.uint64(0x94D049BB133111EB)
        return z ^ (z >> np.uint64(31))

    with np.errstate(over="ignore"):
        base = mix(np.uint64(seed) + golden)
        z = mix(base + (np.arange(2 * count, dtype=np.uint64) + np.uint64(1)) * golden)
    u = (z >> np.uint64(11)).astype(np.float64) * (1.0 / 9007199254740992.0)
    u1, u2 = u[0::2], u[1::2]
    r = np.sqrt(-2.0 * np.log1p(-u1))
    c = (r * np.cos(2 * np.pi * u2) + 1j * r * np.sin(2 * np.pi * u2)).reshape(shape)
    na = abs(n)
    ls = np.zeros(size, dtype=np.int64)
    ms = np.zeros(size, dtype=np.int64)
    k = 0
    for l in range(na, lmax + 1):
        cnt = l + 1 if real else 2 * l + 1
        ls[k:k + cnt] = l
        ms[k:k + cnt] = np.arange(0, l + 1) if real else np.arange(-l, l + 1)
        k += cnt
    c = c * (ls + 1.0) ** (-decay)
    if real:
        c[..., ms == 0] = c[..., ms == 0].real
        if lmax == grid_lmax:
            c[..., -1] = c[..., -1].real
    elif lmax == grid_lmax:
        c[..., -1] = 0
    return c


# ---------------------------------------------------------------------------
# Grid
# ---------------------------------------------------------------------------
def _ptr(t):
    import torch
    if isinstance(t, torch.Tensor):
        if not t.is_cuda:
            raise ValueError("device tensor expected")
        if not t.is_contiguous():
            raise ValueError("contiguous tensor expected")
        return ctypes.c_void_p(t.data_ptr())
    raise TypeError("torch tensor expected")


def _check_coeffs(t, rows, size, dtype, what):
    """Coefficient arrays: a (rows, stride) cuda tensor of `dtype` with stride
    >= size, since the kernels read or write `size` complex entries per row."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what}: coefficients must be a torch tensor")
    if t.dtype != dtype:
        raise ValueError(f"{what}: coefficients must be {dtype}, got {t.dtype}")
    if t.dim() != 2 or t.shape[0] != rows or t.shape[1] < size:
        raise ValueError(f"{what}: coefficients must be ({rows}, >= {size}), got "
                         f"{tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{what}: coefficients must be contiguous")
    if not t.is_cuda:
        raise ValueError(f"{what}: coefficients must be a cuda tensor")


def _check_field(t, rows, ntheta, nphi, dtype, what):
    """Field arrays: a contiguous (rows, ntheta, nphi) cuda tensor of `dtype`."""
    import torch
    if not isinstance(t, torch.Tensor):
        raise ValueError(f"{what}: field must be a torch tensor")
    if t.dtype != dtype:
        raise ValueError(f"{what}: field must be {dtype}, got {t.dtype}")
    if tuple(t.shape) != (rows, ntheta, nphi):
        raise ValueError(f"{what}: field must be ({rows}, {ntheta}, {nphi}), got "
                         f"{tuple(t.shape)}")
    if not t.is_contiguous():
        raise ValueError(f"{what}: field must be contiguous")
    if not t.is_cuda:
        raise ValueError(f"{what}: field must be a cuda tensor")


def _stream(stream):
    import torch
    s = stream if stream is not None else torch.cuda.current_stream()
    return ctypes.c_void_p(s.cuda_stream)


WIGNER_NORM = {"Ortho": 0, "FourPi": 1}
WIGNER_STORAGE = {"ColumnMajor": 0, "RowMajor": 1}
_MRANGE = {"All": 0, "NonNegative": 1}


class Wigner:
    """Wigner d tables, Wigner<double, Norm, MRange, NRange, AngleRange,
    Storage>(lMax, mMax, nMax, theta) (Wigner.h:99-286), computed on the GPU.

    `data` is one float64 device tensor in the reference's layout: a block of
    GSHIndices<MRange>(lmax, mmax, n) (Indexing.h:72-150) per (n, theta_i),
    n over UpperIndices() (NRange All: -nmax..nmax, NonNegative: 0..nmax,
    Single: nmax), blocks ColumnMajor (n outer) or RowMajor (theta outer).
    `self[n, i]` is the block of (n, theta_i) and `self(n, i, l, m)` one
    value d^l_{mn}(theta_i) (Ortho: times sqrt(2l+1) / (2 sqrt(pi))).
    """

    def __init__(self, lmax, mmax, nmax, theta, norm="Ortho", mrange="All", nrange="Single",
                 storage="ColumnMajor", stream=None):
        import torch
        theta = np.ascontiguousarray(np.atleast_1d(np.asarray(theta, dtype=np.float64)))
        self.lmax, self.mmax, self.nmax = int(lmax), int(mmax), int(nmax)
        self.norm, self.mrange, self.nrange, self.storage = norm, mrange, nrange, storage
        self.nmin = {"All": -self.nmax, "NonNegative": 0, "Single": self.nmax}[nrange]
        self.theta = theta
        self._args = (self.lmax, self.mmax, self.nmin, self.nmax, len(theta), _MRANGE[mrange])
        size = _lib().gsh_wigner_size(*self._args)
        if size < 0:
            raise GSHError(1, "Wigner")
        self.data = torch.empty(size, dtype=torch.float64, device="cuda")
        self.recompute(theta, stream)

    def recompute(self, theta, stream=None):
        """Wigner::ReCompute (Wigner.h:142-152): the same table for new angles."""
        theta = np.ascontiguousarray(np.atleast_1d(np.asarray(theta, dtype=np.float64)))
        assert len(theta) == len(self.theta)
        self.theta = theta
        _check(_lib().gsh_wigner_table(*self._args[:5], _dp(theta), self._args[5],
                                       WIGNER_NORM[self.norm], WIGNER_STORAGE[self.storage],
                                       _ptr(self.data), _stream(stream)), "gsh_wigner_table")

    def upper_indices(self):
        return range(self.nmin, self.nmax + 1)

    def offset(self, n, itheta):
        o = _lib().gsh_wigner_offset(*self._args, WIGNER_STORAGE[self.storage], n, itheta)
        if o < 0:
            raise GSHError(1, "Wigner.offset")
        return o

    def index(self, n, l, m):
        """Index(l, m) of GSHIndices<MRange>(lmax, mmax, n) (Indexing.h:54-65, 104-146)."""
        mm = min(self.lmax, self.mmax)
        nabs = abs(n)
        assert nabs <= l <= self.lmax and abs(m) <= min(l, mm)
        if self.mrange == "All":
            if mm >= nabs:
                off = (l * l - nabs * nabs if l <= mm else
                       (mm + 1) ** 2 - nabs * nabs + (l - 1 - mm) * (2 * mm + 1))
            else:
                off = (l - nabs) * (2 * mm + 1)
            return off + m + min(l, mm)
        assert m >= 0
        if mm >= nabs:
            off = (l * (l + 1) // 2 - nabs * (nabs + 1) // 2 if l <= mm else
                   (mm + 1) * (mm + 2) // 2 - nabs * (nabs + 1) // 2 + (l - 1 - mm) * (mm + 1))
        else:
            off = (l - nabs) * (mm + 1)
        return off + m

    def block_size(self, n):
        mm = min(self.lmax, self.mmax)
        return self.index(n, self.lmax, mm) + 1

    def __getitem__(self, key):
        n, itheta = key
        o = self.offset(n, itheta)
        return self.data[o:o + self.block_size(n)]

    def __call__(self, n, itheta, l, m):
        return float(self.data[self.offset(n, itheta) + self.index(n, l, m)])


class GaussLegendreGrid:
    """GaussLegendreGrid<double, MRange, NRange>(lMax, nMax) on one MI355X.

    Mirrors GSHTrans/src/GaussLegendreGrid.h:24-333: colatitudes are the
    Gauss-Legendre points theta = acos(-x) ascending, longitudes are the 2L
    points j*pi/L, fields are row-major (theta, phi).
    """

    def __init__(self, lmax, nmax, mrange="All", nrange="All"):
        self._h = ctypes.c_void_p()
        _check(_lib().gsh_grid_create(lmax, nmax, RANGE[mrange], RANGE[nrange],
                                      ctypes.byref(self._h)), "gsh_grid_create")
        self._lmax, self._nmax = lmax, nmax
        self.mrange, self.nrange = mrange, nrange
        th = np.zeros(lmax + 1)
        w = np.zeros(lmax + 1)
        _check(_lib().gsh_grid_colatitudes(self._h, _dp(th), _dp(w)), "gsh_grid_colatitudes")
        self._theta, self._w = th, w

    def close(self):
        if self._h:
            _lib().gsh_grid_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    @property
    def handle(self):
        return self._h

    # GaussLegendreGrid.h:86-103 / GridBase.h:17-77
    def MaxDegree(self):
        return self._lmax

    def MaxUpperIndex(self):
        return self._nmax

    def MinUpperIndex(self):
        return {"All": -self._nmax, "NonNegative": 0, "Single": self._nmax}[self.nrange]

    def UpperIndices(self):
        return range(self.MinUpperIndex(), self._nmax + 1)

    def CoLatitudes(self):
        return self._theta.copy()

    def CoLatitudeWeights(self):
        return self._w.copy()

    def Longitudes(self):
        L = self._lmax
        return np.arange(2 * L) * (2 * math.pi / (2 * L)) if L > 0 else np.zeros(0)

    def LongitudeWeights(self):
        L = self._lmax
        return np.full(2 * L, 2 * math.pi / (2 * L)) if L > 0 else np.zeros(0)

    def NumberOfCoLatitudes(self):
        return self._lmax + 1

    def NumberOfLongitudes(self):
        return 2 * self._lmax

    def FieldSize(self):
        return self.NumberOfCoLatitudes() * self.NumberOfLongitudes()

    def ComplexCoefficientSize(self, lmax, n=None):
        if n is None:
            lmax, n = self._lmax, lmax
        return coefficient_size(lmax, n, False)

    def RealCoefficientSize(self, lmax, n=None):
        if n is None:
            lmax, n = self._lmax, lmax
        return coefficient_size(lmax, n, True)

    def prepare(self, n):
        _check(_lib().gsh_grid_prepare(self._h, n), "gsh_grid_prepare")

    def timing_enable(self, enable=True):
        """Record HIP events around every kernel launch (resets the totals)."""
        _check(_lib().gsh_timing_enable(self._h, 1 if enable else 0), "gsh_timing_enable")

    def timing_read(self):
        """{kernel: (total_ms, launches)} since timing_enable()."""
        ms = np.zeros(len(KERNELS))
        cnt = (ctypes.c_long * len(KERNELS))()
        _check(_lib().gsh_timing_read(self._h, _dp(ms), cnt, len(KERNELS)), "gsh_timing_read")
        return {k: (float(ms[i]), int(cnt[i])) for i, k in enumerate(KERNELS)}

    # ---- batched device transforms --------------------------------------
    def inverse(self, lmax, n, coeffs, field=None, real=False, precision="fp64",
                stream=None):
        """Synthesis of a batch: coeffs (B, C) complex128 cuda tensor ->
        field (B, ntheta, nphi) float64 (real) or complex128."""
        import torch
        c = coeffs if coeffs.dim() == 2 else coeffs.unsqueeze(0)
        B, C = c.shape
        expect = _lib().gsh_coefficient_size(lmax, n, REAL if real else COMPLEX)
        if expect >= 0 and C != expect:
            raise ValueError("coefficient size mismatch")
        _check_coeffs(c, B, C, torch.complex128, "inverse")
        fdt = torch.float64 if real else torch.complex128
        if field is None:
            field = torch.empty((B, self._lmax + 1, 2 * self._lmax), dtype=fdt, device=c.device)
        _check_field(field if field.dim() == 3 else field.unsqueeze(0), B, self._lmax + 1,
                     2 * self._lmax, fdt, "inverse")
        _check(_lib().gsh_inverse(self._h, lmax, n, REAL if real else COMPLEX,
                                  PRECISION[precision], STORAGE["fp64"], B, _ptr(c), C,
                                  _ptr(field), self.FieldSize(), _stream(stream)),
               "gsh_inverse")
        return field if coeffs.dim() == 2 else field[0]

    def forward(self, lmax, n, field, coeffs=None, real=False, precision="fp64",
                accumulate=False, stream=None):
        """Analysis of a batch: field (B, ntheta, nphi) -> coeffs (B, C)."""
        import torch
        f = field if field.dim() == 3 else field.unsqueeze(0)
        B = f.shape[0]
        _check_field(f, B, self._lmax + 1, 2 * self._lmax,
                     torch.float64 if real else torch.complex128, "forward")
        C = max(int(_lib().gsh_coefficient_size(lmax, n, REAL if real else COMPLEX)), 1)
        if coeffs is None:
            coeffs = torch.zeros((B, C), dtype=torch.complex128, device=f.device)
        co = coeffs if coeffs.dim() == 2 else coeffs.unsqueeze(0)
        _check_coeffs(co, B, C, torch.complex128, "forward")
        _check(_lib().gsh_forward(self._h, lmax, n, REAL if real else COMPLEX,
                                  PRECISION[precision], STORAGE["fp64"], B, _ptr(f),
                                  self.FieldSize(), _ptr(co), co.shape[1],
                                  1 if accumulate else 0, _stream(stream)), "gsh_forward")
        return coeffs if field.dim() == 3 else coeffs[0]

    def forward_product(self, lmax, n, field_a, field_b, coeffs=None, real=False,
                        precision="fp64", accumulate=False, stream=None):
        """Analysis of the pointwise product field_a * field_b (upper index n =
        N0 + N1 of the two factors; CanonicalComponentFieldBinary.h:203-253),
        the product formed inside the ring FFT's loads: no product field is
        written to memory."""
        import torch
        fa = field_a if field_a.dim() == 3 else field_a.unsqueeze(0)
        fb = field_b if field_b.dim() == 3 else field_b.unsqueeze(0)
        B = fa.shape[0]
        fdt = torch.float64 if real else torch.complex128
        _check_field(fa, B, self._lmax + 1, 2 * self._lmax, fdt, "forward_product")
        _check_field(fb, B, self._lmax + 1, 2 * self._lmax, fdt, "forward_product")
        C = max(int(_lib().gsh_coefficient_size(lmax, n, REAL if real else COMPLEX)), 1)
        if coeffs is None:
            coeffs = torch.zeros((B, C), dtype=torch.complex128, device=fa.device)
        co = coeffs if coeffs.dim() == 2 else coeffs.unsqueeze(0)
        _check_coeffs(co, B, C, torch.complex128, "forward_product")
        _check(_lib().gsh_forward_product(self._h, lmax, n, REAL if real else COMPLEX,
                                          PRECISION[precision], B, _ptr(fa), _ptr(fb),
                                          self.FieldSize(), _ptr(co), co.shape[1],
                                          1 if accumulate else 0, _stream(stream)),
               "gsh_forward_product")
        return coeffs if field_a.dim() == 3 else coeffs[0]

    def forward_expr(self, lmax, n, op, field_a, field_b=None, scalar=None, coeffs=None,
                     real=False, precision="fp64", accumulate=False, stream=None):
        """Analysis of a pointwise expression of fields (the expression
        fields of CanonicalComponentFieldOverloads.h:20-266), formed inside
        the ring FFT's loads: op is one of EXPR ("product", "sum",
        "difference", "quotient", "negate", "conj", "real", "imag", "scale",
        "shift"); field_b is the second operand of the first four, scalar the
        operand of "scale" and "shift".  `real` is the kind of the input
        fields; "real" / "imag" of complex fields give real fields (NonNegative
        coefficients).  n is the expression's upper index."""
        import torch
        code = EXPR[op]
        binary = code <= EXPR["quotient"]
        fa = field_a if field_a.dim() == 3 else field_a.unsqueeze(0)
        B = fa.shape[0]
        fdt = torch.float64 if real else torch.complex128
        _check_field(fa, B, self._lmax + 1, 2 * self._lmax, fdt, "forward_expr")
        fb = None
        if binary:
            if field_b is None:
                raise ValueError(f"forward_expr: {op} needs field_b")
            fb = field_b if field_b.dim() == 3 else field_b.unsqueeze(0)
            _check_field(fb, B, self._lmax + 1, 2 * self._lmax, fdt, "forward_expr")
        s = complex(0.0 if scalar is None else scalar)
        if real and s.imag != 0:
            raise ValueError("forward_expr: the scalar of a real field must be real")
        out_real = real or op in ("real", "imag")
        C = max(int(_lib().gsh_coefficient_size(lmax, n, REAL if out_real else COMPLEX)), 1)
        if coeffs is None:
            coeffs = torch.zeros((B, C), dtype=torch.complex128, device=fa.device)
        co = coeffs if coeffs.dim() == 2 else coeffs.unsqueeze(0)
        _check_coeffs(co, B, C, torch.complex128, "forward_expr")
        _check(_lib().gsh_forward_expr(self._h, lmax, n, REAL if real else COMPLEX, code,
                                       PRECISION[precision], B, _ptr(fa),
                                       None if fb is None else _ptr(fb), self.FieldSize(),
                                       s.real, s.imag, _ptr(co), co.shape[1],
                                       1 if accumulate else 0, _stream(stream)),
               "gsh_forward_expr")
        return coeffs if field_a.dim() == 3 else coeffs[0]

    def weights(self, device="cuda"):
        """GridBase::Weights (GridBase.h:60-66) as a (ntheta, nphi) tensor:
        w_i dphi at every point."""
        import torch
        w = torch.from_numpy(self.CoLatitudeWeights() * (2 * np.pi / self.NumberOfLongitudes()))
        return w.to(device)[:, None].expand(self.NumberOfCoLatitudes(),
                                            self.NumberOfLongitudes()).contiguous()

    def project(self, f, device="cuda", dtype=None):
        """GridBase::ProjectFunction (GridBase.h:68-73): f(theta, phi) at every
        point, for f acting elementwise on (ntheta, nphi) tensors."""
        import torch
        th = torch.from_numpy(np.asarray(self.CoLatitudes())).to(device)
        ph = torch.arange(self.NumberOfLongitudes(), dtype=torch.float64,
                          device=device) * (2 * np.pi / self.NumberOfLongitudes())
        out = f(th[:, None].expand(-1, ph.numel()), ph[None, :].expand(th.numel(), -1))
        return out.to(dtype) if dtype is not None else out

    def integrate(self, field, other=None, stream=None):
        """Integrate(u) (ScalarFieldOverloads.h:288-301) of a field or batch,
        or with `other` the L2 inner product Integrate(conj(u) v) (:303-316),
        on the device; returns complex128 (real fields: imaginary part 0)."""
        import torch
        fa = field if field.dim() == 3 else field.unsqueeze(0)
        fb = None if other is None else (other if other.dim() == 3 else other.unsqueeze(0))
        real = not fa.is_complex()
        fdt = torch.float64 if real else torch.complex128
        _check_field(fa, fa.shape[0], self._lmax + 1, 2 * self._lmax, fdt, "integrate")
        if fb is not None:
            _check_field(fb, fa.shape[0], self._lmax + 1, 2 * self._lmax, fdt, "integrate")
        out = torch.zeros(fa.shape[0], dtype=torch.complex128, device=fa.device)
        _check(_lib().gsh_integrate(self._h, REAL if real else COMPLEX, fa.shape[0], _ptr(fa),
                                    None if fb is None else _ptr(fb), self.FieldSize(),
                                    _ptr(out), _stream(stream)), "gsh_integrate")
        return out if field.dim() == 3 else out[0]

    def random_coefficients(self, lmax, n, real=False, seed=0, decay=1.0, batch=None,
                            stream=None):
        """The coefficients of gshtrans_amd.random_coefficients(lmax, n, real,
        seed, decay, grid_lmax=L, batch) generated on the device."""
        import torch
        C = int(_lib().gsh_coefficient_size(lmax, n, REAL if real else COMPLEX))
        out = torch.empty((1 if batch is None else batch, C), dtype=torch.complex128,
                          device="cuda")
        _check(_lib().gsh_random_coefficients(self._h, lmax, n, REAL if real else COMPLEX,
                                              seed, float(decay), out.shape[0], _ptr(out), C,
                                              _stream(stream)), "gsh_random_coefficients")
        return out[0] if batch is None else out

    def product(self, lmax, u, n0, v, n1, real=False, precision="fp64", stream=None):
        """Coefficients (upper index n0 + n1) of the product of the fields of
        expansions u (n0) and v (n1): two syntheses and one fused analysis."""
        fa = self.inverse(lmax, n0, u, real=real, precision=precision, stream=stream)
        fb = self.inverse(lmax, n1, v, real=real, precision=precision, stream=stream)
        return self.forward_product(lmax, n0 + n1, fa, fb, real=real, precision=precision,
                                    stream=stream)

    def adjoint(self, lmax, n, field, coeffs=None, real=False, precision="fp64", stream=None):
        """Adjoint of the synthesis, coeffs = S_n^H(field), for a batch: the
        ring FFT and Legendre contraction of the analysis with unit
        colatitude weights, no 2 pi / N and no (L, L) zeroing; for real
        fields the adjoint of the c2r synthesis (weight 2 on orders
        0 < m < N/2, real part only at m = 0 and N/2).  Overwrites coeffs."""
        import torch
        f = field if field.dim() == 3 else field.unsqueeze(0)
        B = f.shape[0]
        _check_field(f, B, self._lmax + 1, 2 * self._lmax,
                     torch.float64 if real else torch.complex128, "adjoint")
        C = max(int(_lib().gsh_coefficient_size(lmax, n, REAL if real else COMPLEX)), 1)
        if coeffs is None:
            coeffs = torch.empty((B, C), dtype=torch.complex128, device=f.device)
        co = coeffs if coeffs.dim() == 2 else coeffs.unsqueeze(0)
        _check_coeffs(co, B, C, torch.complex128, "adjoint")
        _check(_lib().gsh_adjoint_inverse(self._h, lmax, n, REAL if real else COMPLEX,
                                          PRECISION[precision], STORAGE["fp64"], B, _ptr(f),
                                          self.FieldSize(), _ptr(co), co.shape[1],
                                          _stream(stream)),
               "gsh_adjoint_inverse")
        return coeffs if field.dim() == 3 else coeffs[0]

    # ---- reference-named single-field calls ------------------------------
    def InverseTransformation(self, lmax, n, inp, out):
        """GaussLegendreGrid.h:234: out = synthesis of inp.  Host numpy arrays
        are staged through the device; torch cuda tensors stay resident."""
        real = not np.iscomplexobj(out) if isinstance(out, np.ndarray) else \
            not out.is_complex()
        if isinstance(inp, np.ndarray):
            c = np.ascontiguousarray(inp, dtype=np.complex128)
            if c.size != coefficient_size(lmax, n, real):
                raise ValueError("in.size() != coefficient size")
            if out.size != self.FieldSize() or not out.flags.c_contiguous:
                raise ValueError("out must be a contiguous array of FieldSize()")
            _check(_lib().gsh_inverse_host(self._h, lmax, n, REAL if real else COMPLEX,
                                           _dp(c.view(np.float64)), _dp(out.view(np.float64))),
                   "InverseTransformation")
            return out
        self.inverse(lmax, n, inp.reshape(1, -1),
                     field=out.reshape(1, self._lmax + 1, 2 * self._lmax), real=real)
        return out

    def ForwardTransformation(self, lmax, n, inp, out):
        """GaussLegendreGrid.h:120: out += analysis of inp (the reference
        accumulates into `out`)."""
        real = not np.iscomplexobj(inp) if isinstance(inp, np.ndarray) else \
            not inp.is_complex()
        if isinstance(inp, np.ndarray):
            f = np.ascontiguousarray(inp, dtype=np.float64 if real else np.complex128)
            if f.size != self.FieldSize():
                raise ValueError("in.size() != FieldSize()")
            if out.size != coefficient_size(lmax, n, real) or out.dtype != np.complex128:
                raise ValueError("out must be complex128 of the coefficient size")
            _check(_lib().gsh_forward_host(self._h, lmax, n, REAL if real else COMPLEX,
                                           _dp(f.view(np.float64)), _dp(out.view(np.float64)),
                                           1), "ForwardTransformation")
            return out
        self.forward(lmax, n, inp.reshape(1, self._lmax + 1, 2 * self._lmax),
                     coeffs=out.reshape(1, -1), real=real, accumulate=True)
        return out


class BatchPlan:
    """Many fields of uneven truncation degree and upper index on one grid
    (gsh_plan_*): field i has lmax[i] and n[i].  Coefficients are a (count,
    cstride) complex tensor (complex128 or complex64 for storage="fp32"),
    field i's first coefficient_size(lmax[i], n[i]) entries in the reference
    layout; fields are (count, ntheta, nphi).  The first call in each
    direction autotunes the tile shapes; graph=True captures the launch
    sequence for the given tensors into a HIP graph on first use and replays
    it afterwards.  A graph is keyed on the tensors' addresses and strides,
    so graph=True pays only with stable caller buffers: with field=None or
    coeffs=None every call allocates a new output and captures anew (the
    plan keeps its 8 most recently used graphs)."""

    def __init__(self, grid, lmax, n, real=False, precision="fp64", storage="fp64"):
        import numpy as _np
        self.grid = grid
        self.lmax = _np.ascontiguousarray(lmax, dtype=_np.int32)
        self.n = _np.ascontiguousarray(n, dtype=_np.int32)
        self.real, self.precision, self.storage = real, precision, storage
        self.count = len(self.lmax)
        self.sizes = [coefficient_size(int(l), int(m), real) for l, m in zip(self.lmax, self.n)]
        self.cmax = max(self.sizes)
        self._h = ctypes.c_void_p()
        ip = ctypes.POINTER(ctypes.c_int)
        _check(_lib().gsh_plan_create(grid.handle, self.count,
                                      self.lmax.ctypes.data_as(ip), self.n.ctypes.data_as(ip),
                                      REAL if real else COMPLEX, PRECISION[precision],
                                      STORAGE[storage], ctypes.byref(self._h)),
               "gsh_plan_create")

    def close(self):
        if self._h:
            _lib().gsh_plan_destroy(self._h)
            self._h = ctypes.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def field_dtype(self):
        import torch
        if self.storage == "fp32":
            return torch.float32 if self.real else torch.complex64
        return torch.float64 if self.real else torch.complex128

    def coeff_dtype(self):
        import torch
        return torch.complex64 if self.storage == "fp32" else torch.complex128

    def inverse(self, coeffs, field=None, graph=False, stream=None):
        import torch
        g = self.grid
        _check_coeffs(coeffs, self.count, self.cmax, self.coeff_dtype(), "BatchPlan.inverse")
        if field is None:
            field = torch.empty((self.count, g.NumberOfCoLatitudes(), g.NumberOfLongitudes()),
                                dtype=self.field_dtype(), device=coeffs.device)
        _check_field(field, self.count, g.NumberOfCoLatitudes(), g.NumberOfLongitudes(),
                     self.field_dtype(), "BatchPlan.inverse")
        _check(_lib().gsh_plan_inverse(self._h, _ptr(coeffs), coeffs.shape[1], _ptr(field),
                                       g.FieldSize(), 1 if graph else 0, _stream(stream)),
               "gsh_plan_inverse")
        return field

    def forward(self, field, coeffs=None, graph=False, stream=None):
        import torch
        g = self.grid
        _check_field(field, self.count, g.NumberOfCoLatitudes(), g.NumberOfLongitudes(),
                     self.field_dtype(), "BatchPlan.forward")
        if coeffs is None:
            coeffs = torch.zeros((self.count, self.cmax), dtype=self.coeff_dtype(),
                                 device=field.device)
        _check_coeffs(coeffs, self.count, self.cmax, self.coeff_dtype(), "BatchPlan.forward")
        _check(_lib().gsh_plan_forward(self._h, _ptr(field), self.grid.FieldSize(),
                                       _ptr(coeffs), coeffs.shape[1], 1 if graph else 0,
                                       _stream(stream)), "gsh_plan_forward")
        return coeffs

    def info(self):
        """Launch groups with their tuned choices."""
        cap = 64
        ng = ctypes.c_int()
        arrs = [(ctypes.c_int * cap)() for _ in range(6)]
        _check(_lib().gsh_plan_info(self._h, ctypes.byref(ng), cap, *arrs), "gsh_plan_info")
        keys = ["n", "lmax", "count", "synth_tiles", "fft_inverse", "fft_forward"]
        out = [{k: arrs[j][i] for j, k in enumerate(keys)} for i in range(ng.value)]
        ap = (ctypes.c_int * cap)()
        _check(_lib().gsh_plan_analysis_pass(self._h, ctypes.byref(ng), cap, ap),
               "gsh_plan_analysis_pass")
        for i, d in enumerate(out):
            d["analysis_pass"] = ap[i]
        return out

