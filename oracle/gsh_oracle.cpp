// gsh_oracle.cpp -- CPU restatement of GSHTrans's transform hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: only tests/,
// __graft_entry__.smoke() and bench.py's cpu_baseline leg may load the
// library built from it.  The product (gshtrans_amd/) never links or calls it.
//
// What it restates, step for step and in the same floating-point operation
// order (file:line refer to the reference tree):
//   * WignerDetails::Arguments / WignerMinOrder / MaxOrder / Min/MaxUpperIndex
//       GSHTrans/src/Wigner.h:24-95
//   * Wigner::PreCompute (sqrt tables)              Wigner.h:289-303
//   * Wigner::Compute (seeds, one-term and three-term recurrences, Ortho
//     scaling)                                       Wigner.h:317-486
//   * Wigner::ComputeAll (OpenMP over (n, iTheta))   Wigner.h:305-314
//   * GSHIndices<All|NonNegative> offsets            GSHTrans/src/Indexing.h:22-150
//   * GaussLegendreGrid ctor (nodes, theta = acos(-x), table)
//                                                    GaussLegendreGrid.h:38-53
//   * ForwardTransformation (analysis)               GaussLegendreGrid.h:120-217
//   * InverseTransformation (synthesis)              GaussLegendreGrid.h:234-321
//
// Third-party pieces that are absent from the reference tree and restated from
// their published algorithms:
//   * GaussQuad (github.com/da380/GaussQuad, branch main, unpinned):
//     Gauss-Legendre nodes/weights.  Restated as Newton iteration on P_{L+1}
//     in long double, nodes x ascending (so theta = acos(-x) ascends).
//     Node order and last-bit values are "parity unpinned" against GaussQuad;
//     they are pinned against numpy.polynomial.legendre.leggauss, exact
//     polynomial quadrature and symmetry in tests/.
//   * FFTWpp/FFTW3 (github.com/da380/FFTWpp, branch main): unnormalised DFTs,
//     forward sign e^{-2 pi i jk/N}, backward e^{+2 pi i jk/N}; r2c returns
//     bins 0..N/2, c2r reads bins 0..N/2 of a Hermitian spectrum and ignores
//     the imaginary parts of bins 0 and N/2.  Restated as radix-2 (N a power
//     of two) or Bluestein (other N) with twiddles from long double sin/cos.
//
// Build: oracle/Makefile (g++ -O2 -fopenmp -ffp-contract=off).
#include <omp.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstddef>
#include <cstdint>
#include <cstring>
#include <limits>
#include <memory>
#include <vector>

namespace oracle {

using Int = std::ptrdiff_t;

// Utility.h:8-10 -- (-1)^m, with the reference's `m % 2 ? -1 : 1` semantics.
inline Int MinusOneToPower(Int m) { return m % 2 ? -1 : 1; }

// ---------------------------------------------------------------------------
// Index layouts, Indexing.h:22-150.  MRange: 0 = All, 1 = NonNegative.
// ---------------------------------------------------------------------------
struct Indices {
  Int lMax, mMax, n;
  int mrange;
  Indices(Int lMax_, Int mMax_, Int n_, int mrange_)
      : lMax(lMax_), mMax(std::min(lMax_, mMax_)), n(n_), mrange(mrange_) {}
  Int MinDegree() const { return n < 0 ? -n : n; }
  // Indexing.h:104-131
  Int OffsetForDegree(Int l) const {
    Int nAbs = MinDegree();
    if (mrange == 0) {
      if (mMax >= nAbs) {
        return l <= mMax ? l * l - nAbs * nAbs
                         : (mMax + 1) * (mMax + 1) - nAbs * nAbs +
                               (l - 1 - mMax) * (2 * mMax + 1);
      }
      return (l - nAbs) * (2 * mMax + 1);
    }
    if (mMax >= nAbs) {
      return l <= mMax ? (l * (l + 1)) / 2 - (nAbs * (nAbs + 1)) / 2
                       : ((mMax + 1) * (mMax + 2)) / 2 -
                             (nAbs * (nAbs + 1)) / 2 + (l - 1 - mMax) * (mMax + 1);
    }
    return (l - nAbs) * (mMax + 1);
  }
  // GSHSubIndices, Indexing.h:26-65
  Int SubMaxOrder(Int l) const { return std::min(l, mMax); }
  Int SubMinOrder(Int l) const { return mrange == 0 ? -SubMaxOrder(l) : 0; }
  Int SubSize(Int l) const { return SubMaxOrder(l) - SubMinOrder(l) + 1; }
  Int Index(Int l, Int m) const {
    return OffsetForDegree(l) + (mrange == 0 ? m + SubMaxOrder(l) : m);
  }
  Int Size() const { return OffsetForDegree(lMax) + SubSize(lMax); }
};

// ---------------------------------------------------------------------------
// Wigner.h:24-95
// ---------------------------------------------------------------------------
template <typename Real>
struct Arguments {
  Real logSinHalf, logCosHalf;
  bool atLeft, atRight;
  explicit Arguments(Real theta) {
    const Real half = static_cast<Real>(1) / static_cast<Real>(2);
    logSinHalf = std::sin(half * theta);
    logCosHalf = std::cos(half * theta);
    atLeft = logSinHalf < std::numeric_limits<Real>::min();
    atRight = logCosHalf < std::numeric_limits<Real>::min();
    logSinHalf = atLeft ? static_cast<Real>(0) : std::log(logSinHalf);
    logCosHalf = atRight ? static_cast<Real>(0) : std::log(logCosHalf);
  }
};

template <typename Real>
Real WignerMinOrder(Int l, Int n, const Arguments<Real>& arg) {
  if (l == 0) return static_cast<Real>(1);
  if (arg.atLeft) return n == -l ? static_cast<Real>(1) : static_cast<Real>(0);
  if (arg.atRight) return n == l ? static_cast<Real>(1) : static_cast<Real>(0);
  const Real half = static_cast<Real>(1) / static_cast<Real>(2);
  Real Fl = static_cast<Real>(l);
  Real Fn = static_cast<Real>(n);
  using std::exp;
  using std::lgamma;
  return exp(half * (lgamma(2 * Fl + 1) - lgamma(Fl - Fn + 1) -
                     lgamma(Fl + Fn + 1)) +
             (Fl + Fn) * arg.logSinHalf + (Fl - Fn) * arg.logCosHalf);
}

template <typename Real>
Real WignerMaxOrder(Int l, Int n, const Arguments<Real>& arg) {
  return static_cast<Real>(MinusOneToPower(n + l)) * WignerMinOrder(l, -n, arg);
}
template <typename Real>
Real WignerMinUpperIndex(Int l, Int m, const Arguments<Real>& arg) {
  return WignerMaxOrder(l, -m, arg);
}
template <typename Real>
Real WignerMaxUpperIndex(Int l, Int m, const Arguments<Real>& arg) {
  return WignerMinOrder(l, -m, arg);
}

// Wigner.h:289-303
template <typename Real>
struct SqrtTables {
  std::vector<Real> sqrtInt, sqrtIntInv;
  explicit SqrtTables(Int size) {
    sqrtInt.reserve(size);
    sqrtIntInv.reserve(size);
    for (Int m = 0; m < size; m++) sqrtInt.push_back(std::sqrt(static_cast<Real>(m)));
    for (auto x : sqrtInt) sqrtIntInv.push_back(x > 0 ? 1 / x : 0);
  }
};

// Wigner::Compute, Wigner.h:317-486, for one (n, theta) block written into
// `out` in GSHIndices<MRange>(lMax, mMax, n) order.  norm: 0 = Ortho, 1 = FourPi.
template <typename Real>
void WignerCompute(Int lMax, Int mMaxIn, Int n, int mrange, int norm, Real theta,
                   const SqrtTables<Real>& pre, Real* out) {
  const auto& sqrtInt = pre.sqrtInt;
  const auto& sqrtIntInv = pre.sqrtIntInv;
  Arguments<Real> arg(theta);
  const Real cos = std::cos(theta);
  const Int nAbs = n < 0 ? -n : n;
  Indices d(lMax, mMaxIn, n, mrange);
  const bool all = mrange == 0;

  // l == |n|   (Wigner.h:330-344)
  {
    const Int l = nAbs;
    Int m = d.SubMinOrder(l);
    Real* iter = out + d.OffsetForDegree(l);
    Real* finish = iter + d.SubSize(l);
    if (n >= 0) {
      while (iter != finish) *iter++ = WignerMaxUpperIndex(l, m++, arg);
    } else {
      while (iter != finish) *iter++ = WignerMinUpperIndex(l, m++, arg);
    }
  }

  // l == |n| + 1   (Wigner.h:347-383)
  if (nAbs < lMax) {
    const Int l = nAbs + 1;
    const Int mMin = d.SubMinOrder(l);
    const Int mMax = d.SubMaxOrder(l);
    Int m = mMin;
    const Real* iterMinusOne = out + d.OffsetForDegree(l - 1);
    const Real* finishMinusOne = iterMinusOne + d.SubSize(l - 1);
    Real* iter = out + d.OffsetForDegree(l);
    if (all) {
      if (l <= mMax) {
        *iter++ = WignerMinOrder(l, n, arg);
        m++;
      }
    }
    {
      const Real alpha = static_cast<Real>((2 * l - 1) * l) * cos * sqrtIntInv[l + nAbs];
      const Real beta = static_cast<Real>((n < 0 ? -1 : 1) * (2 * l - 1)) * sqrtIntInv[l + nAbs];
      while (iterMinusOne != finishMinusOne) {
        const Real f1 = (alpha - beta * static_cast<Real>(m)) * sqrtIntInv[l - m] *
                        sqrtIntInv[l + m];
        *iter++ = f1 * *iterMinusOne++;
        m++;
      }
    }
    if (l <= mMax) *iter++ = WignerMaxOrder(l, n, arg);
  }

  // l >= |n| + 2   (Wigner.h:386-472)
  for (Int l = nAbs + 2; l <= lMax; l++) {
    const Int mMin = d.SubMinOrder(l);
    const Int mMax = d.SubMaxOrder(l);
    Int m = mMin;
    const Real* iterMinusTwo = out + d.OffsetForDegree(l - 2);
    const Real* finishMinusTwo = iterMinusTwo + d.SubSize(l - 2);
    const Real* iterMinusOne = out + d.OffsetForDegree(l - 1);
    Real* iter = out + d.OffsetForDegree(l);

    // One-term step used at the boundary orders (Wigner.h:407-410 etc.).
    auto oneTerm = [&](Int mm) {
      return static_cast<Real>(2 * l - 1) *
             (static_cast<Real>(l * (l - 1)) * cos - static_cast<Real>(mm * n)) *
             sqrtIntInv[l - n] * sqrtIntInv[l + n] * sqrtIntInv[l - mm] *
             sqrtIntInv[l + mm] / static_cast<Real>(l - 1);
    };

    if (all) {
      if (l <= mMax) {
        *iter++ = WignerMinOrder(l, n, arg);
        m++;
        const Real f1 = oneTerm(m);
        *iter++ = f1 * (*iterMinusOne++);
        m++;
      }
      if (l == mMax + 1) {
        const Real f1 = oneTerm(m);
        *iter++ = f1 * (*iterMinusOne++);
        m++;
      }
    }

    {
      const Real alpha = static_cast<Real>((2 * l - 1) * l) * cos * sqrtIntInv[l - n] *
                         sqrtIntInv[l + n];
      const Real beta = static_cast<Real>((2 * l - 1) * n) * sqrtIntInv[l - n] *
                        sqrtIntInv[l + n] / static_cast<Real>(l - 1);
      const Real gamma = static_cast<Real>(l) * sqrtInt[l - 1 - n] * sqrtInt[l - 1 + n] *
                         sqrtIntInv[l - n] * sqrtIntInv[l + n] / static_cast<Real>(l - 1);
      while (iterMinusTwo != finishMinusTwo) {
        const Real denom = sqrtIntInv[l - m] * sqrtIntInv[l + m];
        const Real f1 = (alpha - beta * static_cast<Real>(m)) * denom;
        const Real f2 = gamma * sqrtInt[l - 1 - m] * sqrtInt[l - 1 + m] * denom;
        *iter++ = f1 * *iterMinusOne++ - f2 * *iterMinusTwo++;
        m++;
      }
    }

    if (l <= mMax) {
      {
        const Real f1 = oneTerm(m);
        *iter++ = f1 * (*iterMinusOne++);
        m++;
      }
      *iter++ = WignerMaxOrder(l, n, arg);
    }
    if (l == mMax + 1) {
      const Real f1 = oneTerm(m);
      *iter++ = f1 * (*iterMinusOne++);
    }
  }

  // Ortho normalisation (Wigner.h:475-485).
  if (norm == 0) {
    const Real inv_sqrtpi =
        static_cast<Real>(0.564189583547756286948079451560772585844050629328998856844L);
    const Real factor = inv_sqrtpi / static_cast<Real>(2);
    for (Int l = nAbs; l <= lMax; l++) {
      Real* start = out + d.OffsetForDegree(l);
      Real* finish = start + d.SubSize(l);
      for (Real* p = start; p != finish; ++p)
        *p = factor * std::sqrt(static_cast<Real>(2 * l + 1)) * *p;
    }
  }
}

// ---------------------------------------------------------------------------
// Gauss-Legendre nodes and weights (GaussQuad restated; see header).
// x ascending in [-1, 1]; weights sum to 2.
// ---------------------------------------------------------------------------
inline void GaussLegendre(Int npts, std::vector<long double>& x,
                          std::vector<long double>& w) {
  x.assign(npts, 0.0L);
  w.assign(npts, 0.0L);
  const long double pi = 3.141592653589793238462643383279502884L;
  for (Int i = 0; i < (npts + 1) / 2; i++) {
    // Tricomi initial guess for the i-th largest root.
    long double z = std::cos(pi * (static_cast<long double>(i) + 0.75L) /
                             (static_cast<long double>(npts) + 0.5L));
    long double pp = 0;
    for (int it = 0; it < 100; it++) {
      long double p1 = 1, p2 = 0;
      for (Int j = 1; j <= npts; j++) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
      }
      pp = npts * (z * p1 - p2) / (z * z - 1);
      long double dz = p1 / pp;
      z -= dz;
      if (std::fabs(dz) <= 4 * std::numeric_limits<long double>::epsilon() && it > 1) break;
    }
    // Recompute derivative at the converged root for the weight.
    {
      long double p1 = 1, p2 = 0;
      for (Int j = 1; j <= npts; j++) {
        long double p3 = p2;
        p2 = p1;
        p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
      }
      pp = npts * (z * p1 - p2) / (z * z - 1);
    }
    long double wt = 2 / ((1 - z * z) * pp * pp);
    x[npts - 1 - i] = z;
    x[i] = -z;
    w[npts - 1 - i] = wt;
    w[i] = wt;
  }
  if (npts % 2 == 1) x[npts / 2] = 0;  // exact middle node
}

// ---------------------------------------------------------------------------
// FFT (FFTW semantics restated; see header).
// ---------------------------------------------------------------------------
template <typename Real>
struct FFT {
  using C = std::complex<Real>;
  Int N = 0, M = 0;  // M: power-of-two size used internally
  bool pow2 = true;
  std::vector<C> tw;         // twiddles e^{-2 pi i k / M}, k < M/2
  std::vector<C> chirp;      // Bluestein chirp e^{-pi i k^2 / N}, k < N
  std::vector<C> chirpHat;   // FFT of the conjugate chirp filter (size M)
  std::vector<int> rev;

  explicit FFT(Int n) : N(n) {
    pow2 = N > 0 && (N & (N - 1)) == 0;
    M = 1;
    if (pow2) {
      M = N;
    } else {
      while (M < 2 * N - 1) M <<= 1;
    }
    const long double pi = 3.141592653589793238462643383279502884L;
    tw.resize(M / 2 > 0 ? M / 2 : 1);
    for (Int k = 0; k < M / 2; k++) {
      long double a = -2.0L * pi * static_cast<long double>(k) / static_cast<long double>(M);
      tw[k] = C(static_cast<Real>(std::cos(a)), static_cast<Real>(std::sin(a)));
    }
    rev.resize(M);
    int bits = 0;
    while ((Int(1) << bits) < M) bits++;
    for (Int i = 0; i < M; i++) {
      int r = 0;
      for (int b = 0; b < bits; b++)
        if (i & (Int(1) << b)) r |= 1 << (bits - 1 - b);
      rev[i] = r;
    }
    if (!pow2) {
      chirp.resize(N);
      for (Int k = 0; k < N; k++) {
        // k^2 mod 2N keeps the angle argument small.
        long long kk = (static_cast<long long>(k) * k) % (2 * N);
        long double a = -pi * static_cast<long double>(kk) / static_cast<long double>(N);
        chirp[k] = C(static_cast<Real>(std::cos(a)), static_cast<Real>(std::sin(a)));
      }
      chirpHat.assign(M, C(0, 0));
      chirpHat[0] = std::conj(chirp[0]);
      for (Int k = 1; k < N; k++) {
        chirpHat[k] = std::conj(chirp[k]);
        chirpHat[M - k] = std::conj(chirp[k]);
      }
      radix2(chirpHat.data(), false);
    }
  }

  // In-place radix-2 DIT FFT of size M; inverse uses conjugate twiddles.
  void radix2(C* a, bool inverse) const {
    for (Int i = 0; i < M; i++)
      if (i < rev[i]) std::swap(a[i], a[rev[i]]);
    for (Int len = 2; len <= M; len <<= 1) {
      Int half = len / 2, step = M / len;
      for (Int i = 0; i < M; i += len) {
        for (Int j = 0; j < half; j++) {
          C w = tw[j * step];
          if (inverse) w = std::conj(w);
          C u = a[i + j], v = a[i + j + half] * w;
          a[i + j] = u + v;
          a[i + j + half] = u - v;
        }
      }
    }
  }

  // Unnormalised DFT: sign -1 forward, +1 backward.  in/out length N.
  void dft(const C* in, C* out, bool backward) const {
    if (pow2) {
      std::vector<C> a(in, in + N);
      radix2(a.data(), backward);
      std::copy(a.begin(), a.end(), out);
      return;
    }
    // Bluestein: X_k = conj?(chirp_k) * sum_j (x_j chirp_j) conj(chirp_{k-j})
    std::vector<C> a(M, C(0, 0));
    for (Int k = 0; k < N; k++) {
      C ch = backward ? std::conj(chirp[k]) : chirp[k];
      a[k] = in[k] * ch;
    }
    radix2(a.data(), false);
    if (backward) {
      // filter for the backward chirp is conj of the forward filter in time,
      // i.e. chirp (not conj); its transform is conj(chirpHat) reflected.
      std::vector<C> h(M, C(0, 0));
      h[0] = chirp[0];
      for (Int k = 1; k < N; k++) {
        h[k] = chirp[k];
        h[M - k] = chirp[k];
      }
      radix2(h.data(), false);
      for (Int k = 0; k < M; k++) a[k] *= h[k];
    } else {
      for (Int k = 0; k < M; k++) a[k] *= chirpHat[k];
    }
    radix2(a.data(), true);
    const Real invM = static_cast<Real>(1) / static_cast<Real>(M);
    for (Int k = 0; k < N; k++) {
      C ch = backward ? std::conj(chirp[k]) : chirp[k];
      out[k] = a[k] * invM * ch;
    }
  }

  // r2c: real input length N -> bins 0..N/2 (FFTW forward real transform).
  void r2c(const Real* in, C* out) const {
    std::vector<C> a(N), b(N);
    for (Int j = 0; j < N; j++) a[j] = C(in[j], 0);
    dft(a.data(), b.data(), false);
    for (Int k = 0; k <= N / 2; k++) out[k] = b[k];
  }
  // c2r: bins 0..N/2 of a Hermitian spectrum -> real output length N.
  void c2r(const C* in, Real* out) const {
    std::vector<C> a(N), b(N);
    for (Int k = 0; k <= N / 2; k++) a[k] = in[k];
    a[0] = C(in[0].real(), 0);
    if (N % 2 == 0) a[N / 2] = C(in[N / 2].real(), 0);
    for (Int k = 1; k < (N + 1) / 2; k++) a[N - k] = std::conj(in[k]);
    dft(a.data(), b.data(), true);
    for (Int j = 0; j < N; j++) out[j] = b[j].real();
  }
};

// ---------------------------------------------------------------------------
// GaussLegendreGrid, GaussLegendreGrid.h:24-333.
// nrange: 0 = All, 1 = NonNegative, 2 = Single (GridBase.h:17-28).
// ---------------------------------------------------------------------------
template <typename Real>
struct Grid {
  using C = std::complex<Real>;
  Int L, nMax;
  int mrange, nrange;
  bool table;
  std::vector<Real> theta, weight;
  std::unique_ptr<SqrtTables<Real>> pre;
  std::vector<Real> data;          // Wigner table, ColumnMajor over (n, iTheta)
  std::vector<std::size_t> offset;
  std::unique_ptr<FFT<Real>> fft;

  Int MinUpperIndex() const { return nrange == 0 ? -nMax : (nrange == 1 ? 0 : nMax); }
  Int NTheta() const { return L + 1; }
  Int NPhi() const { return 2 * L; }
  Int BlockSize(Int n) const { return Indices(L, L, n, mrange).Size(); }

  Grid(Int L_, Int nMax_, int mrange_, int nrange_, bool table_)
      : L(L_), nMax(nMax_), mrange(mrange_), nrange(nrange_), table(table_) {
    std::vector<long double> x, w;
    GaussLegendre(L + 1, x, w);
    theta.resize(L + 1);
    weight.resize(L + 1);
    for (Int i = 0; i <= L; i++) {
      Real xi = static_cast<Real>(x[i]);
      theta[i] = std::acos(-xi);  // GaussLegendreGrid.h:48
      weight[i] = static_cast<Real>(w[i]);
    }
    Int preSize = L + std::max(L, nMax) + 1;  // Wigner.h:291 with mMax = L
    pre.reset(new SqrtTables<Real>(preSize));
    if (table) {
      // Wigner.h:122-130 then ComputeAll (Wigner.h:305-314)
      std::size_t size = 0;
      for (Int n = MinUpperIndex(); n <= nMax; n++)
        for (Int it = 0; it <= L; it++) {
          offset.push_back(size);
          size += BlockSize(n);
        }
      data.assign(size, 0);
      const Int nN = nMax - MinUpperIndex() + 1;
      const Int total = nN * (L + 1);
#pragma omp parallel for schedule(dynamic)
      for (Int k = 0; k < total; k++) {
        Int n = MinUpperIndex() + k / (L + 1);
        Int it = k % (L + 1);
        WignerCompute<Real>(L, L, n, mrange, 0, theta[it], *pre, &data[offset[k]]);
      }
    }
    if (L > 0) fft.reset(new FFT<Real>(NPhi()));
  }

  // Returns a pointer to the (n, iTheta) block, computing it into `scratch`
  // when no table is held (identical arithmetic).
  const Real* Block(Int n, Int it, std::vector<Real>& scratch) const {
    if (table) return &data[offset[(n - MinUpperIndex()) * (L + 1) + it]];
    scratch.resize(BlockSize(n));
    WignerCompute<Real>(L, L, n, mrange, 0, theta[it], *pre, scratch.data());
    return scratch.data();
  }

  // InverseTransformation, GaussLegendreGrid.h:234-321.
  // in: complex coefficients; out: real (realField) or complex field.
  void Inverse(Int lMax, Int n, bool realField, const C* in, Real* outReal,
               C* outCplx) const {
    const Real inv_sqrtpi =
        static_cast<Real>(0.564189583547756286948079451560772585844050629328998856844L);
    if (lMax == 0) {  // GaussLegendreGrid.h:251-260
      if (realField)
        outReal[0] = std::real(in[0]) * static_cast<Real>(2) / inv_sqrtpi;
      else
        outCplx[0] = in[0] * static_cast<Real>(2) / inv_sqrtpi;
      return;
    }
    const Int nPhi = NPhi();
    std::vector<C> inWork(realField ? nPhi / 2 + 1 : nPhi);
    std::vector<C> cOut(nPhi);
    std::vector<Real> scratch;
    Indices dl(L, L, n, mrange);
    for (Int it = 0; it <= L; it++) {
      std::fill(inWork.begin(), inWork.end(), C(0, 0));
      const Real* wig = Block(n, it, scratch);
      const C* inIter = in;
      const Real* wigIter = wig;
      for (Int l = dl.MinDegree(); l <= L; l++) {
        if (l > lMax) break;  // degrees filtered by l <= lMax (:290-291)
        if (!realField) {
          C* workIter = inWork.data() + (nPhi - l);
          for (Int m = -l; m < 0; m++) *workIter++ += *inIter++ * *wigIter++;
          workIter = inWork.data();
          for (Int m = 0; m <= l; m++) *workIter++ += *inIter++ * *wigIter++;
        } else {
          C* workIter = inWork.data();
          if (mrange == 0) wigIter += l;
          for (Int m = 0; m <= l; m++) *workIter++ += *inIter++ * *wigIter++;
        }
      }
      if (realField) {
        fft->c2r(inWork.data(), outReal + it * nPhi);
      } else {
        fft->dft(inWork.data(), outCplx + it * nPhi, true);
      }
    }
  }

  // ForwardTransformation, GaussLegendreGrid.h:120-217.
  void Forward(Int lMax, Int n, bool realField, const Real* inReal, const C* inCplx,
               C* out) const {
    const Real inv_sqrtpi =
        static_cast<Real>(0.564189583547756286948079451560772585844050629328998856844L);
    if (lMax == 0) {  // :137-140
      if (realField)
        out[0] = C(inReal[0] * inv_sqrtpi / static_cast<Real>(2), 0);
      else
        out[0] = inCplx[0] * inv_sqrtpi / static_cast<Real>(2);
      return;
    }
    const Int nPhi = NPhi();
    const Real pi = static_cast<Real>(3.141592653589793238462643383279502884L);
    const Real scaleFactor = static_cast<Real>(2) * pi / static_cast<Real>(nPhi);
    std::vector<C> outWork(realField ? nPhi / 2 + 1 : nPhi);
    std::vector<Real> scratch;
    Indices dl(L, L, n, mrange);
    const Int outSize = Indices(lMax, lMax, n, realField ? 1 : 0).Size();
    for (Int it = 0; it <= L; it++) {
      if (realField)
        fft->r2c(inReal + it * nPhi, outWork.data());
      else
        fft->dft(inCplx + it * nPhi, outWork.data(), false);
      const Real* wig = Block(n, it, scratch);
      const Real w = weight[it] * scaleFactor;
      C* outIter = out;
      const Real* wigIter = wig;
      for (Int l = dl.MinDegree(); l <= L; l++) {
        if (l > lMax) break;
        if (!realField) {
          const C* workIter = outWork.data() + (nPhi - l);
          for (Int m = -l; m < 0; m++) *outIter++ += *wigIter++ * *workIter++ * w;
          workIter = outWork.data();
          for (Int m = 0; m <= l; m++) *outIter++ += *wigIter++ * *workIter++ * w;
        } else {
          const C* workIter = outWork.data();
          if (mrange == 0) wigIter += l;
          for (Int m = 0; m <= l; m++) *outIter++ += *wigIter++ * *workIter++ * w;
        }
      }
      if (!realField && lMax == L) out[outSize - 1] = C(0, 0);  // :209-215
    }
  }
};

}  // namespace oracle

// ===========================================================================
// C interface (ctypes).  Complex arrays are interleaved (re, im).
// ===========================================================================
using oracle::Int;

extern "C" {

int gsho_version() { return 1; }
int gsho_num_threads() { return omp_get_max_threads(); }

// One Wigner block (Wigner::Compute) for (lMax, mMax, n, mrange, norm, theta).
// out length: GSHIndices<MRange>(lMax, mMax, n).Size().
long gsho_wigner_size(long lMax, long mMax, long n, int mrange) {
  return oracle::Indices(lMax, mMax, n, mrange).Size();
}
int gsho_wigner_block(long lMax, long mMax, long n, int mrange, int norm, double theta,
                      double* out) {
  if (lMax < 0 || mMax < 0 || std::labs(n) > lMax) return 1;
  Int size = lMax + std::max(mMax, std::labs(n)) + 1;
  oracle::SqrtTables<double> pre(size);
  oracle::WignerCompute<double>(lMax, mMax, n, mrange, norm, theta, pre, out);
  return 0;
}
int gsho_wigner_block_ld(long lMax, long mMax, long n, int mrange, int norm,
                         long double theta, long double* out) {
  if (lMax < 0 || mMax < 0 || std::labs(n) > lMax) return 1;
  Int size = lMax + std::max(mMax, std::labs(n)) + 1;
  oracle::SqrtTables<long double> pre(size);
  oracle::WignerCompute<long double>(lMax, mMax, n, mrange, norm, theta, pre, out);
  return 0;
}

// WignerMinOrder (Wigner.h:52-80) exposed for seed tests.
double gsho_wigner_min_order(long l, long n, double theta) {
  oracle::Arguments<double> arg(theta);
  return oracle::WignerMinOrder<double>(l, n, arg);
}

// libstdc++'s std::sph_legendre, the library routine tests/CheckLegendre.h:30
// compares Wigner values against (an independent check, not a restatement).
double gsho_std_sph_legendre(unsigned l, unsigned m, double theta) {
  return std::sph_legendre(l, m, theta);
}

// Gauss-Legendre: theta = acos(-x) ascending, weights.
void gsho_gl_nodes(long L, double* theta, double* w) {
  std::vector<long double> x, wl;
  oracle::GaussLegendre(L + 1, x, wl);
  for (long i = 0; i <= L; i++) {
    double xi = static_cast<double>(x[i]);
    theta[i] = std::acos(-xi);
    w[i] = static_cast<double>(wl[i]);
  }
}
void gsho_gl_nodes_x_ld(long npts, long double* x, long double* w) {
  std::vector<long double> xv, wv;
  oracle::GaussLegendre(npts, xv, wv);
  std::copy(xv.begin(), xv.end(), x);
  std::copy(wv.begin(), wv.end(), w);
}

// FFT helpers (FFTW semantics) for tests of the restated FFT.
int gsho_dft(long N, int backward, const double* in, double* out) {
  oracle::FFT<double> f(N);
  f.dft(reinterpret_cast<const std::complex<double>*>(in),
        reinterpret_cast<std::complex<double>*>(out), backward != 0);
  return 0;
}
int gsho_r2c(long N, const double* in, double* out) {
  oracle::FFT<double> f(N);
  f.r2c(in, reinterpret_cast<std::complex<double>*>(out));
  return 0;
}
int gsho_c2r(long N, const double* in, double* out) {
  oracle::FFT<double> f(N);
  f.c2r(reinterpret_cast<const std::complex<double>*>(in), out);
  return 0;
}

// Grids.  table != 0 holds the reference's precomputed Wigner table
// (OpenMP build); table == 0 recomputes each block per ring with the same
// arithmetic (for memory-light parity runs).
void* gsho_grid_create(long L, long nMax, int mrange, int nrange, int table) {
  if (L < 0 || nMax < 0 || nMax > L) return nullptr;
  return new oracle::Grid<double>(L, nMax, mrange, nrange, table != 0);
}
void gsho_grid_destroy(void* h) { delete static_cast<oracle::Grid<double>*>(h); }
void gsho_grid_nodes(void* h, double* theta, double* w) {
  auto* g = static_cast<oracle::Grid<double>*>(h);
  std::copy(g->theta.begin(), g->theta.end(), theta);
  std::copy(g->weight.begin(), g->weight.end(), w);
}
static bool oracle_check(const oracle::Grid<double>* g, long lMax, long n, int realField) {
  if (lMax < 0 || lMax > g->L) return false;
  if (n < g->MinUpperIndex() || n > g->nMax) return false;
  if (std::labs(n) > lMax) return false;
  if (!realField && g->mrange != 0) return false;
  return true;
}
// out is zeroed by the caller for the analysis (the reference accumulates
// into `out`, GaussLegendreGrid.h:192).
int gsho_inverse(void* h, long lMax, long n, int realField, const double* in, double* out) {
  auto* g = static_cast<oracle::Grid<double>*>(h);
  if (!oracle_check(g, lMax, n, realField)) return 1;
  g->Inverse(lMax, n, realField != 0, reinterpret_cast<const std::complex<double>*>(in),
             out, reinterpret_cast<std::complex<double>*>(out));
  return 0;
}
int gsho_forward(void* h, long lMax, long n, int realField, const double* in, double* out) {
  auto* g = static_cast<oracle::Grid<double>*>(h);
  if (!oracle_check(g, lMax, n, realField)) return 1;
  g->Forward(lMax, n, realField != 0, in, reinterpret_cast<const std::complex<double>*>(in),
             reinterpret_cast<std::complex<double>*>(out));
  return 0;
}

// long double grid (higher-precision cross-check of the double oracle).
void* gsho_grid_create_ld(long L, long nMax, int mrange, int nrange, int table) {
  if (L < 0 || nMax < 0 || nMax > L) return nullptr;
  return new oracle::Grid<long double>(L, nMax, mrange, nrange, table != 0);
}
void gsho_grid_destroy_ld(void* h) { delete static_cast<oracle::Grid<long double>*>(h); }
int gsho_inverse_ld(void* h, long lMax, long n, int realField, const long double* in,
                    long double* out) {
  auto* g = static_cast<oracle::Grid<long double>*>(h);
  if (lMax < 0 || lMax > g->L || std::labs(n) > lMax) return 1;
  g->Inverse(lMax, n, realField != 0, reinterpret_cast<const std::complex<long double>*>(in),
             out, reinterpret_cast<std::complex<long double>*>(out));
  return 0;
}
int gsho_forward_ld(void* h, long lMax, long n, int realField, const long double* in,
                    long double* out) {
  auto* g = static_cast<oracle::Grid<long double>*>(h);
  if (lMax < 0 || lMax > g->L || std::labs(n) > lMax) return 1;
  g->Forward(lMax, n, realField != 0, in,
             reinterpret_cast<const std::complex<long double>*>(in),
             reinterpret_cast<std::complex<long double>*>(out));
  return 0;
}

}  // extern "C"
