// C++ facade test: the reference's own checks written against
// include/GSHTrans (GaussLegendreGrid, CanonicalComponentField,
// CanonicalComponentExpansion, Wigner and its views) running on the GPU
// through the C ABI; C++23, as the reference (multi-argument operator[]).
//   * accessors of the grid (GaussLegendreGrid.h:86-103, GridBase.h:17-113);
//   * the coefficient round trip of tests/CheckCoeff2Coeff.h:74-83 for every
//     upper index, complex and real fields, bar 5e4 eps;
//   * the typed containers feeding the transforms through Data();
//   * the device-resident containers (Device*CanonicalComponent*): the same
//     transforms without host staging, bit-identical to the host-staged
//     calls, in FP64 and in the split-BF16 mode.
// Prints one line per check and "OK" at the end; exits non-zero on failure.
#include <GSHTrans/All>

#include <algorithm>
#include <cmath>
#include <complex>
#include <concepts>
#include <cstdio>
#include <limits>
#include <numeric>
#include <vector>

using namespace GSHTrans;
using Grid = GaussLegendreGrid<double, All, All>;
using RealGrid = GaussLegendreGrid<double, NonNegative, All>;

static int failures = 0;
static void expect(bool ok, const char* what) {
  std::printf("%s %s\n", ok ? "pass" : "FAIL", what);
  std::fflush(stdout);
  if (!ok) failures++;
}

template <typename G, typename Scalar>
static double round_trip(const G& grid, int lMax, int n, bool real, unsigned seed) {
  const auto size = real ? grid.RealCoefficientSize(lMax, n) : grid.ComplexCoefficientSize(lMax, n);
  std::vector<std::complex<double>> c(size), c2(size, 0.0);
  if (real)
    grid.RandomRealCoefficient(lMax, n, c, seed);
  else
    grid.RandomComplexCoefficient(lMax, n, c, seed);
  std::vector<Scalar> f(grid.FieldSize());
  grid.InverseTransformation(lMax, n, c, f);
  grid.ForwardTransformation(lMax, n, f, c2);
  double err = 0;
  for (std::size_t i = 0; i < c.size(); i++) err = std::max(err, std::abs(c2[i] - c[i]));
  return err;
}

// The analysis of a device field expression (formed as the ring FFT loads its
// rings) against the analysis of the same expression formed on the host and
// uploaded: relative L2 of the coefficients.
template <std::ptrdiff_t N, RealOrComplexValued V, typename Expr, typename F>
static double expr_error(const Grid& grid, int L, const Expr& e, F value) {
  CanonicalComponentField<N, Grid, V> h(grid);
  for (int i = 0; i < grid.NumberOfCoLatitudes(); i++)
    for (int j = 0; j < grid.NumberOfLongitudes(); j++) h[i, j] = value(i, j);
  DeviceCanonicalComponentField<N, Grid, V> dh(h);
  DeviceCanonicalComponentExpansion<N, Grid, V> out(grid), ref(grid);
  grid.ForwardTransformation(L, N, e, out);
  grid.ForwardTransformation(L, N, dh, ref);
  grid.Synchronize();
  const auto o = out.ToHost();
  const auto r = ref.ToHost();
  double num = 0, den = 0;
  for (std::size_t k = 0; k < o.Data().size(); k++) {
    num += std::norm(o.Data()[k] - r.Data()[k]);
    den += std::norm(r.Data()[k]);
  }
  return std::sqrt(num / den);
}

int main(int argc, char** argv) {
  // argv[1]: file for the device containers' inputs and outputs (raw doubles),
  // which tests/test_gpu_facade.py holds to the oracle.
  std::FILE* dump = argc > 1 ? std::fopen(argv[1], "wb") : nullptr;
  auto put = [&](const void* p, std::size_t bytes) {
    if (dump) std::fwrite(p, 1, bytes, dump);
  };
  const double eps = std::numeric_limits<double>::epsilon();
  const int L = 32;
  Grid grid(L, 2);
  expect(grid.MaxDegree() == L && grid.MaxUpperIndex() == 2, "MaxDegree / MaxUpperIndex");
  expect(grid.NumberOfCoLatitudes() == L + 1 && grid.NumberOfLongitudes() == 2 * L,
         "grid shape (L+1) x 2L");
  expect(grid.FieldSize() == (L + 1) * 2 * L, "FieldSize");
  expect(grid.ComplexCoefficientSize(L, 2) == (L + 1) * (L + 1) - 4, "ComplexCoefficientSize");
  expect(grid.RealCoefficientSize(L, 2) == (L + 1) * (L + 2) / 2 - 3, "RealCoefficientSize");
  double wsum = 0;
  for (double w : grid.CoLatitudeWeights()) wsum += w;
  expect(std::abs(wsum - 2.0) < 1e-14, "CoLatitudeWeights sum to 2");
  bool asc = true;
  double prev = -1;
  for (double t : grid.CoLatitudes()) {
    asc = asc && t > prev && t > 0 && t < M_PI;
    prev = t;
  }
  expect(asc, "CoLatitudes ascending in (0, pi)");
  double lsum = 0;
  for (double w : grid.LongitudeWeights()) lsum += w;
  expect(std::abs(lsum - 2 * M_PI) < 1e-13, "LongitudeWeights sum to 2 pi");
  int nidx = 0;
  for (auto n : grid.UpperIndices()) nidx += static_cast<int>(n) + 10;
  expect(nidx == 50, "UpperIndices -2..2");

  for (int n = -2; n <= 2; n++) {
    const double e = round_trip<Grid, std::complex<double>>(grid, L, n, false, 10 + n);
    std::printf("  complex n=%d round trip %.2e\n", n, e);
    expect(e < 5e4 * eps, "complex coefficient round trip");
    const double e2 = round_trip<Grid, std::complex<double>>(grid, L - 7, n, false, 20 + n);
    expect(e2 < 5e4 * eps, "complex round trip, lMax < L");
  }
  RealGrid rgrid(L, 2);
  for (int n = -2; n <= 2; n++) {
    const double e = round_trip<RealGrid, double>(rgrid, L, n, true, 30 + n);
    std::printf("  real n=%d round trip %.2e\n", n, e);
    expect(e < 5e4 * eps, "real coefficient round trip");
  }

  // Typed containers.
  ComplexCanonicalComponentExpansion<1, Grid> u(grid);
  grid.RandomComplexCoefficient(L, 1, u.Data(), 7);
  ComplexCanonicalComponentField<1, Grid> f(grid);
  grid.InverseTransformation(L, 1, u.Data(), f.Data());
  ComplexCanonicalComponentExpansion<1, Grid> v(grid);
  grid.ForwardTransformation(L, 1, f.Data(), v.Data());
  double e = 0;
  for (int l = 1; l <= L; l++)
    for (int m = -l; m <= l; m++) e = std::max(e, std::abs(v[l, m] - u[l, m]));
  expect(e < 5e4 * eps, "CanonicalComponentExpansion / Field round trip");
  // Device-resident containers: no host staging, same kernels, same bits.
  {
    DeviceComplexCanonicalComponentExpansion<1, Grid> du(u);
    DeviceComplexCanonicalComponentField<1, Grid> df(grid);
    grid.InverseTransformation(L, 1, du, df);
    DeviceComplexCanonicalComponentExpansion<1, Grid> dv(grid);
    grid.ForwardTransformation(L, 1, df, dv);
    grid.Synchronize();
    const auto fh = df.ToHost();
    const auto vh = dv.ToHost();
    bool same = true;
    for (int i = 0; i < grid.NumberOfCoLatitudes(); i++)
      for (int j = 0; j < grid.NumberOfLongitudes(); j++) same = same && fh[i, j] == f[i, j];
    expect(same, "device synthesis equals the host-staged synthesis bit for bit");
    same = true;
    for (int l = 1; l <= L; l++)
      for (int m = -l; m <= l; m++) same = same && vh[l, m] == v[l, m];
    expect(same, "device analysis equals the host-staged analysis bit for bit");
    // Split-BF16 synthesis on the device against the FP64 field.
    DeviceComplexCanonicalComponentField<1, Grid> dm(grid);
    grid.InverseTransformation(L, 1, du, dm, GSH_BF16X6);
    grid.Synchronize();
    const auto mh = dm.ToHost();
    double num = 0, den = 0;
    for (int i = 0; i < grid.NumberOfCoLatitudes(); i++)
      for (int j = 0; j < grid.NumberOfLongitudes(); j++) {
        num += std::norm(mh[i, j] - f[i, j]);
        den += std::norm(f[i, j]);
      }
    std::printf("  bf16x6 device synthesis relative L2 %.2e\n", std::sqrt(num / den));
    expect(std::sqrt(num / den) < 1e-7, "bf16x6 device synthesis within 1e-7 of FP64");
    // Split-BF16 analysis of the FP64 field (FP32 bins) against the FP64 one.
    DeviceComplexCanonicalComponentExpansion<1, Grid> dw(grid);
    grid.ForwardTransformation(L, 1, df, dw, GSH_BF16X6);
    grid.Synchronize();
    const auto wh = dw.ToHost();
    num = den = 0;
    for (int l = 1; l <= L; l++)
      for (int m = -l; m <= l; m++) {
        num += std::norm(wh[l, m] - v[l, m]);
        den += std::norm(v[l, m]);
      }
    std::printf("  bf16x6 device analysis relative L2 %.2e\n", std::sqrt(num / den));
    expect(std::sqrt(num / den) < 1e-7, "bf16x6 device analysis within 1e-7 of FP64");
    // A real field on a NonNegative grid.
    RealCanonicalComponentExpansion<0, RealGrid> ru(rgrid);
    rgrid.RandomRealCoefficient(L, 0, ru.Data(), 41);
    DeviceRealCanonicalComponentExpansion<0, RealGrid> dru(ru);
    DeviceRealCanonicalComponentField<0, RealGrid> drf(rgrid);
    rgrid.InverseTransformation(L, 0, dru, drf);
    DeviceRealCanonicalComponentExpansion<0, RealGrid> drv(rgrid);
    rgrid.ForwardTransformation(L, 0, drf, drv);
    rgrid.Synchronize();
    const auto rv = drv.ToHost();
    double er = 0;
    for (int l = 0; l <= L; l++)
      for (int m = 0; m <= l; m++) er = std::max(er, std::abs(rv[l, m] - ru[l, m]));
    expect(er < 5e4 * eps, "device real field round trip");
    // Inputs and results of the device containers, in order: complex n = 1
    // coefficients, synthesis, analysis; real n = 0 coefficients, synthesis,
    // analysis.
    const auto rf = drf.ToHost();
    put(u.Data().data(), u.Data().size() * sizeof(u.Data()[0]));
    put(fh.Data().data(), fh.Data().size() * sizeof(fh.Data()[0]));
    put(vh.Data().data(), vh.Data().size() * sizeof(vh.Data()[0]));
    put(ru.Data().data(), ru.Data().size() * sizeof(ru.Data()[0]));
    put(rf.Data().data(), rf.Data().size() * sizeof(rf.Data()[0]));
    put(rv.Data().data(), rv.Data().size() * sizeof(rv.Data()[0]));
    // Size mismatches raise instead of reading past the arrays.
    bool raised = false;
    try {
      grid.InverseTransformation(L - 1, 1, du, df);
    } catch (const Error& err) {
      raised = err.status == GSH_ERR_INVALID_ARGUMENT;
    }
    expect(raised, "device arrays of the wrong size raise");
  }
  // Device field expressions (the reference's expression fields,
  // CanonicalComponentFieldOverloads.h:20-266) analysed without being formed.
  {
    using C = std::complex<double>;
    ComplexCanonicalComponentExpansion<1, Grid> ua(grid), ub(grid);
    ComplexCanonicalComponentExpansion<0, Grid> uc(grid);
    grid.RandomComplexCoefficient(L, 1, ua.Data(), 101);
    grid.RandomComplexCoefficient(L, 1, ub.Data(), 102);
    grid.RandomComplexCoefficient(L, 0, uc.Data(), 103);
    uc[0, 0] = 12.0;  // the divisor stays away from zero
    DeviceComplexCanonicalComponentExpansion<1, Grid> dua(ua), dub(ub);
    DeviceComplexCanonicalComponentExpansion<0, Grid> duc(uc);
    DeviceComplexCanonicalComponentField<1, Grid> fa(grid), fb(grid);
    DeviceComplexCanonicalComponentField<0, Grid> fc(grid);
    grid.InverseTransformation(L, 1, dua, fa);
    grid.InverseTransformation(L, 1, dub, fb);
    grid.InverseTransformation(L, 0, duc, fc);
    grid.Synchronize();
    const auto ha = fa.ToHost(), hb = fb.ToHost();
    const auto hc = fc.ToHost();
    const C s(0.25, -1.5);
    double worst = 0;
    auto note = [&](const char* what, double e) {
      std::printf("  expression %s: %.2e\n", what, e);
      worst = std::max(worst, e);
    };
    note("a * b", expr_error<2, ComplexValued>(grid, L, fa * fb,
                                               [&](int i, int j) { return ha[i, j] * hb[i, j]; }));
    note("a + b", expr_error<1, ComplexValued>(grid, L, fa + fb,
                                               [&](int i, int j) { return ha[i, j] + hb[i, j]; }));
    note("a - b", expr_error<1, ComplexValued>(grid, L, fa - fb,
                                               [&](int i, int j) { return ha[i, j] - hb[i, j]; }));
    note("a / c", expr_error<1, ComplexValued>(grid, L, fa / fc,
                                               [&](int i, int j) { return ha[i, j] / hc[i, j]; }));
    note("-a", expr_error<1, ComplexValued>(grid, L, -fa, [&](int i, int j) { return -ha[i, j]; }));
    note("conj(a)", expr_error<1, ComplexValued>(grid, L, conj(fa),
                                                 [&](int i, int j) { return std::conj(ha[i, j]); }));
    note("real(a)", expr_error<1, RealValued>(grid, L, real(fa),
                                              [&](int i, int j) { return ha[i, j].real(); }));
    note("imag(a)", expr_error<1, RealValued>(grid, L, imag(fa),
                                              [&](int i, int j) { return ha[i, j].imag(); }));
    note("s * a", expr_error<1, ComplexValued>(grid, L, s * fa,
                                               [&](int i, int j) { return s * ha[i, j]; }));
    note("c + s", expr_error<0, ComplexValued>(grid, L, fc + s,
                                               [&](int i, int j) { return hc[i, j] + s; }));
    expect(worst < 1e-13, "device field expressions equal the analysis of the formed field");
  }
  // Wigner tables (Wigner.h:99-286) computed on the GPU: Gauss-Legendre
  // orthonormality of the Ortho values, FourPi = Ortho / (sqrt(2l+1) /
  // (2 sqrt(pi))), RowMajor blocks equal to ColumnMajor ones, NonNegative
  // orders equal to the All table's, and a single-angle table.
  {
    std::vector<double> th(grid.CoLatitudes().begin(), grid.CoLatitudes().end());
    std::vector<double> wt(grid.CoLatitudeWeights().begin(), grid.CoLatitudeWeights().end());
    Wigner<double, Ortho, All, All, Multiple, ColumnMajor> w(L, L, 2, th);
    Wigner<double, FourPi, All, All, Multiple, RowMajor> w4(L, L, 2, th);
    Wigner<double, Ortho, NonNegative, NonNegative, Multiple, ColumnMajor> wp(L, L, 2, th);
    double orth = 0, rel = 0;
    bool same = true;
    for (int n = -2; n <= 2; n++)
      for (int m : {-5, 0, 3, 17})
        for (int l1 = std::max(std::abs(m), std::abs(n)); l1 <= L; l1++)
          for (int l2 = l1; l2 <= L; l2++) {
            double s = 0;
            for (int i = 0; i <= L; i++) s += wt[i] * w[n, i][l1, m] * w[n, i][l2, m];
            orth = std::max(orth, std::abs(2 * M_PI * s - (l1 == l2 ? 1.0 : 0.0)));
          }
    for (int n = -2; n <= 2; n++)
      for (int i = 0; i <= L; i += 7)
        for (int l = std::abs(n); l <= L; l++)
          for (int m = -l; m <= l; m++) {
            const double f = std::sqrt(2.0 * l + 1) / (2 * std::sqrt(M_PI));
            rel = std::max(rel, std::abs(w4[n, i][l, m] * f - w[n, i][l, m]));
            if (n >= 0 && m >= 0) same = same && wp[n, i][l][m] == w[n, i][l][m];
          }
    std::printf("  Wigner orthonormality %.2e, FourPi vs Ortho %.2e\n", orth, rel);
    expect(orth < 1e-13, "Wigner Ortho table orthonormal on the Gauss-Legendre nodes");
    expect(rel < 1e-14, "Wigner FourPi = Ortho / sqrt(2l+1) / (2 sqrt(pi)), RowMajor blocks");
    expect(same, "Wigner NonNegative orders and upper indices equal the All table's");
    Wigner<double, Ortho, All, Single, Single, ColumnMajor> w1(L, 5, 1, th[4]);
    double d1 = 0;
    for (int l = 1; l <= L; l++)
      for (int m = -std::min(l, 5); m <= std::min(l, 5); m++)
        d1 = std::max(d1, std::abs(w1[l][m] - w[1, 4][l, m]));
    std::printf("  single-angle table vs full table %.2e; MaxOrder(3) %ld, MaxOrder(9) %ld\n", d1,
                static_cast<long>(w1.MaxOrder(3)), static_cast<long>(w1.MaxOrder(9)));
    expect(d1 == 0 && w1.MaxOrder(3) == 3 && w1.MaxOrder(9) == 5,
           "single-angle Wigner with mMax < lMax equals the full table's orders");
  }
  // The reference's view interface in the call shapes of its own Wigner
  // tests: degree views d[l] with Orders() and d[l][m] of a one-angle table
  // built by class template argument deduction (tests/CheckLegendre.h:26-35),
  // compared with std::sph_legendre (n = 0: Ortho d^l_{m0}(theta) is
  // Y_lm(theta, 0), Condon-Shortley phase included) for m >= 0, where the
  // library function is defined; and upper-index blocks d[n] whose degree
  // views d1[l].begin() / end() feed std::inner_product, the orthogonality
  // sum_m d^l_{mn} d^l_{mn'} = delta_{nn'} of tests/CheckAdditionTheorem.h:
  // 28-44, with the same bars (1e5 eps absolute, 1e3 eps).
  {
    const double theta = 2.0313;
    auto d = Wigner(300, 300, 0, theta);
    static_assert(std::same_as<decltype(d), Wigner<double, Ortho, All, Single, Single, ColumnMajor>>);
    double err = 0;
    long count = 0;
    for (auto l : d.Degrees(0))
      for (auto m : d[l].Orders()) {
        if (m < 0) continue;
        err = std::max(err, std::abs(d[l][m] - std::sph_legendre(l, m, theta)));
        count++;
      }
    std::printf("  d[l][m] vs std::sph_legendre, %ld values: %.2e\n", count, err);
    expect(count == 301 * 302 / 2 && err < 1e5 * eps, "Wigner degree views against sph_legendre");
    auto dd = Wigner<double, FourPi, All, All>(40, 40, 40, theta);
    double worst = 0;
    for (auto n : dd.UpperIndices()) {
      auto d1 = dd[n];
      for (auto np : dd.UpperIndices()) {
        auto d2 = dd[np];
        for (auto l = std::max(std::abs(n), std::abs(np)); l <= 40; l++) {
          auto sum = std::inner_product(d1[l].begin(), d1[l].end(), d2[l].begin(), 0.0);
          if (n == np) --sum;
          worst = std::max(worst, std::abs(sum));
        }
      }
    }
    std::printf("  addition theorem over d[n][l] views: %.2e\n", worst);
    expect(worst < 1e3 * eps, "Wigner upper-index views satisfy the addition theorem");
    // d[n, iTheta][l, m] = d[n, iTheta][l][m]; GSHView spans the block.
    std::vector<double> th2{0.3, 1.1, 2.9};
    auto dm = Wigner<double, Ortho, All, All, Multiple, RowMajor>(12, 7, 2, th2);
    bool views = true;
    for (auto n : dm.UpperIndices())
      for (int i = 0; i < 3; i++) {
        auto blk = dm[n, i];
        views = views && blk.end() - blk.begin() == blk.Size();
        for (auto l : blk.Degrees())
          for (auto m : blk[l].Orders()) views = views && blk[l, m] == blk[l][m];
      }
    expect(views, "Wigner block views: [l, m] equals [l][m], begin/end span the block");
  }
  // A single mode: Y_{3,2} (n = 0) synthesised then read back.
  ComplexCanonicalComponentExpansion<0, Grid> y(grid);
  y[3, 2] = 1.0;
  ComplexCanonicalComponentField<0, Grid> fy(grid);
  grid.InverseTransformation(L, 0, y.Data(), fy.Data());
  ComplexCanonicalComponentExpansion<0, Grid> y2(grid);
  grid.ForwardTransformation(L, 0, fy.Data(), y2.Data());
  expect(std::abs(y2[3, 2] - 1.0) < 1e-13 && std::abs(y2[3, 1]) < 1e-13, "single mode");
  // The library's errors surface as exceptions (the reference asserts).
  bool threw = false;
  try {
    std::vector<std::complex<double>> c(grid.ComplexCoefficientSize(L, 3));
    std::vector<std::complex<double>> ff(grid.FieldSize());
    grid.InverseTransformation(L, 3, c, ff);
  } catch (const Error& err) {
    threw = err.status == GSH_ERR_UPPER_INDEX;
  }
  expect(threw, "upper index outside the grid raises");
  if (dump) std::fclose(dump);
  if (failures) {
    std::printf("%d FAILED\n", failures);
    return 1;
  }
  std::printf("OK\n");
  return 0;
}
