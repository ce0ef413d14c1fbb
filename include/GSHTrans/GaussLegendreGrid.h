// GSHTrans/GaussLegendreGrid.h -- C++ facade of GaussLegendreGrid over the
// MI355X transform library (C ABI in ../gshtrans_amd.h).
//
// Same class template, constructor and member functions as the reference's
// GaussLegendreGrid<Real, MRange, NRange> (GSHTrans/src/GaussLegendreGrid.h:
// 24-333, with the GridBase accessors of GSHTrans/src/GridBase.h:17-172):
// colatitudes are the Gauss-Legendre points theta = acos(-x) ascending,
// longitudes the 2L points j pi / L, fields row-major (theta, phi), and
// coefficients in GSHIndices order (All for complex fields, NonNegative for
// real fields).  The transforms take contiguous ranges (every caller in the
// reference passes vectors) and run on the GPU; where the reference asserts
// on the arguments, a non-OK status of the library throws GSHTrans::Error.
//
// Differences that the C ABI fixes: Real must be double (the library is
// FP64), and the FFTW planner flag of the reference's constructor is
// accepted and ignored.  The random coefficient generators use a seeded
// counter-based generator instead of an unseeded std::random_device.
//
// Needs C++20 and linking with gshtrans_amd/libgshtrans_hip.so.
#ifndef GSHTRANS_AMD_GAUSS_LEGENDRE_GRID_H
#define GSHTRANS_AMD_GAUSS_LEGENDRE_GRID_H

#include <cassert>
#include <cmath>
#include <complex>
#include <concepts>
#include <cstddef>
#include <cstdint>
#include <memory>
#include <numbers>
#include <ranges>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <vector>

#include "../gshtrans_amd.h"
#include "Indexing.h"

namespace GSHTrans {

template <typename T>
struct IsComplex : std::false_type {};
template <typename T>
struct IsComplex<std::complex<T>> : std::true_type {};

// Thrown when the library reports an error (the reference asserts).
class Error : public std::runtime_error {
 public:
  Error(int status, const std::string& where)
      : std::runtime_error(where + ": " + gsh_status_string(status)), status{status} {}
  int status;
};

inline void Check(int status, const char* where) {
  if (status != GSH_OK) throw Error(status, where);
}

// A field or coefficient array held in device memory
// (DeviceCanonicalComponentField / DeviceCanonicalComponentExpansion in
// DeviceCanonicalComponentField.h): the transforms take its device pointer
// without host staging.
template <typename T>
concept DeviceResident = requires(T& t, const T& ct) {
  requires T::on_device;
  typename T::Scalar;
  { t.DeviceData() } -> std::same_as<typename T::Scalar*>;
  { ct.DeviceData() } -> std::same_as<const typename T::Scalar*>;
  { ct.Size() } -> std::convertible_to<std::ptrdiff_t>;
};

// A pointwise expression of device-resident fields (DeviceFieldExpression in
// DeviceCanonicalComponentField.h), analysed without being formed.
template <typename T>
concept DeviceExpression = requires(const T& e) {
  requires T::device_expression;
  typename T::Scalar;
  { e.Operation() } -> std::convertible_to<int>;
  { e.InputKind() } -> std::convertible_to<int>;
  { e.First() } -> std::convertible_to<const void*>;
  { e.Second() } -> std::convertible_to<const void*>;
};

template <std::floating_point Real_ = double, OrderIndexRange MRange_ = All,
          IndexRange NRange_ = All>
class GaussLegendreGrid {
  static_assert(std::same_as<Real_, double>, "the MI355X library computes in FP64");

 public:
  using Int = std::ptrdiff_t;
  using Real = Real_;
  using Complex = std::complex<Real>;
  using MRange = MRange_;
  using NRange = NRange_;

  GaussLegendreGrid() = default;

  // GaussLegendreGrid.h:38-73.  The planner flag of the reference is
  // accepted for source compatibility and has no effect here.
  template <typename Flag = int>
  GaussLegendreGrid(Int lMax, Int nMax, Flag = Flag{}) : _lMax{lMax}, _nMax{nMax} {
    gsh_grid_t h = nullptr;
    constexpr int mr = std::same_as<MRange, All> ? GSH_ALL : GSH_NONNEGATIVE;
    constexpr int nr = std::same_as<NRange, All>           ? GSH_ALL
                       : std::same_as<NRange, NonNegative> ? GSH_NONNEGATIVE
                                                           : GSH_SINGLE;
    Check(gsh_grid_create(static_cast<int>(lMax), static_cast<int>(nMax), mr, nr, &h),
          "GaussLegendreGrid");
    _h = std::shared_ptr<gsh_grid_s>(h, [](gsh_grid_s* p) { gsh_grid_destroy(p); });
    _theta.resize(lMax + 1);
    _w.resize(lMax + 1);
    Check(gsh_grid_colatitudes(h, _theta.data(), _w.data()), "GaussLegendreGrid");
  }

  // GaussLegendreGrid.h:86-103 and GridBase.h:17-113.
  Int MaxDegree() const { return _lMax; }
  Int MaxUpperIndex() const { return _nMax; }
  Int MinUpperIndex() const {
    if constexpr (std::same_as<NRange, All>) return -_nMax;
    if constexpr (std::same_as<NRange, NonNegative>) return 0;
    if constexpr (std::same_as<NRange, Single>) return _nMax;
  }
  auto UpperIndices() const { return std::views::iota(MinUpperIndex(), _nMax + 1); }
  auto CoLatitudes() const { return std::views::all(_theta); }
  auto CoLatitudeWeights() const { return std::views::all(_w); }
  auto Longitudes() const {
    const Real dPhi = 2 * std::numbers::pi_v<Real> / static_cast<Real>(2 * _lMax);
    return std::views::iota(Int{0}, 2 * _lMax) |
           std::views::transform([dPhi](Int i) { return static_cast<Real>(i) * dPhi; });
  }
  auto LongitudeWeights() const {
    const Real dPhi = 2 * std::numbers::pi_v<Real> / static_cast<Real>(2 * _lMax);
    return std::views::iota(Int{0}, 2 * _lMax) |
           std::views::transform([dPhi](Int) { return dPhi; });
  }
  Int NumberOfCoLatitudes() const { return _lMax + 1; }
  Int NumberOfLongitudes() const { return 2 * _lMax; }
  Int FieldSize() const { return NumberOfCoLatitudes() * NumberOfLongitudes(); }
  Int RealCoefficientSize(Int lMax, Int n) const {
    return GSHIndices<NonNegative>(lMax, lMax, n).Size();
  }
  Int ComplexCoefficientSize(Int lMax, Int n) const {
    return GSHIndices<All>(lMax, lMax, n).Size();
  }
  Int RealCoefficientSize(Int n) const { return RealCoefficientSize(_lMax, n); }
  Int ComplexCoefficientSize(Int n) const { return ComplexCoefficientSize(_lMax, n); }
  Int CoefficientSize(Int lMax, Int n) const { return ComplexCoefficientSize(lMax, n); }
  Int CoefficientSize(Int n) const { return ComplexCoefficientSize(_lMax, n); }
  Int CoefficientSizeNonNegative(Int lMax, Int n) const { return RealCoefficientSize(lMax, n); }
  Int CoefficientSizeNonNegative(Int n) const { return RealCoefficientSize(_lMax, n); }

  // ForwardTransformation (analysis), GaussLegendreGrid.h:108-217: out +=
  // A_n(in) for in of FieldSize() real or complex values and out of the
  // coefficient size, complex.
  template <std::ranges::contiguous_range InRange, std::ranges::contiguous_range OutRange>
    requires std::same_as<std::ranges::range_value_t<OutRange>, Complex> &&
             (std::same_as<std::ranges::range_value_t<InRange>, Real> ||
              std::same_as<std::ranges::range_value_t<InRange>, Complex>)
  void ForwardTransformation(Int lMax, Int n, InRange&& in, OutRange&& out) const {
    constexpr bool real = std::same_as<std::ranges::range_value_t<InRange>, Real>;
    const int kind = real ? GSH_REAL : GSH_COMPLEX;
    assert(static_cast<Int>(std::ranges::size(in)) == FieldSize());
    assert(static_cast<Int>(std::ranges::size(out)) ==
           (real ? RealCoefficientSize(lMax, n) : ComplexCoefficientSize(lMax, n)));
    Check(gsh_forward_host(_h.get(), static_cast<int>(lMax), static_cast<int>(n), kind,
                           reinterpret_cast<const double*>(std::ranges::data(in)),
                           reinterpret_cast<double*>(std::ranges::data(out)), 1),
          "ForwardTransformation");
  }

  // InverseTransformation (synthesis), GaussLegendreGrid.h:222-321.
  template <std::ranges::contiguous_range InRange, std::ranges::contiguous_range OutRange>
    requires std::same_as<std::ranges::range_value_t<InRange>, Complex> &&
             (std::same_as<std::ranges::range_value_t<OutRange>, Real> ||
              std::same_as<std::ranges::range_value_t<OutRange>, Complex>)
  void InverseTransformation(Int lMax, Int n, InRange&& in, OutRange&& out) const {
    constexpr bool real = std::same_as<std::ranges::range_value_t<OutRange>, Real>;
    const int kind = real ? GSH_REAL : GSH_COMPLEX;
    assert(static_cast<Int>(std::ranges::size(out)) == FieldSize());
    assert(static_cast<Int>(std::ranges::size(in)) ==
           (real ? RealCoefficientSize(lMax, n) : ComplexCoefficientSize(lMax, n)));
    Check(gsh_inverse_host(_h.get(), static_cast<int>(lMax), static_cast<int>(n), kind,
                           reinterpret_cast<const double*>(std::ranges::data(in)),
                           reinterpret_cast<double*>(std::ranges::data(out))),
          "InverseTransformation");
  }

  // The same two transforms on device-resident arrays: no host staging, the
  // work queued on `stream` (nullptr: the default stream); `precision` is a
  // gsh_precision (GSH_FP64, the reference's arithmetic, by default).  The
  // analysis accumulates into `out` as the reference does
  // (GaussLegendreGrid.h:192).  Call Synchronize(stream) before reading the
  // results on the host.
  template <DeviceResident In, DeviceResident Out>
    requires std::same_as<typename Out::Scalar, Complex> &&
             (std::same_as<typename In::Scalar, Real> || std::same_as<typename In::Scalar, Complex>)
  void ForwardTransformation(Int lMax, Int n, const In& in, Out& out, int precision = GSH_FP64,
                             void* stream = nullptr) const {
    constexpr bool real = std::same_as<typename In::Scalar, Real>;
    if (in.Size() != FieldSize() ||
        out.Size() != (real ? RealCoefficientSize(lMax, n) : ComplexCoefficientSize(lMax, n)))
      throw Error(GSH_ERR_INVALID_ARGUMENT, "ForwardTransformation");
    Check(gsh_forward(_h.get(), static_cast<int>(lMax), static_cast<int>(n),
                      real ? GSH_REAL : GSH_COMPLEX, precision, GSH_STORE_FP64, 1,
                      in.DeviceData(), 0, out.DeviceData(), 0, 1, stream),
          "ForwardTransformation");
  }
  template <DeviceResident In, DeviceResident Out>
    requires std::same_as<typename In::Scalar, Complex> &&
             (std::same_as<typename Out::Scalar, Real> || std::same_as<typename Out::Scalar, Complex>)
  void InverseTransformation(Int lMax, Int n, const In& in, Out& out, int precision = GSH_FP64,
                             void* stream = nullptr) const {
    constexpr bool real = std::same_as<typename Out::Scalar, Real>;
    if (out.Size() != FieldSize() ||
        in.Size() != (real ? RealCoefficientSize(lMax, n) : ComplexCoefficientSize(lMax, n)))
      throw Error(GSH_ERR_INVALID_ARGUMENT, "InverseTransformation");
    Check(gsh_inverse(_h.get(), static_cast<int>(lMax), static_cast<int>(n),
                      real ? GSH_REAL : GSH_COMPLEX, precision, GSH_STORE_FP64, 1,
                      in.DeviceData(), 0, out.DeviceData(), 0, stream),
          "InverseTransformation");
  }
  // ForwardTransformation of an expression of device-resident fields
  // (a * b, a + b, conj(a), real(a), s * a, ...): the expression is formed as
  // the analysis ring FFT loads its rings, and the result accumulates into
  // `out` as the reference's analysis does.
  template <DeviceExpression Expr, DeviceResident Out>
    requires std::same_as<typename Out::Scalar, Complex>
  void ForwardTransformation(Int lMax, Int n, const Expr& e, Out& out,
                             int precision = GSH_FP64, void* stream = nullptr) const {
    constexpr bool real = std::same_as<typename Expr::Scalar, Real>;
    if (out.Size() != (real ? RealCoefficientSize(lMax, n) : ComplexCoefficientSize(lMax, n)))
      throw Error(GSH_ERR_INVALID_ARGUMENT, "ForwardTransformation");
    const Complex s = e.ScalarOperand();
    Check(gsh_forward_expr(_h.get(), static_cast<int>(lMax), static_cast<int>(n), e.InputKind(),
                           e.Operation(), precision, 1, e.First(), e.Second(), 0, s.real(),
                           s.imag(), out.DeviceData(), 0, 1, stream),
          "ForwardTransformation");
  }
  void Synchronize(void* stream = nullptr) const {
    Check(gsh_synchronize(_h.get(), stream), "Synchronize");
  }

  // GridBase.h:115-172, with a seeded splitmix64 -> Box-Muller generator (the
  // same stream as gshtrans_amd.random_coefficients(..., decay=0)) instead of
  // std::random_device; the reference's masks are kept.
  template <std::ranges::contiguous_range Range>
  void RandomComplexCoefficient(Int lMax, Int n, Range&& range, std::uint64_t seed = 0) const {
    assert(static_cast<Int>(std::ranges::size(range)) == ComplexCoefficientSize(lMax, n));
    Fill(range, seed);
    if (lMax == _lMax) range[GSHIndices<All>(lMax, lMax, n).Index(lMax, lMax)] = 0;
  }
  template <std::ranges::contiguous_range Range>
  void RandomRealCoefficient(Int lMax, Int n, Range&& range, std::uint64_t seed = 0) const {
    assert(static_cast<Int>(std::ranges::size(range)) == RealCoefficientSize(lMax, n));
    Fill(range, seed);
    const auto idx = GSHIndices<NonNegative>(lMax, lMax, n);
    for (Int l = std::abs(n); l <= lMax; l++) range[idx.Index(l, 0)].imag(0);
    if (lMax == _lMax) range[idx.Index(lMax, lMax)].imag(0);
  }

  // The library handle, for the device-pointer entry points (gsh_inverse,
  // gsh_forward, gsh_adjoint_inverse, gsh_plan_*).
  gsh_grid_t Handle() const { return _h.get(); }

 private:
  Int _lMax = 0, _nMax = 0;
  std::shared_ptr<gsh_grid_s> _h;
  std::vector<Real> _theta, _w;

  template <typename Range>
  static void Fill(Range& range, std::uint64_t seed) {
    auto mix = [](std::uint64_t z) {
      z = (z ^ (z >> 30)) * 0xbf58476d1ce4e5b9ULL;
      z = (z ^ (z >> 27)) * 0x94d049bb133111ebULL;
      return z ^ (z >> 31);
    };
    constexpr std::uint64_t golden = 0x9e3779b97f4a7c15ULL;
    const std::uint64_t base = mix(seed + golden);
    const Int size = static_cast<Int>(std::ranges::size(range));
    for (Int i = 0; i < size; i++) {
      const std::uint64_t j = 2 * static_cast<std::uint64_t>(i);
      const Real u1 = static_cast<Real>(mix(base + (j + 1) * golden) >> 11) * 0x1.0p-53;
      const Real u2 = static_cast<Real>(mix(base + (j + 2) * golden) >> 11) * 0x1.0p-53;
      const Real r = std::sqrt(-2 * std::log1p(-u1));
      range[i] = Complex{r * std::cos(2 * std::numbers::pi_v<Real> * u2),
                         r * std::sin(2 * std::numbers::pi_v<Real> * u2)};
    }
  }
};

}  // namespace GSHTrans

#endif
