// GSHTrans/CanonicalComponentField.h -- field and coefficient containers of
// the C++ facade (the reference's CanonicalComponentField,
// GSHTrans/src/CanonicalComponentField/CanonicalComponentField.h:39-125, and
// CanonicalComponentExpansion, GSHTrans/src/CanonicalComponentExpansion/
// CanonicalComponentExpansion.h:21-78): host vectors in the reference's
// layouts whose Data() feeds the grid's transforms.  Element access is the
// reference's f[iTheta, iPhi] and u[l, m] (C++23 multi-argument operator[],
// CanonicalComponentField.h:62-71, CanonicalComponentExpansion.h:63-66), or
// f(iTheta, iPhi) and u(l, m) for C++20 callers.
#ifndef GSHTRANS_AMD_CANONICAL_COMPONENT_FIELD_H
#define GSHTRANS_AMD_CANONICAL_COMPONENT_FIELD_H

#include <cassert>
#include <cstddef>
#include <ranges>
#include <type_traits>
#include <vector>

#include "GaussLegendreGrid.h"

namespace GSHTrans {

template <std::ptrdiff_t N_, typename Grid_, RealOrComplexValued Value_>
class CanonicalComponentField {
 public:
  using Int = std::ptrdiff_t;
  using Value = Value_;
  using Real = typename Grid_::Real;
  using Complex = typename Grid_::Complex;
  using Scalar = std::conditional_t<std::same_as<Value, RealValued>, Real, Complex>;

  explicit CanonicalComponentField(const Grid_& grid)
      : _grid{&grid}, _data(static_cast<std::size_t>(grid.FieldSize())) {}
  template <typename Function>
  CanonicalComponentField(const Grid_& grid, Function&& f) : CanonicalComponentField(grid) {
    Int k = 0;
    for (auto theta : grid.CoLatitudes())
      for (auto phi : grid.Longitudes()) _data[k++] = f(theta, phi);
  }

  static constexpr Int UpperIndex() { return N_; }
  const Grid_& Grid() const { return *_grid; }
  Int NumberOfCoLatitudes() const { return _grid->NumberOfCoLatitudes(); }
  Int NumberOfLongitudes() const { return _grid->NumberOfLongitudes(); }
  Int Size() const { return static_cast<Int>(_data.size()); }
#ifdef __cpp_multidimensional_subscript
  Scalar operator[](Int iTheta, Int iPhi) const { return _data[Index(iTheta, iPhi)]; }
  Scalar& operator[](Int iTheta, Int iPhi) { return _data[Index(iTheta, iPhi)]; }
#endif
  Scalar operator()(Int iTheta, Int iPhi) const { return _data[Index(iTheta, iPhi)]; }
  Scalar& operator()(Int iTheta, Int iPhi) { return _data[Index(iTheta, iPhi)]; }
  auto Data() { return std::views::all(_data); }
  auto Data() const { return std::views::all(_data); }

 private:
  const Grid_* _grid;
  std::vector<Scalar> _data;
  Int Index(Int iTheta, Int iPhi) const {
    assert(iTheta >= 0 && iTheta < NumberOfCoLatitudes() && iPhi >= 0 &&
           iPhi < NumberOfLongitudes());
    return iTheta * NumberOfLongitudes() + iPhi;
  }
};

template <std::ptrdiff_t N, typename Grid>
using RealCanonicalComponentField = CanonicalComponentField<N, Grid, RealValued>;
template <std::ptrdiff_t N, typename Grid>
using ComplexCanonicalComponentField = CanonicalComponentField<N, Grid, ComplexValued>;

// Coefficients c_lm, l = |N| .. lMax, in GSHIndices<All> order for complex
// fields and GSHIndices<NonNegative> order for real fields.
template <std::ptrdiff_t N_, typename Grid_, RealOrComplexValued Value_>
class CanonicalComponentExpansion
    : public GSHIndices<std::conditional_t<std::same_as<Value_, RealValued>, NonNegative, All>> {
 public:
  using Int = std::ptrdiff_t;
  using Value = Value_;
  using MRange = std::conditional_t<std::same_as<Value, RealValued>, NonNegative, All>;
  using Real = typename Grid_::Real;
  using Complex = typename Grid_::Complex;

  explicit CanonicalComponentExpansion(const Grid_& grid)
      : GSHIndices<MRange>(grid.MaxDegree(), grid.MaxDegree(), N_),
        _grid{&grid},
        _data(static_cast<std::size_t>(this->Size())) {}

  static constexpr Int UpperIndex() { return N_; }
  const Grid_& Grid() const { return *_grid; }
#ifdef __cpp_multidimensional_subscript
  Complex operator[](Int l, Int m) const { return _data[this->Index(l, m)]; }
  Complex& operator[](Int l, Int m) { return _data[this->Index(l, m)]; }
#endif
  Complex operator()(Int l, Int m) const { return _data[this->Index(l, m)]; }
  Complex& operator()(Int l, Int m) { return _data[this->Index(l, m)]; }
  auto Data() { return std::views::all(_data); }
  auto Data() const { return std::views::all(_data); }

 private:
  const Grid_* _grid;
  std::vector<Complex> _data;
};

template <std::ptrdiff_t N, typename Grid>
using RealCanonicalComponentExpansion = CanonicalComponentExpansion<N, Grid, RealValued>;
template <std::ptrdiff_t N, typename Grid>
using ComplexCanonicalComponentExpansion = CanonicalComponentExpansion<N, Grid, ComplexValued>;

}  // namespace GSHTrans

#endif
