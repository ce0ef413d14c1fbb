// GSHTrans/Views.h -- views of one GSHIndices-ordered block of values (a
// Wigner table block of one (n, theta), or a coefficient array), with the
// reference's interface (GSHTrans/src/Views.h:11-88):
//
//   GSHView<Scalar, MRange>(lMax, mMax, n, start)   whole block; v[l] is the
//       degree view, v[l, m] (C++23) or v(l, m) the value; begin()/end()
//       span the block
//   GSHSubView<Scalar, MRange>(l, mMax, start)      one degree; s[m] is the
//       value of order m, s.Orders() the orders, begin()/end() its values
//
// ConstGSHView / ConstGSHSubView are the read-only forms (Scalar const).
#ifndef GSHTRANS_AMD_VIEWS_H
#define GSHTRANS_AMD_VIEWS_H

#include <cstddef>
#include <type_traits>

#include "Indexing.h"

namespace GSHTrans {

template <typename Scalar, OrderIndexRange MRange>
class GSHSubView : public GSHSubIndices<MRange> {
 public:
  using Int = std::ptrdiff_t;
  constexpr GSHSubView(Int l, Int mMax, Scalar* start)
      : GSHSubIndices<MRange>(l, mMax), _p{start} {}
  constexpr Scalar* begin() const { return _p; }
  constexpr Scalar* end() const { return _p + this->Size(); }
  constexpr Scalar& operator[](Int m) const { return _p[this->Index(m)]; }

 private:
  Scalar* _p;
};

template <typename Scalar, OrderIndexRange MRange>
class GSHView : public GSHIndices<MRange> {
 public:
  using Int = std::ptrdiff_t;
  constexpr GSHView(Int lMax, Int mMax, Int n, Scalar* start)
      : GSHIndices<MRange>(lMax, mMax, n), _p{start} {}
  constexpr Scalar* begin() const { return _p; }
  constexpr Scalar* end() const { return _p + this->Size(); }
  constexpr GSHSubView<Scalar, MRange> operator[](Int l) const {
    return GSHSubView<Scalar, MRange>(l, this->MaxOrder(), _p + this->OffsetForDegree(l));
  }
#ifdef __cpp_multidimensional_subscript
  constexpr Scalar& operator[](Int l, Int m) const { return _p[this->Index(l, m)]; }
#endif
  constexpr Scalar& operator()(Int l, Int m) const { return _p[this->Index(l, m)]; }

 private:
  Scalar* _p;
};

template <typename Scalar, OrderIndexRange MRange>
using ConstGSHSubView = GSHSubView<const std::remove_const_t<Scalar>, MRange>;
template <typename Scalar, OrderIndexRange MRange>
using ConstGSHView = GSHView<const std::remove_const_t<Scalar>, MRange>;

}  // namespace GSHTrans

#endif
