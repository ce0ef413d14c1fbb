// GSHTrans/Indexing.h -- coefficient index layouts of the C++ facade: the
// reference's GSHSubIndices (one degree's orders) and GSHIndices (degrees
// |n| .. lMax of upper index n, orders |m| <= min(l, mMax)), with the same
// member functions and the same offsets (GSHTrans/src/Indexing.h:22-150).
//
// The layout is fixed by the data the transforms exchange: degree blocks in
// ascending l; within a block, orders ascending from -min(l, mMax) (All) or 0
// (NonNegative).  Indices() lists the (l, m) pairs in storage order.
#ifndef GSHTRANS_AMD_INDEXING_H
#define GSHTRANS_AMD_INDEXING_H

#include <algorithm>
#include <cassert>
#include <concepts>
#include <cstddef>
#include <cstdlib>
#include <ranges>
#include <utility>
#include <vector>

namespace GSHTrans {

// Index-range, normalisation and value tags (GSHTrans/src/Concepts.h:24-54).
struct All {};
struct NonNegative {};
struct Single {};
struct Ortho {};
struct FourPi {};
struct RealValued {};
struct ComplexValued {};

template <typename T>
concept IndexRange =
    std::same_as<T, All> or std::same_as<T, NonNegative> or std::same_as<T, Single>;
template <typename T>
concept OrderIndexRange = std::same_as<T, All> or std::same_as<T, NonNegative>;
template <typename T>
concept RealOrComplexValued = std::same_as<T, RealValued> or std::same_as<T, ComplexValued>;

// Orders of one degree l: -min(l, mMax) .. min(l, mMax) (All) or
// 0 .. min(l, mMax) (NonNegative), stored contiguously in that order.
template <OrderIndexRange MRange>
class GSHSubIndices {
 public:
  using Int = std::ptrdiff_t;
  constexpr GSHSubIndices(Int l, Int mMax) : _l{l}, _top{std::min(l, mMax)} {
    assert(l >= 0 && mMax >= 0);
  }
  constexpr Int Degree() const { return _l; }
  constexpr Int MinOrder() const {
    if constexpr (std::same_as<MRange, All>)
      return -_top;
    else
      return Int{0};
  }
  constexpr Int MaxOrder() const { return _top; }
  constexpr auto Orders() const { return std::views::iota(MinOrder(), MaxOrder() + 1); }
  constexpr auto NegativeOrders() const { return std::views::iota(MinOrder(), Int{0}); }
  constexpr auto NonNegativeOrders() const { return std::views::iota(Int{0}, MaxOrder() + 1); }
  constexpr Int Size() const { return MaxOrder() - MinOrder() + 1; }
  // Position of order m within the degree's block.
  constexpr Int Index(Int m) const {
    assert(m >= MinOrder() && m <= MaxOrder());
    return m - MinOrder();
  }

 private:
  Int _l, _top;
};

// Degrees |n| .. lMax of upper index n, each a GSHSubIndices block.
template <OrderIndexRange MRange>
class GSHIndices {
 public:
  using Int = std::ptrdiff_t;
  GSHIndices() = default;
  constexpr GSHIndices(Int lMax, Int mMax, Int n)
      : _lMax{lMax}, _mMax{std::min(lMax, mMax)}, _n{n} {
    assert(lMax >= 0 && mMax >= 0 && std::abs(n) <= lMax);
  }
  constexpr Int UpperIndex() const { return _n; }
  constexpr Int MaxOrder() const { return _mMax; }
  constexpr Int MinDegree() const { return _n < 0 ? -_n : _n; }
  constexpr Int MaxDegree() const { return _lMax; }
  constexpr auto Degrees() const { return std::views::iota(MinDegree(), MaxDegree() + 1); }
  // Start of degree l's block: the blocks below it hold (l' + 1) or
  // (2 l' + 1) orders while l' <= mMax and mMax + 1 or 2 mMax + 1 after.
  constexpr Int OffsetForDegree(Int l) const {
    assert(l >= MinDegree() && l <= MaxDegree());
    const Int lo = MinDegree(), mm = _mMax;
    if constexpr (std::same_as<MRange, All>) {
      if (mm < lo) return (l - lo) * (2 * mm + 1);
      if (l <= mm) return l * l - lo * lo;
      return (mm + 1) * (mm + 1) - lo * lo + (l - mm - 1) * (2 * mm + 1);
    } else {
      if (mm < lo) return (l - lo) * (mm + 1);
      if (l <= mm) return l * (l + 1) / 2 - lo * (lo + 1) / 2;
      return (mm + 1) * (mm + 2) / 2 - lo * (lo + 1) / 2 + (l - mm - 1) * (mm + 1);
    }
  }
  constexpr Int SizeForDegree(Int l) const { return GSHSubIndices<MRange>(l, _mMax).Size(); }
  constexpr std::pair<Int, GSHSubIndices<MRange>> Index(Int l) const {
    return {OffsetForDegree(l), GSHSubIndices<MRange>(l, _mMax)};
  }
  constexpr Int Index(Int l, Int m) const {
    return OffsetForDegree(l) + GSHSubIndices<MRange>(l, _mMax).Index(m);
  }
  constexpr Int Size() const { return OffsetForDegree(_lMax) + SizeForDegree(_lMax); }
  // The (l, m) pairs in storage order.
  std::vector<std::pair<Int, Int>> Indices() const {
    std::vector<std::pair<Int, Int>> out;
    out.reserve(static_cast<std::size_t>(Size()));
    for (Int l = MinDegree(); l <= _lMax; l++)
      for (Int m : GSHSubIndices<MRange>(l, _mMax).Orders()) out.emplace_back(l, m);
    return out;
  }

 private:
  Int _lMax = 0, _mMax = 0, _n = 0;
};

}  // namespace GSHTrans

#endif
