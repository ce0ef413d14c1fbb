// GSHTrans/Wigner.h -- the reference's Wigner d table class
// (GSHTrans/src/Wigner.h:99-286) over the MI355X library: the same template
// options, constructors, accessors and block layout, the values computed on
// the GPU (gsh_wigner_table_host) into host storage as the reference keeps
// them.
//
//   Norm        Ortho (times sqrt(2l+1) / (2 sqrt(pi)), Wigner.h:475-485) or FourPi
//   MRange      All or NonNegative orders, |m| <= min(l, mMax)
//   NRange      All (-nMax..nMax), NonNegative (0..nMax) or Single (nMax)
//   AngleRange  Single (one angle) or Multiple
//   Storage     ColumnMajor (blocks n outer) or RowMajor (theta outer)
//
// Block access as in the reference (Wigner.h:240-265, C++23):
//   d[n, iTheta]   the (n, theta_i) block, a ConstGSHView: d[n, iTheta][l] is
//                  degree l's view (Orders(), begin()/end(), [m]) and
//                  d[n, iTheta][l, m] the value
//   d[n]           one angle, several upper indices: the block of n
//   d[iTheta]      one upper index, several angles: the block of theta_i
//   d[l]           one angle and one upper index: degree l's view
// operator() takes the same arguments for C++20 callers.  The one-index
// forms with NRange = Single use the table's upper index nMax (the
// reference passes n = 0 there, which is the same block only for nMax = 0).
// MaxOrder(l) returns min(l, mMax) (the reference's std::max there makes
// Orders(l) wrong for l < mMax; its own views use the minimum,
// Indexing.h:24-27).
#ifndef GSHTRANS_AMD_WIGNER_H
#define GSHTRANS_AMD_WIGNER_H

#include <algorithm>
#include <cstddef>
#include <ranges>
#include <vector>

#include "GaussLegendreGrid.h"
#include "Views.h"

namespace GSHTrans {

struct Multiple {};
struct ColumnMajor {};
struct RowMajor {};

template <std::floating_point Real_ = double, typename Norm_ = Ortho,
          OrderIndexRange MRange_ = All, IndexRange NRange_ = Single,
          typename AngleRange_ = Single, typename Storage_ = ColumnMajor>
class Wigner {
  static_assert(std::same_as<Real_, double>, "the MI355X library computes in FP64");
  static_assert(std::same_as<Norm_, Ortho> || std::same_as<Norm_, FourPi>);
  static_assert(std::same_as<AngleRange_, Single> || std::same_as<AngleRange_, Multiple>);
  static_assert(std::same_as<Storage_, ColumnMajor> || std::same_as<Storage_, RowMajor>);

 public:
  using Int = std::ptrdiff_t;
  using Real = Real_;
  using Norm = Norm_;
  using MRange = MRange_;
  using NRange = NRange_;
  using AngleRange = AngleRange_;
  using Storage = Storage_;

  Wigner() = default;

  // Wigner.h:116-133: the table for the angles `theta`.
  template <std::ranges::range Range>
    requires std::floating_point<std::ranges::range_value_t<Range>>
  Wigner(Int lMax, Int mMax, Int nMax, Range&& theta)
      : _lMax{lMax}, _mMax{mMax}, _nMax{nMax}, _nTheta{static_cast<Int>(std::ranges::size(theta))} {
    const long size = gsh_wigner_size(static_cast<int>(_lMax), static_cast<int>(_mMax),
                                      static_cast<int>(MinUpperIndex()), static_cast<int>(_nMax),
                                      static_cast<int>(_nTheta), Mr());
    if (size < 0) throw Error(GSH_ERR_INVALID_ARGUMENT, "Wigner");
    _data.resize(static_cast<std::size_t>(size));
    ReCompute(std::forward<Range>(theta));
  }
  // Wigner.h:135-138: one angle.
  Wigner(Int lMax, Int mMax, Int nMax, Real theta)
    requires std::same_as<AngleRange, Single>
      : Wigner(lMax, mMax, nMax, std::vector<Real>{theta}) {}

  // Wigner.h:142-152: the same table for new angles.
  template <std::ranges::range Range>
    requires std::floating_point<std::ranges::range_value_t<Range>>
  void ReCompute(Range&& theta) {
    std::vector<double> th(std::ranges::begin(theta), std::ranges::end(theta));
    if (static_cast<Int>(th.size()) != _nTheta) throw Error(GSH_ERR_INVALID_ARGUMENT, "ReCompute");
    Check(gsh_wigner_table_host(static_cast<int>(_lMax), static_cast<int>(_mMax),
                                static_cast<int>(MinUpperIndex()), static_cast<int>(_nMax),
                                static_cast<int>(_nTheta), th.data(), Mr(),
                                std::same_as<Norm, Ortho> ? GSH_ORTHO : GSH_FOURPI,
                                std::same_as<Storage, ColumnMajor> ? GSH_COLUMN_MAJOR : GSH_ROW_MAJOR,
                                _data.data()),
          "Wigner");
  }
  void ReCompute(Real theta)
    requires std::same_as<AngleRange, Single>
  {
    ReCompute(std::vector<Real>{theta});
  }

  // Wigner.h:155-232.
  Int MinDegree(Int n) const { return n < 0 ? -n : n; }
  Int MinDegree() const
    requires std::same_as<NRange, Single>
  {
    return MinDegree(_nMax);
  }
  Int MaxDegree() const { return _lMax; }
  auto Degrees(Int n) const { return std::views::iota(MinDegree(n), _lMax + 1); }
  auto Degrees() const
    requires std::same_as<NRange, Single>
  {
    return Degrees(_nMax);
  }
  Int MinOrder(Int l) const {
    if constexpr (std::same_as<MRange, All>) return -std::min(l, _mMax);
    return 0;
  }
  Int MaxOrder() const { return _mMax; }
  Int MaxOrder(Int l) const { return std::min(l, _mMax); }
  auto Orders(Int l) const { return std::views::iota(MinOrder(l), MaxOrder(l) + 1); }
  Int MinUpperIndex() const {
    if constexpr (std::same_as<NRange, All>) return -_nMax;
    if constexpr (std::same_as<NRange, NonNegative>) return 0;
    return _nMax;
  }
  Int MaxUpperIndex() const { return _nMax; }
  auto UpperIndices() const { return std::views::iota(MinUpperIndex(), _nMax + 1); }
  Int NumberOfUpperIndices() const { return _nMax - MinUpperIndex() + 1; }
  Int NumberOfAngles() const { return _nTheta; }
  auto AngleIndices() const { return std::views::iota(Int{0}, _nTheta); }

  // Block of (n, theta_i) (Wigner.h:240-265).
  using Block = ConstGSHView<Real, MRange>;
  Block operator()(Int n, Int iTheta) const {
    const long o = gsh_wigner_offset(static_cast<int>(_lMax), static_cast<int>(_mMax),
                                     static_cast<int>(MinUpperIndex()), static_cast<int>(_nMax),
                                     static_cast<int>(_nTheta), Mr(),
                                     std::same_as<Storage, ColumnMajor> ? GSH_COLUMN_MAJOR
                                                                        : GSH_ROW_MAJOR,
                                     static_cast<int>(n), static_cast<int>(iTheta));
    if (o < 0) throw Error(GSH_ERR_INVALID_ARGUMENT, "Wigner block");
    return Block(_lMax, _mMax, n, _data.data() + o);
  }
  // One index: the block of n (one angle), of theta_k (one upper index).
  Block operator()(Int k) const
    requires(std::same_as<AngleRange, Single> != std::same_as<NRange, Single>)
  {
    if constexpr (std::same_as<AngleRange, Single>) return operator()(k, 0);
    return operator()(_nMax, k);
  }
  // One angle and one upper index: the view of degree l.
  ConstGSHSubView<Real, MRange> operator()(Int l) const
    requires(std::same_as<AngleRange, Single> && std::same_as<NRange, Single>)
  {
    return operator()(_nMax, 0)[l];
  }
#ifdef __cpp_multidimensional_subscript
  Block operator[](Int n, Int iTheta) const { return operator()(n, iTheta); }
#endif
  Block operator[](Int k) const
    requires(std::same_as<AngleRange, Single> != std::same_as<NRange, Single>)
  {
    return operator()(k);
  }
  ConstGSHSubView<Real, MRange> operator[](Int l) const
    requires(std::same_as<AngleRange, Single> && std::same_as<NRange, Single>)
  {
    return operator()(l);
  }
  const std::vector<Real>& Data() const { return _data; }

 private:
  Int _lMax = 0, _mMax = 0, _nMax = 0, _nTheta = 0;
  std::vector<Real> _data;
  static constexpr int Mr() { return std::same_as<MRange, All> ? GSH_ALL : GSH_NONNEGATIVE; }
};

}  // namespace GSHTrans

#endif
