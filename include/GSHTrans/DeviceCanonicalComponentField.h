// GSHTrans/DeviceCanonicalComponentField.h -- device-resident field and
// coefficient containers of the C++ facade: the storage of the reference's
// CanonicalComponentField (GSHTrans/src/CanonicalComponentField/
// CanonicalComponentField.h:74-81, 119-124) and CanonicalComponentExpansion
// (GSHTrans/src/CanonicalComponentExpansion/CanonicalComponentExpansion.h:
// 69-78) in the same layouts, held in device memory of the grid's GPU.  The
// grid's ForwardTransformation / InverseTransformation take them without
// host staging (GaussLegendreGrid::Synchronize before reading results on the
// host); Upload / Download / ToHost copy to and from the host containers.
#ifndef GSHTRANS_AMD_DEVICE_CANONICAL_COMPONENT_FIELD_H
#define GSHTRANS_AMD_DEVICE_CANONICAL_COMPONENT_FIELD_H

#include <cstddef>
#include <type_traits>
#include <utility>

#include "CanonicalComponentField.h"

namespace GSHTrans {

namespace detail {

// Owner of one zero-filled allocation on a grid's GPU (move-only).
class DeviceBuffer {
 public:
  DeviceBuffer() = default;
  DeviceBuffer(gsh_grid_t h, std::size_t bytes) : _h{h}, _bytes{bytes} {
    Check(gsh_device_alloc(h, bytes, &_p), "device allocation");
  }
  DeviceBuffer(const DeviceBuffer&) = delete;
  DeviceBuffer& operator=(const DeviceBuffer&) = delete;
  DeviceBuffer(DeviceBuffer&& o) noexcept
      : _h{o._h}, _p{std::exchange(o._p, nullptr)}, _bytes{o._bytes} {}
  DeviceBuffer& operator=(DeviceBuffer&& o) noexcept {
    std::swap(_h, o._h);
    std::swap(_p, o._p);
    std::swap(_bytes, o._bytes);
    return *this;
  }
  ~DeviceBuffer() {
    if (_p) (void)gsh_device_free(_h, _p);
  }
  void* Get() const { return _p; }
  void Upload(const void* src) { Check(gsh_copy(_h, _p, src, _bytes), "Upload"); }
  void Download(void* dst) const { Check(gsh_copy(_h, dst, _p, _bytes), "Download"); }

 private:
  gsh_grid_t _h = nullptr;
  void* _p = nullptr;
  std::size_t _bytes = 0;
};

}  // namespace detail

template <std::ptrdiff_t N_, typename Grid_, RealOrComplexValued Value_>
class DeviceCanonicalComponentField {
 public:
  static constexpr bool on_device = true;
  using Int = std::ptrdiff_t;
  using Value = Value_;
  using Real = typename Grid_::Real;
  using Complex = typename Grid_::Complex;
  using Scalar = std::conditional_t<std::same_as<Value, RealValued>, Real, Complex>;
  using Host = CanonicalComponentField<N_, Grid_, Value_>;

  // Zero field on the grid's GPU.
  explicit DeviceCanonicalComponentField(const Grid_& grid)
      : _grid{&grid},
        _size{grid.FieldSize()},
        _buf(grid.Handle(), sizeof(Scalar) * static_cast<std::size_t>(_size)) {}
  // A copy of a host field.
  explicit DeviceCanonicalComponentField(const Host& host)
      : DeviceCanonicalComponentField(host.Grid()) {
    Upload(host);
  }

  static constexpr Int UpperIndex() { return N_; }
  const Grid_& Grid() const { return *_grid; }
  Int Size() const { return _size; }
  Scalar* DeviceData() { return static_cast<Scalar*>(_buf.Get()); }
  const Scalar* DeviceData() const { return static_cast<const Scalar*>(_buf.Get()); }
  void Upload(const Host& host) {
    Check(host.Size() == _size ? GSH_OK : GSH_ERR_INVALID_ARGUMENT, "Upload");
    _buf.Upload(host.Data().data());
  }
  void Download(Host& host) const {
    Check(host.Size() == _size ? GSH_OK : GSH_ERR_INVALID_ARGUMENT, "Download");
    _buf.Download(host.Data().data());
  }
  Host ToHost() const {
    Host h(*_grid);
    Download(h);
    return h;
  }

 private:
  const Grid_* _grid;
  Int _size;
  detail::DeviceBuffer _buf;
};

// A pointwise expression of device-resident fields, as the reference builds
// expression fields from fields (CanonicalComponentFieldOverloads.h:20-266:
// sums, differences, products, quotients, negation, conj / real / imag,
// scalar multiples and shifts).  It holds the operands only;
// GaussLegendreGrid::ForwardTransformation analyses it with the expression
// applied as the ring FFT loads its rings (gsh_forward_expr), so no
// expression field is formed in memory.  N_ and Value_ are the expression's
// upper index and value type (real for real / imag of complex fields).
template <std::ptrdiff_t N_, typename Grid_, RealOrComplexValued Value_>
class DeviceFieldExpression {
 public:
  static constexpr bool device_expression = true;
  using Int = std::ptrdiff_t;
  using Value = Value_;
  using Real = typename Grid_::Real;
  using Complex = typename Grid_::Complex;
  using Scalar = std::conditional_t<std::same_as<Value, RealValued>, Real, Complex>;

  DeviceFieldExpression(int op, const Grid_& grid, int input_kind, const void* a,
                        const void* b = nullptr, Complex s = Complex{0})
      : _op{op}, _kind{input_kind}, _grid{&grid}, _a{a}, _b{b}, _s{s} {}
  static constexpr Int UpperIndex() { return N_; }
  const Grid_& Grid() const { return *_grid; }
  int Operation() const { return _op; }
  int InputKind() const { return _kind; }  // GSH_REAL or GSH_COMPLEX operands
  const void* First() const { return _a; }
  const void* Second() const { return _b; }
  Complex ScalarOperand() const { return _s; }

 private:
  int _op, _kind;
  const Grid_* _grid;
  const void* _a;
  const void* _b;
  Complex _s;
};

namespace detail {
template <RealOrComplexValued V>
constexpr int kind_of = std::same_as<V, RealValued> ? GSH_REAL : GSH_COMPLEX;
}  // namespace detail

// CanonicalComponentFieldMultiply (upper index N0 + N1).
template <std::ptrdiff_t N0, std::ptrdiff_t N1, typename G, RealOrComplexValued V>
auto operator*(const DeviceCanonicalComponentField<N0, G, V>& a,
               const DeviceCanonicalComponentField<N1, G, V>& b) {
  return DeviceFieldExpression<N0 + N1, G, V>(GSH_EXPR_PRODUCT, a.Grid(), detail::kind_of<V>,
                                              a.DeviceData(), b.DeviceData());
}
// CanonicalComponentFieldAdd / Subtract (equal upper indices).
template <std::ptrdiff_t N, typename G, RealOrComplexValued V>
auto operator+(const DeviceCanonicalComponentField<N, G, V>& a,
               const DeviceCanonicalComponentField<N, G, V>& b) {
  return DeviceFieldExpression<N, G, V>(GSH_EXPR_SUM, a.Grid(), detail::kind_of<V>,
                                        a.DeviceData(), b.DeviceData());
}
template <std::ptrdiff_t N, typename G, RealOrComplexValued V>
auto operator-(const DeviceCanonicalComponentField<N, G, V>& a,
               const DeviceCanonicalComponentField<N, G, V>& b) {
  return DeviceFieldExpression<N, G, V>(GSH_EXPR_DIFFERENCE, a.Grid(), detail::kind_of<V>,
                                        a.DeviceData(), b.DeviceData());
}
// CanonicalComponentFieldDivide (a divisor of upper index 0).
template <std::ptrdiff_t N, typename G, RealOrComplexValued V>
auto operator/(const DeviceCanonicalComponentField<N, G, V>& a,
               const DeviceCanonicalComponentField<0, G, V>& b) {
  return DeviceFieldExpression<N, G, V>(GSH_EXPR_QUOTIENT, a.Grid(), detail::kind_of<V>,
                                        a.DeviceData(), b.DeviceData());
}
// Negation, conj, real, imag (CanonicalComponentFieldUnary.h).
template <std::ptrdiff_t N, typename G, RealOrComplexValued V>
auto operator-(const DeviceCanonicalComponentField<N, G, V>& a) {
  return DeviceFieldExpression<N, G, V>(GSH_EXPR_NEGATE, a.Grid(), detail::kind_of<V>,
                                        a.DeviceData());
}
template <std::ptrdiff_t N, typename G, RealOrComplexValued V>
auto conj(const DeviceCanonicalComponentField<N, G, V>& a) {
  return DeviceFieldExpression<N, G, V>(GSH_EXPR_CONJ, a.Grid(), detail::kind_of<V>,
                                        a.DeviceData());
}
template <std::ptrdiff_t N, typename G, RealOrComplexValued V>
auto real(const DeviceCanonicalComponentField<N, G, V>& a) {
  return DeviceFieldExpression<N, G, RealValued>(GSH_EXPR_REAL, a.Grid(), detail::kind_of<V>,
                                                 a.DeviceData());
}
template <std::ptrdiff_t N, typename G, RealOrComplexValued V>
auto imag(const DeviceCanonicalComponentField<N, G, V>& a) {
  return DeviceFieldExpression<N, G, RealValued>(GSH_EXPR_IMAG, a.Grid(), detail::kind_of<V>,
                                                 a.DeviceData());
}
// Scalar multiples (any upper index) and shifts (upper index 0), the
// CanonicalComponentFieldUnaryWithScalar forms.
template <std::ptrdiff_t N, typename G, RealOrComplexValued V>
auto operator*(const DeviceCanonicalComponentField<N, G, V>& a,
               typename DeviceCanonicalComponentField<N, G, V>::Scalar s) {
  return DeviceFieldExpression<N, G, V>(GSH_EXPR_SCALE, a.Grid(), detail::kind_of<V>,
                                        a.DeviceData(), nullptr, typename G::Complex(s));
}
template <std::ptrdiff_t N, typename G, RealOrComplexValued V>
auto operator*(typename DeviceCanonicalComponentField<N, G, V>::Scalar s,
               const DeviceCanonicalComponentField<N, G, V>& a) {
  return a * s;
}
template <std::ptrdiff_t N, typename G, RealOrComplexValued V>
auto operator/(const DeviceCanonicalComponentField<N, G, V>& a,
               typename DeviceCanonicalComponentField<N, G, V>::Scalar s) {
  return a * (typename DeviceCanonicalComponentField<N, G, V>::Scalar(1) / s);
}
template <typename G, RealOrComplexValued V>
auto operator+(const DeviceCanonicalComponentField<0, G, V>& a,
               typename DeviceCanonicalComponentField<0, G, V>::Scalar s) {
  return DeviceFieldExpression<0, G, V>(GSH_EXPR_SHIFT, a.Grid(), detail::kind_of<V>,
                                        a.DeviceData(), nullptr, typename G::Complex(s));
}
template <typename G, RealOrComplexValued V>
auto operator+(typename DeviceCanonicalComponentField<0, G, V>::Scalar s,
               const DeviceCanonicalComponentField<0, G, V>& a) {
  return a + s;
}
template <typename G, RealOrComplexValued V>
auto operator-(const DeviceCanonicalComponentField<0, G, V>& a,
               typename DeviceCanonicalComponentField<0, G, V>::Scalar s) {
  return a + (-s);
}

template <std::ptrdiff_t N, typename Grid>
using DeviceRealCanonicalComponentField = DeviceCanonicalComponentField<N, Grid, RealValued>;
template <std::ptrdiff_t N, typename Grid>
using DeviceComplexCanonicalComponentField =
    DeviceCanonicalComponentField<N, Grid, ComplexValued>;

// Coefficients c_lm, l = |N| .. grid lMax, in the layout of the host
// CanonicalComponentExpansion (GSHIndices<All> for complex fields,
// GSHIndices<NonNegative> for real fields), on the grid's GPU.
template <std::ptrdiff_t N_, typename Grid_, RealOrComplexValued Value_>
class DeviceCanonicalComponentExpansion
    : public GSHIndices<std::conditional_t<std::same_as<Value_, RealValued>, NonNegative, All>> {
 public:
  static constexpr bool on_device = true;
  using Int = std::ptrdiff_t;
  using Value = Value_;
  using MRange = std::conditional_t<std::same_as<Value, RealValued>, NonNegative, All>;
  using Real = typename Grid_::Real;
  using Complex = typename Grid_::Complex;
  using Scalar = Complex;
  using Host = CanonicalComponentExpansion<N_, Grid_, Value_>;

  // Zero coefficients on the grid's GPU.
  explicit DeviceCanonicalComponentExpansion(const Grid_& grid)
      : GSHIndices<MRange>(grid.MaxDegree(), grid.MaxDegree(), N_),
        _grid{&grid},
        _buf(grid.Handle(), sizeof(Complex) * static_cast<std::size_t>(this->Size())) {}
  // A copy of host coefficients.
  explicit DeviceCanonicalComponentExpansion(const Host& host)
      : DeviceCanonicalComponentExpansion(host.Grid()) {
    Upload(host);
  }

  static constexpr Int UpperIndex() { return N_; }
  const Grid_& Grid() const { return *_grid; }
  Complex* DeviceData() { return static_cast<Complex*>(_buf.Get()); }
  const Complex* DeviceData() const { return static_cast<const Complex*>(_buf.Get()); }
  void Upload(const Host& host) { _buf.Upload(host.Data().data()); }
  void Download(Host& host) const { _buf.Download(host.Data().data()); }
  Host ToHost() const {
    Host h(*_grid);
    Download(h);
    return h;
  }

 private:
  const Grid_* _grid;
  detail::DeviceBuffer _buf;
};

template <std::ptrdiff_t N, typename Grid>
using DeviceRealCanonicalComponentExpansion =
    DeviceCanonicalComponentExpansion<N, Grid, RealValued>;
template <std::ptrdiff_t N, typename Grid>
using DeviceComplexCanonicalComponentExpansion =
    DeviceCanonicalComponentExpansion<N, Grid, ComplexValued>;

}  // namespace GSHTrans

#endif
