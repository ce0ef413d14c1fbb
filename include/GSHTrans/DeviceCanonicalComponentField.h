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
