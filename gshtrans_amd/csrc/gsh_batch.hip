// gsh_batch.hip -- storage conversion for batch plans with FP32 coefficients:
// the Legendre kernels read and write FP64 coefficients, so a plan with FP32
// storage widens the caller's coefficients into an FP64 staging array before
// a synthesis and narrows the analysis results back afterwards (only the
// first csize[i] entries of field i, the caller's own array).
#include "gsh_internal.h"

namespace gsh {
namespace {

__global__ __launch_bounds__(256) void widen_c32(const float2* __restrict__ in,
                                                 double2* __restrict__ out, long total) {
  const long i = (long)blockIdx.x * 256 + threadIdx.x;
  if (i >= total) return;
  const float2 v = in[i];
  out[i] = make_double2(v.x, v.y);
}

__global__ __launch_bounds__(256) void narrow_c64(const double2* __restrict__ in,
                                                  float2* __restrict__ out, long stride,
                                                  const long* __restrict__ csize) {
  const long j = (long)blockIdx.x * 256 + threadIdx.x;
  const int f = blockIdx.y;
  if (j >= csize[f]) return;
  const double2 v = in[(long)f * stride + j];
  out[(long)f * stride + j] = make_float2((float)v.x, (float)v.y);
}

}  // namespace

hipError_t launch_widen_c32(const void* in, double* out, long total, hipStream_t s) {
  if (total <= 0) return hipSuccess;
  hipLaunchKernelGGL(widen_c32, dim3((total + 255) / 256), dim3(256), 0, s,
                     static_cast<const float2*>(in), reinterpret_cast<double2*>(out), total);
  return hipGetLastError();
}

hipError_t launch_narrow_c64(const double* in, void* out, long stride, const long* csize,
                             long cmax, int count, hipStream_t s) {
  if (count <= 0 || cmax <= 0) return hipSuccess;
  hipLaunchKernelGGL(narrow_c64, dim3((cmax + 255) / 256, count), dim3(256), 0, s,
                     reinterpret_cast<const double2*>(in), static_cast<float2*>(out), stride,
                     csize);
  return hipGetLastError();
}

}  // namespace gsh
