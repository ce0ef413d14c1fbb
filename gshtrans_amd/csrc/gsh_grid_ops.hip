// gsh_grid_ops.hip -- grid helpers on the device: quadrature integration and
// L2 inner products of fields (ScalarFieldOverloads.h:288-316 with the
// weights of GridBase.h:60-66), and the seeded coefficient generator
// (GridBase.h:115-172 with a counter-based stream).
#include "gsh_internal.h"

#include <mutex>

namespace gsh {
namespace {

constexpr int kRingThreads = 256;

// Ring sums of one field: part[f * ntheta + i] = w_i dphi sum_j g(a, b)(i, j)
// with g = a (b null), a b (real fields) or conj(a) b (complex fields).
__global__ __launch_bounds__(kRingThreads) void ring_sums(GridOpParams p) {
  const int i = blockIdx.x, f = blockIdx.y;
  const long base = (long)f * p.fstride + (long)i * p.nphi;
  double sr = 0.0, si = 0.0;
  for (int j = threadIdx.x; j < p.nphi; j += kRingThreads) {
    if (p.real) {
      const double a = static_cast<const double*>(p.fa)[base + j];
      sr += p.fb ? a * static_cast<const double*>(p.fb)[base + j] : a;
    } else {
      const double2 a = static_cast<const double2*>(p.fa)[base + j];
      if (p.fb) {
        const double2 b = static_cast<const double2*>(p.fb)[base + j];
        sr += a.x * b.x + a.y * b.y;  // conj(a) b
        si += a.x * b.y - a.y * b.x;
      } else {
        sr += a.x;
        si += a.y;
      }
    }
  }
  // Wavefront sums, then the 4 wavefronts of the workgroup in a fixed order.
  for (int o = 32; o > 0; o >>= 1) {
    sr += __shfl_down(sr, o, 64);
    si += __shfl_down(si, o, 64);
  }
  __shared__ double2 ws[kRingThreads / 64];
  if ((threadIdx.x & 63) == 0) ws[threadIdx.x >> 6] = make_double2(sr, si);
  __syncthreads();
  if (threadIdx.x == 0) {
    double tr = 0.0, ti = 0.0;
    for (int k = 0; k < kRingThreads / 64; k++) {
      tr += ws[k].x;
      ti += ws[k].y;
    }
    const double w = p.wq[i] * p.dphi;
    p.part[(long)f * p.ntheta + i] = make_double2(tr * w, ti * w);
  }
}

// out[f] = sum over the colatitudes of the ring sums, in colatitude order.
__global__ void theta_sums(GridOpParams p) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= p.batch) return;
  double tr = 0.0, ti = 0.0;
  for (int i = 0; i < p.ntheta; i++) {
    const double2 v = p.part[(long)f * p.ntheta + i];
    tr += v.x;
    ti += v.y;
  }
  p.out[2 * f] = tr;
  p.out[2 * f + 1] = ti;
}

__device__ __forceinline__ unsigned long long mix64(unsigned long long z) {
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

// Coefficient k of field f: the counter-based splitmix64 -> Box-Muller stream
// (element e = f size + k uses counters 2e + 1 and 2e + 2 past the seed's
// base), scaled by (l+1)^-decay, with the reference's masks.
__global__ __launch_bounds__(256) void random_coeffs(RandomParams p) {
  const long e = (long)blockIdx.x * blockDim.x + threadIdx.x;
  if (e >= (long)p.batch * p.size) return;
  const int f = static_cast<int>(e / p.size);
  const long k = e - (long)f * p.size;
  constexpr unsigned long long golden = 0x9E3779B97F4A7C15ULL;
  const unsigned long long base = mix64(p.seed + golden);
  const unsigned long long z1 = mix64(base + (unsigned long long)(2 * e + 1) * golden);
  const unsigned long long z2 = mix64(base + (unsigned long long)(2 * e + 2) * golden);
  const double u1 = (double)(z1 >> 11) * (1.0 / 9007199254740992.0);
  const double u2 = (double)(z2 >> 11) * (1.0 / 9007199254740992.0);
  const double r = sqrt(-2.0 * log1p(-u1));
  double s, c;
  sincos(2.0 * 3.14159265358979323846 * u2, &s, &c);
  // Degree and order of entry k (GSHIndices, Indexing.h:104-131).
  const long na = p.nabs;
  long l, m;
  if (p.real) {
    const long t = k + na * (na + 1) / 2;  // = l (l + 1) / 2 + m
    l = (long)((sqrt(8.0 * (double)t + 1.0) - 1.0) / 2.0);
    while (l * (l + 1) / 2 > t) l--;
    while ((l + 1) * (l + 2) / 2 <= t) l++;
    m = t - l * (l + 1) / 2;
  } else {
    const long t = k + na * na;  // = l^2 + l + m
    l = (long)sqrt((double)t);
    while (l * l > t) l--;
    while ((l + 1) * (l + 1) <= t) l++;
    m = t - l * l - l;
  }
  const double scale = p.decay == 0.0 ? 1.0 : pow((double)(l + 1), -p.decay);
  double re = r * c * scale, im = r * s * scale;
  const bool last = k == p.size - 1 && p.lmax == p.grid_lmax;
  if (p.real) {
    if (m == 0 || last) im = 0.0;
  } else if (last) {
    re = im = 0.0;
  }
  double* o = p.out + 2 * ((long)f * p.cstride + k);
  o[0] = re;
  o[1] = im;
}

}  // namespace

hipError_t launch_integrate(const GridOpParams& p, hipStream_t s) {
  hipLaunchKernelGGL(ring_sums, dim3(p.ntheta, p.batch), dim3(kRingThreads), 0, s, p);
  hipLaunchKernelGGL(theta_sums, dim3((p.batch + 63) / 64), dim3(64), 0, s, p);
  return hipGetLastError();
}

// Scratch of the allocation-free integration (gsh_integrate_noalloc): the
// ring sums part[f * ntheta + i].  Every device that loads this library holds
// its own copy until the process ends.  In host code `scratch` names a
// host-side stand-in, so kernels receive only the device address looked up
// below, never the array's name.
__device__ __align__(256) double2 scratch[kNoAllocRingSums];

namespace {
std::mutex scratch_mu;
double2* scratch_dev[kMaxDevices] = {};
}  // namespace

hipError_t noalloc_scratch_init(int device) {
  if (device < 0 || device >= kMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(scratch_mu);
  if (scratch_dev[device]) return hipSuccess;
  int cur = -1;
  hipError_t e = hipGetDevice(&cur);
  if (e != hipSuccess) return e;
  if (cur != device) return hipErrorInvalidDevice;  // the lookup needs `device` current
  void* ptr = nullptr;
  e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
  if (e != hipSuccess) return e;
  if (!ptr) return hipErrorInvalidSymbol;
  scratch_dev[device] = static_cast<double2*>(ptr);
  return hipSuccess;
}

double2* noalloc_scratch(int device) {
  if (device < 0 || device >= kMaxDevices) return nullptr;
  std::lock_guard<std::mutex> lock(scratch_mu);
  return scratch_dev[device];
}

hipError_t launch_random_coeffs(const RandomParams& p, hipStream_t s) {
  const long total = (long)p.batch * p.size;
  hipLaunchKernelGGL(random_coeffs, dim3((total + 255) / 256), dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace gsh
