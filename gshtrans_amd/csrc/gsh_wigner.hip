// gsh_wigner.hip -- Wigner d tables on the GPU in the reference's layout
// (Wigner<Real, Norm, MRange, NRange, AngleRange, Storage>, Wigner.h:99-286):
// one block of GSHIndices<MRange>(lMax, mMax, n) per (n, theta_i)
// (Indexing.h:72-150), the blocks in ColumnMajor (n outer) or RowMajor
// (theta outer) order (Wigner.h:280-286).
//
// A wavefront takes 64 consecutive orders of one (n, theta): each lane runs
// its order's FP64 recurrence over the degrees from the seed at l0 = max(|m|,
// |n|), so for every degree the 64 orders it writes are consecutive entries
// of the block (512 contiguous bytes).  The recurrence works on the
// Ortho-normalised values (Wigner.h:475-485); FourPi tables divide the Ortho
// factor out per degree.
#include "gsh_internal.h"

namespace gsh {
namespace {

// Offset of degree l in GSHIndices<MRange>(lMax, mMax, n) with mMax already
// min(lMax, mMax) (GSHIndices::OffsetForDegree, Indexing.h:104-131).
__device__ __forceinline__ long degree_offset(long l, long mmax, long nabs, int all) {
  if (all) {
    if (mmax >= nabs)
      return l <= mmax ? l * l - nabs * nabs
                       : (mmax + 1) * (mmax + 1) - nabs * nabs + (l - 1 - mmax) * (2 * mmax + 1);
    return (l - nabs) * (2 * mmax + 1);
  }
  if (mmax >= nabs)
    return l <= mmax ? (l * (l + 1)) / 2 - (nabs * (nabs + 1)) / 2
                     : ((mmax + 1) * (mmax + 2)) / 2 - (nabs * (nabs + 1)) / 2 +
                           (l - 1 - mmax) * (mmax + 1);
  return (l - nabs) * (mmax + 1);
}

__global__ __launch_bounds__(256) void wigner_table(WignerParams p) {
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int it = blockIdx.y, ni = blockIdx.z;
  const int mb = blockIdx.x * 4 + wave;  // block of 64 orders
  if (64 * mb >= p.nm) return;
  const int mi = 64 * mb + lane;
  const bool valid = mi < p.nm;
  const int m = p.mlo + (valid ? mi : 64 * mb);
  const int n = p.nmin + ni;
  const int nabs = abs(n);
  const int l0 = max(abs(m), nabs);
  // First degree of the wavefront: the smallest l0 of its orders.
  const int mfirst = p.mlo + 64 * mb, mlast = p.mlo + min(64 * mb + 63, p.nm - 1);
  const int mabs_min = (mfirst <= 0 && mlast >= 0) ? 0 : min(abs(mfirst), abs(mlast));
  const int lstart = max(mabs_min, nabs);
  const double cs = p.cosv[it];
  double e = valid ? p.seeds[((long)ni * p.ntheta + it) * p.nm + mi] : 0.0;
  double ep = 0.0;
  double* out = p.out + p.blockoff[(long)ni * p.ntheta + it];
  const double4* rec = p.rec + (long)ni * (p.lmax + 1) * p.nm + mi;
  for (int l = lstart; l <= p.lmax; l++) {
    if (valid && l > l0) {
      const double4 q = rec[(long)l * p.nm];
      const double en = fma(fma(q.x, cs, -q.y), e, -(q.z * ep));
      ep = e;
      e = en;
    }
    if (valid && l >= l0) {
      const long off = degree_offset(l, p.mmax, nabs, p.all) + (p.all ? m + min(l, p.mmax) : m);
      out[off] = p.fourpi ? e * p.inv_ortho[l] : e;
    }
  }
}

}  // namespace

hipError_t launch_wigner_table(const WignerParams& p, hipStream_t s) {
  const dim3 grid((p.nm + 255) / 256, p.ntheta, p.nn);
  hipLaunchKernelGGL(wigner_table, grid, dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace gsh
