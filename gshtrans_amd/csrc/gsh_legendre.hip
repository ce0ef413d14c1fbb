// gsh_legendre.hip -- fused Wigner-d recurrence + Legendre contraction on
// gfx950 FP64 MFMA (v_mfma_f64_16x16x4_f64).
//
// For one order m and upper index n the reference sums, per colatitude ring,
//   synthesis  F_m(theta_i) = sum_l c_lm D^l_{mn}(theta_i)
//              (GSHTrans/src/GaussLegendreGrid.h:288-312)
//   analysis   c_lm = sum_i D^l_{mn}(theta_i) w_i (2 pi / N) X_m(theta_i)
//              (GaussLegendreGrid.h:181-207)
// reading D from a precomputed O(L^3) table (Wigner.h:122-130, 305-314).
// Here D is generated in registers by the three-term recurrence in l
// (Wigner.h:427-445), one lane per colatitude, and never touches HBM:
//   e_l = (RA_l cos(theta) - RB_l) e_{l-1} - RC_l e_{l-2},   e_{l0-1} = 0,
// with e_{l0} the Ortho-normalised seed of Wigner.h:52-95 (computed on the
// host with the reference's libm expression) and RA/RB/RC the recurrence
// coefficients with the Ortho factor sqrt(2l+1)/(2 sqrt(pi)) folded in.
//
// The sums over l (synthesis) or theta (analysis) are dense contractions per
// order m: (theta x l) . (l x columns) and (l x theta) . (theta x columns),
// columns = (real, imaginary) x fields.  Recurrence lanes hold one theta
// each; a per-wave LDS transpose turns them into MFMA operands.
#include "gsh_internal.h"

namespace gsh {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Order handled by work index mw: complex fields take m = 0, -1, 1, -2, 2, ...
// (longest l-ranges first); real fields take m = 0, 1, 2, ...
__device__ __forceinline__ int order_of(int mw, int real) {
  return real ? mw : ((mw & 1) ? -((mw + 1) >> 1) : (mw >> 1));
}

// Coefficient index of (l, m): GSHIndices<All>(lmax, lmax, n).Index(l, m) =
// l^2 - n^2 + m + l, or GSHIndices<NonNegative>: l(l+1)/2 - |n|(|n|+1)/2 + m
// (Indexing.h:104-146).
__device__ __forceinline__ long coef_index(int l, int m, int nabs, int real) {
  return real ? (long)l * (l + 1) / 2 - (long)nabs * (nabs + 1) / 2 + m
              : (long)l * l - (long)nabs * nabs + m + l;
}

// One recurrence step to degree l (uniform coefficients, per-lane cos).
__device__ __forceinline__ void rec_step(const LegParams& p, long ridx, double cs, double& e0,
                                         double& e1) {
  const double a = p.RA[ridx], b = p.RB[ridx], c = p.RC[ridx];
  const double t = fma(a, cs, -b);
  const double en = fma(t, e0, -(c * e1));
  e1 = e0;
  e0 = en;
}

// ---------------------------------------------------------------------------
// Synthesis.  Workgroup = 4 waves = 256 colatitudes x (16*T) columns for one
// order.  Per 4 degrees: each lane advances its recurrence 4 steps, the wave
// transposes the 4 x 64 block through LDS into four 16x4 A operands (one per
// 16-colatitude group), the coefficients of those 4 degrees form the 4x16 B
// operands, and 4*T MFMAs accumulate F[theta][col].
// ---------------------------------------------------------------------------
template <int T>
__global__ __launch_bounds__(256) void leg_synth_f64(LegParams p) {
  __shared__ double tr[4][4][66];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  int blk = blockIdx.x;
  const int tsb = blk % p.n_tsb;
  blk /= p.n_tsb;
  const int cg = blk % p.n_cg;
  const int mw = blk / p.n_cg;
  const int m = order_of(mw, p.real);
  const int slot = m - p.m_first;
  const int l0 = max(abs(m), p.nabs);
  const int t = tsb * 256 + wave * 64 + lane;
  const bool tv = t < p.ntheta;
  const double cs = tv ? p.cosv[t] : 0.0;
  double e0 = tv ? p.seeds[(long)(m + p.L) * p.ntheta + t] : 0.0;
  double e1 = 0.0;
  const long rbase = p.roff[m + p.L] - l0;

  // B-operand column of this lane for each column tile.
  int fld[T], part[T];
  bool cv[T];
#pragma unroll
  for (int j = 0; j < T; j++) {
    const int col = cg * 16 * T + 16 * j + (lane & 15);
    cv[j] = col < p.ncol;
    fld[j] = col >> 1;
    part[j] = col & 1;
  }
  const int krow = lane >> 4;

  d4 acc[4][T];
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int j = 0; j < T; j++) acc[g][j] = d4{0.0, 0.0, 0.0, 0.0};

  // Prefetched B operands for the current 4 degrees.
  double B[T];
  auto loadB = [&](int l, double* out) {
    const int lk = l + krow;
    const bool lv = lk <= p.lmax;
    const long ci = coef_index(lk, m, p.nabs, p.real);
#pragma unroll
    for (int j = 0; j < T; j++)
      out[j] = (lv && cv[j]) ? p.cin[((long)fld[j] * p.cstride + ci) * 2 + part[j]] : 0.0;
  };
  loadB(l0, B);

  for (int l = l0; l <= p.lmax; l += 4) {
    double Bn[T];
    loadB(l + 4, Bn);
    double v[4];
#pragma unroll
    for (int k = 0; k < 4; k++) {
      const int lk = l + k;
      v[k] = (lk <= p.lmax) ? e0 : 0.0;
      if (lk + 1 <= p.lmax) rec_step(p, rbase + lk + 1, cs, e0, e1);
    }
#pragma unroll
    for (int k = 0; k < 4; k++) tr[wave][k][lane] = v[k];
    wave_sync();
    double A[4];
#pragma unroll
    for (int g = 0; g < 4; g++) A[g] = tr[wave][krow][g * 16 + (lane & 15)];
    wave_sync();
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int j = 0; j < T; j++)
        acc[g][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(A[g], B[j], acc[g][j], 0, 0, 0);
#pragma unroll
    for (int j = 0; j < T; j++) B[j] = Bn[j];
  }

  // C layout of the 16x16x4 f64 MFMA: lane holds rows lane/16 + 4r, column lane%16.
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int j = 0; j < T; j++) {
      const int col = cg * 16 * T + 16 * j + (lane & 15);
      if (col >= p.ncol) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = tsb * 256 + wave * 64 + g * 16 + krow + 4 * r;
        if (row < p.ntheta) p.F[((long)row * p.nslot + slot) * p.ncol + col] = acc[g][j][r];
      }
    }
}

// ---------------------------------------------------------------------------
// Analysis.  Workgroup = 4 waves for one (order, 16-column group), covering
// every colatitude: wave w owns theta blocks w, w+4, ... of 64 lanes.  Per
// chunk of 64 degrees, each wave walks its blocks: loads the block's B
// operands (X'[theta][col], 16 theta-quads), advances each block's
// recurrence 16 steps per degree tile, transposes through LDS into 16x4 A
// operands (16 degrees x 4 colatitudes) and accumulates C[l][col].  The four
// waves' partial sums are then added in LDS in a fixed order.
// ---------------------------------------------------------------------------
constexpr int kAnalMaxBlocks = 9;   // theta blocks per wave: L <= 2303
constexpr int kAnalLT = 4;          // 16-degree tiles per chunk

__global__ __launch_bounds__(256) void leg_anal_f64(LegParams p) {
  __shared__ double tr[4][16][66];
  __shared__ double red[4][kAnalLT][4][64];
  const int lane = threadIdx.x & 63;
  const int wave = threadIdx.x >> 6;
  const int cg = blockIdx.x % p.n_cg;
  const int mw = blockIdx.x / p.n_cg;
  const int m = order_of(mw, p.real);
  const int l0 = max(abs(m), p.nabs);
  const long rbase = p.roff[m + p.L] - l0;
  const int slot = m - p.m_first;
  const int krow = lane >> 4;
  const int col = cg * 16 + (lane & 15);
  const bool cv = col < p.ncol;

  // Per-block recurrence state.
  double cs[kAnalMaxBlocks], e0[kAnalMaxBlocks], e1[kAnalMaxBlocks];
  int nb = 0;
#pragma unroll
  for (int bi = 0; bi < kAnalMaxBlocks; bi++) {
    const int b = wave + 4 * bi;
    const int t = b * 64 + lane;
    const bool tv = (b < p.nblk) && (t < p.ntheta);
    cs[bi] = tv ? p.cosv[t] : 0.0;
    e0[bi] = tv ? p.seeds[(long)(m + p.L) * p.ntheta + t] : 0.0;
    e1[bi] = 0.0;
    if (b < p.nblk) nb = bi + 1;
  }

  for (int lc = l0; lc <= p.lmax; lc += 16 * kAnalLT) {
    d4 acc[kAnalLT];
#pragma unroll
    for (int lt = 0; lt < kAnalLT; lt++) acc[lt] = d4{0.0, 0.0, 0.0, 0.0};

#pragma unroll
    for (int bi = 0; bi < kAnalMaxBlocks; bi++) {
      if (bi >= nb) continue;
      const int b = wave + 4 * bi;
      // B operands: theta = b*64 + 4q + krow, column col.
      double Bq[16];
      const int nq = min(16, (p.ntheta - b * 64 + 3) >> 2);
#pragma unroll
      for (int q = 0; q < 16; q++) {
        const int t = b * 64 + 4 * q + krow;
        Bq[q] = (cv && t < p.ntheta) ? p.F[((long)t * p.nslot + slot) * p.ncol + col] : 0.0;
      }
#pragma unroll
      for (int lt = 0; lt < kAnalLT; lt++) {
        const int lb = lc + 16 * lt;
        if (lb > p.lmax) continue;
#pragma unroll
        for (int k = 0; k < 16; k++) {
          const int lk = lb + k;
          tr[wave][k][lane] = (lk <= p.lmax) ? e0[bi] : 0.0;
          if (lk + 1 <= p.lmax) rec_step(p, rbase + lk + 1, cs[bi], e0[bi], e1[bi]);
        }
        wave_sync();
#pragma unroll
        for (int q = 0; q < 16; q++) {
          if (q >= nq) continue;
          const double A = tr[wave][lane & 15][4 * q + krow];
          acc[lt] = __builtin_amdgcn_mfma_f64_16x16x4f64(A, Bq[q], acc[lt], 0, 0, 0);
        }
        wave_sync();
      }
    }

    // Cross-wave reduction in a fixed order, then write the coefficients.
#pragma unroll
    for (int lt = 0; lt < kAnalLT; lt++)
#pragma unroll
      for (int r = 0; r < 4; r++) red[wave][lt][r][lane] = acc[lt][r];
    __syncthreads();
    {
      // Thread -> (lt, r) pair with lt = wave.
      const int lt = wave;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const double s = ((red[0][lt][r][lane] + red[1][lt][r][lane]) + red[2][lt][r][lane]) +
                         red[3][lt][r][lane];
        const int l = lc + 16 * lt + krow + 4 * r;
        if (cv && l <= p.lmax) {
          const int f = col >> 1, pt = col & 1;
          const long ci = ((long)f * p.cstride + coef_index(l, m, p.nabs, p.real)) * 2 + pt;
          double v = s;
          if (p.accumulate) v += p.cout[ci];
          if (p.zero_last && l == p.lmax && m == p.lmax) v = 0.0;
          p.cout[ci] = v;
        }
      }
    }
    __syncthreads();
  }
}

}  // namespace

hipError_t launch_leg_synth_f64(const LegParams& p0, hipStream_t s) {
  LegParams p = p0;
  p.n_tsb = (p.ntheta + 255) / 256;
  const int T = p.ncol > 16 ? 2 : 1;
  p.n_cg = (p.ncol + 16 * T - 1) / (16 * T);
  const long nblocks = (long)p.nslot * p.n_cg * p.n_tsb;
  if (T == 2)
    hipLaunchKernelGGL(leg_synth_f64<2>, dim3(nblocks), dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL(leg_synth_f64<1>, dim3(nblocks), dim3(256), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_leg_anal_f64(const LegParams& p0, hipStream_t s) {
  LegParams p = p0;
  p.n_cg = (p.ncol + 15) / 16;
  p.nblk = (p.ntheta + 63) / 64;
  if ((p.nblk + 3) / 4 > kAnalMaxBlocks) return hipErrorInvalidValue;
  const long nblocks = (long)p.nslot * p.n_cg;
  hipLaunchKernelGGL(leg_anal_f64, dim3(nblocks), dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace gsh
