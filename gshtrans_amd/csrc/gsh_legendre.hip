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
//   e_l = (a_l cos(theta) - b_l) e_{l-1} - c_l e_{l-2},   e_{l0-1} = 0,
// with e_{l0} the Ortho-normalised seed of Wigner.h:52-95 (computed on the
// host with the reference's libm expression) and (a, b, c) the recurrence
// coefficients with the Ortho factor sqrt(2l+1)/(2 sqrt(pi)) folded in.
//
// The sums over l (synthesis) or theta (analysis) are dense contractions per
// order m: (theta x l) . (l x columns) and (l x theta) . (theta x columns),
// columns = (real, imaginary) x fields.  Recurrence lanes hold one theta
// each; a per-wave LDS transpose turns them into MFMA operands.  The loop
// bodies are branch-free (coefficient tables are zero-padded past L and
// loads past lmax are masked by selects), which keeps the MFMA accumulators
// resident in AGPRs across iterations.
#include "gsh_internal.h"

namespace gsh {
namespace {

typedef double d4 __attribute__((ext_vector_type(4)));
// Recurrence tables are read through the constant address space: they are
// never written while a kernel runs, and uniform loads from it become scalar
// (s_load) loads even when the kernel stores to global memory in the loop.
typedef const __attribute__((address_space(4))) d4 rec_t;

typedef float f4 __attribute__((ext_vector_type(4)));

// The contraction's MFMA and accumulator: FP64 (v_mfma_f64_16x16x4f64; the
// lane holding column lane%16 holds C rows lane/16 + 4r) or, for the FP32
// precision mode, v_mfma_f32_16x16x4f32 on FP32-rounded operands with FP32
// accumulation (C rows 4 (lane/16) + r).  A operand: row lane%16, k
// lane/16; B operand: k lane/16, column lane%16 (both instructions).
template <bool F32M>
struct Mma {
  typedef d4 acc_t;
  static __device__ __forceinline__ acc_t zero() { return d4{0.0, 0.0, 0.0, 0.0}; }
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int q, int r) { return q + 4 * r; }
};
template <>
struct Mma<true> {
  typedef f4 acc_t;
  static __device__ __forceinline__ acc_t zero() { return f4{0.f, 0.f, 0.f, 0.f}; }
  static __device__ __forceinline__ acc_t mma(double a, double b, acc_t c) {
    return __builtin_amdgcn_mfma_f32_16x16x4f32((float)a, (float)b, c, 0, 0, 0);
  }
  static __device__ __forceinline__ int row(int q, int r) { return 4 * q + r; }
};

__device__ __forceinline__ rec_t* rec_ptr(const double4* rec, long off) {
  return (rec_t*)(rec) + off;
}

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

// One recurrence step with uniform coefficients q and per-lane cos.
__device__ __forceinline__ void rec_step(const d4 q, double cs, double& e0, double& e1) {
  const double t = fma(q.x, cs, -q.y);
  const double en = fma(t, e0, -(q.z * e1));
  e1 = e0;
  e0 = en;
}

// ---------------------------------------------------------------------------
// Synthesis.  Workgroup = 4 waves = 256 colatitudes x (16*T) columns for one
// order.  Per 4 degrees: each lane advances its recurrence 4 steps, the wave
// transposes the 4 x 64 block through LDS into four 16x4 A operands (one per
// 16-colatitude group), the coefficients of those 4 degrees form the 4x16 B
// operands, and 4*T MFMAs accumulate F[theta][col].  Coefficients of the
// next 4 degrees are loaded one chunk ahead.
// ---------------------------------------------------------------------------
template <int T, int REAL, bool F32M>
__global__ __launch_bounds__(256) void leg_synth_f64(LegParams p) {
  typedef Mma<F32M> M;
  __shared__ double tr[4][4][66];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  // Work item: (order, column group, colatitude super-block).
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int mw = it.x;
  const int tsb = it.y & 0xffff;
  const int cg = it.y >> 16;
  const int m = order_of(mw, REAL);
  const int slot = m - p.m_first;
  const int l0 = max(abs(m), p.nabs);
  const int lmax = p.lmax;
  const int t = tsb * 256 + wave * 64 + lane;
  const bool tv = t < p.ntheta;
  // Branch-free loads at a clamped colatitude keep the control flow uniform.
  const int tc = min(t, p.ntheta - 1);
  const double cs0 = p.cosv[tc], s0 = p.seeds[(long)(m + p.L) * p.ntheta + tc];
  const double cs = tv ? cs0 : 0.0;
  double e0 = tv ? s0 : 0.0;
  double e1 = 0.0;
  const int krow = lane >> 4;
  // Recurrence coefficients of one 8-degree iteration (entries l+1 .. l+8,
  // 256 bytes) are staged through a per-wave LDS slot: lanes 0..15 fetch the
  // next iteration's 16 bytes each one iteration ahead, so the steps never
  // wait on scalar loads.  Entries up to L + kRecPad exist.
  __shared__ uint4 rsw[4][16];
  const uint4* recq = reinterpret_cast<const uint4*>(p.rec + (ld_uniform(p.roff, m + p.L) - l0));
  auto fetch = [&](int l) { return lane < 16 ? recq[2 * (l + 1) + lane] : make_uint4(0, 0, 0, 0); };
  const d4* rl = reinterpret_cast<const d4*>(&rsw[wave][0]);  // rl[j]: entry l + 1 + j

  // B-operand column of this lane per column tile.  Every load reads a valid
  // address (column and degree clamped); out-of-range values are zeroed by a
  // select when the operand is used, so loads stay in flight across a chunk.
  // Each column's own truncation degree lf masks its degrees; a column with
  // lf < l0 (no degree of this order) reads element 0 and is masked out.
  const double* cptr[T];
  bool cv[T];
  int lf[T];
#pragma unroll
  for (int j = 0; j < T; j++) {
    const int col = cg * 16 * T + 16 * j + (lane & 15);
    cv[j] = col < p.ncol;
    const int cc = cv[j] ? col : 0;
    lf[j] = field_lmax(p, cc >> 1);
    cptr[j] = p.cin + (long)field_index(p.fmap, cc >> 1) * p.cstride * 2 + (cc & 1);
  }
  // Degree loop bound of the workgroup: the largest lf of its column groups.
  int lwg = lmax;
  if (p.cg_lmax) {
    lwg = 0;
#pragma unroll
    for (int j = 0; j < T; j++)
      if ((cg * T + j) * 16 < p.ncol) lwg = max(lwg, ld_uniform(p.cg_lmax, cg * T + j));
  }
  auto loadB = [&](int l, double* out) {
#pragma unroll
    for (int j = 0; j < T; j++) {
      const int lk = min(l + krow, lf[j]);
      const long ci = lf[j] >= l0 ? 2 * coef_index(lk, m, p.nabs, REAL) : 0;
      out[j] = cptr[j][ci];
    }
  };

  typename M::acc_t acc[4][T];
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int j = 0; j < T; j++) acc[g][j] = M::zero();

  // One 4-degree chunk: recurrence, LDS transpose, B mask, 4*T MFMAs; j0 is
  // the chunk's first entry in the staged coefficients.
  auto chunk = [&](int l, int j0, double* B) {
    const d4 q1 = rl[j0], q2 = rl[j0 + 1], q3 = rl[j0 + 2], q4 = rl[j0 + 3];
    tr[wave][0][lane] = e0;
    rec_step(q1, cs, e0, e1);
    tr[wave][1][lane] = e0;
    rec_step(q2, cs, e0, e1);
    tr[wave][2][lane] = e0;
    rec_step(q3, cs, e0, e1);
    tr[wave][3][lane] = e0;
    rec_step(q4, cs, e0, e1);
    wave_sync();
    double A[4];
#pragma unroll
    for (int g = 0; g < 4; g++) A[g] = tr[wave][krow][g * 16 + (lane & 15)];
    wave_sync();
    bool lv[T];
#pragma unroll
    for (int j = 0; j < T; j++) lv[j] = cv[j] && l + krow <= lf[j];
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int j = 0; j < T; j++)
        acc[g][j] = M::mma(A[g], lv[j] ? B[j] : 0.0, acc[g][j]);
  };

  // Chunks are processed in pairs with ping-pong operand buffers, so each
  // load has a full chunk of work to hide behind; an odd final chunk is
  // paired with an all-zero one.
  double B0[T], B1[T];
  loadB(l0, B0);
  uint4 rq = fetch(l0);
  // With ck_restart the state is taken from the checkpoint table every kPass
  // degrees (checkpoint (l - l0) / kCk at degree l <= L), loaded a pass ahead.
  const int nck = (p.L - l0 + kCk) / kCk;  // checkpoints 1 .. nck - 1 exist
  const double2* ckm = p.ck_restart ? p.ck + ld_uniform(p.ckoff, m + p.L) * (long)p.ntheta + tc
                                    : nullptr;
  double2 ckn = make_double2(0.0, 0.0);
  if (ckm && kPass / kCk < nck) ckn = ckm[(long)(kPass / kCk - 1) * p.ntheta];
  for (int l = l0; l <= lwg; l += 8) {
    if (ckm && l > l0 && ((l - l0) & (kPass - 1)) == 0) {
      const int k = (l - l0) / kCk;
      if (k < nck) {
        e0 = tv ? ckn.x : 0.0;
        e1 = tv ? ckn.y : 0.0;
        if (k + kPass / kCk < nck) ckn = ckm[(long)(k + kPass / kCk - 1) * p.ntheta];
      }
    }
    if (lane < 16) rsw[wave][lane] = rq;
    wave_sync();
    rq = fetch(l + 8);
    loadB(l + 4, B1);
    chunk(l, 0, B0);
    loadB(l + 8, B0);
    chunk(l + 4, 4, B1);
  }

  // C layout: the lane holds column lane%16, rows M::row(lane/16, r).
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int j = 0; j < T; j++) {
      const int col = cg * 16 * T + 16 * j + (lane & 15);
      if (col >= p.ncol) continue;
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = tsb * 256 + wave * 64 + g * 16 + M::row(krow, r);
        if (row < p.ntheta)
          p.F[f_index(row, slot, col, p.nslot, (p.ncol + 15) >> 4)] = acc[g][j][r];
      }
    }
}

// ---------------------------------------------------------------------------
// Analysis.  Workgroup = 4 waves for one (order, 16-column group), covering
// every colatitude.  The colatitudes are cut into quads (the K=4 of one
// MFMA); each wave owns MB full blocks of 16 quads (64 lanes), wave w taking
// blocks w, w+4, ...; the leftover quads (fewer than 64) are dealt round-robin
// to the waves as a "tail" block.  Per chunk of 64 degrees each wave walks
// its blocks: the block's B operands (X'[theta][col], one per quad) sit in
// registers (the next block's are loaded while the current one runs), the
// block's recurrence advances 16 steps per 16-degree tile, an LDS transpose
// forms 16x4 A operands (16 degrees x 4 colatitudes), and MFMAs accumulate
// C[l][col].  The tail block accumulates into its own registers so the main
// accumulators never cross a branch.  The four waves' partial sums are then
// added in LDS in a fixed order.  Degrees past lmax are computed but never
// written; colatitudes past the grid have zero seeds, so their D vanish.
// ---------------------------------------------------------------------------
constexpr int kAnalLT = 4;  // 16-degree tiles per chunk

template <int MB, bool F32M>
__global__ __launch_bounds__(256) void leg_anal_f64(LegParams p) {
  typedef Mma<F32M> M;
  __shared__ double tr[4][16][66];  // also holds the cross-wave reduction
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  // Work item: (order, pass, column group).
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int mw = it.x;
  const int pass = it.y & 0xffff;
  const int cg = it.y >> 16;
  const int m = order_of(mw, p.real);
  const int l0 = max(abs(m), p.nabs);
  const int lmax = p.lmax;
  const int slot = m - p.m_first;
  const int krow = lane >> 4;
  const int col = cg * 16 + (lane & 15);
  const int lc = l0 + kPass * pass;  // first degree of the pass
  // Recurrence coefficients of the pass (entries lc+1 .. lc+64, 2 KiB),
  // staged in LDS.  Entries up to L + kRecPad exist.
  __shared__ uint4 rsa[128];
  const uint4* recq = reinterpret_cast<const uint4*>(p.rec + (ld_uniform(p.roff, m + p.L) - l0));
  const int tid = threadIdx.x;
  if (tid < 128) rsa[tid] = recq[2 * (lc + 1) + tid];
  __syncthreads();
  const d4* rl = reinterpret_cast<const d4*>(&rsa[0]);  // entry lc + 1 + j
  // Recurrence state at degree lc: the seed for pass 0, else the checkpoint.
  const double2* ckp =
      p.ck + (ld_uniform(p.ckoff, m + p.L) + (kPass / kCk) * pass - 1) * (long)p.ntheta;
  auto state = [&](int t, double& e0_, double& e1_) {
    if (pass == 0) {
      e0_ = p.seeds[(long)(m + p.L) * p.ntheta + t];
      e1_ = 0.0;
    } else {
      const double2 s2 = ckp[t];
      e0_ = s2.x;
      e1_ = s2.y;
    }
  };
  const bool cv = col < p.ncol;
  const int lf = field_lmax(p, cv ? col >> 1 : 0);  // this column's truncation degree
  const long cbase = (long)field_index(p.fmap, cv ? col >> 1 : 0) * p.cstride;
  const int lwg = p.cg_lmax ? ld_uniform(p.cg_lmax, cg) : lmax;  // degree loop bound
  const int nt = p.ntheta;
  // B operands of this (order, column group): quad-major, 64 lanes per quad,
  // read with raw buffer loads (one lane offset register, scalar quad
  // offsets; quads past the end read as zero).
  const __amdgpu_buffer_rsrc_t fa = __builtin_amdgcn_make_buffer_rsrc(
      (void*)(p.F + ((long)slot * p.n_cg + cg) * p.nquad * 64), (short)0, p.nquad * 512,
      0x00020000);
  const int voff = lane * 8;
  auto ldB = [&](int quad) -> double {
    return __builtin_bit_cast(double, __builtin_amdgcn_raw_buffer_load_b64(fa, voff, quad * 512, 0));
  };

  // Full blocks: block bi of this wave starts at colatitude 64*(wave + 4*bi).
  double cs[MB > 0 ? MB : 1], e0[MB > 0 ? MB : 1], e1[MB > 0 ? MB : 1];
#pragma unroll
  for (int bi = 0; bi < MB; bi++) {
    const int t = 64 * (wave + 4 * bi) + lane;
    cs[bi] = p.cosv[t];
    state(t, e0[bi], e1[bi]);
  }
  // Tail: quads 4*w + 16*i + {0..3} after the full blocks; lane j covers
  // quad i = j/4, colatitude tail0 + 16*(j/4) + 4*wave + j%4.
  const int tail0 = 256 * MB;
  const int ntailq = (nt - tail0 + 3) >> 2;
  const int tq = (ntailq - wave + 3) >> 2;  // tail quads of this wave
  const int tt = tail0 + 16 * (lane >> 2) + 4 * wave + (lane & 3);
  const bool ttv = tt < nt;
  const int ttc = min(tt, nt - 1);
  double e0s, e1s;
  state(ttc, e0s, e1s);
  const double cs0 = p.cosv[ttc];
  double cst = ttv ? cs0 : 0.0;
  double e0t = ttv ? e0s : 0.0;
  double e1t = ttv ? e1s : 0.0;

  auto loadBlk = [&](int bi, double* Bq) {
#pragma unroll
    for (int q = 0; q < 16; q++) Bq[q] = ldB(16 * (wave + 4 * bi) + q);
  };
  // 16 recurrence steps of one block into the transpose buffer: tile lt of
  // the pass, staged coefficients rl (entry lc + 1 + j at rl[j]).
  auto steps16 = [&](const d4* rl, int lt, double cs_, double& e0_, double& e1_) {
    const d4* r = rl + 16 * lt;
#pragma unroll
    for (int k = 0; k < 16; k++) {
      tr[wave][k][lane] = e0_;
      rec_step(r[k], cs_, e0_, e1_);
    }
  };

  (void)lwg;
  {
    typename M::acc_t acc[kAnalLT];
#pragma unroll
    for (int lt = 0; lt < kAnalLT; lt++) acc[lt] = M::zero();

    double Ba[16], Bb[16];
    if (MB > 0) loadBlk(0, Ba);
#pragma unroll
    for (int bi = 0; bi < MB; bi++) {
      double* Bcur = (bi & 1) ? Bb : Ba;
      double* Bnext = (bi & 1) ? Ba : Bb;
      if (bi + 1 < MB) loadBlk(bi + 1, Bnext);
#pragma unroll
      for (int lt = 0; lt < kAnalLT; lt++) {
        steps16(rl, lt, cs[bi], e0[bi], e1[bi]);
        wave_sync();
#pragma unroll
        for (int q = 0; q < 16; q++) {
          const double A = tr[wave][lane & 15][4 * q + krow];
          acc[lt] = M::mma(A, Bcur[q], acc[lt]);
        }
        wave_sync();
      }
    }

    if (tq > 0) {
      typename M::acc_t acct[kAnalLT];
#pragma unroll
      for (int lt = 0; lt < kAnalLT; lt++) acct[lt] = M::zero();
#pragma unroll
      for (int lt = 0; lt < kAnalLT; lt++) {
        steps16(rl, lt, cst, e0t, e1t);
        wave_sync();
        for (int i = 0; i < tq; i++) {
          const int q = (tail0 >> 2) + 4 * i + wave;  // global quad index
          const double Bv = ldB(q);
          const double A = tr[wave][lane & 15][4 * i + krow];
          acct[lt] = M::mma(A, Bv, acct[lt]);
        }
        wave_sync();
      }
#pragma unroll
      for (int lt = 0; lt < kAnalLT; lt++) acc[lt] += acct[lt];
    }

    // Cross-wave reduction in a fixed order; wave w sums degree tile w.
    double* red = &tr[0][0][0];
    __syncthreads();
#pragma unroll
    for (int lt = 0; lt < kAnalLT; lt++)
#pragma unroll
      for (int r = 0; r < 4; r++) red[((wave * kAnalLT + lt) * 4 + r) * 64 + lane] = acc[lt][r];
    __syncthreads();
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int lt = wave;
      double s = red[((0 * kAnalLT + lt) * 4 + r) * 64 + lane];
      s += red[((1 * kAnalLT + lt) * 4 + r) * 64 + lane];
      s += red[((2 * kAnalLT + lt) * 4 + r) * 64 + lane];
      s += red[((3 * kAnalLT + lt) * 4 + r) * 64 + lane];
      const int l = lc + 16 * lt + M::row(krow, r);
      if (cv && l <= lf) {
        const long ci = (cbase + coef_index(l, m, p.nabs, p.real)) * 2 + (col & 1);
        double v = s;
        if (p.accumulate) v += p.cout[ci];
        if (p.zero_last && lf == p.L && l == lf && m == lf) v = 0.0;
        p.cout[ci] = v;
      }
    }
  }
}

template <int MB>
hipError_t launch_anal(const LegParams& p, hipStream_t s, bool f32m) {
  const dim3 grid(8 * p.items_per);
  if (f32m)
    hipLaunchKernelGGL((leg_anal_f64<MB, true>), grid, dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((leg_anal_f64<MB, false>), grid, dim3(256), 0, s, p);
  return hipGetLastError();
}

// Recurrence checkpoints: order m (blockIdx.y = m + L) runs its recurrence
// for 256 colatitudes per workgroup from the seed at l0 and stores the state
// (e_l, e_{l-1}) at l = l0 + kCk k, k = 1 .. nck - 1, with the same
// arithmetic as the transform kernels.  The coefficients are staged in LDS
// 64 degrees at a time (all threads of the workgroup share the order).
//
// Scaling against underflow: a seed below the normal range (far from the
// turning point, high degree) starts scaled, e = sseed = seed * 2^k; the
// recurrence is linear and homogeneous, so the scaled state runs the same
// steps, and whenever |e_l| passes 2^600 both values are divided by 2^600
// (exactly) and k drops by as much, until k = 0.  The stored states are the
// unscaled ones, e * 2^-k: zero while they are below the double range, and
// accurate once the order's values rise toward significance.  Lanes with
// k = 0 take exactly the steps of the transform kernels.
__global__ __launch_bounds__(256) void rec_checkpoints(LegParams p) {
  const int mi = blockIdx.y;
  const int m = mi - p.L;
  const int l0 = max(abs(m), p.nabs);
  const int nt = p.ntheta;
  const int t = blockIdx.x * 256 + threadIdx.x;
  const int tc = min(t, nt - 1);
  const double cs = p.cosv[tc];
  double e0 = p.seeds[(long)mi * nt + tc], e1 = 0.0;
  int kk = p.skexp ? p.skexp[(long)mi * nt + tc] : 0;
  if (kk > 0) e0 = p.sseed[(long)mi * nt + tc];
  const bool scaled = kk > 0;
  bool rose = false;  // a scaled lane reached 2^-60
  const int nck = (p.L - l0 + kCk) / kCk;
  __shared__ uint4 rs[128];
  const uint4* recq = reinterpret_cast<const uint4*>(p.rec + (ld_uniform(p.roff, mi) - l0));
  double2* out = const_cast<double2*>(p.ck) + ld_uniform(p.ckoff, mi) * (long)nt;
  // Entries up to L + kRecPad exist, so staging 64 entries past any
  // checkpoint's start stays in bounds.
  for (int k = 1; k < nck; k++) {
    const int lp = l0 + kCk * (k - 1);
    const int j0 = (k - 1) % (64 / kCk) * kCk;  // position in the staged 64
    if (j0 == 0) {
      __syncthreads();
      if (threadIdx.x < 128) rs[threadIdx.x] = recq[2 * (lp + 1) + threadIdx.x];
      __syncthreads();
    }
    const d4* rl = reinterpret_cast<const d4*>(rs) + j0;
#pragma unroll
    for (int j = 0; j < kCk; j++) {
      rec_step(rl[j], cs, e0, e1);
      if (kk > 0 && fabs(e0) > 0x1p+600) {
        const int r = min(kk, 600);
        e0 = ldexp(e0, -r);
        e1 = ldexp(e1, -r);
        kk -= r;
      }
    }
    const double s0 = kk > 0 ? ldexp(e0, -kk) : e0;
    const double s1 = kk > 0 ? ldexp(e1, -kk) : e1;
    rose = rose || (scaled && fabs(s0) >= 0x1p-60);
    if (t < nt) out[(long)(k - 1) * nt + t] = make_double2(s0, s1);
  }
  if (p.ck_flag && t < nt && rose) atomicOr(p.ck_flag, 1);
}

template <bool F32M>
hipError_t launch_synth(const LegParams& p, int T, hipStream_t s) {
  const dim3 nblocks(8 * p.items_per);  // work list: (order, column group, theta block)
  if (T == 2 && p.real)
    hipLaunchKernelGGL((leg_synth_f64<2, 1, F32M>), nblocks, dim3(256), 0, s, p);
  else if (T == 2)
    hipLaunchKernelGGL((leg_synth_f64<2, 0, F32M>), nblocks, dim3(256), 0, s, p);
  else if (p.real)
    hipLaunchKernelGGL((leg_synth_f64<1, 1, F32M>), nblocks, dim3(256), 0, s, p);
  else
    hipLaunchKernelGGL((leg_synth_f64<1, 0, F32M>), nblocks, dim3(256), 0, s, p);
  return hipGetLastError();
}

}  // namespace

hipError_t launch_rec_checkpoints(const LegParams& p, hipStream_t s) {
  hipLaunchKernelGGL(rec_checkpoints, dim3((p.ntheta + 255) / 256, 2 * p.L + 1), dim3(256), 0, s,
                     p);
  return hipGetLastError();
}

hipError_t launch_leg_synth_f64(const LegParams& p0, hipStream_t s, int tiles, bool f32m) {
  LegParams p = p0;
  p.n_tsb = (p.ntheta + 255) / 256;
  const int T = tiles == 1 || tiles == 2 ? tiles : (p.ncol > 16 ? 2 : 1);
  p.n_cg = (p.ncol + 16 * T - 1) / (16 * T);
  return f32m ? launch_synth<true>(p, T, s) : launch_synth<false>(p, T, s);
}

hipError_t launch_leg_anal_f64(const LegParams& p0, hipStream_t s, bool f32m) {
  if (!p0.ck || !p0.ckoff) return hipErrorInvalidValue;  // checkpoint table required
  LegParams p = p0;
  p.n_cg = (p.ncol + 15) / 16;
  p.nquad = (p.ntheta + 3) / 4;
  const int mb = p.nquad / 64;  // full 16-quad blocks per wave
  switch (mb) {
    case 0: return launch_anal<0>(p, s, f32m);
    case 1: return launch_anal<1>(p, s, f32m);
    case 2: return launch_anal<2>(p, s, f32m);
    case 3: return launch_anal<3>(p, s, f32m);
    case 4: return launch_anal<4>(p, s, f32m);
    case 5: return launch_anal<5>(p, s, f32m);
    case 6: return launch_anal<6>(p, s, f32m);
    case 7: return launch_anal<7>(p, s, f32m);
    case 8: return launch_anal<8>(p, s, f32m);
    case 9: return launch_anal<9>(p, s, f32m);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gsh
