// gsh_fft_wave.hip -- register-resident ring FFTs, one wavefront per ring,
// for ring lengths N = 64 R with R in {8, 16, 32} (L = 256, 512, 1024).
//
// Same transforms as gsh_fft.hip (GaussLegendreGrid.h:147-174 analysis,
// :266-319 synthesis: unnormalised DFTs of the longitude rings, bins k = m for
// m >= 0 and k = N + m for m < 0, c2r/r2c semantics for real fields), laid
// out for HBM: every ring is read and written as R rows of 64 consecutive
// points (1 KiB per wavefront instruction for complex rings), and the order
// bins are gathered from / scattered to the Legendre work buffers with whole
// 128-byte lines per workgroup (8 wavefronts = 8 complex or 16 real fields of
// one colatitude, i.e. one 16- or 32-column group).
//
// Four-step decomposition, n = 64 n1 + lane, k = k1 + R (c + 8 d):
//   A   per lane: DFT_R over n1 (registers), times W_N^(lane k1)
//   T1  transpose through LDS: lane = b + 8 (k1 % 8), b = lane % 8
//   B1  per lane: DFT_8 over a = lane_in / 8 (registers), times W_64^(b c)
//   T2  transpose through LDS: lane = k1 + R (c % P), P = 64 / R
//   B2  per lane: DFT_8 over b (registers)
// which leaves X[lane + 64 t] in register t.  Each transpose moves at most
// 16 KiB per wavefront per round (two rounds for R = 32); the LDS addresses
// are XOR-swizzled so that every 16-lane phase of a 16-byte access touches
// all 64 banks once.  Real fields are transformed two at a time as the
// complex ring x_a + i x_b.
#include <type_traits>

#include "gsh_internal.h"

namespace gsh {
namespace {

// LDS per wavefront (double2): one transpose round of R / H rows of 64.
constexpr int lds_per_wave(int R) { return (R >= 32 ? R / 2 : R) * 64; }
// Wavefronts (rings) per workgroup: 8 complex or 16 real fields of one
// colatitude, so the workgroup gathers and scatters the order bins of F / FA
// in whole pieces of 8 rings.  At N = 2048 (256 VGPRs per wave for 32
// complex doubles per lane, two wavefronts per SIMD) such a workgroup fills
// a CU and its load, transform and store phases run one after the other;
// 4-wavefront workgroups run two per CU, whose phases overlap.  They move
// the bins in pieces of 4 rings, which the XCD pairing of the grid
// (xcd_pair_map) sends through one L2 with the other 4 rings of the 8:
// bf16x6 at lMax=1024, B=64, synthesis 0.38 -> 0.32 ms (real pairs) and
// 0.69 -> 0.64 ms (complex), real-pair analysis 0.52 -> 0.50 ms; the
// complex analysis, whose FP32 bins would go out in 32-byte pieces, stays
// at 8 (0.83 against 1.07 ms).
constexpr int synth_waves(int R) { return R >= 32 ? 4 : 8; }
constexpr int anal_waves(int R, int real) { return R >= 32 && real ? 4 : 8; }

// cos(pi k / 16) and sin(pi k / 16), k = 0..16.
constexpr double kCos16[17] = {
    1.0,                     0.98078528040323044913, 0.92387953251128675613,
    0.83146961230254523708,  0.70710678118654752440, 0.55557023301960222474,
    0.38268343236508977173,  0.19509032201612826785, 0.0,
    -0.19509032201612826785, -0.38268343236508977173, -0.55557023301960222474,
    -0.70710678118654752440, -0.83146961230254523708, -0.92387953251128675613,
    -0.98078528040323044913, -1.0};
constexpr double kSin16[17] = {
    0.0,                    0.19509032201612826785, 0.38268343236508977173,
    0.55557023301960222474, 0.70710678118654752440, 0.83146961230254523708,
    0.92387953251128675613, 0.98078528040323044913, 1.0,
    0.98078528040323044913, 0.92387953251128675613, 0.83146961230254523708,
    0.70710678118654752440, 0.55557023301960222474, 0.38268343236508977173,
    0.19509032201612826785, 0.0};

__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ double2 cmulc(double2 a, double2 b) {
  return make_double2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}

// a * exp(i SIGN pi K / 16); K is a constant after unrolling.
template <int SIGN>
__device__ __forceinline__ double2 rot16(double2 a, int K) {
  if (K == 0) return a;
  if (K == 8) return SIGN > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
  // exp(i SIGN pi / 4) = h (1 + SIGN i), exp(i SIGN 3 pi / 4) = h (-1 + SIGN i)
  const double h = kCos16[4];
  if (K == 4) return make_double2(h * (a.x - SIGN * a.y), h * (a.y + SIGN * a.x));
  if (K == 12) return make_double2(h * (-a.x - SIGN * a.y), h * (SIGN * a.x - a.y));
  const double c = kCos16[K], s = SIGN * kSin16[K];
  return make_double2(a.x * c - a.y * s, a.x * s + a.y * c);
}

constexpr int bitrev(int x, int bits) {
  int r = 0;
  for (int i = 0; i < bits; i++) r |= ((x >> i) & 1) << (bits - 1 - i);
  return r;
}
constexpr int log2c(int x) { return x <= 1 ? 0 : 1 + log2c(x >> 1); }

// In-register DFT of length R (radix-2 decimation in frequency); on return
// v[bitrev(k)] holds X[k] = sum_j v_in[j] exp(i SIGN 2 pi j k / R).
template <int R, int SIGN>
__device__ __forceinline__ void dft_reg(double2* v) {
#pragma unroll
  for (int h = R / 2; h >= 1; h >>= 1) {
#pragma unroll
    for (int base = 0; base < R; base += 2 * h) {
#pragma unroll
      for (int j = 0; j < h; j++) {
        const double2 a = v[base + j], b = v[base + j + h];
        v[base + j] = cadd(a, b);
        // twiddle exp(i SIGN 2 pi j / (2h)) = exp(i SIGN pi (16 j / h) / 16)
        v[base + j + h] = rot16<SIGN>(csub(a, b), 16 * j / h);
      }
    }
  }
}

__device__ __forceinline__ void wave_sync() {
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Transform the ring held as v[n1] = x[64 n1 + lane]; on return v[t] =
// X[lane + 64 t] = sum_n x[n] exp(i SIGN 2 pi n (lane + 64 t) / N).
// twa[k1 * 64 + lane] = exp(-2 pi i lane k1 / (TS N)); tw[j] = exp(-2 pi i
// j / (TS N)): the tables of a TS times longer ring, read at stride TS.
template <int R, int SIGN, int TS = 1>
__device__ __forceinline__ void wave_fft(double2 (&v)[R], double2* lds,
                                         const double2* __restrict__ twa,
                                         const double2* __restrict__ tw, int lane) {
  constexpr int G = R / 8;        // DFT-8 groups per lane in stage B
  constexpr int P = 64 / R;       // lanes per k1 in the output
  constexpr int H = R >= 32 ? 2 : 1;  // transpose rounds
  constexpr int LR = log2c(R);

  // Stage A.  The twiddles W^(lane k1), W = exp(-2 pi i / (TS N)), from the
  // table at k1 = 1 and at multiples of 8 only: W^(lane (8 q + j)) =
  // W^(8 q lane) (W^lane)^j, j < 8 (at most 7 rounded products per twiddle),
  // so a lane reads R/8 + 1 table entries instead of R - 1.
  dft_reg<R, SIGN>(v);
  {
    double2 wj[8];
    wj[1] = twa[TS * 64 + lane];
#pragma unroll
    for (int j = 2; j < 8; j++) wj[j] = cmul(wj[j - 1], wj[1]);
#pragma unroll
    for (int q = 0; q < R / 8; q++) {
      const double2 wq = q == 0 ? make_double2(1.0, 0.0) : twa[(TS * 8 * q) * 64 + lane];
#pragma unroll
      for (int j = 0; j < 8; j++) {
        const int k1 = 8 * q + j;
        if (k1 == 0) continue;
        const double2 w = q == 0 ? wj[j] : (j == 0 ? wq : cmul(wq, wj[j]));
        v[bitrev(k1, LR)] = SIGN < 0 ? cmul(v[bitrev(k1, LR)], w) : cmulc(v[bitrev(k1, LR)], w);
      }
    }
  }

  // T1 + B1: lane (b, cp) = (lane % 8, lane / 8) gathers k1 = 8 g + cp,
  // a = 0..7 from rows k1 of the stage-A output.
  const int b = lane & 7, cp = lane >> 3;
  double2 u[G][8];
#pragma unroll
  for (int h = 0; h < H; h++) {
    if (h > 0) wave_sync();
#pragma unroll
    for (int kl = 0; kl < R / H; kl++) {
      const int k1 = h * (R / H) + kl;
      lds[kl * 64 + (lane ^ (8 * (k1 & 1)))] = v[bitrev(k1, LR)];
    }
    wave_sync();
#pragma unroll
    for (int gl = 0; gl < G / H; gl++)
#pragma unroll
      for (int a = 0; a < 8; a++)
        u[h * (G / H) + gl][a] = lds[(8 * gl + cp) * 64 + ((b + 8 * a) ^ (8 * (cp & 1)))];
  }
#pragma unroll
  for (int g = 0; g < G; g++) dft_reg<8, SIGN>(u[g]);
  {
    // exp(-2 pi i b c / 64) as the c-th power of the table's exp(-2 pi i b / 64).
    const double2 w1 = tw[b * R * TS];
    double2 w = w1;
#pragma unroll
    for (int c = 1; c < 8; c++) {
      if (c > 1) w = cmul(w, w1);
#pragma unroll
      for (int g = 0; g < G; g++)
        u[g][bitrev(c, 3)] = SIGN < 0 ? cmul(u[g][bitrev(c, 3)], w) : cmulc(u[g][bitrev(c, 3)], w);
    }
  }

  // T2 + B2: lane (k1o, cm) = (lane % R, lane / R) gathers c = cm + P cq,
  // all b.  Element (c, k1, b) sits at ((c - c0) R + k1) 8 + (b ^ sw(k1, c)).
  const int k1o = lane % R, cm = lane / R;
  double2 o[G][8];
#pragma unroll
  for (int h = 0; h < H; h++) {
    wave_sync();
#pragma unroll
    for (int g = 0; g < G; g++)
#pragma unroll
      for (int cl = 0; cl < 8 / H; cl++) {
        const int c = h * (8 / H) + cl;
        const int k1 = 8 * g + cp;
        const int sw = ((k1 >> 1) ^ ((c & 1) << 2)) & 7;
        lds[(cl * R + k1) * 8 + (b ^ sw)] = u[g][bitrev(c, 3)];
      }
    wave_sync();
#pragma unroll
    for (int cql = 0; cql < G / H; cql++) {
      const int cq = h * (G / H) + cql;
      const int c = cm + P * cq;
      const int cl = c - h * (8 / H);
      const int sw = ((k1o >> 1) ^ ((c & 1) << 2)) & 7;
#pragma unroll
      for (int bb = 0; bb < 8; bb++) o[cq][bb] = lds[(cl * R + k1o) * 8 + (bb ^ sw)];
    }
  }
#pragma unroll
  for (int cq = 0; cq < G; cq++) dft_reg<8, SIGN>(o[cq]);
  // X[lane + 64 (cq + G d)] = o[cq][bitrev(d)]
#pragma unroll
  for (int cq = 0; cq < G; cq++)
#pragma unroll
    for (int d = 0; d < 8; d++) v[cq + G * d] = o[cq][bitrev(d, 3)];
}

// Field element access for FP64 or FP32 storage (index in elements).
template <int F32>
__device__ __forceinline__ double2 ld_cplx(const void* b, long i) {
  if (F32) {
    const float2 v = static_cast<const float2*>(b)[i];
    return make_double2(v.x, v.y);
  }
  return static_cast<const double2*>(b)[i];
}
template <int F32>
__device__ __forceinline__ double ld_real(const void* b, long i) {
  if (F32) return static_cast<const float*>(b)[i];
  return static_cast<const double*>(b)[i];
}
template <int F32>
__device__ __forceinline__ void st_cplx(void* b, long i, double2 v) {
  if (F32)
    static_cast<float2*>(b)[i] = make_float2((float)v.x, (float)v.y);
  else
    static_cast<double2*>(b)[i] = v;
}
template <int F32>
__device__ __forceinline__ void st_real(void* b, long i, double v) {
  if (F32)
    static_cast<float*>(b)[i] = (float)v;
  else
    static_cast<double*>(b)[i] = v;
}

// Two consecutive real samples (elements i, i + 1) of a real field.
template <int F32>
__device__ __forceinline__ double2 ld_pair(const void* b, long i) {
  if (F32) {
    const float2 v = *reinterpret_cast<const float2*>(static_cast<const float*>(b) + i);
    return make_double2(v.x, v.y);
  }
  return *reinterpret_cast<const double2*>(static_cast<const double*>(b) + i);
}

// Re (op kExprReal) or Im (kExprImag) of a complex sample, and of the two
// consecutive complex samples i, i + 1 (FP64 storage).
__device__ __forceinline__ double cplx_part(double2 z, int op) {
  return op == kExprImag ? z.y : z.x;
}
__device__ __forceinline__ double2 cplx_part2(const void* b, long i, int op) {
  const double2 z0 = static_cast<const double2*>(b)[i];
  const double2 z1 = static_cast<const double2*>(b)[i + 1];
  return make_double2(cplx_part(z0, op), cplx_part(z1, op));
}
// A real expression of two consecutive samples of each operand.
__device__ __forceinline__ double2 expr_pair(int op, double2 a, double2 b, double s) {
  return make_double2(expr_real(op, a.x, b.x, s), expr_real(op, a.y, b.y, s));
}

__device__ __forceinline__ double2 ld2(const double* p) {
  return *reinterpret_cast<const double2*>(p);
}
__device__ __forceinline__ void st2(double* p, double2 v) {
  *reinterpret_cast<double2*>(p) = v;
}

// Order bins are exchanged between the rings of a workgroup through LDS, so
// that every global access of F / FA covers whole 128-byte lines.  Round r of
// H = N / LPW rounds covers LPW = lds_per_wave(R) bins (ring bins k = r LPW +
// b for complex rings; bins k = r LPW / 2 + b, xa and xb apart, when the
// analysis separates a real pair).  With W rings per workgroup, thread tid
// moves bins b = 64 i + tid / W of ring c = tid % W; ring c's entry b sits at
// lds[c LPW + (b ^ (16 / W) c)], which spreads every 16-lane phase of the
// b128 accesses over all 64 banks, both for the W rings x 16 / W bins of a
// cooperative access and for the 16 consecutive bins a ring reads or writes.
template <int W>
__device__ __forceinline__ int xch(int c, int b) { return b ^ ((16 / W) * c); }

// Workgroup -> (field group, colatitude).  Units of 8 workgroups -- field
// groups 2j and 2j+1 at the 4 colatitudes of a quad -- go to the XCDs in
// turn (workgroup b runs on XCD b % 8), each unit's workgroups back to back
// on its XCD: the pieces that the two field groups and neighbouring
// colatitudes write or read of the same lines of F / FA meet in that XCD's
// L2.  nfg is even; workgroups past the grid return false.
__device__ __forceinline__ bool xcd_pair_map(int nfg, int ntheta, int& fg, int& theta) {
  const int b = blockIdx.x;
  const int x = b & 7, sq = b >> 3;
  const int u = ((sq >> 3) << 3) | x;  // unit
  const int np = nfg >> 1;
  const int quad = u / np;
  fg = 2 * (u - quad * np) + (sq & 1);
  theta = 4 * quad + ((sq >> 1) & 3);
  return theta < ntheta;
}
// Grid of xcd_pair_map: whole units for every XCD.
inline unsigned xcd_pair_grid(int nfg, int ntheta) {
  const long units = (long)(nfg >> 1) * ((ntheta + 3) >> 2);
  return static_cast<unsigned>((units + 7) / 8 * 64);
}

// Synthesis: order bins of F[theta][slot][col] -> backward DFT -> field ring.
// Every wave stays to the end (the bin exchange synchronises the workgroup);
// a wave without a field transforms zeros and stores nothing.  All global
// loads of the exchange are issued before its first barrier.
//  - Complex rings: bin k holds order m = k (k <= lmax) or m = k - N
//    (k >= N - lmax), one load per bin; bin N/2 holds both when lmax = N/2,
//    the second added by the threads of bin offset 0 (N/2 is a multiple of
//    64).
//  - Real pairs: the exchange moves the slots kk < N/2 (xa, xb apart, BH
//    per round); register n1 of bin k = 64 n1 + lane takes slot kk = k for
//    k < N/2 and the conjugates of slot N - k for k > N/2 (c2r symmetry);
//    slot N/2 (lane 0, register R/2) is loaded by the ring itself.  c2r
//    ignores Im of bins 0 and N/2.
template <int R, int REAL, int F32, int FA32>
__global__ __launch_bounds__(64 * synth_waves(R), 2) void fft_synth_w(FFTParams p) {
  constexpr int W = synth_waves(R);
  constexpr int LPW = lds_per_wave(R);
  constexpr int N = 64 * R;
  constexpr int H = N / LPW;      // exchange rounds
  constexpr int CPT = LPW / 64;   // complex rings: bins per thread and round
  constexpr int BH = LPW / 2;     // real pairs: slots per round
  constexpr int CPH = BH / 64;
  static_assert((W == 8 || W == 4) && LPW % 128 == 0, "exchange of 4 or 8 rings");
  __shared__ double2 lds_all[W * LPW];
  const int lane = threadIdx.x & 63;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  int fg, theta;
  if (!xcd_pair_map(p.nfg, p.ntheta, fg, theta)) return;
  const int f0 = (fg * W + wave) * (REAL ? 2 : 1);
  const bool fv = f0 < p.nfield;
  const bool f1v = REAL && f0 + 1 < p.nfield;
  const int lmax = p.lmax, m_first = p.m_first;
  const int ncg16 = (p.ncol + 15) >> 4;
  // This thread's share of the cooperative gather: ring c, bins 64 i + bq.
  const int c = tid % W, bq = tid / W;
  const int cf = (fg * W + c) * (REAL ? 2 : 1);
  const bool cv = cf < p.nfield, cv1 = REAL && cf + 1 < p.nfield;
  // Bin (re, im) at element i of F: FP64, or FP32 (FA32: the same indices in
  // a float array) from the split-BF16 synthesis.  FP32 bins are loaded
  // unconditionally (an absent bin reads element f00, slot 0 and column 0 of
  // this colatitude, and is zeroed when stored to LDS) and stay raw until
  // then, so that no load waits for a conversion and all of them are in
  // flight together.
  using Bin = typename std::conditional<FA32 != 0, float2, double2>::type;
  const long fc = f_index(theta, 0, 2 * cf, p.nslot, ncg16);
  const long f00 = f_index(theta, 0, 0, p.nslot, ncg16);
  auto ldF = [&](bool ok, long i) -> Bin {
    if constexpr (FA32) {
      return *reinterpret_cast<const float2*>(reinterpret_cast<const float*>(p.F) + (ok ? i : f00));
    } else {
      double2 x = make_double2(0.0, 0.0);
      if (ok) x = ld2(p.F + i);
      return x;
    }
  };
  // A loaded bin as stored: FP32 bins that were not asked for are zeros.
  auto d2 = [](bool ok, const Bin& x) {
    if constexpr (FA32)
      return ok ? make_double2((double)x.x, (double)x.y) : make_double2(0.0, 0.0);
    else
      return x;
  };
  double2 v[R];
  if (!REAL) {
    Bin st[H][CPT];
    bool sok[H][CPT];
#pragma unroll
    for (int r = 0; r < H; r++)
#pragma unroll
      for (int i = 0; i < CPT; i++) {
        const int k = r * LPW + 64 * i + bq;
        const bool lo = k <= lmax;
        sok[r][i] = cv && (lo || k >= N - lmax);
        st[r][i] = ldF(sok[r][i], fc + 16 * (long)(lo ? k - m_first : k - N - m_first));
      }
    constexpr int RH = (N / 2) / LPW, IH = ((N / 2) % LPW) / 64;  // bin N/2
    const bool both = cv && bq == 0 && N / 2 <= lmax;
    const Bin xm = ldF(both, fc + 16 * (long)(-N / 2 - m_first));
#pragma unroll
    for (int r = 0; r < H; r++) {
      if (r > 0) __syncthreads();  // round r - 1 read back
#pragma unroll
      for (int i = 0; i < CPT; i++) {
        double2 x = d2(sok[r][i], st[r][i]);
        if (r == RH && i == IH) x = cadd(x, d2(both, xm));
        lds_all[c * LPW + xch<W>(c, 64 * i + bq)] = x;
      }
      __syncthreads();
#pragma unroll
      for (int j = 0; j < CPT; j++)
        v[r * CPT + j] = lds_all[wave * LPW + xch<W>(wave, 64 * j + lane)];
    }
  } else {
    Bin sa[H][CPH], sb[H][CPH];
    bool aok[H][CPH], bok[H][CPH];
#pragma unroll
    for (int r = 0; r < H; r++)
#pragma unroll
      for (int i = 0; i < CPH; i++) {
        const int kk = r * BH + 64 * i + bq;
        aok[r][i] = cv && kk <= lmax;
        bok[r][i] = cv1 && kk <= lmax;
        sa[r][i] = ldF(aok[r][i], fc + 16 * (long)kk);
        sb[r][i] = ldF(bok[r][i], fc + 16 * (long)kk + 2);
      }
    // Slot N/2 of this ring's pair (register R/2, lane 0).
    const long fw = f_index(theta, 0, 2 * f0, p.nslot, ncg16);
    const bool nok = fv && lane == 0 && N / 2 <= lmax;
    const Bin na = ldF(nok, fw + 16 * (long)(N / 2));
    const Bin nb = ldF(nok && f1v, fw + 16 * (long)(N / 2) + 2);
#pragma unroll
    for (int r = 0; r < H; r++) {
      if (r > 0) __syncthreads();
#pragma unroll
      for (int i = 0; i < CPH; i++) {
        lds_all[c * LPW + xch<W>(c, 64 * i + bq)] = d2(aok[r][i], sa[r][i]);
        lds_all[c * LPW + BH + xch<W>(c, 64 * i + bq)] = d2(bok[r][i], sb[r][i]);
      }
      __syncthreads();
      const double2* lw = lds_all + wave * LPW;
#pragma unroll
      for (int j = 0; j < CPH; j++) {  // k = kk = r BH + 64 j + lane < N/2
        double2 xa = lw[xch<W>(wave, 64 * j + lane)];
        double2 xb = lw[BH + xch<W>(wave, 64 * j + lane)];
        if (r == 0 && j == 0 && lane == 0) {
          xa.y = 0.0;
          xb.y = 0.0;
        }
        v[r * CPH + j] = make_double2(xa.x - xb.y, xa.y + xb.x);  // xa + i xb
      }
      // k = 64 n1 + lane > N/2 with slot N - k in this round.  Register
      // R - r CPH (r > 0) already holds its lanes >= 1 from round r - 1; the
      // others are first touched here (lane 0 of register R/2: slot N/2).
#pragma unroll
      for (int n1 = R - (r + 1) * CPH; n1 <= R - r * CPH && n1 < R; n1++) {
        const int kk = N - 64 * n1 - lane;
        const bool in = kk >= r * BH && kk < (r + 1) * BH;
        const double2 xa = lw[xch<W>(wave, in ? kk - r * BH : 0)];
        const double2 xb = lw[BH + xch<W>(wave, in ? kk - r * BH : 0)];
        const double2 X = make_double2(xa.x + xb.y, xb.x - xa.y);  // conj xa + i conj xb
        const double2 keep = n1 == R - r * CPH ? v[n1] : make_double2(0.0, 0.0);
        v[n1] = in ? X : keep;
      }
    }
    if (lane == 0) v[R / 2] = make_double2(d2(nok, na).x, d2(nok && f1v, nb).x);
  }
  // Each wave now works in its own LDS region only.
  wave_fft<R, +1>(v, lds_all + wave * LPW, p.twa, p.tw, lane);
  if (!fv) return;
  const long oa = (long)field_index(p.fmap, f0) * p.fstride + (long)theta * N + lane;
  if (!REAL) {
#pragma unroll
    for (int t = 0; t < R; t++) st_cplx<F32>(p.fout, oa + 64 * t, v[t]);
  } else {
#pragma unroll
    for (int t = 0; t < R; t++) st_real<F32>(p.fout, oa + 64 * t, v[t].x);
    if (f1v) {
      const long ob = (long)field_index(p.fmap, f0 + 1) * p.fstride + (long)theta * N + lane;
#pragma unroll
      for (int t = 0; t < R; t++) st_real<F32>(p.fout, ob + 64 * t, v[t].y);
    }
  }
}

// Separation of a real pair transformed as Z = FFT(x_a + i x_b), weighted by
// w: X_a[k] = (Z[k] + conj Z[N-k]) / 2, X_b[k] = (Z[k] - conj Z[N-k]) / (2 i)
// for k = lane + 64 t <= N/2.  Z[N - k] sits in lane (64 - lane) % 64,
// register R-1-t (lane 0: its own register (R - t) % R).  All lanes take
// part (ds_bpermute).
template <int R>
__device__ __forceinline__ void separate(const double2 (&v)[R], int t, int lane, double w,
                                         bool adjoint, double2& xa, double2& xb) {
  constexpr int N = 64 * R;
  const int src = ((64 - lane) & 63) << 2;
  const double2 q = v[R - 1 - t];
  const int2 qx = __builtin_bit_cast(int2, q.x), qy = __builtin_bit_cast(int2, q.y);
  double2 zp;
  zp.x = __builtin_bit_cast(double, make_int2(__builtin_amdgcn_ds_bpermute(src, qx.x),
                                              __builtin_amdgcn_ds_bpermute(src, qx.y)));
  zp.y = __builtin_bit_cast(double, make_int2(__builtin_amdgcn_ds_bpermute(src, qy.x),
                                              __builtin_amdgcn_ds_bpermute(src, qy.y)));
  if (lane == 0) zp = v[(R - t) % R];
  const int k = lane + 64 * t;
  const double2 z = v[t];
  // Adjoint of the c2r synthesis: weight 2 for 0 < k < N/2, real part only
  // at k = 0 and N/2.
  const bool edge = k == 0 || 2 * k == N;
  const double hw = (adjoint && !edge ? 1.0 : 0.5) * w;
  const double im = adjoint && edge ? 0.0 : 1.0;
  xa = make_double2((z.x + zp.x) * hw, (z.y - zp.y) * hw * im);
  xb = make_double2((z.y + zp.y) * hw, (zp.x - z.x) * hw * im);
}

// One weighted bin into FA: FP64, or FP32 (the same indices in a float
// array) for the split-BF16 modes.
template <int FA32>
__device__ __forceinline__ void st_bin(double* F, long i, double2 X) {
  if (FA32)
    *reinterpret_cast<float2*>(reinterpret_cast<float*>(F) + i) = make_float2((float)X.x, (float)X.y);
  else
    st2(F + i, X);
}

// Analysis: field ring -> forward DFT -> weighted bins into the analysis
// layout FA (fa_index), exchanged through LDS as in the synthesis; the last
// ring row also zeroes the padding colatitudes of the last quad.  A wave
// without a field transforms zeros and stores nothing.
template <int R, int REAL, int F32, int FA32>
__global__ __launch_bounds__(64 * anal_waves(R, REAL), 2) void fft_anal_w(FFTParams p) {
  constexpr int W = anal_waves(R, REAL);
  constexpr int LPW = lds_per_wave(R);
  constexpr int N = 64 * R;
  constexpr int H = N / LPW;      // exchange rounds
  constexpr int CPT = LPW / 64;   // complex rings: bins per thread and round
  constexpr int BH = LPW / 2;     // real pairs: bins per round (xa, xb apart)
  constexpr int CPH = BH / 64;
  static_assert((W == 8 || W == 4) && LPW % 128 == 0, "exchange of 4 or 8 rings");
  __shared__ double2 lds_all[W * LPW];
  const int lane = threadIdx.x & 63;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  int fg, theta;
  if (!xcd_pair_map(p.nfg, p.ntheta, fg, theta)) return;
  const int f0 = (fg * W + wave) * (REAL ? 2 : 1);
  const bool fv = f0 < p.nfield;
  const bool f1v = REAL && f0 + 1 < p.nfield;
  const int lmax = p.lmax, m_first = p.m_first;
  double2 v[R];
  double2* lw = lds_all + wave * LPW;
  if constexpr (REAL && R >= 16) {
    // Real pairs: each lane loads the even / odd samples x[2m], x[2m + 1]
    // (m = 64 a + lane) of both rings with one 16-byte access per ring, the
    // two half-length rings are transformed, and X[k] = E[k] + W_N^k O[k],
    // X[k + N/2] = E[k] - W_N^k O[k] (k = lane + 64 t < N/2).
    constexpr int R2 = R / 2;
    double2 e[R2], o[R2];
    if (fv && !p.src_cplx && !p.op) {
      // Plain fields: every load issued before the first use (a branch
      // between the loads would make each one wait for the last).  A missing
      // second field reads the first and is zeroed after the loads.
      const long ia = (long)field_index(p.fmap, f0) * p.fstride + (long)theta * N + 2 * lane;
      const long ib =
          f1v ? (long)field_index(p.fmap, f0 + 1) * p.fstride + (long)theta * N + 2 * lane : ia;
#pragma unroll
      for (int a = 0; a < R2; a++) {
        const double2 pa = ld_pair<F32>(p.fin, ia + 128 * a);
        const double2 pb = ld_pair<F32>(p.fin, ib + 128 * a);
        e[a] = make_double2(pa.x, pb.x);
        o[a] = make_double2(pa.y, pb.y);
      }
      if (!f1v) {
#pragma unroll
        for (int a = 0; a < R2; a++) e[a].y = o[a].y = 0.0;
      }
    } else if (fv) {
      const long ia = (long)field_index(p.fmap, f0) * p.fstride + (long)theta * N + 2 * lane;
      const long ib =
          f1v ? (long)field_index(p.fmap, f0 + 1) * p.fstride + (long)theta * N + 2 * lane : ia;
#pragma unroll
      for (int a = 0; a < R2; a++) {
        double2 pa, pb;
        if (p.src_cplx) {  // Re or Im of complex fields (samples 2m, 2m + 1)
          pa = cplx_part2(p.fin, ia + 128 * a, p.op);
          pb = cplx_part2(p.fin, ib + 128 * a, p.op);
        } else {
          pa = ld_pair<F32>(p.fin, ia + 128 * a);
          pb = ld_pair<F32>(p.fin, ib + 128 * a);
          if (p.op) {  // pointwise expression of the fields
            double2 qa = make_double2(0.0, 0.0), qb = qa;
            if (expr_binary(p.op)) {
              qa = ld_pair<F32>(p.fin2, ia + 128 * a);
              qb = ld_pair<F32>(p.fin2, ib + 128 * a);
            }
            pa = expr_pair(p.op, pa, qa, p.scalar.x);
            pb = expr_pair(p.op, pb, qb, p.scalar.x);
          }
        }
        if (!f1v) pb = make_double2(0.0, 0.0);
        e[a] = make_double2(pa.x, pb.x);
        o[a] = make_double2(pa.y, pb.y);
      }
    } else {
#pragma unroll
      for (int a = 0; a < R2; a++) e[a] = o[a] = make_double2(0.0, 0.0);
    }
    wave_fft<R2, -1, 2>(e, lw, p.twa, p.tw, lane);
    wave_fft<R2, -1, 2>(o, lw, p.twa, p.tw, lane);
    // W_N^(lane + 64 t) = W_N^lane exp(-2 pi i t / R) (N = 64 R), the second
    // factor a constant rotation by -pi (32 t / R) / 16.
    static_assert(R <= 32, "rotation in steps of pi / 16");
    const double2 wl = p.tw[lane];
#pragma unroll
    for (int t = 0; t < R2; t++) {
      const double2 wo = cmul(o[t], rot16<-1>(wl, (32 / R) * t));
      v[t] = cadd(e[t], wo);
      v[t + R2] = csub(e[t], wo);
    }
  } else {
  if (fv) {
    const long ia = (long)field_index(p.fmap, f0) * p.fstride + (long)theta * N + lane;
    const long ib =
        REAL && f1v ? (long)field_index(p.fmap, f0 + 1) * p.fstride + (long)theta * N + lane : ia;
    if (!REAL && !p.op) {
      // Plain fields: every load issued before the first use.
#pragma unroll
      for (int n1 = 0; n1 < R; n1++) v[n1] = ld_cplx<F32>(p.fin, ia + 64 * n1);
    } else if (!REAL) {
#pragma unroll
      for (int n1 = 0; n1 < R; n1++) {
        double2 x = ld_cplx<F32>(p.fin, ia + 64 * n1);
        if (p.op) {  // pointwise expression of the fields
          const double2 y =
              expr_binary(p.op) ? ld_cplx<F32>(p.fin2, ia + 64 * n1) : make_double2(0.0, 0.0);
          x = expr_cplx(p.op, x, y, p.scalar);
        }
        v[n1] = x;
      }
    } else {
#pragma unroll
      for (int n1 = 0; n1 < R; n1++) {
        double xa, xb;
        if (p.src_cplx) {  // Re or Im of complex fields
          xa = cplx_part(ld_cplx<F32>(p.fin, ia + 64 * n1), p.op);
          xb = cplx_part(ld_cplx<F32>(p.fin, ib + 64 * n1), p.op);
        } else {
          xa = ld_real<F32>(p.fin, ia + 64 * n1);
          xb = ld_real<F32>(p.fin, ib + 64 * n1);
          if (p.op) {  // pointwise expression of the fields
            const bool two = expr_binary(p.op);
            xa = expr_real(p.op, xa, two ? ld_real<F32>(p.fin2, ia + 64 * n1) : 0.0, p.scalar.x);
            xb = expr_real(p.op, xb, two ? ld_real<F32>(p.fin2, ib + 64 * n1) : 0.0, p.scalar.x);
          }
        }
        v[n1] = make_double2(xa, f1v ? xb : 0.0);
      }
    }
  } else {
#pragma unroll
    for (int n1 = 0; n1 < R; n1++) v[n1] = make_double2(0.0, 0.0);
  }
  wave_fft<R, -1>(v, lw, p.twa, p.tw, lane);
  }
  const double w = p.wq[theta];
  const int n_cg = p.n_cg, nquad = p.nquad;
  // This thread's share of the cooperative scatter: ring c, bins 64 i + bq.
  const int c = tid % W, bq = tid / W;
  const int cf = (fg * W + c) * (REAL ? 2 : 1);
  const bool cv = cf < p.nfield, cv1 = REAL && cf + 1 < p.nfield;
  if (!REAL) {
    // Bin k goes to order k (k <= lmax) or k - N (k >= N - lmax), at
    // fa_index(slot, 2 cf, theta) = slot * (n_cg nquad 64) + that of slot 0;
    // bin N/2 (k = N/2, lane bq = 0 of round RH, piece IH) goes to both when
    // lmax = N/2.  One store per bin, one branch.
    constexpr int RH = (N / 2) / LPW, IH = ((N / 2) % LPW) / 64;
    const long sst = (long)n_cg * nquad * 64;
    const long cb0 = fa_index(0, 2 * cf, theta, n_cg, nquad);
#pragma unroll
    for (int r = 0; r < H; r++) {
      if (r > 0) __syncthreads();  // round r - 1 stored
#pragma unroll
      for (int j = 0; j < CPT; j++)
        lw[xch<W>(wave, 64 * j + lane)] = make_double2(v[r * CPT + j].x * w, v[r * CPT + j].y * w);
      __syncthreads();
#pragma unroll
      for (int i = 0; i < CPT; i++) {
        const int b = 64 * i + bq;
        const int k = r * LPW + b;
        const double2 X = lds_all[c * LPW + xch<W>(c, b)];
        const bool lo = k <= lmax;
        const int slot = lo ? k - m_first : k - N - m_first;
        if (cv && (lo || k >= N - lmax)) st_bin<FA32>(p.F, cb0 + slot * sst, X);
        if (r == RH && i == IH && cv && bq == 0 && N / 2 <= lmax)
          st_bin<FA32>(p.F, cb0 + (-N / 2 - m_first) * sst, X);
      }
    }
  } else {
#pragma unroll
    for (int r = 0; r < H; r++) {
      if (r > 0) __syncthreads();
#pragma unroll
      for (int j = 0; j < CPH; j++) {
        double2 xa, xb;
        separate<R>(v, r * CPH + j, lane, w, p.adjoint, xa, xb);
        lw[xch<W>(wave, 64 * j + lane)] = xa;
        lw[BH + xch<W>(wave, 64 * j + lane)] = xb;
      }
      __syncthreads();
#pragma unroll
      for (int i = 0; i < CPH; i++) {
        const int b = 64 * i + bq;
        const int k = r * BH + b;
        if (cv && k <= lmax) {
          st_bin<FA32>(p.F, fa_index(k, 2 * cf, theta, n_cg, nquad), lds_all[c * LPW + xch<W>(c, b)]);
          if (cv1)
            st_bin<FA32>(p.F, fa_index(k, 2 * cf + 2, theta, n_cg, nquad),
                         lds_all[c * LPW + BH + xch<W>(c, b)]);
        }
      }
    }
    // Bin N/2 (lane 0, register R/2) exists only when lmax = N/2.
    if (fv && N / 2 <= lmax) {
      double2 xa, xb;
      separate<R>(v, R / 2, lane, w, p.adjoint, xa, xb);
      if (lane == 0) {
        st_bin<FA32>(p.F, fa_index(N / 2, 2 * f0, theta, n_cg, nquad), xa);
        if (f1v) st_bin<FA32>(p.F, fa_index(N / 2, 2 * f0 + 2, theta, n_cg, nquad), xb);
      }
    }
  }
  if (fv && theta == p.ntheta - 1) {
    const int ncw = REAL && f1v ? 4 : 2;
    for (int tp = theta + 1; tp < 4 * nquad; tp++)
      for (int s = lane; s < p.nslot; s += 64)
        for (int cc = 0; cc < ncw; cc++) {
          const long zi = fa_index(s, 2 * f0 + cc, tp, n_cg, nquad);
          if (FA32)
            reinterpret_cast<float*>(p.F)[zi] = 0.0f;
          else
            p.F[zi] = 0.0;
        }
  }
}

template <int R>
hipError_t launch_synth_r(const FFTParams& p, hipStream_t s) {
  // one workgroup per (field group, colatitude), in xcd_pair_map's order
  const dim3 grid(xcd_pair_grid(p.nfg, p.ntheta)), block(64 * synth_waves(R));
  if (p.fa32) {
    if (p.real && p.f32)
      hipLaunchKernelGGL((fft_synth_w<R, 1, 1, 1>), grid, block, 0, s, p);
    else if (p.real)
      hipLaunchKernelGGL((fft_synth_w<R, 1, 0, 1>), grid, block, 0, s, p);
    else if (p.f32)
      hipLaunchKernelGGL((fft_synth_w<R, 0, 1, 1>), grid, block, 0, s, p);
    else
      hipLaunchKernelGGL((fft_synth_w<R, 0, 0, 1>), grid, block, 0, s, p);
  } else if (p.real && p.f32)
    hipLaunchKernelGGL((fft_synth_w<R, 1, 1, 0>), grid, block, 0, s, p);
  else if (p.real)
    hipLaunchKernelGGL((fft_synth_w<R, 1, 0, 0>), grid, block, 0, s, p);
  else if (p.f32)
    hipLaunchKernelGGL((fft_synth_w<R, 0, 1, 0>), grid, block, 0, s, p);
  else
    hipLaunchKernelGGL((fft_synth_w<R, 0, 0, 0>), grid, block, 0, s, p);
  return hipGetLastError();
}

template <int R>
hipError_t launch_anal_r(const FFTParams& p, hipStream_t s) {
  // one workgroup per (field group, colatitude), in xcd_pair_map's order
  const dim3 grid(xcd_pair_grid(p.nfg, p.ntheta)), block(64 * (p.real ? anal_waves(R, 1) : anal_waves(R, 0)));
  if (p.fa32) {
    if (p.real && p.f32)
      hipLaunchKernelGGL((fft_anal_w<R, 1, 1, 1>), grid, block, 0, s, p);
    else if (p.real)
      hipLaunchKernelGGL((fft_anal_w<R, 1, 0, 1>), grid, block, 0, s, p);
    else if (p.f32)
      hipLaunchKernelGGL((fft_anal_w<R, 0, 1, 1>), grid, block, 0, s, p);
    else
      hipLaunchKernelGGL((fft_anal_w<R, 0, 0, 1>), grid, block, 0, s, p);
  } else if (p.real && p.f32)
    hipLaunchKernelGGL((fft_anal_w<R, 1, 1, 0>), grid, block, 0, s, p);
  else if (p.real)
    hipLaunchKernelGGL((fft_anal_w<R, 1, 0, 0>), grid, block, 0, s, p);
  else if (p.f32)
    hipLaunchKernelGGL((fft_anal_w<R, 0, 1, 0>), grid, block, 0, s, p);
  else
    hipLaunchKernelGGL((fft_anal_w<R, 0, 0, 0>), grid, block, 0, s, p);
  return hipGetLastError();
}

}  // namespace

bool wave_fft_supported(int N) { return N == 512 || N == 1024 || N == 2048; }

// Field groups of a launch (W rings per workgroup, 2 real fields per ring),
// rounded up to an even count for xcd_pair_map (the extra group's waves
// transform zeros and store nothing).
static int field_groups(const FFTParams& p, int W) {
  const int per = W * (p.real ? 2 : 1);
  return ((p.nfield + per - 1) / per + 1) / 2 * 2;
}

hipError_t launch_fft_synth_wave(const FFTParams& p0, hipStream_t s) {
  FFTParams p = p0;
  p.nfg = field_groups(p, synth_waves(p.N / 64));
  switch (p.N) {
    case 512: return launch_synth_r<8>(p, s);
    case 1024: return launch_synth_r<16>(p, s);
    case 2048: return launch_synth_r<32>(p, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fft_anal_wave(const FFTParams& p0, hipStream_t s) {
  FFTParams p = p0;
  p.nfg = field_groups(p, anal_waves(p.N / 64, p.real));
  p.n_cg = (p.ncol + 15) / 16;
  p.nquad = (p.ntheta + 3) / 4;
  switch (p.N) {
    case 512: return launch_anal_r<8>(p, s);
    case 1024: return launch_anal_r<16>(p, s);
    case 2048: return launch_anal_r<32>(p, s);
    default: return hipErrorInvalidValue;
  }
}

}  // namespace gsh
