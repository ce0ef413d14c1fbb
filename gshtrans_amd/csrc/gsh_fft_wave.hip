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
#include "gsh_internal.h"

namespace gsh {
namespace {

// LDS per wavefront (double2): one transpose round of R / H rows of 64.
constexpr int lds_per_wave(int R) { return (R >= 32 ? R / 2 : R) * 64; }
// Wavefronts (rings) per workgroup: 8 complex or 16 real fields of one
// colatitude, so the workgroup gathers and scatters the order bins of F / FA
// in whole 128-byte lines (4-wavefront workgroups, two per CU, measured
// slower: their 64-byte pieces cost more than the overlap gained).  At most
// 256 VGPRs (two wavefronts per SIMD), as the launch bounds require.
constexpr int waves_for(int) { return 8; }

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
// twa[k1 * 64 + lane] = exp(-2 pi i lane k1 / N); tw[j] = exp(-2 pi i j / N).
template <int R, int SIGN>
__device__ __forceinline__ void wave_fft(double2 (&v)[R], double2* lds,
                                         const double2* __restrict__ twa,
                                         const double2* __restrict__ tw, int lane) {
  constexpr int G = R / 8;        // DFT-8 groups per lane in stage B
  constexpr int P = 64 / R;       // lanes per k1 in the output
  constexpr int H = R >= 32 ? 2 : 1;  // transpose rounds
  constexpr int LR = log2c(R);

  // Stage A.
  dft_reg<R, SIGN>(v);
#pragma unroll
  for (int k1 = 1; k1 < R; k1++) {
    double2 w = twa[k1 * 64 + lane];
    v[bitrev(k1, LR)] = SIGN < 0 ? cmul(v[bitrev(k1, LR)], w) : cmulc(v[bitrev(k1, LR)], w);
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
#pragma unroll
  for (int c = 1; c < 8; c++) {
    const double2 w = tw[b * c * R];  // exp(-2 pi i b c / 64)
#pragma unroll
    for (int g = 0; g < G; g++)
      u[g][bitrev(c, 3)] = SIGN < 0 ? cmul(u[g][bitrev(c, 3)], w) : cmulc(u[g][bitrev(c, 3)], w);
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

__device__ __forceinline__ double2 ld2(const double* p) {
  return *reinterpret_cast<const double2*>(p);
}
__device__ __forceinline__ void st2(double* p, double2 v) {
  *reinterpret_cast<double2*>(p) = v;
}

// Synthesis: order bins of F[theta][slot][col] -> backward DFT -> field ring.
template <int R, int REAL, int F32>
__global__ __launch_bounds__(64 * waves_for(R), 2) void fft_synth_w(FFTParams p) {
  constexpr int W = waves_for(R);
  __shared__ double2 lds_all[W * lds_per_wave(R)];
  constexpr int N = 64 * R;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int fg = blockIdx.x;
  const int theta = blockIdx.y;
  const int f0 = (fg * W + wave) * (REAL ? 2 : 1);
  if (f0 >= p.nfield) return;
  const bool f1v = REAL && f0 + 1 < p.nfield;
  const int lmax = p.lmax;
  // Order bins of this field: slot s at Frow[16 s] (F[theta][col/16][slot][16]).
  const double* Frow = p.F + f_index(theta, 0, 2 * f0, p.nslot, (p.ncol + 15) >> 4);
  double2 v[R];
#pragma unroll
  for (int n1 = 0; n1 < R; n1++) {
    const int k = 64 * n1 + lane;
    double2 X = make_double2(0.0, 0.0);
    if (!REAL) {
      if (k <= lmax) X = ld2(Frow + 16 * (long)(k - p.m_first));
      if (k >= N - lmax) X = cadd(X, ld2(Frow + 16 * (long)(k - N - p.m_first)));
    } else {
      const int kk = k <= N / 2 ? k : N - k;
      if (kk <= lmax) {
        double2 xa = ld2(Frow + 16 * (long)kk);
        double2 xb = f1v ? ld2(Frow + 16 * (long)kk + 2) : make_double2(0.0, 0.0);
        // c2r reads bins 0..N/2 only and ignores Im of bins 0 and N/2.
        if (kk == 0 || kk == N / 2) {
          xa.y = 0.0;
          xb.y = 0.0;
        } else if (k > N / 2) {
          xa.y = -xa.y;
          xb.y = -xb.y;
        }
        X = make_double2(xa.x - xb.y, xa.y + xb.x);  // xa + i xb
      }
    }
    v[n1] = X;
  }
  wave_fft<R, +1>(v, lds_all + wave * lds_per_wave(R), p.twa, p.tw, lane);
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

// Analysis: field ring -> forward DFT -> weighted bins into the analysis
// layout FA (fa_index); the last ring also zeroes the padding colatitudes
// of the last quad.
template <int R, int REAL, int F32>
__global__ __launch_bounds__(64 * waves_for(R), 2) void fft_anal_w(FFTParams p) {
  constexpr int W = waves_for(R);
  __shared__ double2 lds_all[W * lds_per_wave(R)];
  constexpr int N = 64 * R;
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const int fg = blockIdx.x;
  const int theta = blockIdx.y;
  const int f0 = (fg * W + wave) * (REAL ? 2 : 1);
  if (f0 >= p.nfield) return;
  const bool f1v = REAL && f0 + 1 < p.nfield;
  const int lmax = p.lmax;
  double2 v[R];
  const long ia = (long)field_index(p.fmap, f0) * p.fstride + (long)theta * N + lane;
  if (!REAL) {
#pragma unroll
    for (int n1 = 0; n1 < R; n1++) v[n1] = ld_cplx<F32>(p.fin, ia + 64 * n1);
  } else {
    const long ib =
        f1v ? (long)field_index(p.fmap, f0 + 1) * p.fstride + (long)theta * N + lane : ia;
#pragma unroll
    for (int n1 = 0; n1 < R; n1++) {
      const double xa = ld_real<F32>(p.fin, ia + 64 * n1);
      const double xb = ld_real<F32>(p.fin, ib + 64 * n1);
      v[n1] = make_double2(xa, f1v ? xb : 0.0);
    }
  }
  wave_fft<R, -1>(v, lds_all + wave * lds_per_wave(R), p.twa, p.tw, lane);
  const double w = p.wq[theta];
  const int n_cg = p.n_cg, nquad = p.nquad;
  if (!REAL) {
    const int col = 2 * f0;
#pragma unroll
    for (int t = 0; t < R; t++) {
      const int k = lane + 64 * t;
      const double2 X = make_double2(v[t].x * w, v[t].y * w);
      if (k <= lmax) st2(p.F + fa_index(k - p.m_first, col, theta, n_cg, nquad), X);
      if (k >= N - lmax) st2(p.F + fa_index(k - N - p.m_first, col, theta, n_cg, nquad), X);
    }
  } else {
    // Z = FFT(x_a + i x_b): X_a[k] = (Z[k] + conj Z[N-k]) / 2,
    // X_b[k] = (Z[k] - conj Z[N-k]) / (2 i), k <= N/2.  Z[N - k] with
    // k = lane + 64 t sits in lane (64 - lane) % 64, register R-1-t (lane 0:
    // its own register (R - t) % R).
    const int src = ((64 - lane) & 63) << 2;
    const int col = 2 * f0;
#pragma unroll
    for (int t = 0; t <= R / 2; t++) {
      const double2 q = v[R - 1 - t];
      const int2 qx = __builtin_bit_cast(int2, q.x), qy = __builtin_bit_cast(int2, q.y);
      double2 zp;
      zp.x = __builtin_bit_cast(double, make_int2(__builtin_amdgcn_ds_bpermute(src, qx.x),
                                                  __builtin_amdgcn_ds_bpermute(src, qx.y)));
      zp.y = __builtin_bit_cast(double, make_int2(__builtin_amdgcn_ds_bpermute(src, qy.x),
                                                  __builtin_amdgcn_ds_bpermute(src, qy.y)));
      if (lane == 0) zp = v[(R - t) % R];
      const int k = lane + 64 * t;
      if (k <= lmax) {
        const double2 z = v[t];
        // Adjoint of the c2r synthesis: weight 2 for 0 < k < N/2, real
        // part only at k = 0 and N/2.
        const bool edge = k == 0 || 2 * k == N;
        const double hw = (p.adjoint && !edge ? 1.0 : 0.5) * w;
        const double im = p.adjoint && edge ? 0.0 : 1.0;
        const double2 xa = make_double2((z.x + zp.x) * hw, (z.y - zp.y) * hw * im);
        const double2 xb = make_double2((z.y + zp.y) * hw, (zp.x - z.x) * hw * im);
        st2(p.F + fa_index(k, col, theta, n_cg, nquad), xa);
        if (f1v) st2(p.F + fa_index(k, col + 2, theta, n_cg, nquad), xb);
      }
    }
  }
  if (theta == p.ntheta - 1) {
    const int ncw = REAL && f1v ? 4 : 2;
    for (int tp = theta + 1; tp < 4 * nquad; tp++)
      for (int s = lane; s < p.nslot; s += 64)
        for (int c = 0; c < ncw; c++) p.F[fa_index(s, 2 * f0 + c, tp, n_cg, nquad)] = 0.0;
  }
}

template <int R>
hipError_t launch_synth_r(const FFTParams& p, hipStream_t s) {
  const dim3 grid(p.nfg, p.ntheta), block(64 * waves_for(R));
  if (p.real && p.f32)
    hipLaunchKernelGGL((fft_synth_w<R, 1, 1>), grid, block, 0, s, p);
  else if (p.real)
    hipLaunchKernelGGL((fft_synth_w<R, 1, 0>), grid, block, 0, s, p);
  else if (p.f32)
    hipLaunchKernelGGL((fft_synth_w<R, 0, 1>), grid, block, 0, s, p);
  else
    hipLaunchKernelGGL((fft_synth_w<R, 0, 0>), grid, block, 0, s, p);
  return hipGetLastError();
}

template <int R>
hipError_t launch_anal_r(const FFTParams& p, hipStream_t s) {
  const dim3 grid(p.nfg, p.ntheta), block(64 * waves_for(R));
  if (p.real && p.f32)
    hipLaunchKernelGGL((fft_anal_w<R, 1, 1>), grid, block, 0, s, p);
  else if (p.real)
    hipLaunchKernelGGL((fft_anal_w<R, 1, 0>), grid, block, 0, s, p);
  else if (p.f32)
    hipLaunchKernelGGL((fft_anal_w<R, 0, 1>), grid, block, 0, s, p);
  else
    hipLaunchKernelGGL((fft_anal_w<R, 0, 0>), grid, block, 0, s, p);
  return hipGetLastError();
}

}  // namespace

bool wave_fft_supported(int N) { return N == 512 || N == 1024 || N == 2048; }

// Field groups of a launch: waves_for(R) rings per workgroup, 2 real fields per ring.
static int field_groups(const FFTParams& p) {
  const int per = waves_for(p.N / 64) * (p.real ? 2 : 1);
  return (p.nfield + per - 1) / per;
}

hipError_t launch_fft_synth_wave(const FFTParams& p0, hipStream_t s) {
  FFTParams p = p0;
  p.nfg = field_groups(p);
  switch (p.N) {
    case 512: return launch_synth_r<8>(p, s);
    case 1024: return launch_synth_r<16>(p, s);
    case 2048: return launch_synth_r<32>(p, s);
    default: return hipErrorInvalidValue;
  }
}

hipError_t launch_fft_anal_wave(const FFTParams& p0, hipStream_t s) {
  FFTParams p = p0;
  p.nfg = field_groups(p);
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
