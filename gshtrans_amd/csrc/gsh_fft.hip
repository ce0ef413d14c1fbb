// gsh_fft.hip -- batched longitude-ring FFTs for the Gauss-Legendre grid.
//
// The reference runs one FFTW plan per ring (GaussLegendreGrid.h:147-174 for
// the analysis, :266-319 for the synthesis): an unnormalised c2c/c2r backward
// transform of the order bins F_m placed at k = m (m >= 0) and k = N + m
// (m < 0), and an unnormalised c2c/r2c forward transform of each ring.  Here
// each workgroup loads `rings` rings of N = 2L points into LDS (64 KiB of
// complex values, 128 KiB for transforms of 8192 points), runs a Stockham radix-4 (+ one radix-2) autosort FFT in
// place (values go through registers between barriers), and writes the rings
// back.  A ring length that is not a power of two is transformed with
// Bluestein's chirp-z identity jk = (j^2 + k^2 - (k-j)^2)/2: chirp multiply,
// power-of-two FFT of length M >= 2N-1, multiply by the transformed chirp
// filter, inverse FFT, chirp multiply (Bluestein 1970).  The synthesis kernel also assembles the bins from
// the Legendre work buffer F[theta][slot][col]; the analysis kernel scatters
// X_m * w_theta * 2 pi / N back into F for the Legendre contraction.
#include "gsh_internal.h"

namespace gsh {
namespace {

constexpr int kFFTThreads = 256;
// Complex values per workgroup: CAP = 4096 (64 KiB of LDS) holds every ring
// of up to 4096 points, or its Bluestein convolution, several rings at a
// time; CAP = 8192 (128 KiB) takes the 8192-point transforms of the longer
// rings, one at a time.
constexpr int kFFTElems = 4096;
constexpr int kFFTElemsMax = 8192;

__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
  return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
__device__ __forceinline__ double2 cadd(double2 a, double2 b) {
  return make_double2(a.x + b.x, a.y + b.y);
}
__device__ __forceinline__ double2 csub(double2 a, double2 b) {
  return make_double2(a.x - b.x, a.y - b.y);
}
// multiply by s*i (s = +1 or -1)
__device__ __forceinline__ double2 cmuli(double2 a, int s) {
  return s > 0 ? make_double2(-a.y, a.x) : make_double2(a.y, -a.x);
}

// Stockham autosort passes over `rings` rings of length N held in buf
// (ring q at buf[q*N]).  sign = -1 forward, +1 backward.
template <int SIGN, int CAP>
__device__ void stockham(double2* buf, int N, int rings, const double2* __restrict__ tw) {
  int logN = 31 - __clz(N);
  int Ns = 1;
  // radix-4 passes
  while (logN >= 2) {
    const int quarter = N >> 2;
    const int total = rings * quarter;  // butterflies
    constexpr int kPer = CAP / 4 / kFFTThreads;
    double2 v[kPer][4];
    int dst[kPer];
#pragma unroll
    for (int i = 0; i < kPer; i++) {
      const int bi = threadIdx.x + i * kFFTThreads;
      dst[i] = -1;
      if (bi >= total) continue;
      const int q = bi / quarter;
      const int j = bi - q * quarter;
      const int k = j & (Ns - 1);
      const double2* x = buf + q * N;
      double2 a0 = x[j], a1 = x[j + quarter], a2 = x[j + 2 * quarter], a3 = x[j + 3 * quarter];
      if (Ns > 1) {
        const int step = N / (Ns * 4);
        double2 w1 = tw[k * step], w2 = tw[2 * k * step], w3 = tw[3 * k * step];
        if (SIGN > 0) {
          w1.y = -w1.y;
          w2.y = -w2.y;
          w3.y = -w3.y;
        }
        a1 = cmul(a1, w1);
        a2 = cmul(a2, w2);
        a3 = cmul(a3, w3);
      }
      const double2 t0 = cadd(a0, a2), t1 = csub(a0, a2);
      const double2 t2 = cadd(a1, a3), t3 = cmuli(csub(a1, a3), SIGN);
      v[i][0] = cadd(t0, t2);
      v[i][1] = cadd(t1, t3);
      v[i][2] = csub(t0, t2);
      v[i][3] = csub(t1, t3);
      dst[i] = q * N + (j / Ns) * Ns * 4 + k;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPer; i++) {
      if (dst[i] < 0) continue;
      buf[dst[i]] = v[i][0];
      buf[dst[i] + Ns] = v[i][1];
      buf[dst[i] + 2 * Ns] = v[i][2];
      buf[dst[i] + 3 * Ns] = v[i][3];
    }
    __syncthreads();
    Ns *= 4;
    logN -= 2;
  }
  if (logN == 1) {
    const int half = N >> 1;
    const int total = rings * half;
    constexpr int kPer = CAP / 2 / kFFTThreads;
    double2 v[kPer][2];
    int dst[kPer];
#pragma unroll
    for (int i = 0; i < kPer; i++) {
      const int bi = threadIdx.x + i * kFFTThreads;
      dst[i] = -1;
      if (bi >= total) continue;
      const int q = bi / half;
      const int j = bi - q * half;
      const int k = j & (Ns - 1);
      const double2* x = buf + q * N;
      double2 a0 = x[j], a1 = x[j + half];
      if (Ns > 1) {
        double2 w1 = tw[k * (N / (Ns * 2))];
        if (SIGN > 0) w1.y = -w1.y;
        a1 = cmul(a1, w1);
      }
      v[i][0] = cadd(a0, a1);
      v[i][1] = csub(a0, a1);
      dst[i] = q * N + (j / Ns) * Ns * 2 + k;
    }
    __syncthreads();
#pragma unroll
    for (int i = 0; i < kPer; i++) {
      if (dst[i] < 0) continue;
      buf[dst[i]] = v[i][0];
      buf[dst[i] + Ns] = v[i][1];
    }
    __syncthreads();
  }
}

__device__ __forceinline__ double2 cconj(double2 a) { return make_double2(a.x, -a.y); }

// Unnormalised DFT of length N of each ring (ring q at buf[q*M], zero from N
// to M when Bluestein is used).  SIGN = -1 forward, +1 backward.
template <int SIGN, int CAP>
__device__ void ring_dft(double2* buf, const FFTParams& p) {
  const int N = p.N, M = p.M, R = p.rings;
  if (M == N) {
    stockham<SIGN, CAP>(buf, N, R, p.tw);
    return;
  }
  for (int idx = threadIdx.x; idx < R * N; idx += kFFTThreads) {
    const int q = idx / N, j = idx - q * N;
    const double2 w = SIGN < 0 ? p.chirp[j] : cconj(p.chirp[j]);
    buf[q * M + j] = cmul(buf[q * M + j], w);
  }
  __syncthreads();
  stockham<-1, CAP>(buf, M, R, p.tw);
  const double2* h = SIGN < 0 ? p.hf : p.hb;
  for (int idx = threadIdx.x; idx < R * M; idx += kFFTThreads)
    buf[idx] = cmul(buf[idx], h[idx & (M - 1)]);
  __syncthreads();
  stockham<+1, CAP>(buf, M, R, p.tw);
  for (int idx = threadIdx.x; idx < R * N; idx += kFFTThreads) {
    const int q = idx / N, j = idx - q * N;
    const double2 w = SIGN < 0 ? p.chirp[j] : cconj(p.chirp[j]);
    buf[q * M + j] = cmul(buf[q * M + j], w);
  }
  __syncthreads();
}

// Synthesis: bins from F, backward FFT, write field rings.
template <int CAP>
__global__ __launch_bounds__(kFFTThreads) void fft_synth(FFTParams p) {
  __shared__ double2 buf[CAP];
  const int theta = blockIdx.x / p.nfg;
  const int fg = blockIdx.x % p.nfg;
  const int N = p.N, M = p.M, R = p.rings;
  const int ncg16 = (p.ncol + 15) >> 4;
  // Fill: consecutive threads read consecutive fields of one bin.
  for (int idx = threadIdx.x; idx < R * M; idx += kFFTThreads) {
    const int q = idx % R;
    const int k = idx / R;
    const int f = fg * R + q;
    double2 X = make_double2(0.0, 0.0);
    if (f < p.nfield && k < N) {
      // slot s of this field at Fp[16 s] (F[theta][col/16][slot][16])
      const double* Fp = p.F + f_index(theta, 0, 2 * f, p.nslot, ncg16);
      if (!p.real) {
        if (k <= p.lmax) {
          const int s = k - p.m_first;
          X = make_double2(Fp[16 * (long)s], Fp[16 * (long)s + 1]);
        }
        if (k >= N - p.lmax) {
          const int s = (k - N) - p.m_first;
          X = cadd(X, make_double2(Fp[16 * (long)s], Fp[16 * (long)s + 1]));
        }
      } else {
        // Hermitian spectrum from bins 0..N/2 (c2r semantics).
        const int kk = k <= N / 2 ? k : N - k;
        if (kk <= p.lmax) {
          const int s = kk - p.m_first;
          X = make_double2(Fp[16 * (long)s], Fp[16 * (long)s + 1]);
          if (kk == 0 || kk == N / 2) X.y = 0.0;
          else if (k > N / 2) X.y = -X.y;
        }
      }
    }
    buf[q * M + k] = X;
  }
  __syncthreads();
  ring_dft<+1, CAP>(buf, p);
  for (int idx = threadIdx.x; idx < R * N; idx += kFFTThreads) {
    const int q = idx / N;
    const int j = idx - q * N;
    const int f = fg * R + q;
    if (f >= p.nfield) continue;
    const long o = (long)(p.fmap ? p.fmap[f] : f) * p.fstride + (long)theta * N + j;
    const double2 x = buf[q * M + j];
    if (p.real && p.f32)
      static_cast<float*>(p.fout)[o] = (float)x.x;
    else if (p.real)
      static_cast<double*>(p.fout)[o] = x.x;
    else if (p.f32)
      static_cast<float2*>(p.fout)[o] = make_float2((float)x.x, (float)x.y);
    else
      static_cast<double2*>(p.fout)[o] = x;
  }
}

// Analysis: load rings, forward FFT, scatter weighted bins into F.
template <int CAP>
__global__ __launch_bounds__(kFFTThreads) void fft_anal(FFTParams p) {
  __shared__ double2 buf[CAP];
  const int theta = blockIdx.x / p.nfg;
  const int fg = blockIdx.x % p.nfg;
  const int N = p.N, M = p.M, R = p.rings;
  for (int idx = threadIdx.x; idx < R * M; idx += kFFTThreads) {
    const int q = idx / M;
    const int j = idx - q * M;
    const int f = fg * R + q;
    double2 x = make_double2(0.0, 0.0);
    if (f < p.nfield && j < N) {
      const long o = (long)(p.fmap ? p.fmap[f] : f) * p.fstride + (long)theta * N + j;
      auto load = [&](const void* b) {
        if (p.real && p.f32) return make_double2(static_cast<const float*>(b)[o], 0.0);
        if (p.real) return make_double2(static_cast<const double*>(b)[o], 0.0);
        if (p.f32) {
          const float2 y = static_cast<const float2*>(b)[o];
          return make_double2(y.x, y.y);
        }
        return static_cast<const double2*>(b)[o];
      };
      if (p.src_cplx) {  // Re or Im of a complex field, transformed as a real ring
        const double2 z = static_cast<const double2*>(p.fin)[o];
        x.x = p.op == kExprImag ? z.y : z.x;
      } else {
        x = load(p.fin);
        if (p.op) {  // pointwise expression of the fields
          const double2 y = expr_binary(p.op) ? load(p.fin2) : make_double2(0.0, 0.0);
          x = p.real ? make_double2(expr_real(p.op, x.x, y.x, p.scalar.x), 0.0)
                     : expr_cplx(p.op, x, y, p.scalar);
        }
      }
    }
    buf[idx] = x;
  }
  __syncthreads();
  ring_dft<-1, CAP>(buf, p);
  const double w = p.wq[theta];
  for (int idx = threadIdx.x; idx < R * p.nslot; idx += kFFTThreads) {
    const int q = idx % R;
    const int s = idx / R;
    const int f = fg * R + q;
    if (f >= p.nfield) continue;
    const int m = s + p.m_first;
    const int k = m >= 0 ? m : N + m;
    double2 X = buf[q * M + k];
    if (p.adjoint && p.real) {
      // c2r bin weights: 2 for 0 < k < N/2, real part only at k = 0, N/2
      if (k == 0 || 2 * k == N) X.y = 0.0;
      else X = make_double2(2.0 * X.x, 2.0 * X.y);
    }
    const long fi = fa_index(s, 2 * f, theta, p.n_cg, p.nquad);
    if (p.fa32) {
      float* Fp = reinterpret_cast<float*>(p.F) + fi;
      Fp[0] = (float)(X.x * w);
      Fp[1] = (float)(X.y * w);
    } else {
      p.F[fi] = X.x * w;
      p.F[fi + 1] = X.y * w;
    }
    // The last ring also zeroes the padding rows of the last colatitude
    // quad, which the Legendre contraction reads as B operands.
    for (int tp = theta + 1; theta == p.ntheta - 1 && tp < 4 * p.nquad; tp++) {
      const long zi = fa_index(s, 2 * f, tp, p.n_cg, p.nquad);
      if (p.fa32) {
        reinterpret_cast<float*>(p.F)[zi] = 0.0f;
        reinterpret_cast<float*>(p.F)[zi + 1] = 0.0f;
      } else {
        p.F[zi] = 0.0;
        p.F[zi + 1] = 0.0;
      }
    }
  }
}

// lmax == 0 branches of the reference act on element 0 only
// (GaussLegendreGrid.h:137-140 and :251-260): the synthesis writes
// out[0] = c[0] 2 sqrt(pi), the analysis assigns c[0] = f[0] / (2 sqrt(pi))
// (even though its general branch accumulates).  mode 2 is the adjoint of
// that synthesis: c[0] = 2 sqrt(pi) f[0] (real part only for real fields).
__global__ void lmax0_kernel(int mode, int real, int nfield, const double* in, long istride,
                             double* out, long ostride) {
  const int f = blockIdx.x * blockDim.x + threadIdx.x;
  if (f >= nfield) return;
  const double inv_sqrtpi = 0.56418958354775628695;  // 1/sqrt(pi)
  if (mode == 1) {
    // in: complex coefficients; out: real or complex field
    const double* c = in + 2 * (long)f * istride;
    if (real) {
      out[(long)f * ostride] = c[0] * 2.0 / inv_sqrtpi;
    } else {
      out[2 * (long)f * ostride] = c[0] * 2.0 / inv_sqrtpi;
      out[2 * (long)f * ostride + 1] = c[1] * 2.0 / inv_sqrtpi;
    }
  } else {
    // in: field; out: complex coefficients
    double re, im;
    if (real) {
      re = in[(long)f * istride];
      im = 0.0;
    } else {
      re = in[2 * (long)f * istride];
      im = in[2 * (long)f * istride + 1];
    }
    const double sc = mode == 2 ? 2.0 / inv_sqrtpi : inv_sqrtpi / 2.0;
    double* c = out + 2 * (long)f * ostride;
    c[0] = re * sc;
    c[1] = im * sc;
  }
}

}  // namespace

hipError_t launch_fft_synth(const FFTParams& p0, hipStream_t s) {
  FFTParams p = p0;
  if (p.M > kFFTElemsMax) return hipErrorInvalidValue;
  if (p.fa32) return hipErrorInvalidValue;  // FP32 bins: wavefront kernels only
  const bool big = p.M > kFFTElems;
  p.rings = (big ? kFFTElemsMax : kFFTElems) / p.M;
  p.nfg = (p.nfield + p.rings - 1) / p.rings;
  const dim3 grid((long)p.ntheta * p.nfg);
  if (big)
    hipLaunchKernelGGL((fft_synth<kFFTElemsMax>), grid, dim3(kFFTThreads), 0, s, p);
  else
    hipLaunchKernelGGL((fft_synth<kFFTElems>), grid, dim3(kFFTThreads), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_fft_anal(const FFTParams& p0, hipStream_t s) {
  FFTParams p = p0;
  p.n_cg = (p.ncol + 15) / 16;
  p.nquad = (p.ntheta + 3) / 4;
  if (p.M > kFFTElemsMax) return hipErrorInvalidValue;
  const bool big = p.M > kFFTElems;
  p.rings = (big ? kFFTElemsMax : kFFTElems) / p.M;
  p.nfg = (p.nfield + p.rings - 1) / p.rings;
  const dim3 grid((long)p.ntheta * p.nfg);
  if (big)
    hipLaunchKernelGGL((fft_anal<kFFTElemsMax>), grid, dim3(kFFTThreads), 0, s, p);
  else
    hipLaunchKernelGGL((fft_anal<kFFTElems>), grid, dim3(kFFTThreads), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_lmax0(int mode, int real, int nfield, const double* in, long istride,
                        double* out, long ostride, hipStream_t s) {
  hipLaunchKernelGGL(lmax0_kernel, dim3((nfield + 255) / 256), dim3(256), 0, s, mode, real,
                     nfield, in, istride, out, ostride);
  return hipGetLastError();
}

}  // namespace gsh
