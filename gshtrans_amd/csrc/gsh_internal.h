// gsh_internal.h -- shared declarations between the host layer (gsh_capi.cpp)
// and the HIP kernels (gsh_legendre.hip, gsh_fft.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace gsh {

// Zero recurrence entries kept after degree L for every order, so that
// fixed-size chunks of the l-loop may step past L without bounds checks.
constexpr int kRecPad = 64;
// Colatitudes past the last whole 256-row super-block that the split-BF16
// synthesis leaves to its tail kernel (more get a super-block item).
constexpr int kSynthTailRows = 2;

// Analysis passes: the degree range of an order is cut into passes of
// kPass degrees that run in separate workgroups; the recurrence state at the
// start of every pass k >= 1 comes from a checkpoint table computed once per
// grid and upper index (rec_checkpoints), which holds the state every kCk
// degrees (the split-BF16 analysis starts its recurrences at each of them).
constexpr int kPass = 64;
constexpr int kCk = 16;
// Passes of the wide split-BF16 analysis (leg_anal_bf16_w): 128 degrees.
constexpr int kWideAnalPass = 128;
// Floats of the split-BF16 analysis buffer FA32 kept before its first block
// (32 rows of 16 columns, 2 KiB: the alignment of the buffer is kept).
constexpr int kFAFrontPad = 512;
static_assert(kPass % kCk == 0, "passes start at checkpoints");

// Parameters of one Legendre-contraction launch (synthesis, analysis or
// adjoint) for a batch of fields sharing (lmax, n, value kind).
//
// Work layouts in HBM.  slot indexes the order m (slot = m - m_first;
// m_first = -lmax for complex fields, 0 for real fields) and
// col = 2*field + part (part 0 = real, 1 = imaginary).
//   synthesis  F[theta][col/16][slot][16]: the Legendre kernel writes
//              F_m(theta) in 128-byte rows of 16 columns, and the ring FFT of
//              one colatitude and 16-column group reads one contiguous block
//              of every order; see f_index()
//   analysis   FA[slot][cg][quad][64]: the ring FFT writes the weighted bins
//              X'_m(theta) in the lane order of the 4x16 MFMA B operand,
//              quad = theta/4, cg = col/16, lane = 16*(theta%4) + col%16,
//              so one colatitude quad of one column group is 512 contiguous
//              bytes; see fa_index().  The split-BF16 modes keep FA in FP32
//              (the same indices in a float array): their split rounds every
//              operand to FP32 first.
struct LegParams {
  int L;          // grid degree; ntheta = L + 1
  int ntheta;
  int lmax;       // truncation degree of this call
  int n, nabs;    // upper index
  int real;       // 1: real field, orders m >= 0, NonNegative coefficient layout
  int m_first;    // order of slot 0
  int nslot;      // number of orders
  int nfield;     // batch
  int ncol;       // 2 * nfield
  // coefficients: complex, field-major, cstride complex elements apart
  const double* cin;
  double* cout;
  long cstride;
  // F work buffer
  double* F;
  int f32out;     // split-BF16 synthesis: F written as float elements (same
                  // indices; the wavefront ring FFT reads them)
  // per-grid / per-n tables
  const double* cosv;   // cos(theta_t), t < ntheta
  const double* seeds;  // Ortho-normalised d^{l0}_{m n}(theta_t) at [(m+L)*ntheta + t]
  // Seeds below the normal double range, in extended range (rec_checkpoints
  // only): where skexp[i] = k > 0 the seed of entry i is sseed[i] * 2^-k.
  // Null when every seed of the upper index is a normal double or zero.
  const double* sseed;
  const int* skexp;
  int* ck_flag;     // rec_checkpoints: set to 1 when a lane that started scaled
                    // reaches a state of magnitude 2^-60 or more before L
  int ck_restart;   // FP64 synthesis: take the state from the checkpoint
                    // table every kPass degrees (the table's scaled states
                    // carry the orders whose seeds underflow)
  // Recurrence coefficients {a, b, c, 0} per (m, l):
  //   e_l = (a cos - b) e_{l-1} - c e_{l-2}
  // rec[roff[m + L] + (l - l0)], l0 <= l <= L + kRecPad, zero beyond L.
  const double4* rec;
  const long* roff;
  int accumulate;       // analysis: coefficients += result
  int zero_last;        // analysis: zero the (L, L) coefficient of complex fields
                        // whose truncation degree is L
  int n_tsb;            // synthesis: theta super-blocks of 256
  int n_cg;             // column groups
  int nquad;            // analysis: colatitude quads, ceil(ntheta / 4)
  // Mixed precision (split-BF16 MFMA):
  //   synthesis: bs holds the split coefficients, block (cso[slot] + chunk,
  //   tile, split) = 64 lanes x 8 bf16, see coef_split in gsh_mixed.hip;
  //   analysis: FA32 holds the bins FA in FP32.
  const uint4* bs;
  uint4* bs_out;
  const float* FA32;
  const long* cso;      // per slot: first chunk of that order in bs
  int ntile;            // 16-column tiles, ceil(ncol / 16)
  int nkc;              // analysis: 32-colatitude chunks, ceil(ntheta / 32)
  // Batches of uneven truncation (null pointers: every field has lmax and
  // column field f is the caller's field f):
  const int* fmap;      // column field f -> index of the field in the caller's arrays
  const int* f_lmax;    // truncation degree of column field f (<= lmax)
  const int* cg_lmax;   // max of f_lmax over each 16-column group
  // Analysis work list: items[(b % 8) * items_per + b / 8] = (mw, pass |
  // group << 16) for workgroup b, x = ~0u for padding; see work_item().
  const uint2* items;
  int items_per;
  // Recurrence checkpoints: state (e_l, e_{l-1}) at l = l0 + kCk k of order
  // m, k >= 1, colatitude t: ck[(ckoff[m + L] + k - 1) * ntheta + t], for
  // every l0 + kCk k <= L.
  const double2* ck;
  const long* ckoff;
  // Split-BF16 synthesis of the tail_rows colatitudes tail_t0 .. (0: none),
  // from the d values tail_dv[r * tail_dstride + 32 block + j] of the degree
  // of entry j of each split block and the per-block partial sums
  // tail_part[r * tail_pstride + 16 block + column] that coef_split adds.
  const double* tail_dv;
  double* tail_part;
  long tail_dstride;
  long tail_pstride;
  int tail_rows;
  int tail_t0;
  // Split-BF16 synthesis of upper index 0 over the northern colatitudes
  // (leg_synth_bf16_s): the split blocks go by parity class, two per
  // 64-degree chunk c of an order, block cso[slot] + 2 c + k holding the
  // degrees l0 + 64 c + 2 j + ((k + l0 + m) % 2), j = 0..31, of class k
  // (k = (l + m) % 2).  0: blocks of 32 consecutive degrees.
  int sym;
  // Split-BF16 synthesis of complex fields with orders m and -m paired
  // (leg_synth_bf16_pm): coef_split multiplies the coefficients of orders
  // m < 0 by (-1)^(l+n), tail_dvals gives those orders the d values of order
  // |m| at the same rows, and tail_reduce writes their sums to the mirrored
  // rows ntheta-1-t.
  int pm;
};



// Truncation degree and caller index of column field f.
__device__ __forceinline__ int field_lmax(const LegParams& p, int f) {
  return p.f_lmax ? p.f_lmax[f] : p.lmax;
}
__device__ __forceinline__ int field_index(const int* fmap, int f) {
  return fmap ? fmap[f] : f;
}

// Index of (theta, slot, col) in the synthesis work layout F (ncg16 =
// ceil(ncol / 16) column groups).
__host__ __device__ inline long f_index(int theta, int slot, int col, int nslot, int ncg16) {
  return (((long)theta * ncg16 + (col >> 4)) * nslot + slot) * 16 + (col & 15);
}

// Load of a wave-uniform table entry through the constant address space, so
// that it becomes a scalar load into an SGPR (the tables are written only by
// the host).
template <class T>
__device__ __forceinline__ T ld_uniform(const T* p, long i) {
  return ((const __attribute__((address_space(4))) T*)p)[i];
}

// Work item of this workgroup.  Workgroups are dispatched to the device's 8
// XCDs round-robin (workgroup b on XCD b % 8), so reading the list as
// items[(b % 8) * per + b / 8] runs consecutive items on one XCD at about the
// same time, where they share that XCD's L2.
__device__ __forceinline__ uint2 work_item(const uint2* items, int per) {
  const int b = blockIdx.x;
  const unsigned long long v = ld_uniform(reinterpret_cast<const unsigned long long*>(items),
                                          (long)(b & 7) * per + (b >> 3));
  return make_uint2(static_cast<unsigned>(v), static_cast<unsigned>(v >> 32));
}

// Index of (slot, col, theta) in the analysis work layout FA.
__host__ __device__ inline long fa_index(int slot, int col, int theta, int n_cg, int nquad) {
  return (((long)slot * n_cg + (col >> 4)) * nquad + (theta >> 2)) * 64 + 16 * (theta & 3) +
         (col & 15);
}

// Pointwise field expressions that the analysis ring FFTs form as they load
// a ring (gsh_forward_expr): codes of include/gshtrans_amd.h's
// gsh_field_expr.  a and b are samples of the two input fields, s the
// scalar (its real part for real fields).
enum ExprOp {
  kExprNone = 0,
  kExprProduct = 1,     // a b
  kExprSum = 2,         // a + b
  kExprDifference = 3,  // a - b
  kExprQuotient = 4,    // a / b
  kExprNegate = 5,      // -a
  kExprConj = 6,        // conj(a)
  kExprReal = 7,        // Re a (a real field)
  kExprImag = 8,        // Im a (a real field)
  kExprScale = 9,       // s a
  kExprShift = 10       // a + s
};
__host__ __device__ inline bool expr_binary(int op) {
  return op >= kExprProduct && op <= kExprQuotient;
}
// Complex samples.  The quotient scales by the larger part of b (Smith's
// method), as complex division without overflow of |b|^2 does.
__device__ inline double2 expr_cplx(int op, double2 a, double2 b, double2 s) {
  switch (op) {
    case kExprProduct: return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
    case kExprSum: return make_double2(a.x + b.x, a.y + b.y);
    case kExprDifference: return make_double2(a.x - b.x, a.y - b.y);
    case kExprQuotient:
      if (fabs(b.x) >= fabs(b.y)) {
        const double r = b.y / b.x, d = b.x + b.y * r;
        return make_double2((a.x + a.y * r) / d, (a.y - a.x * r) / d);
      } else {
        const double r = b.x / b.y, d = b.x * r + b.y;
        return make_double2((a.x * r + a.y) / d, (a.y * r - a.x) / d);
      }
    case kExprNegate: return make_double2(-a.x, -a.y);
    case kExprConj: return make_double2(a.x, -a.y);
    case kExprReal: return make_double2(a.x, 0.0);
    case kExprImag: return make_double2(a.y, 0.0);
    case kExprScale: return make_double2(s.x * a.x - s.y * a.y, s.x * a.y + s.y * a.x);
    case kExprShift: return make_double2(a.x + s.x, a.y + s.y);
    default: return a;
  }
}
// Real samples (conj and Re are the identity, Im is zero).
__device__ inline double expr_real(int op, double a, double b, double s) {
  switch (op) {
    case kExprProduct: return a * b;
    case kExprSum: return a + b;
    case kExprDifference: return a - b;
    case kExprQuotient: return a / b;
    case kExprNegate: return -a;
    case kExprImag: return 0.0;
    case kExprScale: return s * a;
    case kExprShift: return a + s;
    default: return a;
  }
}

// Parameters of one ring-FFT launch.
struct FFTParams {
  int N;            // ring length 2L
  int M;            // transform length in LDS: N if N is a power of two, else
                    // the Bluestein convolution length (power of two >= 2N-1)
  int L, ntheta, lmax, nslot, ncol, nfield, real, m_first;
  int rings;        // rings per workgroup, rings * M <= 4096
  int nfg;          // field groups
  int n_cg;         // analysis: 16-column groups, ceil(ncol / 16)
  int nquad;        // analysis: colatitude quads
  const double2* tw;  // tw[k] = exp(-2 pi i k / M), k < M
  const double2* twa; // wavefront FFT: twa[k1 * 64 + j] = exp(-2 pi i j k1 / N), k1 < N / 64
  const double2* chirp;  // Bluestein: exp(-i pi k^2 / N), k < N
  const double2* hf;     // Bluestein: FFT_M of the forward filter / M
  const double2* hb;     // Bluestein: FFT_M of the backward filter / M
  const double* wq;   // analysis weights w_t * 2 pi / N (all 1 for the adjoint)
  int adjoint;        // adjoint of the synthesis: real fields take the c2r
                      // bin weights (2 for 0 < k < N/2, Re only at 0, N/2)
  double* F;          // bins: FA (analysis) or F (synthesis); float elements
                      // when fa32
  int fa32;           // bins in FP32 (split-BF16 modes; the synthesis only
                      // through the wavefront kernels)
  const void* fin;    // analysis input field
  const void* fin2;   // analysis of a pointwise expression of two fields: the
                      // second operand (same layout as fin), or null
  int op;             // ExprOp applied to each loaded sample (kExprNone: none)
  int src_cplx;       // real analysis of Re / Im of a complex field: fin is
                      // complex (kExprReal / kExprImag), the rings real
  double2 scalar;     // kExprScale / kExprShift operand
  void* fout;         // synthesis output field
  long fstride;       // elements between fields
  int f32;            // fields stored in FP32 (float / complex float)
  const int* fmap;    // column field f -> index of the field in the caller's arrays
};

// tiles: 16-column tiles per workgroup (1 or 2; 0 picks 2 when ncol > 16);
// f32m: FP32 MFMA on FP32-rounded operands with FP32 accumulation (the
// recurrence and weights stay FP64).
hipError_t launch_leg_synth_f64(const LegParams& p, hipStream_t s, int tiles = 0,
                                bool f32m = false);
hipError_t launch_leg_anal_f64(const LegParams& p, hipStream_t s, bool f32m = false);
// Recurrence checkpoints of every order of upper index p.n (uses L, ntheta,
// nabs, cosv, seeds, rec, roff, ck, ckoff).
hipError_t launch_rec_checkpoints(const LegParams& p, hipStream_t s);
hipError_t launch_fft_synth(const FFTParams& p, hipStream_t s);
hipError_t launch_fft_anal(const FFTParams& p, hipStream_t s);
// Wavefront-per-ring FFTs (gsh_fft_wave.hip) for N in {512, 1024, 2048}.
bool wave_fft_supported(int N);
hipError_t launch_fft_synth_wave(const FFTParams& p, hipStream_t s);
hipError_t launch_fft_anal_wave(const FFTParams& p, hipStream_t s);
// Mixed precision kernels (gsh_mixed.hip); ns = 2 (bf16x3) or 3 (bf16x6).
hipError_t launch_coef_split(const LegParams& p, int ns, hipStream_t s);
hipError_t launch_leg_synth_bf16(const LegParams& p, int ns, hipStream_t s);
// Split-BF16 synthesis of the p.tail_rows colatitudes from p.tail_t0 (at
// most kSynthTailRows): their d values before coef_split, which then adds the
// partial sums, and the sums over the blocks into F after it.
hipError_t launch_tail_dvals(const LegParams& p, hipStream_t s);
hipError_t launch_tail_reduce(const LegParams& p, hipStream_t s);
// form: 0 passes of 64 degrees, 1 passes of 128 degrees (work list family
// 3); 2 and 3 the same over the northern colatitudes (upper index 0,
// symmetric grids); 4 passes of 64 degrees of orders m and -m together
// (complex fields; work list family 5, orders m >= 0).
hipError_t launch_leg_anal_bf16(const LegParams& p, int ns, hipStream_t s, int form = 0);

// FP32 <-> FP64 coefficient staging of batch plans (gsh_batch.hip); strides
// and sizes in complex elements.
hipError_t launch_widen_c32(const void* in, double* out, long total, hipStream_t s);
hipError_t launch_narrow_c64(const double* in, void* out, long stride, const long* csize,
                             long cmax, int count, hipStream_t s);

// Wigner d tables (gsh_wigner.hip).  Orders m = mlo .. mlo + nm - 1 (mmax =
// min(lmax, caller's mMax), mlo = -mmax for MRange All, 0 for NonNegative),
// upper indices n = nmin .. nmin + nn - 1, ntheta angles.
struct WignerParams {
  int lmax, mmax, mlo, nm, nmin, nn, ntheta;
  int all;     // MRange All (else NonNegative)
  int fourpi;  // FourPi normalisation (else Ortho)
  const double* cosv;      // cos(theta_i)
  const double* seeds;     // Ortho value at l0 of (n, theta_i, m): [(ni * ntheta + i) * nm + mi]
  const double4* rec;      // recurrence coefficients of (n, l, m): [(ni * (lmax + 1) + l) * nm + mi]
  const double* inv_ortho; // 2 sqrt(pi) / sqrt(2l + 1), l <= lmax
  const long* blockoff;    // first entry of block (n, theta_i): [ni * ntheta + i]
  double* out;
};
hipError_t launch_wigner_table(const WignerParams& p, hipStream_t s);

// Grid helpers (gsh_grid_ops.hip).
struct GridOpParams {
  int ntheta, nphi, batch, real;
  long fstride;             // elements between fields
  const void* fa;           // fields (double or complex double)
  const void* fb;           // second factor of the inner product, or null
  const double* wq;         // quadrature weights w_i
  double dphi;              // longitude weight 2 pi / nphi
  double2* part;            // ring sums [f * ntheta + i]
  double* out;              // complex results, interleaved
};
hipError_t launch_integrate(const GridOpParams& p, hipStream_t s);
// Allocation-free integration: ring sums of at most kNoAllocRingSums (field,
// colatitude) pairs in a file-scope device array (4 MiB per device).
// noalloc_scratch_init looks up the calling thread's current device's copy
// (once per device; gsh_grid_create calls it), noalloc_scratch returns that
// address (null before the lookup).
constexpr long kNoAllocRingSums = 262144;
constexpr int kMaxDevices = 64;
hipError_t noalloc_scratch_init(int device);
double2* noalloc_scratch(int device);
struct RandomParams {
  int lmax, grid_lmax, nabs, real, batch;
  long size, cstride;       // coefficients per field, complex elements between fields
  unsigned long long seed;
  double decay;
  double* out;
};
hipError_t launch_random_coeffs(const RandomParams& p, hipStream_t s);

// lmax == 0: mode 1 synthesis, 0 analysis, 2 adjoint of the synthesis.
hipError_t launch_lmax0(int mode, int real, int nfield, const double* in, long istride,
                        double* out, long ostride, hipStream_t s);

}  // namespace gsh
