// gsh_mixed.hip -- mixed-precision Legendre contraction: split-BF16 MFMA
// products (v_mfma_f32_16x16x32_bf16) with the Wigner recurrence, its seeds
// and the quadrature weights kept in FP64.
//
// Splitting.  A double x is rounded to float f, then f = h + m + l with h, m,
// l bf16 (each the round-to-nearest bf16 of the remainder).  bf16x6 keeps
// the products h*h, h*m, m*h, h*l, m*m, l*h of the two operands (every term
// down to 2^-16 of the leading one); bf16x3 splits in two and keeps h*h,
// h*m, m*h.
//
// Accumulation.  Each MFMA rounds its 32-product sum into an FP32
// accumulator.  Per 32-term chunk the small cross products go in first and
// the leading h*h product last, so the only rounding at the size of the chunk
// sum is the last one; the chunk result is then added into an FP64
// accumulator.  The error therefore stays near one FP32 rounding of each
// chunk sum instead of growing with the number of chunks.
//
// Sharing.  Generating and splitting D costs ~10 VALU instructions per value,
// more than the MFMAs need when a value feeds few columns.  A workgroup of 8
// waves therefore covers 256 colatitudes x 128 columns of one order: waves
// 0-3 run the recurrence for 64 colatitudes each and write the bf16 splits
// to LDS (double-buffered, one barrier per 32-degree chunk); every wave
// multiplies them into its own 128 colatitudes x 32 columns.
#include "gsh_internal.h"

namespace gsh {
namespace {

typedef float f4 __attribute__((ext_vector_type(4)));
typedef double d4 __attribute__((ext_vector_type(4)));
typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef __bf16 bf16x2 __attribute__((ext_vector_type(2)));
typedef float f2v __attribute__((ext_vector_type(2)));
typedef const __attribute__((address_space(4))) d4 rec_t;

__device__ __forceinline__ rec_t* rec_ptr(const double4* rec, long off) {
  return (rec_t*)(rec) + off;
}

// Recurrence coefficients of one degree, read with a scalar load: every lane
// of a wave asks for the same entry, so the values arrive in SGPRs and the
// recurrence takes them as scalar operands (no LDS traffic, no VGPRs).
__device__ __forceinline__ d4 rec_entry(rec_t* rq, int e) { return rq[e]; }

__device__ __forceinline__ void rec_step(const d4 q, double cs, double& e0, double& e1) {
  const double t = fma(q.x, cs, -q.y);
  const double en = fma(t, e0, -(q.z * e1));
  e1 = e0;
  e0 = en;
}

__device__ __forceinline__ int order_of(int mw, int real) {
  return real ? mw : ((mw & 1) ? -((mw + 1) >> 1) : (mw >> 1));
}

__device__ __forceinline__ long coef_index(int l, int m, int nabs, int real) {
  return real ? (long)l * (l + 1) / 2 - (long)nabs * (nabs + 1) / 2 + m
              : (long)l * l - (long)nabs * nabs + m + l;
}

// Split two floats (consecutive K entries) into NS bf16 planes; plane s of
// the pair is returned packed in one dword (element 0 in the low half).  One
// packed round-to-nearest conversion per plane; the residuals subtract the
// planes' values, read back by placing each bf16 in a float's high half.
template <int NS>
__device__ __forceinline__ void split_pair_f(float f0, float f1, unsigned* out) {
#pragma unroll
  for (int s = 0; s < NS; s++) {
    const bf16x2 b = __builtin_convertvector((f2v){f0, f1}, bf16x2);
    const unsigned u = __builtin_bit_cast(unsigned, b);
    out[s] = u;
    if (s + 1 < NS) {
      f0 -= __builtin_bit_cast(float, u << 16);
      f1 -= __builtin_bit_cast(float, u & 0xffff0000u);
    }
  }
}
template <int NS>
__device__ __forceinline__ void split_pair(double x0, double x1, unsigned* out) {
  float f0 = (float)x0, f1 = (float)x1;
  // Keep the f64->f32 rounding separate from the bf16 rounding below.
  asm volatile("" : "+v"(f0), "+v"(f1));
  split_pair_f<NS>(f0, f1, out);
}

__device__ __forceinline__ bf16x8 as_bf8(uint4 v) { return __builtin_bit_cast(bf16x8, v); }

__device__ __forceinline__ f4 mfma(uint4 a, uint4 b, f4 c) {
  return __builtin_amdgcn_mfma_f32_16x16x32_bf16(as_bf8(a), as_bf8(b), c, 0, 0, 0);
}

// Split products of NA A operands (rows) with NB B operands (columns), NS
// planes each, as NA NB independent MFMA chains advanced one product at a
// time (so consecutive MFMAs never depend on each other); within each chain
// the cross products go first and the leading product last.
template <int NS, int NA, int NB = 2>
__device__ __forceinline__ void split_mma(const uint4 (&A)[NA][NS], const uint4 (&B)[NB][NS],
                                          f4 (&c)[NA][NB]) {
  constexpr int np = NS == 3 ? 6 : 3;
  constexpr int pa[6] = {1, 0, 2, 0, 1, 0};  // plane of A in product q
  constexpr int pb[6] = {1, 2, 0, 1, 0, 0};  // plane of B in product q
#pragma unroll
  for (int i = 0; i < NA; i++)
#pragma unroll
    for (int j = 0; j < NB; j++) c[i][j] = f4{0.f, 0.f, 0.f, 0.f};
#pragma unroll
  for (int q = 6 - np; q < 6; q++)
#pragma unroll
    for (int i = 0; i < NA; i++)
#pragma unroll
      for (int j = 0; j < NB; j++) c[i][j] = mfma(A[i][pa[q]], B[j][pb[q]], c[i][j]);
}

__device__ __forceinline__ void flush(d4& acc, const f4 c) {
  acc[0] += (double)c[0];
  acc[1] += (double)c[1];
  acc[2] += (double)c[2];
  acc[3] += (double)c[3];
}

// Analysis epilogues.  Column 16 tile + lane % 16 of the sums is the real
// (even column) or imaginary (odd column) part of field col / 2, so the two
// lanes of a field hold halves of the same coefficients.  trade_halves takes
// rows r and r + 1 of the pair and returns the whole coefficient of row r to
// the even lane and that of row r + 1 to the odd lane, which each store 16
// bytes (store_coef) instead of 8.
__device__ __forceinline__ double2 trade_halves(double a0, double a1, bool odd) {
  const double recv = __shfl_xor(odd ? a0 : a1, 1);
  return odd ? make_double2(recv, a1) : make_double2(a0, recv);
}

// Writes (adds, under p.accumulate) coefficient (l, m) of the field whose
// coefficients start at cbase; under p.zero_last the coefficient (L, L) of a
// field truncated at L is zero.
__device__ __forceinline__ void store_coef(const LegParams& p, long cbase, int l, int m, int lf,
                                           int real, double2 v) {
  double2* c = reinterpret_cast<double2*>(p.cout) + cbase + coef_index(l, m, p.nabs, real);
  if (p.accumulate) {
    const double2 o = *c;
    v = make_double2(v.x + o.x, v.y + o.y);
  }
  if (p.zero_last && lf == p.L && l == lf && m == lf) v = make_double2(0.0, 0.0);
  *c = v;
}

// ---------------------------------------------------------------------------
// Coefficient split: synthesis B operands.  Block (cso[slot] + chunk, tile,
// s) holds, for lane = 16*kr + c, the bf16 plane s of the coefficients of
// degrees l0 + 32*chunk + 8*kr + e (e = 0..7) and column 16*tile + c; degrees
// past the column's truncation degree and columns past ncol are zero.
//
// A workgroup of 8 waves writes the blocks (slot 8 gs + w, chunk, tile) of
// waves w = 0..7.  The degrees they need lie in a window of 40 from lw0, the
// smallest first degree l0 + 32 chunk of the 8 orders (l0 = max(|m|, |n|)
// changes by at most one from one order to the next).  The workgroup first
// reads that window for its 8 fields and 8 orders into LDS -- for each
// (degree, field) the 8 orders are 128 contiguous bytes of the coefficient
// array -- and each wave then takes its order's 32 degrees from LDS.
// ---------------------------------------------------------------------------
constexpr int kSplitRow = 8 * 16 + 2;  // doubles per window degree (8 orders x 16 columns + pad)

// SYM: the parity layout of p.sym (64-degree chunks, two blocks each, entry j
// of class k's block being degree l0 + 64 chunk + 2 j + ((k + l0 + m) % 2)).
template <int NS, int REAL, int SYM>
__global__ __launch_bounds__(512) void coef_split(LegParams p) {
  constexpr int CW = SYM ? 64 : 32;     // degrees per chunk
  constexpr int kSplitWin = CW + 8;     // degree window of a workgroup
  constexpr int ST = SYM ? 2 : 1;       // degrees between consecutive entries
  __shared__ double win[kSplitWin * kSplitRow];
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int gs = it.x;
  const int chunk = it.y & 0xffff;
  const int tile = it.y >> 16;
  const int tid = threadIdx.x;
  const int lane = tid & 63;
  const int wave = __builtin_amdgcn_readfirstlane(tid >> 6);  // wave-uniform
  const int s0 = 8 * gs;
  int lw0 = 1 << 30;
#pragma unroll
  for (int w = 0; w < 8; w++) {
    const int sl = s0 + w;
    if (sl < p.nslot) lw0 = min(lw0, max(abs(sl + p.m_first), p.nabs) + CW * chunk);
  }
  // Window: piece (l_off, f, mo) = (tid + 512 r) / 64, (tid / 8) % 8, tid % 8.
  {
    const int mo = tid & 7, f = (tid >> 3) & 7;
    const int sl = s0 + mo;
    const int m = sl + p.m_first;
    const int fld = 8 * tile + f;  // field of columns 2 fld, 2 fld + 1
    const bool ok = sl < p.nslot && fld < p.nfield;
    const int lf = ok ? field_lmax(p, fld) : -1;
    const double* cp = p.cin + (ok ? (long)field_index(p.fmap, fld) * p.cstride * 2 : 0);
    // Every load is issued unconditionally (degrees outside the order or the
    // field read its first coefficient and are zeroed after the loads), so
    // none waits for the one before.  Paired orders: -m takes (-1)^(l+n)
    // c_{l,-m} (the values of order |m| at the mirrored colatitude multiply
    // it).
    double2 v[kSplitWin / 8];
#pragma unroll
    for (int r = 0; r < kSplitWin / 8; r++) {
      const int l = lw0 + 8 * r + (tid >> 6);
      const bool in = l <= lf && l >= abs(m) && l >= p.nabs;
      v[r] = *reinterpret_cast<const double2*>(cp + (in ? 2 * coef_index(l, m, p.nabs, REAL) : 0));
    }
#pragma unroll
    for (int r = 0; r < kSplitWin / 8; r++) {
      const int l = lw0 + 8 * r + (tid >> 6);
      const bool in = l <= lf && l >= abs(m) && l >= p.nabs;
      const double sg = p.pm && m < 0 && ((l + p.n) & 1) ? -1.0 : 1.0;
      v[r] = in ? make_double2(sg * v[r].x, sg * v[r].y) : make_double2(0.0, 0.0);
    }
#pragma unroll
    for (int r = 0; r < kSplitWin / 8; r++)
      *reinterpret_cast<double2*>(&win[(8 * r + (tid >> 6)) * kSplitRow + 16 * mo + 2 * f]) = v[r];
  }
  __syncthreads();
  const int sl = s0 + wave;
  if (sl >= p.nslot) return;
  const int m = sl + p.m_first;
  const int l0 = max(abs(m), p.nabs);
  if (chunk >= (p.lmax - l0 + CW) / CW) return;  // no such block for this order
  const int kr = lane >> 4, c = lane & 15;
#pragma unroll
  for (int k = 0; k < (SYM ? 2 : 1); k++) {
    const long bc = ld_uniform(p.cso, sl) + (SYM ? 2 * chunk + k : chunk);  // split block
    const long b = bc * p.ntile + tile;                                    // output block
    // Window offset of entry 8 kr (e = 0); entry 8 kr + e is ST e further.
    const int lo = l0 + CW * chunk + (SYM ? (k + l0 + m) & 1 : 0) + 8 * ST * kr - lw0;
    unsigned planes[NS][4];
#pragma unroll
    for (int e2 = 0; e2 < 4; e2++) {
      const double x0 = win[(lo + ST * 2 * e2) * kSplitRow + 16 * wave + c];
      const double x1 = win[(lo + ST * (2 * e2 + 1)) * kSplitRow + 16 * wave + c];
      unsigned o[NS];
      split_pair<NS>(x0, x1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) planes[s][e2] = o[s];
    }
#pragma unroll
    for (int s = 0; s < NS; s++)
      p.bs_out[(b * NS + s) * 64 + lane] =
          make_uint4(planes[s][0], planes[s][1], planes[s][2], planes[s][3]);
    // Tail colatitudes: this block's share of their sums, from the FP64
    // coefficients, one partial per (row, block, column); tail_reduce adds
    // the blocks up.
    for (int r = 0; r < p.tail_rows; r++) {
      const double2* dq =
          reinterpret_cast<const double2*>(p.tail_dv + r * p.tail_dstride + bc * 32 + 8 * kr);
      double acc = 0.0;
#pragma unroll
      for (int e2 = 0; e2 < 4; e2++) {
        const double2 d = dq[e2];
        acc = fma(d.x, win[(lo + ST * 2 * e2) * kSplitRow + 16 * wave + c], acc);
        acc = fma(d.y, win[(lo + ST * (2 * e2 + 1)) * kSplitRow + 16 * wave + c], acc);
      }
      acc += __shfl_xor(acc, 16);
      acc += __shfl_xor(acc, 32);
      if (lane < 16) p.tail_part[r * p.tail_pstride + b * 16 + c] = acc;
    }
  }
}

// ---------------------------------------------------------------------------
// Synthesis.  Workgroup = 8 waves: (256 colatitudes of super-block tsb) x
// (128 columns of column super-group csg) for one order m, walking the
// degrees in 32-degree K chunks whose bf16 planes of D are double-buffered in
// LDS.  While the waves multiply chunk c they produce chunk c + 1: wave w
// runs the recurrence of colatitude block w % 4 (64 colatitudes) for the 16
// degrees of half w / 4, starting from the checkpoint there, so every wave
// does the same share.  Wave w multiplies colatitude half h = w / 4 (8 MFMA
// row groups) with column tiles 2 (w % 4) and 2 (w % 4) + 1.  Each row
// group's MFMAs share a scheduling region with one recurrence pair and the
// LDS loads of the next group's A operands, so the VALU work, the MFMAs and
// the LDS latency overlap.
// ---------------------------------------------------------------------------
template <int NS, int REAL>
__global__ __launch_bounds__(512) void leg_synth_bf16(LegParams p) {
  static_assert(kCk == 16, "two 16-degree halves per 32-degree chunk");
  __shared__ uint4 trb[2][NS][4][256];
  // B operands of a chunk (8 column tiles x NS planes x 64 lanes) and the
  // recurrence coefficients of a chunk (32 entries, 1 KiB), double-buffered.
  __shared__ uint4 bsh[2][8][NS][64];
  __shared__ uint4 rsc[2][64];
  const int lane = threadIdx.x & 63;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  // Work item: (order, column super-group, colatitude super-block).
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int mw = it.x;
  const int tsb = it.y & 0xffff;
  const int csg = it.y >> 16;
  const int m = order_of(mw, REAL);
  const int slot = m - p.m_first;
  const int l0 = max(abs(m), p.nabs);
  const int nt = p.ntheta;
  const int nchunk = (p.lmax - l0 + 32) >> 5;
  const int rows = min(256, nt - tsb * 256);  // valid colatitudes of the item

  // Producer share: colatitude block pb, degree half ph.
  const int pb = wave & 3, ph = wave >> 2;
  const int tpc = min(tsb * 256 + 64 * pb + lane, nt - 1);  // past ntheta: the last one
  const double pcs = p.cosv[tpc];
  const int nck = (p.L - l0 + kCk) / kCk;  // checkpoints 1 .. nck-1 exist
  // State at checkpoint k >= 1; past L it clamps to the last checkpoint
  // (those degrees only meet zero coefficients, and the values stay finite).
  // An order without checkpoints (nck < 2, so a single chunk) reads entry 0
  // of the table instead and never uses it.
  const double2* ckm =
      (nck > 1 ? p.ck + ld_uniform(p.ckoff, m + p.L) * (long)nt : p.ck) + tpc;
  auto ck_state = [&](int k) { return ckm[(long)(nck > 1 ? min(k, nck - 1) - 1 : 0) * nt]; };
  double pe0, pe1;
  {
    // Chunk 0: the seed for half 0, checkpoint 1 for half 1.
    const double seed = p.seeds[(long)(m + p.L) * nt + tpc];
    const double2 c1 = ck_state(1);
    pe0 = ph == 0 || nck < 2 ? seed : c1.x;
    pe1 = ph == 0 || nck < 2 ? 0.0 : c1.y;
  }

  const uint4* recq = reinterpret_cast<const uint4*>(p.rec + (ld_uniform(p.roff, m + p.L) - l0));
  // Entries up to L + kRecPad exist: chunk c's entries l0 + 32c + 1 .. + 32
  // stay in the table for every c <= nchunk.
  auto fetch_rec = [&](int c) { return recq[2 * (l0 + 32 * c + 1) + (tid & 63)]; };

  // B operand staging: piece j of this thread is uint4 tid + 512 j of the
  // chunk's 8 contiguous tiles (ntile is a multiple of 8), in the source and
  // in the LDS buffer alike.
  const uint4* bbase = p.bs + (ld_uniform(p.cso, slot) * (long)p.ntile + csg * 8) * NS * 64;
  const int cstride = p.ntile * NS * 64;  // uint4 per chunk
  struct Pieces {
    uint4 v[NS];
    uint4 r;
  };
  auto fetch = [&](int chunk) {
    const uint4* cb = bbase + (long)min(chunk, nchunk - 1) * cstride;
    Pieces q;
#pragma unroll
    for (int j = 0; j < NS; j++) q.v[j] = cb[tid + 512 * j];
    q.r = fetch_rec(min(chunk + 1, nchunk));
    return q;
  };
  auto store = [&](int buf, const Pieces q) {
#pragma unroll
    for (int j = 0; j < NS; j++) (&bsh[buf][0][0][0])[512 * j + tid] = q.v[j];
    if (tid < 64) rsc[buf ^ 1][tid] = q.r;
  };

  // Consumer role.
  const int h = wave >> 2;
  const int tile0 = csg * 8 + (wave & 3) * 2;
  d4 acc[8][2];
#pragma unroll
  for (int g = 0; g < 8; g++)
#pragma unroll
    for (int t = 0; t < 2; t++) acc[g][t] = d4{0.0, 0.0, 0.0, 0.0};

  // One recurrence pair (steps 2i, 2i+1 of this wave's 16 degrees of a
  // chunk) into the packed planes pl; every 4 pairs fill the uint4 of each
  // plane that holds degrees 8 kr .. 8 kr + 7 of the colatitude (kr =
  // 2 ph + i / 4), stored whole (lanes 16 bytes apart: conflict-free).
  unsigned pl[NS][4];
  auto produce_pair = [&](int rbuf, int tbuf, int i) {
    const d4* rl = reinterpret_cast<const d4*>(&rsc[rbuf][0]) + 16 * ph;
    const double v0 = pe0;
    rec_step(rl[2 * i], pcs, pe0, pe1);
    const double v1 = pe0;
    rec_step(rl[2 * i + 1], pcs, pe0, pe1);
    unsigned o[NS];
    split_pair<NS>(v0, v1, o);
#pragma unroll
    for (int s = 0; s < NS; s++) pl[s][i & 3] = o[s];
    if ((i & 3) == 3) {
#pragma unroll
      for (int s = 0; s < NS; s++)
        trb[tbuf][s][2 * ph + (i >> 2)][64 * pb + lane] =
            make_uint4(pl[s][0], pl[s][1], pl[s][2], pl[s][3]);
    }
  };
  // A operands of row group g (one uint4 per plane; always in bounds), and
  // row group g's MFMAs (two independent chains) with their flush to FP64.
  // A group whose rows hold no colatitude is skipped.  The A operands of
  // group g + 1 are loaded while group g multiplies (ping-pong registers, so
  // the LDS latency hides behind the MFMAs).
  auto loadA = [&](int buf, int g, uint4 (&A)[1][NS]) {
#pragma unroll
    for (int s = 0; s < NS; s++) A[0][s] = trb[buf][s][lane >> 4][h * 128 + g * 16 + (lane & 15)];
  };
  auto consume_group = [&](int g, const uint4 (&A)[1][NS], const uint4 (&B)[2][NS]) {
    if (h * 128 + g * 16 < rows) {
      f4 c[1][2];
      split_mma<NS, 1>(A, B, c);
      flush(acc[g][0], c[0][0]);
      flush(acc[g][1], c[0][1]);
    }
  };
  auto loadB = [&](int buf, uint4 (&B)[2][NS]) {
#pragma unroll
    for (int s = 0; s < NS; s++) {
      B[0][s] = bsh[buf][(wave & 3) * 2][s][lane];
      B[1][s] = bsh[buf][(wave & 3) * 2 + 1][s][lane];
    }
  };

  // Prologue: coefficients and B operands of chunk 0, production of chunk 0,
  // coefficients of chunk 1.
  if (tid < 64) rsc[0][tid] = fetch_rec(0);
  {
    const Pieces q = fetch(0);
#pragma unroll
    for (int j = 0; j < NS; j++) (&bsh[0][0][0][0])[512 * j + tid] = q.v[j];
    if (tid < 64) rsc[1][tid] = q.r;
  }
  __syncthreads();
#pragma unroll
  for (int i = 0; i < 8; i++) produce_pair(0, 0, i);
  __syncthreads();

  // Chunks 0 .. nchunk-2 with the production of the next one.  Producer
  // blocks and row groups past ntheta (the last super-block only) are
  // skipped with uniform branches; the branches also bound the scheduling
  // regions, which keeps the register pressure of the unrolled chunk low.
  const bool pv = 64 * pb < rows;  // this wave's producer block holds colatitudes
  // The state of chunk c + 1 (checkpoint 2 (c + 1) + ph, k >= 2) is loaded
  // one chunk before its use.
  double2 nx = ck_state(2 + ph);
  for (int c = 0; c + 1 < nchunk; c++) {
    const int buf = c & 1;
    const Pieces q = fetch(c + 1);
    pe0 = nx.x;
    pe1 = nx.y;
    nx = ck_state(2 * (c + 2) + ph);
    uint4 B[2][NS];
    loadB(buf, B);
    uint4 Ab[2][1][NS];
    loadA(buf, 0, Ab[0]);
    // Region g: load group g + 1's A operands, produce pair g, multiply
    // group g.
#pragma unroll
    for (int g = 0; g < 8; g++) {
      if (g < 7) loadA(buf, g + 1, Ab[(g + 1) & 1]);
      if (pv) produce_pair(buf ^ 1, buf ^ 1, g);
      consume_group(g, Ab[g & 1], B);
      __builtin_amdgcn_sched_barrier(0);
    }
    store(buf ^ 1, q);  // buffers last read in chunk c - 1
    __syncthreads();
  }
  {
    const int buf = (nchunk - 1) & 1;
    uint4 B[2][NS];
    loadB(buf, B);
    uint4 Ab[2][1][NS];
    loadA(buf, 0, Ab[0]);
#pragma unroll
    for (int g = 0; g < 8; g++) {
      if (g < 7) loadA(buf, g + 1, Ab[(g + 1) & 1]);
      consume_group(g, Ab[g & 1], B);
    }
  }

  // C layout of the 16x16x32 MFMA: lane holds rows 4*(lane/16)+r, column
  // lane%16.  F is linear in the colatitude (stride rs), so every store is
  // one base pointer plus a uniform offset.
  const int ncg16 = (p.ncol + 15) >> 4;
  const long rs = (long)ncg16 * p.nslot * 16;
  const int row0 = tsb * 256 + h * 128 + 4 * (lane >> 4);
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int col = (tile0 + t) * 16 + (lane & 15);
    if (col >= p.ncol) continue;
    const long fi = f_index(row0, slot, col, p.nslot, ncg16);
    if (p.f32out) {
      float* fp = reinterpret_cast<float*>(p.F) + fi;
#pragma unroll
      for (int g = 0; g < 8; g++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (row0 + g * 16 + r < nt) fp[(g * 16 + r) * rs] = (float)acc[g][t][r];
    } else {
      double* fp = p.F + fi;
#pragma unroll
      for (int g = 0; g < 8; g++)
#pragma unroll
        for (int r = 0; r < 4; r++)
          if (row0 + g * 16 + r < nt) fp[(g * 16 + r) * rs] = acc[g][t][r];
    }
  }
}

// ---------------------------------------------------------------------------
// Synthesis of upper index 0 over the northern colatitudes.  Gauss-Legendre
// nodes are symmetric about the equator and d^l_{m0}(pi - theta) =
// (-1)^(l+m) d^l_{m0}(theta), so with E and O the sums over the degrees of
// even and of odd l + m at a northern colatitude t, F(t) = E + O and
// F(nt-1-t) = E - O: half the recurrence values and MFMAs of leg_synth_bf16.
// The coefficients come in the parity layout (coef_split with sym: two
// 32-degree blocks per 64-degree chunk, one per class).  Workgroup = 8 waves
// for (128 northern colatitudes of block tsb) x (128 columns of super-group
// csg) of one order m, walking the blocks (K steps of one class each), each
// wave loading its B operands straight from the split coefficients one block
// ahead, the bf16 planes of D double-buffered in LDS per 64-degree chunk
// (one barrier per chunk).  While the waves multiply chunk c's
// two blocks they produce chunk c + 1: wave w runs the recurrence of
// colatitude block w % 2 (64 colatitudes) for the 16 degrees of quarter
// w / 2 from the checkpoint there, one recurrence pair per row group, and
// packs the values of each class into that class's planes (consecutive
// entries of a class are two degrees apart).  Wave w multiplies colatitude
// half h = w / 4 (4 row groups) with column tiles 2 (w % 4), 2 (w % 4) + 1,
// keeping E and O apart.  The equator (odd ntheta) is a tail row.
// ---------------------------------------------------------------------------
template <int NS, int REAL>
__global__ __launch_bounds__(512) void leg_synth_bf16_s(LegParams p) {
  static_assert(kCk == 16, "four 16-degree quarters per 64-degree chunk");
  // Planes of a chunk: class k, entries 8 kr .. 8 kr + 7, 128 colatitudes.
  __shared__ uint4 trb[2][NS][2][4][128];
  // Recurrence coefficients of a chunk (64 entries, 2 KiB).
  __shared__ uint4 rsc[2][128];
  const int lane = threadIdx.x & 63;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  // Work item: (order, column super-group, northern block of 128).
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int mw = it.x;
  const int tsb = it.y & 0xffff;
  const int csg = it.y >> 16;
  const int m = order_of(mw, REAL);
  const int slot = m - p.m_first;
  const int l0 = max(abs(m), p.nabs);
  const int nt = p.ntheta;
  const int nn = nt >> 1;                    // northern colatitudes (the equator apart)
  const int nC = (p.lmax - l0 + 64) >> 6;    // 64-degree chunks
  const int nb = 2 * nC;                     // blocks: block 2 c + k is chunk c's class k
  const int rows = min(128, nn - tsb * 128);  // valid colatitudes of the item
  const int par0 = (l0 + m) & 1;             // class of the even offsets from l0

  // Producer share: colatitude block pb, degree quarter pq of a chunk.
  const int pb = wave & 1, pq = wave >> 1;
  const int tpc = min(tsb * 128 + 64 * pb + lane, nn - 1);  // past the north: the last one
  const double pcs = p.cosv[tpc];
  const int nck = (p.L - l0 + kCk) / kCk;  // checkpoints 1 .. nck-1 exist
  // State at checkpoint k >= 1; past L it clamps to the last checkpoint.  An
  // order without checkpoints reads entry 0 of the table and never uses it.
  const double2* ckm =
      (nck > 1 ? p.ck + ld_uniform(p.ckoff, m + p.L) * (long)nt : p.ck) + tpc;
  auto ck_state = [&](int k) { return ckm[(long)(nck > 1 ? min(k, nck - 1) - 1 : 0) * nt]; };
  double pe0, pe1;
  {
    // Chunk 0: the seed for quarter 0, checkpoint pq for the others.
    const double seed = p.seeds[(long)(m + p.L) * nt + tpc];
    const double2 c1 = ck_state(max(pq, 1));
    pe0 = pq == 0 || nck < 2 ? seed : c1.x;
    pe1 = pq == 0 || nck < 2 ? 0.0 : c1.y;
  }

  const uint4* recq = reinterpret_cast<const uint4*>(p.rec + (ld_uniform(p.roff, m + p.L) - l0));
  // Entries up to L + kRecPad exist: chunk c's entries l0 + 64 c + 1 .. + 64
  // for every c < nC.
  auto fetch_rec = [&](int c) { return recq[2 * (l0 + 64 * min(c, nC - 1) + 1) + (tid & 127)]; };

  // B operands of block b (class b % 2 of chunk b / 2): each wave loads its
  // two column tiles' planes straight from the split coefficients (16 bytes
  // a lane, 1 KiB a wave instruction) one block ahead.
  const uint4* bbase = p.bs + (ld_uniform(p.cso, slot) * (long)p.ntile + csg * 8) * NS * 64;
  const int bstride = p.ntile * NS * 64;  // uint4 per block
  const int toff = (wave & 3) * 2 * NS * 64 + lane;
  auto fetchB = [&](int b, uint4 (&B)[2][NS]) {
    const uint4* cb = bbase + (long)min(b, nb - 1) * bstride + toff;
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int s = 0; s < NS; s++) B[t][s] = cb[(t * NS + s) * 64];
  };

  // Consumer role.
  const int h = wave >> 2;
  const int tile0 = csg * 8 + (wave & 3) * 2;
  d4 acc[4][2][2];  // row group, column tile, class
#pragma unroll
  for (int g = 0; g < 4; g++)
#pragma unroll
    for (int t = 0; t < 2; t++)
#pragma unroll
      for (int k = 0; k < 2; k++) acc[g][t][k] = d4{0.0, 0.0, 0.0, 0.0};

  // Production: pair i (0..7) of this wave's quarter is steps 2 i, 2 i + 1,
  // degrees d = 16 pq + 2 i (+1) of the chunk.  An even pair keeps its two
  // values; the odd pair after it stores dword i / 2 of the even offsets
  // (d, d + 2: class par0) and of the odd offsets (d + 1, d + 3) of entries
  // 8 pq .. + 7 of each class.
  double hold0 = 0.0, hold1 = 0.0;
  auto produce_pair = [&](int rbuf, int tbuf, int i) {
    const d4* rl = reinterpret_cast<const d4*>(&rsc[rbuf][0]) + 16 * pq;
    const double v0 = pe0;
    rec_step(rl[2 * i], pcs, pe0, pe1);
    const double v1 = pe0;
    rec_step(rl[2 * i + 1], pcs, pe0, pe1);
    if ((i & 1) == 0) {
      hold0 = v0;
      hold1 = v1;
    } else {
      unsigned o0[NS], o1[NS];
      split_pair<NS>(hold0, v0, o0);
      split_pair<NS>(hold1, v1, o1);
#pragma unroll
      for (int s = 0; s < NS; s++) {
        reinterpret_cast<unsigned*>(&trb[tbuf][s][par0][pq][64 * pb + lane])[i >> 1] = o0[s];
        reinterpret_cast<unsigned*>(&trb[tbuf][s][par0 ^ 1][pq][64 * pb + lane])[i >> 1] = o1[s];
      }
    }
  };
  // A operands of row group g, class k (one uint4 per plane); the group's
  // MFMAs (two independent chains) with their flush to FP64.
  auto loadA = [&](int abuf, int k, int g, uint4 (&A)[1][NS]) {
#pragma unroll
    for (int s = 0; s < NS; s++) A[0][s] = trb[abuf][s][k][lane >> 4][h * 64 + g * 16 + (lane & 15)];
  };
  auto consume_group = [&](int g, int k, const uint4 (&A)[1][NS], const uint4 (&B)[2][NS]) {
    if (h * 64 + g * 16 < rows) {
      f4 c[1][2];
      split_mma<NS, 1>(A, B, c);
      flush(acc[g][0][k], c[0][0]);
      flush(acc[g][1][k], c[0][1]);
    }
  };

  // Prologue: coefficients of chunks 0 and 1, B operands of block 0,
  // production of chunk 0.
  if (tid < 128) {
    rsc[0][tid] = fetch_rec(0);
    rsc[1][tid] = fetch_rec(1);
  }
  uint4 Bc[2][NS];
  fetchB(0, Bc);
  __syncthreads();
  const bool pv = 64 * pb < rows;  // this wave's producer block holds colatitudes
  if (pv) {
#pragma unroll
    for (int i = 0; i < 8; i++) produce_pair(0, 0, i);
  }
  __syncthreads();

  // Block 2 c + k of chunk c (k a constant at each call): the B operands of
  // the next block fetched into registers, production pairs 4 k .. 4 k + 3
  // of chunk c + 1 (none in the last chunk).  One barrier per chunk, after
  // its class-1 block.  The state of chunk c + 1 (checkpoint 4 (c + 1) + pq)
  // is loaded one chunk before its use.
  double2 nx = ck_state(4 + pq);
  auto block = [&](int c, const int k, const bool next) {
    const int b = 2 * c + k;
    const int abuf = c & 1;
    const bool prod = pv && c + 1 < nC;
    uint4 Bn[2][NS];
    if (next) fetchB(b + 1, Bn);
    uint4 qr = make_uint4(0u, 0u, 0u, 0u);
    if (k == 0) {
      pe0 = nx.x;
      pe1 = nx.y;
      nx = ck_state(4 * (c + 2) + pq);
    } else if (tid < 128) {
      qr = fetch_rec(c + 2);
    }
    uint4 Ab[2][1][NS];
    loadA(abuf, k, 0, Ab[0]);
#pragma unroll
    for (int g = 0; g < 4; g++) {
      if (g < 3) loadA(abuf, k, g + 1, Ab[(g + 1) & 1]);
      if (prod) produce_pair(abuf ^ 1, abuf ^ 1, 4 * k + g);
      consume_group(g, k, Ab[g & 1], Bc);
      __builtin_amdgcn_sched_barrier(0);
    }
    if (next) {
#pragma unroll
      for (int t = 0; t < 2; t++)
#pragma unroll
        for (int s = 0; s < NS; s++) Bc[t][s] = Bn[t][s];
    }
    if (k == 1) {
      // Chunk c + 2's coefficients (rsc[c & 1] was last read in chunk c - 1).
      if (tid < 128) rsc[abuf][tid] = qr;
      __syncthreads();
    }
  };
  for (int c = 0; c + 1 < nC; c++) {
    block(c, 0, true);
    block(c, 1, true);
  }
  block(nC - 1, 0, true);
  block(nC - 1, 1, false);

  // F(t) = E + O at the northern t, F(nt-1-t) = E - O at its mirror (C
  // layout of the 16x16x32 MFMA: lane holds rows 4 (lane / 16) + r, column
  // lane % 16).
  const int ncg16 = (p.ncol + 15) >> 4;
  const long rs = (long)ncg16 * p.nslot * 16;
  const int row0 = tsb * 128 + h * 64 + 4 * (lane >> 4);
#pragma unroll
  for (int t = 0; t < 2; t++) {
    const int col = (tile0 + t) * 16 + (lane & 15);
    if (col >= p.ncol) continue;
    const long fi = f_index(row0, slot, col, p.nslot, ncg16);
    const long fm = f_index(nt - 1 - row0, slot, col, p.nslot, ncg16);
#pragma unroll
    for (int g = 0; g < 4; g++)
#pragma unroll
      for (int r = 0; r < 4; r++) {
        const int row = row0 + g * 16 + r;
        if (row >= nn) continue;
        const double e = acc[g][t][0][r], o = acc[g][t][1][r];
        const long dn = (long)(g * 16 + r) * rs;  // rows past row0 (north) / before nt-1-row0 (south)
        if (p.f32out) {
          reinterpret_cast<float*>(p.F)[fi + dn] = (float)(e + o);
          reinterpret_cast<float*>(p.F)[fm - dn] = (float)(e - o);
        } else {
          p.F[fi + dn] = e + o;
          p.F[fm - dn] = e - o;
        }
      }
  }
}

// ---------------------------------------------------------------------------
// Synthesis of complex fields, orders m and -m in one workgroup.  The
// Gauss-Legendre nodes are symmetric about the equator and d^l_{-m,n}(theta)
// = (-1)^(l+n) d^l_{m,n}(pi - theta), so
//   F_{-m}(nt-1-t) = sum_l d^l_{m,n}(theta_t) (-1)^(l+n) c_{l,-m}:
// the recurrence values of order m >= 0 at colatitude t serve order -m at
// the mirrored row, against its coefficients with the sign folded in
// (coef_split with p.pm).  Workgroup = 8 waves for (128 colatitudes of block
// tsb) x (128 columns of super-group csg) of orders m and -m: the same MFMAs
// per recurrence value as leg_synth_bf16 over twice the columns, so half the
// recurrence work and coefficient reads per MFMA.  The degrees go in
// 64-degree chunks (bf16 planes of D double-buffered in LDS), each taken as
// two K blocks of 32 degrees.  Wave w multiplies column tile w of both
// orders over all 128 colatitudes (8 row groups; each A operand feeds the
// two orders' MFMA chains), loading its B operands straight from the split
// coefficients one block ahead (no LDS staging, one barrier per chunk).
// While the waves multiply chunk c they produce chunk c + 1: wave w runs
// the recurrence of colatitude block w % 2 (64 colatitudes) for the 16
// degrees of quarter w / 2 from the checkpoint there, four recurrence pairs
// per K block.  Blocks past the order's last 32-degree chunk are skipped;
// order -0's sums are not stored.
// ---------------------------------------------------------------------------
template <int NS>
__global__ __launch_bounds__(512) void leg_synth_bf16_pm(LegParams p) {
  static_assert(kCk == 16, "four 16-degree quarters per 64-degree chunk");
  // Planes of a chunk: degree half dh, entries 8 kr .. 8 kr + 7, 128 colatitudes.
  __shared__ uint4 trb[2][NS][2][4][128];
  // Recurrence coefficients of a chunk (64 entries, 2 KiB).
  __shared__ uint4 rsc[2][128];
  const int lane = threadIdx.x & 63;
  const int tid = threadIdx.x;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  // Work item: (order m >= 0, column super-group, block of 128 colatitudes).
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int m = it.x;
  const int tsb = it.y & 0xffff;
  const int csg = it.y >> 16;
  const int slot = m - p.m_first, slotn = -m - p.m_first;
  const int l0 = max(m, p.nabs);
  const int nt = p.ntheta;
  const int nC = (p.lmax - l0 + 64) >> 6;   // 64-degree chunks
  const int n32 = (p.lmax - l0 + 32) >> 5;  // 32-degree coefficient blocks of each order
  const int rows = min(128, nt - tsb * 128);  // valid colatitudes of the item

  // Producer share: colatitude block pb, degree quarter pq of a chunk.
  const int pb = wave & 1, pq = wave >> 1;
  const int tpc = min(tsb * 128 + 64 * pb + lane, nt - 1);  // past ntheta: the last one
  const double pcs = p.cosv[tpc];
  const int nck = (p.L - l0 + kCk) / kCk;  // checkpoints 1 .. nck-1 exist
  const double2* ckm =
      (nck > 1 ? p.ck + ld_uniform(p.ckoff, m + p.L) * (long)nt : p.ck) + tpc;
  auto ck_state = [&](int k) { return ckm[(long)(nck > 1 ? min(k, nck - 1) - 1 : 0) * nt]; };
  double pe0, pe1;
  {
    // Chunk 0: the seed for quarter 0, checkpoint pq for the others.
    const double seed = p.seeds[(long)(m + p.L) * nt + tpc];
    const double2 c1 = ck_state(max(pq, 1));
    pe0 = pq == 0 || nck < 2 ? seed : c1.x;
    pe1 = pq == 0 || nck < 2 ? 0.0 : c1.y;
  }

  const uint4* recq = reinterpret_cast<const uint4*>(p.rec + (ld_uniform(p.roff, m + p.L) - l0));
  auto fetch_rec = [&](int c) { return recq[2 * (l0 + 64 * min(c, nC - 1) + 1) + (tid & 127)]; };

  // B operands of block b = 4 c + 2 dh + o: coefficient block 2 c + dh
  // (clamped to the last) of order m (o = 0) or -m (o = 1); each wave loads
  // its two column tiles' planes straight from the split buffer (16 bytes a
  // lane, 1 KiB a wave instruction) one block ahead.
  const uint4* bbp = p.bs + (ld_uniform(p.cso, slot) * (long)p.ntile + csg * 8) * NS * 64;
  const uint4* bbn = p.bs + (ld_uniform(p.cso, slotn) * (long)p.ntile + csg * 8) * NS * 64;
  const int bstride = p.ntile * NS * 64;  // uint4 per coefficient block
  const int toff = wave * NS * 64 + lane;  // column tile csg * 8 + wave
  // B operands of degree half dh of chunk c for both orders: B[o] of order m
  // (o = 0) or -m (o = 1).
  auto fetchB = [&](int c, int dh, uint4 (&B)[2][NS]) {
    const long off = (long)min(2 * c + dh, n32 - 1) * bstride + toff;
#pragma unroll
    for (int s = 0; s < NS; s++) {
      B[0][s] = bbp[off + s * 64];
      B[1][s] = bbn[off + s * 64];
    }
  };

  // Consumer role: column tile w of both orders, all 128 colatitudes.
  const int tile = csg * 8 + wave;
  d4 acc[8][2];  // row group, order
#pragma unroll
  for (int g = 0; g < 8; g++)
#pragma unroll
    for (int o = 0; o < 2; o++) acc[g][o] = d4{0.0, 0.0, 0.0, 0.0};

  // Production: pair i (0..7) of this wave's quarter is steps 2 i, 2 i + 1,
  // degrees 16 pq + 2 i (+1) of the chunk; every 4 pairs fill the uint4 of
  // each plane holding entries 8 kr .. 8 kr + 7 of degree half pq / 2 (kr =
  // 2 (pq % 2) + i / 4), stored whole.
  auto produce_pair = [&](int rbuf, int tbuf, int i) {
    const d4* rl = reinterpret_cast<const d4*>(&rsc[rbuf][0]) + 16 * pq;
    const double v0 = pe0;
    rec_step(rl[2 * i], pcs, pe0, pe1);
    const double v1 = pe0;
    rec_step(rl[2 * i + 1], pcs, pe0, pe1);
    unsigned o[NS];
    split_pair<NS>(v0, v1, o);
#pragma unroll
    for (int s = 0; s < NS; s++)
      reinterpret_cast<unsigned*>(
          &trb[tbuf][s][pq >> 1][2 * (pq & 1) + (i >> 2)][64 * pb + lane])[i & 3] = o[s];
  };
  auto loadA = [&](int abuf, int dh, int g, uint4 (&A)[1][NS]) {
#pragma unroll
    for (int s = 0; s < NS; s++) A[0][s] = trb[abuf][s][dh][lane >> 4][g * 16 + (lane & 15)];
  };
  auto consume_group = [&](int g, const uint4 (&A)[1][NS], const uint4 (&B)[2][NS]) {
    if (g * 16 < rows) {
      f4 c[1][2];
      split_mma<NS, 1>(A, B, c);
      flush(acc[g][0], c[0][0]);
      flush(acc[g][1], c[0][1]);
    }
  };

  // Prologue: coefficients of chunks 0 and 1, B operands of block 0,
  // production of chunk 0.
  if (tid < 128) {
    rsc[0][tid] = fetch_rec(0);
    rsc[1][tid] = fetch_rec(1);
  }
  uint4 Bc[2][NS];
  fetchB(0, 0, Bc);
  __syncthreads();
  const bool pv = 64 * pb < rows;  // this wave's producer block holds colatitudes
  if (pv) {
#pragma unroll
    for (int i = 0; i < 8; i++) produce_pair(0, 0, i);
  }
  __syncthreads();

  // Block 2 c + dh of chunk c (dh a constant at each call): degree half dh
  // for both orders, the B operands of the next block fetched into
  // registers, production pairs 4 dh .. 4 dh + 3 of chunk c + 1 (none in the
  // last chunk).  One barrier per chunk: the planes of chunk c + 1 and the
  // coefficients of chunk c + 2 are complete, and chunk c's planes free,
  // after it.  The state of chunk c + 1 (checkpoint 4 (c + 1) + pq) is
  // loaded one chunk before its use.
  double2 nx = ck_state(4 + pq);
  auto block = [&](int c, const int dh, const bool next) {
    const int abuf = c & 1;
    const bool prod = pv && c + 1 < nC;
    const bool live = 2 * c + dh < n32;
    uint4 Bn[2][NS];
    if (next) fetchB(c + dh, dh ^ 1, Bn);
    uint4 qr = make_uint4(0u, 0u, 0u, 0u);
    if (dh == 0) {
      pe0 = nx.x;
      pe1 = nx.y;
      nx = ck_state(4 * (c + 2) + pq);
    } else if (tid < 128) {
      qr = fetch_rec(c + 2);
    }
    if (live) {
      uint4 Ab[2][1][NS];
      loadA(abuf, dh, 0, Ab[0]);
#pragma unroll
      for (int g = 0; g < 8; g++) {
        if (g < 7) loadA(abuf, dh, g + 1, Ab[(g + 1) & 1]);
        if (prod && (g & 1)) produce_pair(abuf ^ 1, abuf ^ 1, 4 * dh + (g >> 1));
        consume_group(g, Ab[g & 1], Bc);
        __builtin_amdgcn_sched_barrier(0);
      }
    } else if (prod) {
#pragma unroll
      for (int i = 0; i < 4; i++) produce_pair(abuf ^ 1, abuf ^ 1, 4 * dh + i);
    }
    if (next) {
#pragma unroll
      for (int o = 0; o < 2; o++)
#pragma unroll
        for (int s = 0; s < NS; s++) Bc[o][s] = Bn[o][s];
    }
    if (dh == 1) {
      // Chunk c + 2's coefficients (rsc[c & 1] was last read in chunk c - 1).
      if (tid < 128) rsc[abuf][tid] = qr;
      __syncthreads();
    }
  };
  for (int c = 0; c + 1 < nC; c++) {
    block(c, 0, true);
    block(c, 1, true);
  }
  block(nC - 1, 0, true);
  block(nC - 1, 1, false);

  // Order m at rows t, order -m at rows nt-1-t (C layout of the 16x16x32
  // MFMA: lane holds rows 4 (lane / 16) + r, column lane % 16).
  const int ncg16 = (p.ncol + 15) >> 4;
  const long rs = (long)ncg16 * p.nslot * 16;
  const int row0 = tsb * 128 + 4 * (lane >> 4);
  const int col = tile * 16 + (lane & 15);
  if (col >= p.ncol) return;
  const long fi = f_index(row0, slot, col, p.nslot, ncg16);
  const long fm = f_index(nt - 1 - row0, slotn, col, p.nslot, ncg16);
#pragma unroll
  for (int g = 0; g < 8; g++)
#pragma unroll
    for (int r = 0; r < 4; r++) {
      const int row = row0 + g * 16 + r;
      if (row >= nt) continue;
      const long dn = (long)(g * 16 + r) * rs;  // rows past row0 / before nt-1-row0
      const double a = acc[g][0][r], an = acc[g][1][r];
      if (p.f32out) {
        reinterpret_cast<float*>(p.F)[fi + dn] = (float)a;
        if (m > 0) reinterpret_cast<float*>(p.F)[fm - dn] = (float)an;
      } else {
        p.F[fi + dn] = a;
        if (m > 0) p.F[fm - dn] = an;
      }
    }
}

// ---------------------------------------------------------------------------
// Analysis.  Workgroup = 8 waves for one order m, one pass of 64 degrees (4
// MFMA row tiles) and 128 columns (column super-group csg), covering every
// colatitude in groups of 128 (two 64-lane blocks = four 32-colatitude K
// chunks); the bf16 planes D[s][l][theta] of a group are double-buffered in
// LDS.  While the waves multiply group g they produce group g + 1: wave w
// runs the recurrence of block half w % 2 for the 16 degrees
// lp + 16 (w / 2) .. + 15, starting from the checkpoint there, so every wave
// does the same share with no branches.  Wave w multiplies the 4 degree
// tiles with column tile w (so one wave loads each B operand block), and
// ends a pass with complete sums over all colatitudes, which it writes
// directly.
// ---------------------------------------------------------------------------
template <int NS, int REAL>
__global__ __launch_bounds__(512) void leg_anal_bf16(LegParams p) {
  static_assert(kPass == 64 && kCk == 16, "4 degree quarters of 16 per pass");
  // Rows of 128 colatitudes padded to 136 (272 bytes): the 16 lanes of a
  // 16-byte A-operand read hit 16 consecutive rows, which then start 4 banks
  // apart and cover all 64 banks once (256-byte rows would all map to the
  // same 4 banks).
  __shared__ __bf16 Dl[2][NS][64][136];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  // Work item: (order, pass, column super-group).
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int mw = it.x;
  const int pass = it.y & 0xffff;
  const int csg = it.y >> 16;
  const int m = order_of(mw, REAL);
  const int slot = m - p.m_first;
  const int l0 = max(abs(m), p.nabs);
  const int nt = p.ntheta;
  const int ngrp = (nt + 127) >> 7;
  const int lp = l0 + kPass * pass;  // first degree of the pass
  // Recurrence coefficients of the pass: entries lp+1 .. lp+64 (entry l for
  // degree l; entries up to L + kRecPad exist).
  rec_t* rq = rec_ptr(p.rec, ld_uniform(p.roff, m + p.L) - l0);

  // Producer share of this wave: block half bh, degree quarter q.  Its
  // recurrence starts at degree lp + 16 q from checkpoint kq (the seed for
  // kq = 0).  Quarters past L (only garbage rows, never written) clamp to
  // the last checkpoint so every load stays in the table.
  const int bh = wave & 1, q = wave >> 1;
  const int nck = (p.L - l0 + kCk) / kCk;  // checkpoints 1 .. nck-1 exist
  const int kq = min((kPass / kCk) * pass + q, nck - 1);
  const double2* ckp = kq == 0 ? p.ck : p.ck + (ld_uniform(p.ckoff, m + p.L) + kq - 1) * (long)nt;
  const double* sdp = p.seeds + (long)(m + p.L) * nt;
  // The state and cosine of group g's colatitude are loaded into registers
  // one group before they are used, after the previous production (so the
  // loads land in registers just freed and nothing waits on them until the
  // next production).  Colatitudes past ntheta load the last one's values:
  // their products meet zero B operands, and the values stay finite.
  double nx_seed, nx_cs;
  double2 nx_ck;
  auto prefetch = [&](int g) {
    const int tc = min((2 * g + bh) * 64 + lane, nt - 1);
    nx_seed = sdp[tc];
    nx_ck = ckp[tc];
    nx_cs = p.cosv[tc];
  };
  // Production of this wave's 16 degrees x 64 colatitudes of a group into
  // buffer buf, in four parts of 4 degrees (one per K chunk of the group
  // multiplied meanwhile, so the scheduler can interleave their VALU work
  // with that chunk's MFMAs).  begin() takes the prefetched state.
  double pe0, pe1, pcs;
  auto begin = [&]() {
    pcs = nx_cs;
    pe0 = kq == 0 ? nx_seed : nx_ck.x;
    pe1 = kq == 0 ? 0.0 : nx_ck.y;
  };
  auto produce4 = [&](int buf, int part) {
    const int eq = lp + 16 * q + 1;  // entry of step 0, wave-uniform
#pragma unroll
    for (int k = 4 * part; k < 4 * part + 4; k += 2) {
      const double v0 = pe0;
      rec_step(rec_entry(rq, eq + k), pcs, pe0, pe1);
      const double v1 = pe0;
      rec_step(rec_entry(rq, eq + k + 1), pcs, pe0, pe1);
      unsigned o[NS];
      split_pair<NS>(v0, v1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const bf16x2 pr = __builtin_bit_cast(bf16x2, o[s]);
        Dl[buf][s][16 * q + k][bh * 64 + lane] = pr[0];
        Dl[buf][s][16 * q + k + 1][bh * 64 + lane] = pr[1];
      }
    }
  };

  // Consumer role.
  // Wave w multiplies all 4 degree tiles with column tile w, so every B
  // operand block is loaded by one wave of the workgroup (the A operands
  // come from LDS).  B comes straight from the FP32 bins FA32: lane (q4, c) =
  // (lane / 16, lane % 16) of K chunk kc takes colatitudes 32 kc + 8 q4 + e
  // (e = 0..7) of column 16 tile + c, i.e. FA32 entries 16 e past its base
  // in quad 8 kc + 2 q4 (each e 64 contiguous bytes per 16 lanes), split
  // into bf16 planes just before the chunk's MFMAs.  A tile past the last
  // column group reads the last one: its outputs are never stored, and a
  // column only meets its own B values, so nothing else needs those columns
  // zero.  Reads of the last block run up to 8 quads past it (the FA buffer
  // is padded for that); colatitudes from ntheta on are zeroed at the split
  // of the last group's chunks.  Callers pass chunks below nkc only.
  struct RawB {
    float x[8];
  };
  const int tile = csg * 8 + wave;
  const float* fb =
      p.FA32 + (((long)slot * p.n_cg + min(tile, p.n_cg - 1)) * p.nquad + 2 * (lane >> 4)) * 64 +
      (lane & 15);
  auto loadB = [&](int kc, RawB& X) {
#pragma unroll
    for (int e = 0; e < 8; e++) X.x[e] = fb[(long)kc * 512 + 16 * e];
  };
  auto splitB = [&](const RawB& X, uint4 (&B)[1][NS], int kc, bool last) {
    unsigned pl[NS][4];
    const int th = 32 * kc + 8 * (lane >> 4);  // colatitude of e = 0
#pragma unroll
    for (int e2 = 0; e2 < 4; e2++) {
      float x0 = X.x[2 * e2], x1 = X.x[2 * e2 + 1];
      if (last) {
        x0 = th + 2 * e2 < nt ? x0 : 0.0f;
        x1 = th + 2 * e2 + 1 < nt ? x1 : 0.0f;
      }
      unsigned o[NS];
      split_pair_f<NS>(x0, x1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) pl[s][e2] = o[s];
    }
#pragma unroll
    for (int s = 0; s < NS; s++) B[0][s] = make_uint4(pl[s][0], pl[s][1], pl[s][2], pl[s][3]);
  };

  d4 acc[4][1];
#pragma unroll
  for (int i = 0; i < 4; i++) acc[i][0] = d4{0.0, 0.0, 0.0, 0.0};
  // One K chunk kc of buffer buf.  After this chunk's MFMAs are issued, the
  // chains of the previous chunk (cp, complete by then; zero before the
  // first) join a pair: an odd chunk kc keeps them in pp, an even one adds
  // them to pp in FP32 and flushes the pair sum to FP64 (one FP32 rounding
  // per pair of 32-term chunk sums, half the FP64 work of flushing every
  // chunk).  Chunk indices of a group are 4 g + j, so the parity is j's.
  f4 cp[4][1], pp[4][1];
#pragma unroll
  for (int i = 0; i < 4; i++) cp[i][0] = pp[i][0] = f4{0.f, 0.f, 0.f, 0.f};
  auto step = [&](int buf, int kc, const RawB& X, bool last, bool odd) {
    uint4 B[1][NS];
    splitB(X, B, kc, last);
    uint4 A[4][NS];
#pragma unroll
    for (int lt = 0; lt < 4; lt++)
#pragma unroll
      for (int s = 0; s < NS; s++)
        A[lt][s] = *reinterpret_cast<const uint4*>(
            &Dl[buf][s][lt * 16 + (lane & 15)][(kc & 3) * 32 + 8 * (lane >> 4)]);
    f4 c[4][1];
    split_mma<NS, 4, 1>(A, B, c);
#pragma unroll
    for (int i = 0; i < 4; i++) {
      if (odd)
        pp[i][0] = cp[i][0];
      else
        flush(acc[i][0], pp[i][0] + cp[i][0]);
      cp[i][0] = c[i][0];
    }
  };

  prefetch(0);
  begin();
#pragma unroll
  for (int part = 0; part < 4; part++) produce4(0, part);
  prefetch(1);
  __syncthreads();
  // B operands: chunk c sits in buffer X[c & 3], loaded kAhead chunks before
  // its MFMAs (into the buffer chunk c - 4 + kAhead used).  Loads clamp to
  // the last chunk nkc - 1; what they bring past it is never multiplied.
  // Every group but the last has four chunks below nkc (nt > 128 (ngrp-1)),
  // and the last group's first chunk is below nkc too; only the last group's
  // chunks hold colatitudes past ntheta.  The scheduling barriers keep each
  // chunk's loads, production part and MFMAs together.
  constexpr int kAhead = 3;
  const int kcl = p.nkc - 1;
  RawB X[4];
#pragma unroll
  for (int j = 0; j < kAhead; j++) loadB(min(j, kcl), X[j]);
  for (int g = 0; g + 1 < ngrp; g++) {
    const int buf = g & 1;
    const int kc0 = 4 * g;
    begin();
    prefetch(g + 2);  // a whole group ahead of its use
#pragma unroll
    for (int j = 0; j < 4; j++) {
      loadB(min(kc0 + j + kAhead, kcl), X[(j + kAhead) & 3]);
      produce4(buf ^ 1, j);
      step(buf, kc0 + j, X[j], false, j & 1);
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
  }
  {
    // Last group: chunks 4 (ngrp-1) .. nkc-1.
    const int buf = (ngrp - 1) & 1;
    const int kc0 = 4 * (ngrp - 1);
    step(buf, kc0, X[0], true, false);
    if (kc0 + 1 < p.nkc) {
      step(buf, kc0 + 1, X[1], true, true);
      if (kc0 + 2 < p.nkc) {
        step(buf, kc0 + 2, X[2], true, false);
        if (kc0 + 3 < p.nkc) {
          loadB(kc0 + 3, X[3]);
          step(buf, kc0 + 3, X[3], true, true);
        }
      }
    }
    // The last chunk nkc - 1: an odd one completes the pair in pp, an even
    // one (its predecessor's pair already flushed) stands alone.
    const bool lodd = (p.nkc - 1) & 1;
#pragma unroll
    for (int i = 0; i < 4; i++) flush(acc[i][0], lodd ? pp[i][0] + cp[i][0] : cp[i][0]);
  }

  // Rows l = lp + 16*lt + 4*(lane/16) + r, column 16*tile + lane%16.
  const int col = tile * 16 + (lane & 15);
  const bool odd = lane & 1;
  const bool cv = col < p.ncol;  // the same for both lanes of a field (ncol is even)
  const int lf = cv ? field_lmax(p, col >> 1) : -1;
  const long cbase = cv ? (long)field_index(p.fmap, col >> 1) * p.cstride : 0;
#pragma unroll
  for (int lt = 0; lt < 4; lt++)
#pragma unroll
    for (int rp = 0; rp < 2; rp++) {
      const double2 v = trade_halves(acc[lt][0][2 * rp], acc[lt][0][2 * rp + 1], odd);
      const int l = lp + 16 * lt + 4 * (lane >> 4) + 2 * rp + odd;
      if (l <= lf) store_coef(p, cbase, l, m, lf, REAL, v);
    }
}

// ---------------------------------------------------------------------------
// Analysis of complex fields, orders m and -m in one workgroup.  The
// Gauss-Legendre nodes and weights are symmetric about the equator and
// d^l_{-m,n}(theta) = (-1)^(l+n) d^l_{m,n}(pi - theta), so
//   c_{l,-m} = (-1)^(l+n) sum_t d^l_{m,n}(theta_t) FA_{-m}(nt-1-t):
// the recurrence values of order m >= 0 serve both orders.  Items, production
// and chunk order are leg_anal_bf16's (order m, a 64-degree pass, 128
// columns, colatitudes in groups of 128); wave w multiplies the 4 degree
// tiles with column tile w of order m and column tile w of order -m, whose
// B operand reads the bins of row nt-1-theta for K index theta.  Per MFMA,
// half the A-operand LDS reads and half the recurrence work of leg_anal_bf16.
// The +m sums are leg_anal_bf16's bits; the -m sums group the colatitudes
// into chunks from the other pole.  Order 0 has no partner: its second tile
// repeats the first and is not stored.
// ---------------------------------------------------------------------------
template <int NS>
__global__ __launch_bounds__(512) void leg_anal_bf16_pm(LegParams p) {
  static_assert(kPass == 64 && kCk == 16, "4 degree quarters of 16 per pass");
  __shared__ __bf16 Dl[2][NS][64][136];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  // Work item: (order m >= 0, pass, column super-group).
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int m = it.x;
  const int pass = it.y & 0xffff;
  const int csg = it.y >> 16;
  const int slot = m - p.m_first, slotn = -m - p.m_first;
  const int l0 = max(m, p.nabs);
  const int nt = p.ntheta;
  const int ngrp = (nt + 127) >> 7;
  const int lp = l0 + kPass * pass;
  rec_t* rq = rec_ptr(p.rec, ld_uniform(p.roff, m + p.L) - l0);

  // Production (leg_anal_bf16's): block half bh, degree quarter q.
  const int bh = wave & 1, q = wave >> 1;
  const int nck = (p.L - l0 + kCk) / kCk;
  const int kq = min((kPass / kCk) * pass + q, nck - 1);
  const double2* ckp = kq == 0 ? p.ck : p.ck + (ld_uniform(p.ckoff, m + p.L) + kq - 1) * (long)nt;
  const double* sdp = p.seeds + (long)(m + p.L) * nt;
  double nx_seed, nx_cs;
  double2 nx_ck;
  auto prefetch = [&](int g) {
    const int tc = min((2 * g + bh) * 64 + lane, nt - 1);
    nx_seed = sdp[tc];
    nx_ck = ckp[tc];
    nx_cs = p.cosv[tc];
  };
  double pe0, pe1, pcs;
  auto begin = [&]() {
    pcs = nx_cs;
    pe0 = kq == 0 ? nx_seed : nx_ck.x;
    pe1 = kq == 0 ? 0.0 : nx_ck.y;
  };
  auto produce4 = [&](int buf, int part) {
    const int eq = lp + 16 * q + 1;
#pragma unroll
    for (int k = 4 * part; k < 4 * part + 4; k += 2) {
      const double v0 = pe0;
      rec_step(rec_entry(rq, eq + k), pcs, pe0, pe1);
      const double v1 = pe0;
      rec_step(rec_entry(rq, eq + k + 1), pcs, pe0, pe1);
      unsigned o[NS];
      split_pair<NS>(v0, v1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const bf16x2 pr = __builtin_bit_cast(bf16x2, o[s]);
        Dl[buf][s][16 * q + k][bh * 64 + lane] = pr[0];
        Dl[buf][s][16 * q + k + 1][bh * 64 + lane] = pr[1];
      }
    }
  };

  // B operands: lane (q4, c) of K chunk kc takes colatitudes theta = 32 kc +
  // 8 q4 + e (e = 0..7) of column 16 tile + c.  Order m reads FA row theta
  // (leg_anal_bf16's addresses); order -m reads row nt-1-theta, 16 floats
  // before the previous e within the (slot, column group) block, whose rows
  // are 16 contiguous floats each.  Rows below 0 (theta >= ntheta: the last
  // chunk only, at most 31 rows) read the previous block or the buffer's
  // front pad of kFAFrontPad floats; the split zeroes them.
  struct RawB {
    float x[8];
  };
  const int tile = csg * 8 + wave;
  const int tcl = min(tile, p.n_cg - 1);
  const float* fb =
      p.FA32 + (((long)slot * p.n_cg + tcl) * p.nquad + 2 * (lane >> 4)) * 64 + (lane & 15);
  const float* fbn = p.FA32 + ((long)slotn * p.n_cg + tcl) * p.nquad * 64 + (lane & 15);
  const int rn0 = nt - 1 - 8 * (lane >> 4);  // row of order -m at kc = 0, e = 0
  auto loadB = [&](int kc, RawB (&X)[2]) {
#pragma unroll
    for (int e = 0; e < 8; e++) X[0].x[e] = fb[(long)kc * 512 + 16 * e];
    const float* rn = fbn + 16 * (rn0 - 32 * kc);
#pragma unroll
    for (int e = 0; e < 8; e++) X[1].x[e] = rn[-16 * e];
  };
  auto splitB = [&](const RawB& X, uint4 (&B)[NS], int kc, bool last) {
    unsigned pl[NS][4];
    const int th = 32 * kc + 8 * (lane >> 4);
#pragma unroll
    for (int e2 = 0; e2 < 4; e2++) {
      float x0 = X.x[2 * e2], x1 = X.x[2 * e2 + 1];
      if (last) {
        x0 = th + 2 * e2 < nt ? x0 : 0.0f;
        x1 = th + 2 * e2 + 1 < nt ? x1 : 0.0f;
      }
      unsigned o[NS];
      split_pair_f<NS>(x0, x1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) pl[s][e2] = o[s];
    }
#pragma unroll
    for (int s = 0; s < NS; s++) B[s] = make_uint4(pl[s][0], pl[s][1], pl[s][2], pl[s][3]);
  };

  d4 acc[4][2];
#pragma unroll
  for (int i = 0; i < 4; i++) acc[i][0] = acc[i][1] = d4{0.0, 0.0, 0.0, 0.0};
  // Chunk sums flushed in pairs, as in leg_anal_bf16 (per tile and order).
  f4 cp[4][2], pp[4][2];
#pragma unroll
  for (int i = 0; i < 4; i++)
#pragma unroll
    for (int j = 0; j < 2; j++) cp[i][j] = pp[i][j] = f4{0.f, 0.f, 0.f, 0.f};
  auto step = [&](int buf, int kc, const RawB (&X)[2], bool last, bool odd) {
    uint4 B[2][NS];
    splitB(X[0], B[0], kc, last);
    splitB(X[1], B[1], kc, last);
#pragma unroll
    for (int half = 0; half < 2; half++) {
      uint4 A[2][NS];
#pragma unroll
      for (int lt = 0; lt < 2; lt++)
#pragma unroll
        for (int s = 0; s < NS; s++)
          A[lt][s] = *reinterpret_cast<const uint4*>(
              &Dl[buf][s][(2 * half + lt) * 16 + (lane & 15)][(kc & 3) * 32 + 8 * (lane >> 4)]);
      f4 c[2][2];
      split_mma<NS, 2, 2>(A, B, c);
#pragma unroll
      for (int i = 0; i < 2; i++)
#pragma unroll
        for (int j = 0; j < 2; j++) {
          const int t = 2 * half + i;
          if (odd)
            pp[t][j] = cp[t][j];
          else
            flush(acc[t][j], pp[t][j] + cp[t][j]);
          cp[t][j] = c[i][j];
        }
    }
  };

  prefetch(0);
  begin();
#pragma unroll
  for (int part = 0; part < 4; part++) produce4(0, part);
  prefetch(1);
  __syncthreads();
  // B operands one chunk ahead: chunk c in X[c & 1] (loads clamp to the
  // last chunk nkc - 1).
  const int kcl = p.nkc - 1;
  RawB X[2][2];
  loadB(0, X[0]);
  for (int g = 0; g + 1 < ngrp; g++) {
    const int buf = g & 1;
    const int kc0 = 4 * g;
    begin();
    prefetch(g + 2);
#pragma unroll
    for (int j = 0; j < 4; j++) {
      loadB(min(kc0 + j + 1, kcl), X[(j + 1) & 1]);
      produce4(buf ^ 1, j);
      step(buf, kc0 + j, X[j & 1], false, j & 1);
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
  }
  {
    // Last group: chunks 4 (ngrp-1) .. nkc-1 (X[0] holds the first).
    const int buf = (ngrp - 1) & 1;
    const int kc0 = 4 * (ngrp - 1);
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (kc0 + j < p.nkc) {
        if (kc0 + j + 1 < p.nkc) loadB(kc0 + j + 1, X[(j + 1) & 1]);
        step(buf, kc0 + j, X[j & 1], true, j & 1);
      }
    const bool lodd = (p.nkc - 1) & 1;
#pragma unroll
    for (int i = 0; i < 4; i++)
#pragma unroll
      for (int j = 0; j < 2; j++) flush(acc[i][j], lodd ? pp[i][j] + cp[i][j] : cp[i][j]);
  }

  // Rows l = lp + 16*lt + 4*(lane/16) + r, column 16*tile + lane%16; order
  // -m's sums take the sign (-1)^(l+n).
  const int col = tile * 16 + (lane & 15);
  const bool odd = lane & 1;
  const bool cv = col < p.ncol;  // the same for both lanes of a field (ncol is even)
  const int lf = cv ? field_lmax(p, col >> 1) : -1;
  const long cbase = cv ? (long)field_index(p.fmap, col >> 1) * p.cstride : 0;
#pragma unroll
  for (int j = 0; j < 2; j++) {
    if (j == 1 && m == 0) break;
#pragma unroll
    for (int lt = 0; lt < 4; lt++)
#pragma unroll
      for (int rp = 0; rp < 2; rp++) {
        double2 v = trade_halves(acc[lt][j][2 * rp], acc[lt][j][2 * rp + 1], odd);
        const int l = lp + 16 * lt + 4 * (lane >> 4) + 2 * rp + odd;
        if (j == 1 && ((l + p.n) & 1)) v = make_double2(-v.x, -v.y);
        if (l <= lf) store_coef(p, cbase, l, j ? -m : m, lf, 0, v);
      }
  }
}

// ---------------------------------------------------------------------------
// Analysis for upper index 0 over the northern colatitudes.  Gauss-Legendre
// nodes are symmetric about the equator and d^l_{m0}(pi - theta) =
// (-1)^(l+m) d^l_{m0}(theta), so sum_t d_l(theta_t) FA(t) runs over the
// northern t (the equator once) against FA(t) + FA(nt-1-t) for even l + m and
// FA(t) - FA(nt-1-t) for odd l + m: half the colatitudes, recurrence values
// and MFMAs of leg_anal_bf16, whose items, production and chunk order this
// kernel keeps.  The degrees of a pass are stored in LDS by parity class
// (tiles 0, 1: l + m even; tiles 2, 3: odd), each class multiplied with its
// own B operand; the sums FA(t) +- FA(nt-1-t) of the FP32 bins round once to
// FP32 before the split, and each chunk's sums are flushed to FP64 on their
// own, as in leg_anal_bf16_ws (the same bits).
// ---------------------------------------------------------------------------
template <int NS, int REAL>
__global__ __launch_bounds__(512) void leg_anal_bf16_s(LegParams p) {
  static_assert(kPass == 64 && kCk == 16, "4 degree quarters of 16 per pass");
  // Rows of 128 colatitudes padded to 136 (272 bytes): the 16 lanes of a
  // 16-byte A-operand read hit 16 consecutive rows, which then start 4 banks
  // apart and cover all 64 banks once (256-byte rows would all map to the
  // same 4 banks).
  __shared__ __bf16 Dl[2][NS][64][136];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  // Work item: (order, pass, column super-group).
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int mw = it.x;
  const int pass = it.y & 0xffff;
  const int csg = it.y >> 16;
  const int m = order_of(mw, REAL);
  const int slot = m - p.m_first;
  const int l0 = max(abs(m), p.nabs);
  const int nt = p.ntheta;
  const int nts = (nt + 1) >> 1;      // northern colatitudes, the equator included
  const int ngrp = (nts + 127) >> 7;
  const int nkc = (nts + 31) >> 5;    // 32-colatitude K chunks of the north
  const int lp = l0 + kPass * pass;  // first degree of the pass
  const int par0 = (lp + m) & 1;     // parity of l + m at the pass's first degree
  // Recurrence coefficients of the pass: entries lp+1 .. lp+64 (entry l for
  // degree l; entries up to L + kRecPad exist).
  rec_t* rq = rec_ptr(p.rec, ld_uniform(p.roff, m + p.L) - l0);

  // Producer share of this wave: block half bh, degree quarter q.  Its
  // recurrence starts at degree lp + 16 q from checkpoint kq (the seed for
  // kq = 0).  Quarters past L (only garbage rows, never written) clamp to
  // the last checkpoint so every load stays in the table.
  const int bh = wave & 1, q = wave >> 1;
  const int nck = (p.L - l0 + kCk) / kCk;  // checkpoints 1 .. nck-1 exist
  const int kq = min((kPass / kCk) * pass + q, nck - 1);
  const double2* ckp = kq == 0 ? p.ck : p.ck + (ld_uniform(p.ckoff, m + p.L) + kq - 1) * (long)nt;
  const double* sdp = p.seeds + (long)(m + p.L) * nt;
  // The state and cosine of group g's colatitude are loaded into registers
  // one group before they are used, after the previous production (so the
  // loads land in registers just freed and nothing waits on them until the
  // next production).  Colatitudes past ntheta load the last one's values:
  // their products meet zero B operands, and the values stay finite.
  double nx_seed, nx_cs;
  double2 nx_ck;
  auto prefetch = [&](int g) {
    const int tc = min((2 * g + bh) * 64 + lane, nts - 1);
    nx_seed = sdp[tc];
    nx_ck = ckp[tc];
    nx_cs = p.cosv[tc];
  };
  // Production of this wave's 16 degrees x 64 colatitudes of a group into
  // buffer buf, in four parts of 4 degrees (one per K chunk of the group
  // multiplied meanwhile, so the scheduler can interleave their VALU work
  // with that chunk's MFMAs).  begin() takes the prefetched state.
  double pe0, pe1, pcs;
  auto begin = [&]() {
    pcs = nx_cs;
    pe0 = kq == 0 ? nx_seed : nx_ck.x;
    pe1 = kq == 0 ? 0.0 : nx_ck.y;
  };
  // Degree lp + d sits in LDS row 32 c + d / 2 of its parity class c
  // ((d + par0) % 2; class 0: l + m even): tiles 0, 1 hold class 0, tiles
  // 2, 3 class 1.
  auto row_of = [&](int d) { return 32 * ((d + par0) & 1) + (d >> 1); };
  auto produce4 = [&](int buf, int part) {
    const int eq = lp + 16 * q + 1;  // entry of step 0, wave-uniform
#pragma unroll
    for (int k = 4 * part; k < 4 * part + 4; k += 2) {
      const double v0 = pe0;
      rec_step(rec_entry(rq, eq + k), pcs, pe0, pe1);
      const double v1 = pe0;
      rec_step(rec_entry(rq, eq + k + 1), pcs, pe0, pe1);
      unsigned o[NS];
      split_pair<NS>(v0, v1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const bf16x2 pr = __builtin_bit_cast(bf16x2, o[s]);
        Dl[buf][s][row_of(16 * q + k)][bh * 64 + lane] = pr[0];
        Dl[buf][s][row_of(16 * q + k + 1)][bh * 64 + lane] = pr[1];
      }
    }
  };

  // Consumer role.
  // Wave w multiplies all 4 degree tiles with column tile w, so every B
  // operand block is loaded by one wave of the workgroup (the A operands
  // come from LDS).  B comes straight from the FP32 bins FA32: lane (q4, c) =
  // (lane / 16, lane % 16) of K chunk kc takes colatitudes 32 kc + 8 q4 + e
  // (e = 0..7) of column 16 tile + c, i.e. FA32 entries 16 e past its base
  // in quad 8 kc + 2 q4 (each e 64 contiguous bytes per 16 lanes), split
  // into bf16 planes just before the chunk's MFMAs.  A tile past the last
  // column group reads the last one: its outputs are never stored, and a
  // column only meets its own B values, so nothing else needs those columns
  // zero.  Reads of the last block run up to 8 quads past it (the FA buffer
  // is padded for that); colatitudes from ntheta on are zeroed at the split
  // of the last group's chunks.  Callers pass chunks below nkc only.
  struct RawB {
    float x[8];
  };
  const int tile = csg * 8 + wave;
  const float* ft = p.FA32 + ((long)slot * p.n_cg + min(tile, p.n_cg - 1)) * p.nquad * 64 + (lane & 15);
  const float* fb = ft + 2 * (lane >> 4) * 64;
  // Chunk kc's colatitudes t = 32 kc + 8 (lane / 16) + e (X) and their
  // mirrors nt - 1 - t (Y), 16 floats apart per row within the (slot, column
  // group) block.  A mirror row below 0 (t past the north, the last chunk
  // only, at most 31 rows) reads the previous block or the front pad of FA;
  // combine zeroes it.
  auto loadB = [&](int kc, RawB& X, RawB& Y) {
    const float* fm = ft + 16 * (nt - 1 - 32 * kc - 8 * (lane >> 4));
#pragma unroll
    for (int e = 0; e < 8; e++) {
      X.x[e] = fb[(long)kc * 512 + 16 * e];
      Y.x[e] = fm[-16 * e];
    }
  };
  // B operands of a chunk: FA(t) + FA(nt - 1 - t) (sign +1) or
  // FA(t) - FA(nt - 1 - t) (sign -1) for northern t below the equator, FA(t)
  // (+1) and 0 (-1) at the equator, 0 past the north.  Only the chunks of
  // the last group (`last`) hold the equator or rows past the north.
  auto combine = [&](const RawB& X, const RawB& Y, int kc, float sign, RawB& Z, bool last) {
    const int th = 32 * kc + 8 * (lane >> 4);
#pragma unroll
    for (int e = 0; e < 8; e++) {
      if (!last) {
        Z.x[e] = X.x[e] + sign * Y.x[e];
      } else {
        const int t = th + e;
        const float y = 2 * t == nt - 1 ? (sign > 0.f ? 0.0f : X.x[e]) : Y.x[e];
        Z.x[e] = t < nts ? X.x[e] + sign * y : 0.0f;
      }
    }
  };
  auto splitB = [&](const RawB& X, uint4 (&B)[1][NS], int kc, bool last) {
    unsigned pl[NS][4];
    const int th = 32 * kc + 8 * (lane >> 4);  // colatitude of e = 0
#pragma unroll
    for (int e2 = 0; e2 < 4; e2++) {
      float x0 = X.x[2 * e2], x1 = X.x[2 * e2 + 1];
      if (last) {
        x0 = th + 2 * e2 < nt ? x0 : 0.0f;
        x1 = th + 2 * e2 + 1 < nt ? x1 : 0.0f;
      }
      unsigned o[NS];
      split_pair_f<NS>(x0, x1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) pl[s][e2] = o[s];
    }
#pragma unroll
    for (int s = 0; s < NS; s++) B[0][s] = make_uint4(pl[s][0], pl[s][1], pl[s][2], pl[s][3]);
  };

  d4 acc[4][1];
#pragma unroll
  for (int i = 0; i < 4; i++) acc[i][0] = d4{0.0, 0.0, 0.0, 0.0};
  // Each chunk's sums (cp) are flushed to FP64 after the next chunk's MFMAs
  // are issued, one chunk at a time (no pairs), in the order of
  // leg_anal_bf16_ws, so both pass lengths give the same bits.
  f4 cp[4][1];
#pragma unroll
  for (int i = 0; i < 4; i++) cp[i][0] = f4{0.f, 0.f, 0.f, 0.f};
  auto step = [&](int buf, int kc, const RawB& X, const RawB& Y, bool last) {
    f4 c[4][1];
#pragma unroll
    for (int h = 0; h < 2; h++) {  // parity class h: tiles 2 h, 2 h + 1
      RawB Z;
      combine(X, Y, kc, h == 0 ? 1.0f : -1.0f, Z, last);
      uint4 B[1][NS];
      splitB(Z, B, kc, false);
      uint4 A[2][NS];
#pragma unroll
      for (int lt = 0; lt < 2; lt++)
#pragma unroll
        for (int s = 0; s < NS; s++)
          A[lt][s] = *reinterpret_cast<const uint4*>(
              &Dl[buf][s][(2 * h + lt) * 16 + (lane & 15)][(kc & 3) * 32 + 8 * (lane >> 4)]);
      f4 ch[2][1];
      split_mma<NS, 2, 1>(A, B, ch);
      c[2 * h][0] = ch[0][0];
      c[2 * h + 1][0] = ch[1][0];
    }
#pragma unroll
    for (int i = 0; i < 4; i++) {
      flush(acc[i][0], cp[i][0]);
      cp[i][0] = c[i][0];
    }
  };

  prefetch(0);
  begin();
#pragma unroll
  for (int part = 0; part < 4; part++) produce4(0, part);
  prefetch(1);
  __syncthreads();
  // B operands: chunk c sits in buffer X[c & 3], loaded kAhead chunks before
  // its MFMAs (into the buffer chunk c - 4 + kAhead used).  Loads clamp to
  // the last chunk nkc - 1; what they bring past it is never multiplied.
  // Every group but the last has four chunks below nkc (nt > 128 (ngrp-1)),
  // and the last group's first chunk is below nkc too; only the last group's
  // chunks hold colatitudes past ntheta.  The scheduling barriers keep each
  // chunk's loads, production part and MFMAs together.
  constexpr int kAhead = 3;
  const int kcl = nkc - 1;
  RawB X[4], Y[4];
#pragma unroll
  for (int j = 0; j < kAhead; j++) loadB(min(j, kcl), X[j], Y[j]);
  for (int g = 0; g + 1 < ngrp; g++) {
    const int buf = g & 1;
    const int kc0 = 4 * g;
    begin();
    prefetch(g + 2);  // a whole group ahead of its use
#pragma unroll
    for (int j = 0; j < 4; j++) {
      loadB(min(kc0 + j + kAhead, kcl), X[(j + kAhead) & 3], Y[(j + kAhead) & 3]);
      produce4(buf ^ 1, j);
      step(buf, kc0 + j, X[j], Y[j], false);
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
  }
  {
    // Last group: chunks 4 (ngrp-1) .. nkc-1.
    const int buf = (ngrp - 1) & 1;
    const int kc0 = 4 * (ngrp - 1);
    step(buf, kc0, X[0], Y[0], true);
    if (kc0 + 1 < nkc) {
      step(buf, kc0 + 1, X[1], Y[1], true);
      if (kc0 + 2 < nkc) {
        step(buf, kc0 + 2, X[2], Y[2], true);
        if (kc0 + 3 < nkc) {
          loadB(kc0 + 3, X[3], Y[3]);
          step(buf, kc0 + 3, X[3], Y[3], true);
        }
      }
    }
#pragma unroll
    for (int i = 0; i < 4; i++) flush(acc[i][0], cp[i][0]);
  }

  // Row 16 lt + 4 (lane / 16) + r of class lt / 2 is degree lp + d, d = 2 i
  // + (lt / 2 + par0) % 2 with i its index in the class; column 16 tile +
  // lane % 16.
  const int col = tile * 16 + (lane & 15);
  const bool odd = lane & 1;
  const bool cv = col < p.ncol;  // the same for both lanes of a field (ncol is even)
  const int lf = cv ? field_lmax(p, col >> 1) : -1;
  const long cbase = cv ? (long)field_index(p.fmap, col >> 1) * p.cstride : 0;
#pragma unroll
  for (int lt = 0; lt < 4; lt++)
#pragma unroll
    for (int rp = 0; rp < 2; rp++) {
      const double2 v = trade_halves(acc[lt][0][2 * rp], acc[lt][0][2 * rp + 1], odd);
      const int i = 16 * (lt & 1) + 4 * (lane >> 4) + 2 * rp + odd;
      const int l = lp + 2 * i + (((lt >> 1) + par0) & 1);
      if (l <= lf) store_coef(p, cbase, l, m, lf, REAL, v);
    }
}

// ---------------------------------------------------------------------------
// Analysis over 128-degree passes (kWidePass): the same contraction as
// leg_anal_bf16 with eight degree tiles per wave, so each split B operand
// (the VALU-heavy part of a K chunk) feeds twice as many MFMAs.  Colatitudes
// are walked in groups of 64 (two 32-colatitude K chunks) so that the bf16
// planes of 128 degrees still fit the LDS double buffer.  Wave w produces
// the 16 degrees lp + 16 w .. + 15 of each group (one colatitude per lane)
// from checkpoint 8 pass + w, and multiplies all 8 degree tiles with column
// tile w in two halves of 4.  Chunk sums are flushed to FP64 in pairs, as
// in leg_anal_bf16 and in the same order, so the results are the same bits.
// ---------------------------------------------------------------------------
constexpr int kWidePass = kWideAnalPass;

template <int NS, int REAL>
__global__ __launch_bounds__(512) void leg_anal_bf16_w(LegParams p) {
  static_assert(kCk == 16 && kWidePass == 8 * kCk, "8 degree tiles of 16 per pass");
  // Rows of 64 colatitudes padded to 72 (144 bytes, 36 banks apart): the 16
  // rows of a 16-byte A-operand read start at 16 different multiples of 4
  // banks and cover all 64 banks once.
  __shared__ __bf16 Dl[2][NS][kWidePass][72];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int mw = it.x;
  const int pass = it.y & 0xffff;
  const int csg = it.y >> 16;
  const int m = order_of(mw, REAL);
  const int slot = m - p.m_first;
  const int l0 = max(abs(m), p.nabs);
  const int nt = p.ntheta;
  const int ngrp = (nt + 63) >> 6;
  const int lp = l0 + kWidePass * pass;  // first degree of the pass
  // Recurrence coefficients of the pass: entries lp+1 .. lp+128 (entry l for
  // degree l).  Entries up to L + kRecPad exist: kRecPad covers the 128
  // entries only from a pass start at most L - 63, so the last pass's tail
  // reads clamp.
  rec_t* rq = rec_ptr(p.rec, ld_uniform(p.roff, m + p.L) - l0);

  // Producer share: degrees lp + 16 w .. + 15 of every colatitude of a group,
  // from checkpoint kq (the seed for kq = 0; past L clamped to the last).
  const int nck = (p.L - l0 + kCk) / kCk;  // checkpoints 1 .. nck-1 exist
  const int kq = min((kWidePass / kCk) * pass + wave, nck - 1);
  const double2* ckp = kq == 0 ? p.ck : p.ck + (ld_uniform(p.ckoff, m + p.L) + kq - 1) * (long)nt;
  const double* sdp = p.seeds + (long)(m + p.L) * nt;
  double nx_seed, nx_cs;
  double2 nx_ck;
  auto prefetch = [&](int g) {
    const int tc = min(64 * g + lane, nt - 1);
    nx_seed = sdp[tc];
    nx_ck = ckp[tc];
    nx_cs = p.cosv[tc];
  };
  double pe0, pe1, pcs;
  auto begin = [&]() {
    pcs = nx_cs;
    pe0 = kq == 0 ? nx_seed : nx_ck.x;
    pe1 = kq == 0 ? 0.0 : nx_ck.y;
  };
  // Half `part` (8 degrees) of this wave's production for a group.
  auto produce8 = [&](int buf, int part) {
    const int ew = lp + 16 * wave + 1;  // entry of step 0, wave-uniform
#pragma unroll
    for (int k = 8 * part; k < 8 * part + 8; k += 2) {
      const double v0 = pe0;
      rec_step(rec_entry(rq, min(ew + k, p.L + kRecPad)), pcs, pe0, pe1);
      const double v1 = pe0;
      rec_step(rec_entry(rq, min(ew + k + 1, p.L + kRecPad)), pcs, pe0, pe1);
      unsigned o[NS];
      split_pair<NS>(v0, v1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const bf16x2 pr = __builtin_bit_cast(bf16x2, o[s]);
        Dl[buf][s][16 * wave + k][lane] = pr[0];
        Dl[buf][s][16 * wave + k + 1][lane] = pr[1];
      }
    }
  };

  // Consumer: column tile w, B straight from the FP32 bins as in
  // leg_anal_bf16 (colatitudes from ntheta on zeroed at the split).
  struct RawB {
    float x[8];
  };
  const int tile = csg * 8 + wave;
  const float* fb =
      p.FA32 + (((long)slot * p.n_cg + min(tile, p.n_cg - 1)) * p.nquad + 2 * (lane >> 4)) * 64 +
      (lane & 15);
  auto loadB = [&](int kc, RawB& X) {
#pragma unroll
    for (int e = 0; e < 8; e++) X.x[e] = fb[(long)kc * 512 + 16 * e];
  };
  auto splitB = [&](const RawB& X, uint4 (&B)[1][NS], int kc, bool last) {
    unsigned pl[NS][4];
    const int th = 32 * kc + 8 * (lane >> 4);
#pragma unroll
    for (int e2 = 0; e2 < 4; e2++) {
      float x0 = X.x[2 * e2], x1 = X.x[2 * e2 + 1];
      if (last) {
        x0 = th + 2 * e2 < nt ? x0 : 0.0f;
        x1 = th + 2 * e2 + 1 < nt ? x1 : 0.0f;
      }
      unsigned o[NS];
      split_pair_f<NS>(x0, x1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) pl[s][e2] = o[s];
    }
#pragma unroll
    for (int s = 0; s < NS; s++) B[0][s] = make_uint4(pl[s][0], pl[s][1], pl[s][2], pl[s][3]);
  };

  d4 acc[8][1];
#pragma unroll
  for (int i = 0; i < 8; i++) acc[i][0] = d4{0.0, 0.0, 0.0, 0.0};
  // Pairs of chunk sums: an odd chunk keeps its predecessor's sums in pp,
  // an even one adds the pair in FP32 and flushes it to FP64.
  f4 cp[8], pp[8];
#pragma unroll
  for (int i = 0; i < 8; i++) cp[i] = pp[i] = f4{0.f, 0.f, 0.f, 0.f};
  auto step = [&](int buf, int kc, const RawB& X, bool last, bool odd) {
    uint4 B[1][NS];
    splitB(X, B, kc, last);
#pragma unroll
    for (int half = 0; half < 2; half++) {
      uint4 A[4][NS];
#pragma unroll
      for (int lt = 0; lt < 4; lt++)
#pragma unroll
        for (int s = 0; s < NS; s++)
          A[lt][s] = *reinterpret_cast<const uint4*>(
              &Dl[buf][s][(4 * half + lt) * 16 + (lane & 15)][(kc & 1) * 32 + 8 * (lane >> 4)]);
      f4 c[4][1];
      split_mma<NS, 4, 1>(A, B, c);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int t = 4 * half + i;
        if (odd)
          pp[t] = cp[t];
        else
          flush(acc[t][0], pp[t] + cp[t]);
        cp[t] = c[i][0];
      }
    }
  };

  prefetch(0);
  begin();
  produce8(0, 0);
  produce8(0, 1);
  prefetch(1);
  __syncthreads();
  // B operands one chunk ahead: chunk c in X[c & 1].
  const int kcl = p.nkc - 1;
  RawB X[2];
  loadB(0, X[0]);
  for (int g = 0; g + 1 < ngrp; g++) {
    const int buf = g & 1;
    const int kc0 = 2 * g;
    begin();
    prefetch(g + 2);
#pragma unroll
    for (int j = 0; j < 2; j++) {
      loadB(min(kc0 + j + 1, kcl), X[(j + 1) & 1]);
      produce8(buf ^ 1, j);
      step(buf, kc0 + j, X[j], false, j & 1);
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
  }
  {
    // Last group: chunks 2 (ngrp-1) .. nkc-1 (one or two).
    const int buf = (ngrp - 1) & 1;
    const int kc0 = 2 * (ngrp - 1);
    if (kc0 + 1 < p.nkc) loadB(kc0 + 1, X[1]);
    step(buf, kc0, X[0], true, false);
    if (kc0 + 1 < p.nkc) step(buf, kc0 + 1, X[1], true, true);
    const bool lodd = (p.nkc - 1) & 1;
#pragma unroll
    for (int i = 0; i < 8; i++) flush(acc[i][0], lodd ? pp[i] + cp[i] : cp[i]);
  }

  const int col = tile * 16 + (lane & 15);
  const bool odd = lane & 1;
  const bool cv = col < p.ncol;  // the same for both lanes of a field (ncol is even)
  const int lf = cv ? field_lmax(p, col >> 1) : -1;
  const long cbase = cv ? (long)field_index(p.fmap, col >> 1) * p.cstride : 0;
#pragma unroll
  for (int lt = 0; lt < 8; lt++)
#pragma unroll
    for (int rp = 0; rp < 2; rp++) {
      const double2 v = trade_halves(acc[lt][0][2 * rp], acc[lt][0][2 * rp + 1], odd);
      const int l = lp + 16 * lt + 4 * (lane >> 4) + 2 * rp + odd;
      if (l <= lf) store_coef(p, cbase, l, m, lf, REAL, v);
    }
}

// ---------------------------------------------------------------------------
// Analysis for upper index 0 over the northern colatitudes, 128-degree
// passes.  Gauss-Legendre nodes are symmetric about the equator and
// d^l_{m0}(pi - theta) = (-1)^(l+m) d^l_{m0}(theta), so sum_t d_l(theta_t)
// FA(t) runs over the northern t (the equator once) against FA(t) +
// FA(nt-1-t) for even l + m and FA(t) - FA(nt-1-t) for odd l + m: half the
// colatitudes, recurrence values, A-operand reads and MFMAs of
// leg_anal_bf16_w, whose items, production and chunk order this kernel
// keeps.  The degrees of a pass sit in LDS by parity class (tiles 0-3:
// l + m even, 4-7: odd), each class multiplied with its own B operand; the
// sums FA(t) +- FA(nt-1-t) of the FP32 bins round once to FP32 before the
// split, and each chunk's sums are flushed to FP64 on their own (no pairs).
// ---------------------------------------------------------------------------
template <int NS, int REAL>
__global__ __launch_bounds__(512) void leg_anal_bf16_ws(LegParams p) {
  static_assert(kCk == 16 && kWidePass == 8 * kCk, "8 degree tiles of 16 per pass");
  // Rows of 64 colatitudes padded to 72 (144 bytes, 36 banks apart): the 16
  // rows of a 16-byte A-operand read start at 16 different multiples of 4
  // banks and cover all 64 banks once.
  __shared__ __bf16 Dl[2][NS][kWidePass][72];
  const int lane = threadIdx.x & 63;
  const int wave = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);  // wave-uniform
  const uint2 it = work_item(p.items, p.items_per);
  if (it.x == ~0u) return;
  const int mw = it.x;
  const int pass = it.y & 0xffff;
  const int csg = it.y >> 16;
  const int m = order_of(mw, REAL);
  const int slot = m - p.m_first;
  const int l0 = max(abs(m), p.nabs);
  const int nt = p.ntheta;
  const int nts = (nt + 1) >> 1;  // northern colatitudes, the equator included
  const int ngrp = (nts + 63) >> 6;
  const int nkc = (nts + 31) >> 5;
  const int lp = l0 + kWidePass * pass;  // first degree of the pass
  const int par0 = (lp + m) & 1;         // parity of l + m at the pass's first degree
  // Recurrence coefficients of the pass: entries lp+1 .. lp+128 (entry l for
  // degree l).  Entries up to L + kRecPad exist: kRecPad covers the 128
  // entries only from a pass start at most L - 63, so the last pass's tail
  // reads clamp.
  rec_t* rq = rec_ptr(p.rec, ld_uniform(p.roff, m + p.L) - l0);

  // Producer share: degrees lp + 16 w .. + 15 of every colatitude of a group,
  // from checkpoint kq (the seed for kq = 0; past L clamped to the last).
  const int nck = (p.L - l0 + kCk) / kCk;  // checkpoints 1 .. nck-1 exist
  const int kq = min((kWidePass / kCk) * pass + wave, nck - 1);
  const double2* ckp = kq == 0 ? p.ck : p.ck + (ld_uniform(p.ckoff, m + p.L) + kq - 1) * (long)nt;
  const double* sdp = p.seeds + (long)(m + p.L) * nt;
  double nx_seed, nx_cs;
  double2 nx_ck;
  auto prefetch = [&](int g) {
    const int tc = min(64 * g + lane, nts - 1);
    nx_seed = sdp[tc];
    nx_ck = ckp[tc];
    nx_cs = p.cosv[tc];
  };
  double pe0, pe1, pcs;
  auto begin = [&]() {
    pcs = nx_cs;
    pe0 = kq == 0 ? nx_seed : nx_ck.x;
    pe1 = kq == 0 ? 0.0 : nx_ck.y;
  };
  // Degree lp + d sits in LDS row 64 c + d / 2 of its parity class c
  // ((d + par0) % 2; class 0: l + m even): tiles 0-3 hold class 0, tiles
  // 4-7 class 1.
  auto row_of = [&](int d) { return 64 * ((d + par0) & 1) + (d >> 1); };
  // Half `part` (8 degrees) of this wave's production for a group.
  auto produce8 = [&](int buf, int part) {
    const int ew = lp + 16 * wave + 1;  // entry of step 0, wave-uniform
#pragma unroll
    for (int k = 8 * part; k < 8 * part + 8; k += 2) {
      const double v0 = pe0;
      rec_step(rec_entry(rq, min(ew + k, p.L + kRecPad)), pcs, pe0, pe1);
      const double v1 = pe0;
      rec_step(rec_entry(rq, min(ew + k + 1, p.L + kRecPad)), pcs, pe0, pe1);
      unsigned o[NS];
      split_pair<NS>(v0, v1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) {
        const bf16x2 pr = __builtin_bit_cast(bf16x2, o[s]);
        Dl[buf][s][row_of(16 * wave + k)][lane] = pr[0];
        Dl[buf][s][row_of(16 * wave + k + 1)][lane] = pr[1];
      }
    }
  };

  // Consumer: column tile w, B straight from the FP32 bins as in
  // leg_anal_bf16 (colatitudes from ntheta on zeroed at the split).
  struct RawB {
    float x[8];
  };
  const int tile = csg * 8 + wave;
  const float* ft = p.FA32 + ((long)slot * p.n_cg + min(tile, p.n_cg - 1)) * p.nquad * 64 + (lane & 15);
  const float* fb = ft + 2 * (lane >> 4) * 64;
  // Chunk kc's colatitudes t = 32 kc + 8 (lane / 16) + e (X) and their
  // mirrors nt - 1 - t (Y), 16 floats apart per row within the (slot, column
  // group) block.  A mirror row below 0 (t past the north, the last chunk
  // only, at most 31 rows) reads the previous block or the front pad of FA;
  // combine zeroes it.
  auto loadB = [&](int kc, RawB& X, RawB& Y) {
    const float* fm = ft + 16 * (nt - 1 - 32 * kc - 8 * (lane >> 4));
#pragma unroll
    for (int e = 0; e < 8; e++) {
      X.x[e] = fb[(long)kc * 512 + 16 * e];
      Y.x[e] = fm[-16 * e];
    }
  };
  // B operands of a chunk: FA(t) + FA(nt - 1 - t) (sign +1) or
  // FA(t) - FA(nt - 1 - t) (sign -1) for northern t below the equator, FA(t)
  // (+1) and 0 (-1) at the equator, 0 past the north.  Only the chunks of
  // the last group (`last`) hold the equator or rows past the north.
  auto combine = [&](const RawB& X, const RawB& Y, int kc, float sign, RawB& Z, bool last) {
    const int th = 32 * kc + 8 * (lane >> 4);
#pragma unroll
    for (int e = 0; e < 8; e++) {
      if (!last) {
        Z.x[e] = X.x[e] + sign * Y.x[e];
      } else {
        const int t = th + e;
        const float y = 2 * t == nt - 1 ? (sign > 0.f ? 0.0f : X.x[e]) : Y.x[e];
        Z.x[e] = t < nts ? X.x[e] + sign * y : 0.0f;
      }
    }
  };
  auto splitB = [&](const RawB& X, uint4 (&B)[1][NS], int kc, bool last) {
    unsigned pl[NS][4];
    const int th = 32 * kc + 8 * (lane >> 4);
#pragma unroll
    for (int e2 = 0; e2 < 4; e2++) {
      float x0 = X.x[2 * e2], x1 = X.x[2 * e2 + 1];
      if (last) {
        x0 = th + 2 * e2 < nt ? x0 : 0.0f;
        x1 = th + 2 * e2 + 1 < nt ? x1 : 0.0f;
      }
      unsigned o[NS];
      split_pair_f<NS>(x0, x1, o);
#pragma unroll
      for (int s = 0; s < NS; s++) pl[s][e2] = o[s];
    }
#pragma unroll
    for (int s = 0; s < NS; s++) B[0][s] = make_uint4(pl[s][0], pl[s][1], pl[s][2], pl[s][3]);
  };

  d4 acc[8][1];
#pragma unroll
  for (int i = 0; i < 8; i++) acc[i][0] = d4{0.0, 0.0, 0.0, 0.0};
  // Each chunk's sums (cp) are flushed to FP64 after the next chunk's MFMAs
  // are issued (no pairs: the registers of a second set hold the mirror
  // rows' bins instead).
  f4 cp[8];
#pragma unroll
  for (int i = 0; i < 8; i++) cp[i] = f4{0.f, 0.f, 0.f, 0.f};
  auto step = [&](int buf, int kc, const RawB& X, const RawB& Y, bool last) {
#pragma unroll
    for (int half = 0; half < 2; half++) {  // parity class half: tiles 4 half .. + 3
      uint4 B[1][NS];
      {
        RawB Z;
        combine(X, Y, kc, half == 0 ? 1.0f : -1.0f, Z, last);
        splitB(Z, B, kc, false);
      }
      uint4 A[4][NS];
#pragma unroll
      for (int lt = 0; lt < 4; lt++)
#pragma unroll
        for (int s = 0; s < NS; s++)
          A[lt][s] = *reinterpret_cast<const uint4*>(
              &Dl[buf][s][(4 * half + lt) * 16 + (lane & 15)][(kc & 1) * 32 + 8 * (lane >> 4)]);
      f4 c[4][1];
      split_mma<NS, 4, 1>(A, B, c);
#pragma unroll
      for (int i = 0; i < 4; i++) {
        const int t = 4 * half + i;
        flush(acc[t][0], cp[t]);
        cp[t] = c[i][0];
      }
    }
  };

  prefetch(0);
  begin();
  produce8(0, 0);
  produce8(0, 1);
  prefetch(1);
  __syncthreads();
  // B operands one chunk ahead: chunk c in X[c & 1], its mirrors in Y[c & 1].
  const int kcl = nkc - 1;
  RawB X[2], Y[2];
  loadB(0, X[0], Y[0]);
  for (int g = 0; g + 1 < ngrp; g++) {
    const int buf = g & 1;
    const int kc0 = 2 * g;
    begin();
    prefetch(g + 2);
#pragma unroll
    for (int j = 0; j < 2; j++) {
      loadB(min(kc0 + j + 1, kcl), X[(j + 1) & 1], Y[(j + 1) & 1]);
      produce8(buf ^ 1, j);
      step(buf, kc0 + j, X[j], Y[j], false);
      __builtin_amdgcn_sched_barrier(0);
    }
    __syncthreads();
  }
  {
    // Last group: chunks 2 (ngrp-1) .. nkc-1 (one or two).
    const int buf = (ngrp - 1) & 1;
    const int kc0 = 2 * (ngrp - 1);
    if (kc0 + 1 < nkc) loadB(kc0 + 1, X[1], Y[1]);
    step(buf, kc0, X[0], Y[0], true);
    if (kc0 + 1 < nkc) step(buf, kc0 + 1, X[1], Y[1], true);
#pragma unroll
    for (int i = 0; i < 8; i++) flush(acc[i][0], cp[i]);
  }

  const int col = tile * 16 + (lane & 15);
  const bool odd = lane & 1;
  const bool cv = col < p.ncol;  // the same for both lanes of a field (ncol is even)
  const int lf = cv ? field_lmax(p, col >> 1) : -1;
  const long cbase = cv ? (long)field_index(p.fmap, col >> 1) * p.cstride : 0;
  // Row 16 lt + 4 (lane / 16) + r of class lt / 4 is degree lp + d, d =
  // 2 i + (lt / 4 + par0) % 2 with i its index in the class.
#pragma unroll
  for (int lt = 0; lt < 8; lt++)
#pragma unroll
    for (int rp = 0; rp < 2; rp++) {
      const double2 v = trade_halves(acc[lt][0][2 * rp], acc[lt][0][2 * rp + 1], odd);
      const int i = 16 * (lt & 3) + 4 * (lane >> 4) + 2 * rp + odd;
      const int l = lp + 2 * i + (((lt >> 2) + par0) & 1);
      if (l <= lf) store_coef(p, cbase, l, m, lf, REAL, v);
    }
}


// ---------------------------------------------------------------------------
// Synthesis of the last colatitudes t0 .. ntheta - 1 (t0 = ntheta -
// tail_rows) when ntheta is only a row or two past a multiple of 256.  A
// leg_synth_bf16 item holding one row would still walk every degree chunk,
// bound by the latency of its B operand loads while it occupies a whole CU.
// Instead coef_split, which holds every FP64 coefficient in LDS anyway, adds
// each block's partial sums for these rows, and tail_reduce sums them.
//
// tail_dvals: the FP64 values d_l(theta_t) of row t0 + blockIdx.y for every
// order and degree of the split layout.  Lane s of the wave of a slot runs
// the 16-degree segment s from checkpoint s (the seed for s = 0; segments
// that start past L meet only zero coefficients and are written as zeros).
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(64) void tail_dvals(LegParams p, double* dv) {
  const int slot = blockIdx.x;
  const int t = p.tail_t0 + blockIdx.y;
  // Paired orders (p.pm): -m takes the values of order |m| at row t, which
  // multiply its sign-folded coefficients into its sum at row ntheta-1-t.
  const int m = p.pm ? abs(slot + p.m_first) : slot + p.m_first;
  const int l0 = max(abs(m), p.nabs);
  const int cw = p.sym ? 64 : 32;  // degrees per chunk of the split layout
  const int nchunk = (p.lmax - l0 + cw) / cw;
  const int nck = (p.L - l0 + kCk) / kCk;  // checkpoints 1 .. nck-1 exist
  const double cs = p.cosv[t];
  const d4* rl = reinterpret_cast<const d4*>(p.rec + (ld_uniform(p.roff, m + p.L) - l0));
  const double2* ckm = p.ck + ld_uniform(p.ckoff, m + p.L) * (long)p.ntheta + t;
  double* out = dv + blockIdx.y * p.tail_dstride + ld_uniform(p.cso, slot) * 32;
  // Entry of degree l0 + d: d itself, or in the parity layout block
  // 2 (d / 64) + k (k = (d + l0 + m) % 2), entry (d % 64) / 2.
  auto entry = [&](int d) {
    return p.sym ? 32 * (2 * (d >> 6) + ((d + l0 + m) & 1)) + ((d & 63) >> 1) : d;
  };
  for (int sg = threadIdx.x; sg < nchunk * cw / kCk; sg += 64) {
    double e0 = 0.0, e1 = 0.0;
    if (sg == 0) {
      e0 = p.seeds[(long)(m + p.L) * p.ntheta + t];
    } else if (sg < nck) {
      const double2 st = ckm[(long)(sg - 1) * p.ntheta];
      e0 = st.x;
      e1 = st.y;
    }
    const int ls = l0 + kCk * sg;  // entries up to ls + 15 <= lmax + 31 <= L + kRecPad
#pragma unroll
    for (int j = 0; j < kCk; j++) {
      out[entry(kCk * sg + j)] = e0;
      if (j + 1 < kCk) rec_step(rl[ls + j + 1], cs, e0, e1);
    }
  }
}

// tail_reduce: row t0 + blockIdx.y of order slot blockIdx.x, column by
// thread, summed over the chunks in order.
__global__ __launch_bounds__(256) void tail_reduce(LegParams p) {
  const int slot = blockIdx.x;
  const int t = p.tail_t0 + blockIdx.y;
  const int m = slot + p.m_first;
  const int l0 = max(abs(m), p.nabs);
  // Split blocks of the order: one per 32-degree chunk, or two per 64-degree
  // chunk in the parity layout.
  const int nchunk = p.sym ? 2 * ((p.lmax - l0 + 64) >> 6) : (p.lmax - l0 + 32) >> 5;
  const long cstride = (long)p.ntile * 16;  // doubles per block
  const double* pp = p.tail_part + blockIdx.y * p.tail_pstride + ld_uniform(p.cso, slot) * cstride;
  const int ncg16 = (p.ncol + 15) >> 4;
  for (int col = threadIdx.x; col < p.ncol; col += 256) {
    double sum = 0.0;
    for (int c = 0; c < nchunk; c++) sum += pp[c * cstride + col];
    // Paired orders: the sums of orders m < 0 belong to row ntheta-1-t.
    const long fi = f_index(p.pm && m < 0 ? p.ntheta - 1 - t : t, slot, col, p.nslot, ncg16);
    if (p.f32out)
      reinterpret_cast<float*>(p.F)[fi] = (float)sum;
    else
      p.F[fi] = sum;
  }
}

}  // namespace

template <int SYM>
hipError_t launch_coef_split_t(const LegParams& p, int ns, hipStream_t s) {
  const dim3 grid(8 * p.items_per);  // one item (8 orders) per workgroup
  if (ns == 3 && p.real)
    hipLaunchKernelGGL((coef_split<3, 1, SYM>), grid, dim3(512), 0, s, p);
  else if (ns == 3)
    hipLaunchKernelGGL((coef_split<3, 0, SYM>), grid, dim3(512), 0, s, p);
  else if (p.real)
    hipLaunchKernelGGL((coef_split<2, 1, SYM>), grid, dim3(512), 0, s, p);
  else
    hipLaunchKernelGGL((coef_split<2, 0, SYM>), grid, dim3(512), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_coef_split(const LegParams& p, int ns, hipStream_t s) {
  return p.sym ? launch_coef_split_t<1>(p, ns, s) : launch_coef_split_t<0>(p, ns, s);
}

hipError_t launch_leg_synth_bf16(const LegParams& p0, int ns, hipStream_t s) {
  if (!p0.ck || !p0.ckoff) return hipErrorInvalidValue;  // checkpoint table required
  LegParams p = p0;
  p.n_tsb = (p.ntheta + 255) / 256;
  p.n_cg = (p.ntile + 7) / 8;  // column super-groups of 8 tiles
  const dim3 grid(8 * p.items_per);  // work list: (order, column super-group, theta block)
  if (p.sym) {
    if (p.nabs != 0) return hipErrorInvalidValue;  // symmetry of n = 0 only
    if (ns == 3 && p.real)
      hipLaunchKernelGGL((leg_synth_bf16_s<3, 1>), grid, dim3(512), 0, s, p);
    else if (ns == 3)
      hipLaunchKernelGGL((leg_synth_bf16_s<3, 0>), grid, dim3(512), 0, s, p);
    else if (p.real)
      hipLaunchKernelGGL((leg_synth_bf16_s<2, 1>), grid, dim3(512), 0, s, p);
    else
      hipLaunchKernelGGL((leg_synth_bf16_s<2, 0>), grid, dim3(512), 0, s, p);
    return hipGetLastError();
  }
  if (p.pm) {  // orders m and -m together: complex fields
    if (p.real) return hipErrorInvalidValue;
    if (ns == 3)
      hipLaunchKernelGGL((leg_synth_bf16_pm<3>), grid, dim3(512), 0, s, p);
    else
      hipLaunchKernelGGL((leg_synth_bf16_pm<2>), grid, dim3(512), 0, s, p);
    return hipGetLastError();
  }
  if (ns == 3 && p.real)
    hipLaunchKernelGGL((leg_synth_bf16<3, 1>), grid, dim3(512), 0, s, p);
  else if (ns == 3)
    hipLaunchKernelGGL((leg_synth_bf16<3, 0>), grid, dim3(512), 0, s, p);
  else if (p.real)
    hipLaunchKernelGGL((leg_synth_bf16<2, 1>), grid, dim3(512), 0, s, p);
  else
    hipLaunchKernelGGL((leg_synth_bf16<2, 0>), grid, dim3(512), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_tail_dvals(const LegParams& p, hipStream_t s) {
  if (p.tail_rows < 1 || p.tail_rows > kSynthTailRows || !p.tail_dv || !p.ck || !p.ckoff)
    return hipErrorInvalidValue;
  hipLaunchKernelGGL(tail_dvals, dim3(p.nslot, p.tail_rows), dim3(64), 0, s, p,
                     const_cast<double*>(p.tail_dv));
  return hipGetLastError();
}

hipError_t launch_tail_reduce(const LegParams& p, hipStream_t s) {
  if (p.tail_rows < 1 || p.tail_rows > kSynthTailRows || !p.tail_part) return hipErrorInvalidValue;
  hipLaunchKernelGGL(tail_reduce, dim3(p.nslot, p.tail_rows), dim3(256), 0, s, p);
  return hipGetLastError();
}

template <int NS>
hipError_t launch_anal_bf16_t(const LegParams& p0, hipStream_t s, bool sym) {
  LegParams p = p0;
  p.n_cg = (p.ncol + 15) / 16;  // the layout FA
  p.nquad = (p.ntheta + 3) / 4;
  p.nkc = (p.ntheta + 31) / 32;
  const dim3 grid(8 * p.items_per);  // work list: (order, pass, column super-group)
  if (sym && p.real)
    hipLaunchKernelGGL((leg_anal_bf16_s<NS, 1>), grid, dim3(512), 0, s, p);
  else if (sym)
    hipLaunchKernelGGL((leg_anal_bf16_s<NS, 0>), grid, dim3(512), 0, s, p);
  else if (p.real)
    hipLaunchKernelGGL((leg_anal_bf16<NS, 1>), grid, dim3(512), 0, s, p);
  else
    hipLaunchKernelGGL((leg_anal_bf16<NS, 0>), grid, dim3(512), 0, s, p);
  return hipGetLastError();
}

template <int NS>
hipError_t launch_anal_bf16_w(const LegParams& p0, hipStream_t s, bool sym) {
  LegParams p = p0;
  p.n_cg = (p.ncol + 15) / 16;
  p.nquad = (p.ntheta + 3) / 4;
  p.nkc = (p.ntheta + 31) / 32;
  const dim3 grid(8 * p.items_per);  // work list: (order, 128-degree pass, column super-group)
  if (sym && p.real)
    hipLaunchKernelGGL((leg_anal_bf16_ws<NS, 1>), grid, dim3(512), 0, s, p);
  else if (sym)
    hipLaunchKernelGGL((leg_anal_bf16_ws<NS, 0>), grid, dim3(512), 0, s, p);
  else if (p.real)
    hipLaunchKernelGGL((leg_anal_bf16_w<NS, 1>), grid, dim3(512), 0, s, p);
  else
    hipLaunchKernelGGL((leg_anal_bf16_w<NS, 0>), grid, dim3(512), 0, s, p);
  return hipGetLastError();
}

hipError_t launch_leg_anal_bf16(const LegParams& p, int ns, hipStream_t s, int form) {
  if (!p.ck || !p.ckoff) return hipErrorInvalidValue;  // checkpoint table required
  if (!p.FA32) return hipErrorInvalidValue;             // FP32 bins required
  if ((form == 2 || form == 3) && p.nabs != 0) return hipErrorInvalidValue;  // n = 0 only
  if (form == 4) {  // orders m and -m together: complex fields
    if (p.real) return hipErrorInvalidValue;
    LegParams q = p;
    q.n_cg = (q.ncol + 15) / 16;
    q.nquad = (q.ntheta + 3) / 4;
    q.nkc = (q.ntheta + 31) / 32;
    const dim3 grid(8 * q.items_per);  // work list: (order m >= 0, pass, column super-group)
    if (ns == 3)
      hipLaunchKernelGGL((leg_anal_bf16_pm<3>), grid, dim3(512), 0, s, q);
    else
      hipLaunchKernelGGL((leg_anal_bf16_pm<2>), grid, dim3(512), 0, s, q);
    return hipGetLastError();
  }
  if (form == 1 || form == 3)
    return ns == 3 ? launch_anal_bf16_w<3>(p, s, form == 3) : launch_anal_bf16_w<2>(p, s, form == 3);
  return ns == 3 ? launch_anal_bf16_t<3>(p, s, form == 2) : launch_anal_bf16_t<2>(p, s, form == 2);
}

}  // namespace gsh
