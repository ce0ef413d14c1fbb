// gsh_capi.cpp -- host layer of the MI355X GSH transform library: the C ABI
// declared in include/gshtrans_amd.h, grid set-up (Gauss-Legendre nodes,
// per-upper-index recurrence data) and launch orchestration.
//
// Reference counterparts (paths relative to the reference tree):
//   GaussLegendreGrid ctor            GSHTrans/src/GaussLegendreGrid.h:38-73
//   Wigner seeds (closed forms)       GSHTrans/src/Wigner.h:24-95, 329-344
//   Wigner recurrence coefficients    Wigner.h:347-472 (restated per order m)
//   Forward/InverseTransformation     GaussLegendreGrid.h:108-321
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <complex>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <limits>
#include <map>
#include <memory>
#include <mutex>
#include <string>
#include <thread>
#include <tuple>
#include <utility>
#include <vector>

#include "../../include/gshtrans_amd.h"
#include "gsh_internal.h"

namespace {

using gsh::FFTParams;
using gsh::LegParams;

constexpr long double kPiL = 3.141592653589793238462643383279502884L;

// ---------------------------------------------------------------------------
// Host math
// ---------------------------------------------------------------------------

// Gauss-Legendre nodes x ascending on [-1, 1] and weights, by Newton's method
// on P_{npts} in long double (the quadrature GaussLegendreGrid.h:47 obtains
// from GaussQuad), so theta = acos(-x) ascends (GaussLegendreGrid.h:48).
void gauss_legendre(int npts, std::vector<long double>& x, std::vector<long double>& w) {
  x.assign(npts, 0.0L);
  w.assign(npts, 0.0L);
  auto legendre = [npts](long double z, long double& p1, long double& p2) {
    p1 = 1;
    p2 = 0;
    for (int j = 1; j <= npts; j++) {
      long double p3 = p2;
      p2 = p1;
      p1 = ((2 * j - 1) * z * p2 - (j - 1) * p3) / j;
    }
  };
  for (int i = 0; i < (npts + 1) / 2; i++) {
    long double z = std::cos(kPiL * (static_cast<long double>(i) + 0.75L) /
                             (static_cast<long double>(npts) + 0.5L));
    long double p1, p2;
    for (int it = 0; it < 100; it++) {
      legendre(z, p1, p2);
      long double pp = npts * (z * p1 - p2) / (z * z - 1);
      long double dz = p1 / pp;
      z -= dz;
      if (std::fabs(dz) <= 4 * std::numeric_limits<long double>::epsilon() && it > 1) break;
    }
    legendre(z, p1, p2);
    long double pp = npts * (z * p1 - p2) / (z * z - 1);
    long double wt = 2 / ((1 - z * z) * pp * pp);
    x[npts - 1 - i] = z;
    x[i] = -z;
    w[npts - 1 - i] = wt;
    w[i] = wt;
  }
  if (npts % 2 == 1) x[npts / 2] = 0;
}

// WignerDetails::Arguments (Wigner.h:24-50).
struct SeedArgs {
  double logSinHalf, logCosHalf;
  bool atLeft, atRight;
  explicit SeedArgs(double theta) {
    const double half = 1.0 / 2.0;
    logSinHalf = std::sin(half * theta);
    logCosHalf = std::cos(half * theta);
    atLeft = logSinHalf < std::numeric_limits<double>::min();
    atRight = logCosHalf < std::numeric_limits<double>::min();
    logSinHalf = atLeft ? 0.0 : std::log(logSinHalf);
    logCosHalf = atRight ? 0.0 : std::log(logCosHalf);
  }
};

// WignerMinOrder (Wigner.h:52-80), evaluated with the same libm expression so
// the seeds agree with the reference's to the last bit.
double min_order(long l, long n, const SeedArgs& a) {
  if (l == 0) return 1.0;
  if (a.atLeft) return n == -l ? 1.0 : 0.0;
  if (a.atRight) return n == l ? 1.0 : 0.0;
  const double half = 1.0 / 2.0;
  const double Fl = static_cast<double>(l), Fn = static_cast<double>(n);
  return std::exp(half * (std::lgamma(2 * Fl + 1) - std::lgamma(Fl - Fn + 1) -
                          std::lgamma(Fl + Fn + 1)) +
                  (Fl + Fn) * a.logSinHalf + (Fl - Fn) * a.logCosHalf);
}
double sign_pow(long k) { return (k % 2) ? -1.0 : 1.0; }  // Utility.h:8-10
double max_order(long l, long n, const SeedArgs& a) { return sign_pow(n + l) * min_order(l, -n, a); }

// d^{l0}_{m n}(theta) at the lowest degree l0 = max(|m|, |n|) of order m,
// picking the same closed form the reference's Compute uses there
// (Wigner.h:330-344 at l = |n|; :358-363 and :398-404 for m = -l; :379-382
// and :458-459 for m = l).
double seed_value(long m, long n, const SeedArgs& a) {
  const long nabs = std::labs(n);
  const long mabs = std::labs(m);
  if (mabs <= nabs) {
    const long l = nabs;
    return n >= 0 ? min_order(l, -m, a)          // WignerMaxUpperIndex
                  : max_order(l, -m, a);         // WignerMinUpperIndex
  }
  const long l = mabs;
  return m < 0 ? min_order(l, n, a) : max_order(l, n, a);
}

// Ortho factor applied as in Wigner.h:475-485.
double ortho_scale(long l, double v) {
  const double factor = 0.564189583547756286948079451560772585844050629328998856844 / 2.0;
  return factor * std::sqrt(static_cast<double>(2 * l + 1)) * v;
}

// log2 |WignerMinOrder(l, n)| and its sign (false: exactly zero), in long
// double: the closed form of min_order without the final exp, so values far
// below the double range keep their exponent.
bool min_order_log2(long l, long n, long double lsin, long double lcos, bool atLeft,
                    bool atRight, long double* lg2) {
  if (l == 0 || (atLeft && n == -l) || (atRight && n == l)) {
    *lg2 = 0.0L;
    return true;
  }
  if (atLeft || atRight) return false;
  const long double Fl = l, Fn = n;
  const long double ln = 0.5L * (std::lgamma(2 * Fl + 1) - std::lgamma(Fl - Fn + 1) -
                                 std::lgamma(Fl + Fn + 1)) +
                         (Fl + Fn) * lsin + (Fl - Fn) * lcos;
  *lg2 = ln / std::log(2.0L);
  return true;
}

// The Ortho-normalised seed of order m as s * 2^-k with s a normal double:
// k = 0 and s = seed (the double value every kernel starts from) while the
// seed is at least 2^-960; below that, where the double expression gives a
// subnormal or zero, the same closed form in long double logarithms, scaled
// to 2^-960 <= |s| < 2^-959.  Seeds of exactly zero (the pole cases) stay 0.
void seed_extended(long m, long n, double theta, double seed, double* s, int* k) {
  *s = seed;
  *k = 0;
  if (seed == 0.0 ? false : std::fabs(seed) >= 0x1p-960) return;
  const long double half = static_cast<long double>(theta) / 2;
  const long double sh = std::sin(half), ch = std::cos(half);
  const bool atLeft = sh < std::numeric_limits<double>::min();
  const bool atRight = ch < std::numeric_limits<double>::min();
  const long double lsin = atLeft ? 0.0L : std::log(sh), lcos = atRight ? 0.0L : std::log(ch);
  const long nabs = std::labs(n), mabs = std::labs(m);
  long l, nn;
  double sign = 1.0;
  if (mabs <= nabs) {  // seed_value's choice of closed form
    l = nabs;
    if (n >= 0) {
      nn = -m;
    } else {
      sign = sign_pow(-m + l);
      nn = m;
    }
  } else {
    l = mabs;
    if (m < 0) {
      nn = n;
    } else {
      sign = sign_pow(n + l);
      nn = -n;
    }
  }
  long double lg2;
  if (!min_order_log2(l, nn, lsin, lcos, atLeft, atRight, &lg2)) return;  // exactly zero
  const long double factor = 0.5L / std::sqrt(3.14159265358979323846264338327950288L);
  lg2 += std::log2(factor * std::sqrt(static_cast<long double>(2 * l + 1)));
  const long double e = std::floor(lg2);
  if (e >= -960.0L) return;  // normal after all: keep the double seed
  *k = static_cast<int>(std::min<long double>(-960.0L - e, 1e9L));
  *s = sign * std::ldexp(static_cast<double>(std::exp2(lg2 - e)), -960);
}

// Recurrence coefficients {a, b, c} of the Ortho-normalised values of order
// m at degree l > l0 = max(|m|, |n|), in long double:
//   d_l = [(2l-1)(l(l-1)cos - m n) d_{l-1} - l sqrt(((l-1)^2-n^2)((l-1)^2-m^2)) d_{l-2}]
//         / ((l-1) sqrt((l^2-n^2)(l^2-m^2)))            (Wigner.h:407-445)
// with the l = |n|+1 one-term form (Wigner.h:368-376) as its n m / (l-1) -> sign(n) m
// limit, and the Ortho ratio sqrt(2l+1) folded in:
//   e_l = (a cos - b) e_{l-1} - c e_{l-2}.
double4 rec_coef(long l, long m, long n) {
  const long l0 = std::max(std::labs(m), std::labs(n));
  const long double ll = l, mm = m, nn = n;
  const long double den = std::sqrt((ll * ll - nn * nn) * (ll * ll - mm * mm));
  const long double r1 = std::sqrt((2 * ll + 1) / (2 * ll - 1));
  const long double a = (2 * ll - 1) * ll / den;
  long double b = 0;
  if (m * n != 0) b = (2 * ll - 1) * mm * nn / ((ll - 1) * den);
  long double c = 0;
  if (l >= l0 + 2) {
    const long double r2 = std::sqrt((2 * ll + 1) / (2 * ll - 3));
    c = ll * std::sqrt(((ll - 1) * (ll - 1) - nn * nn) * ((ll - 1) * (ll - 1) - mm * mm)) /
        ((ll - 1) * den) * r2;
  }
  return make_double4(static_cast<double>(a * r1), static_cast<double>(b * r1),
                      static_cast<double>(c), 0.0);
}

template <typename F>
void parallel_for(long n, F f) {
  unsigned nt = std::max(1u, std::min(32u, std::thread::hardware_concurrency()));
  if (n < 256) nt = 1;
  std::vector<std::thread> th;
  for (unsigned t = 0; t < nt; t++)
    th.emplace_back([=, &f]() {
      for (long i = t; i < n; i += nt) f(i);
    });
  for (auto& x : th) x.join();
}

// ---------------------------------------------------------------------------
// Grid state
// ---------------------------------------------------------------------------
struct NData {
  double* seeds = nullptr;  // [(m+L)*ntheta + t]
  // Seeds below the normal range in extended form (see seed_extended), or
  // null when there are none: sseed * 2^-skexp.
  double* sseed = nullptr;
  int* skexp = nullptr;
  bool ck_restart = false;  // some such order rises to significance: the FP64
                            // synthesis restarts from the checkpoints
  double4* rec = nullptr;  // {a, b, c, 0} per (m, l), see gsh_internal.h
  long* roff = nullptr;
  // Recurrence checkpoints at the start of every analysis pass k >= 1
  // (built on the first analysis with this upper index).
  double2* ck = nullptr;
  long* ckoff = nullptr;
};

// Analysis work list on the device (see gsh::work_item).
struct WorkList {
  uint2* d = nullptr;
  int per = 0;
};

#define GSH_HIP(call)                                  \
  do {                                                 \
    hipError_t e_ = (call);                            \
    if (e_ != hipSuccess) return GSH_ERR_DEVICE;       \
  } while (0)

}  // namespace

struct gsh_plan_s;

struct gsh_grid_s {
  int L = 0, nmax = 0, mrange = 0, nrange = 0, device = 0;
  std::vector<double> theta, weight;
  double* d_cos = nullptr;
  double* d_wq = nullptr;
  double* d_one = nullptr;     // unit colatitude weights (adjoint of the synthesis)
  double2* d_tw = nullptr;
  double2* d_twa = nullptr;    // wavefront FFT twiddles (N in {512, 1024, 2048})
  double2* d_chirp = nullptr;  // Bluestein tables (ring length not a power of two)
  double2* d_hf = nullptr;
  double2* d_hb = nullptr;
  int M = 0;                   // FFT length in LDS
  std::map<int, NData> ndata;
  double* work = nullptr;
  size_t work_bytes = 0;
  void* work2 = nullptr;  // split-BF16 operands of the mixed-precision path
  size_t work2_bytes = 0;
  // Per (lmax, n, value kind): first 32-degree chunk of each order in the
  // split coefficient buffer (device), and the total chunk count.
  std::map<long, std::pair<long*, long>> cso;
  // Analysis work lists by (kernel family, lmax, n, real, per-group degree bounds).
  std::map<std::tuple<int, int, int, int, std::vector<int>>, WorkList> worklists;
  double* stage = nullptr;  // host-call staging
  size_t stage_bytes = 0;
  // Incremented whenever work or work2 is (re)allocated: a captured plan
  // graph embeds those pointers and is valid only for the generation it was
  // captured in.
  unsigned long long work_gen = 0;
  // Live batch plans on this grid (gsh_grid_destroy releases their device
  // resources and detaches them).
  std::vector<gsh_plan_s*> plans;
  std::mutex mu;
  // Kernel timing (gsh_timing_*): event pairs per kernel id.
  bool timing = false;
  std::vector<std::pair<hipEvent_t, hipEvent_t>> pending[GSH_KERNEL_COUNT];
  std::vector<hipEvent_t> event_pool;
  double total_ms[GSH_KERNEL_COUNT] = {};
  long launches[GSH_KERNEL_COUNT] = {};

  int min_upper() const { return nrange == GSH_ALL ? -nmax : (nrange == GSH_NONNEGATIVE ? 0 : nmax); }
  int ntheta() const { return L + 1; }
  int nphi() const { return 2 * L; }
};

namespace {

int ensure_ndata(gsh_grid_s* g, int n, NData** out) {
  auto it = g->ndata.find(n);
  if (it != g->ndata.end()) {
    *out = &it->second;
    return GSH_OK;
  }
  const int L = g->L, nt = g->ntheta();
  const long nabs = std::labs(n);
  const long norders = 2L * L + 1;
  // Seeds, Ortho-normalised.
  std::vector<double> seeds(norders * nt), sseed(norders * nt);
  std::vector<int> skexp(norders * nt);
  std::vector<char> any_scaled(nt, 0);
  parallel_for(nt, [&](long t) {
    SeedArgs a(g->theta[t]);
    for (long m = -L; m <= L; m++) {
      const long l0 = std::max(std::labs(m), nabs);
      const long i = (m + L) * nt + t;
      seeds[i] = ortho_scale(l0, seed_value(m, n, a));
      seed_extended(m, n, g->theta[t], seeds[i], &sseed[i], &skexp[i]);
      if (skexp[i] > 0) any_scaled[t] = 1;
    }
  });
  const bool scaled = std::find(any_scaled.begin(), any_scaled.end(), 1) != any_scaled.end();
  // Recurrence coefficients per (m, l), l0 < l <= L (rec_coef).
  std::vector<long> roff(norders);
  long total = 0;
  for (long m = -L; m <= L; m++) {
    roff[m + L] = total;
    total += L - std::max(std::labs(m), nabs) + 1 + gsh::kRecPad;
  }
  std::vector<double4> rec(total, make_double4(0.0, 0.0, 0.0, 0.0));
  parallel_for(norders, [&](long mi) {
    const long m = mi - L;
    const long l0 = std::max(std::labs(m), nabs);
    for (long l = l0 + 1; l <= L; l++) rec[roff[mi] + (l - l0)] = rec_coef(l, m, n);
  });
  NData d;
  GSH_HIP(hipSetDevice(g->device));
  // Every partial allocation is released on failure.
  auto release = [&](int status) {
    if (d.seeds) (void)hipFree(d.seeds);
    if (d.rec) (void)hipFree(d.rec);
    if (d.roff) (void)hipFree(d.roff);
    if (d.sseed) (void)hipFree(d.sseed);
    if (d.skexp) (void)hipFree(d.skexp);
    return status;
  };
  if (hipMalloc(&d.seeds, seeds.size() * sizeof(double)) != hipSuccess ||
      hipMalloc(&d.rec, total * sizeof(double4)) != hipSuccess ||
      hipMalloc(&d.roff, norders * sizeof(long)) != hipSuccess ||
      (scaled && (hipMalloc(&d.sseed, sseed.size() * sizeof(double)) != hipSuccess ||
                  hipMalloc(&d.skexp, skexp.size() * sizeof(int)) != hipSuccess)))
    return release(GSH_ERR_OUT_OF_MEMORY);
  if (hipMemcpy(d.seeds, seeds.data(), seeds.size() * sizeof(double), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(d.rec, rec.data(), total * sizeof(double4), hipMemcpyHostToDevice) !=
          hipSuccess ||
      hipMemcpy(d.roff, roff.data(), norders * sizeof(long), hipMemcpyHostToDevice) != hipSuccess ||
      (scaled && (hipMemcpy(d.sseed, sseed.data(), sseed.size() * sizeof(double),
                            hipMemcpyHostToDevice) != hipSuccess ||
                  hipMemcpy(d.skexp, skexp.data(), skexp.size() * sizeof(int),
                            hipMemcpyHostToDevice) != hipSuccess)))
    return release(GSH_ERR_DEVICE);
  auto res = g->ndata.emplace(n, d);
  *out = &res.first->second;
  return GSH_OK;
}

hipEvent_t take_event(gsh_grid_s* g) {
  if (!g->event_pool.empty()) {
    hipEvent_t e = g->event_pool.back();
    g->event_pool.pop_back();
    return e;
  }
  hipEvent_t e = nullptr;
  if (hipEventCreate(&e) != hipSuccess) return nullptr;
  return e;
}

// Runs `launch` and, when timing is on, brackets it with events on stream s.
template <typename F>
hipError_t timed(gsh_grid_s* g, int kid, hipStream_t s, F launch) {
  if (!g->timing) return launch();
  hipEvent_t a = take_event(g), b = take_event(g);
  if (a) (void)hipEventRecord(a, s);
  hipError_t e = launch();
  if (b) (void)hipEventRecord(b, s);
  if (a && b) g->pending[kid].emplace_back(a, b);
  return e;
}

int ensure_work(gsh_grid_s* g, size_t bytes) {
  if (bytes <= g->work_bytes) return GSH_OK;
  if (g->work) {
    GSH_HIP(hipDeviceSynchronize());
    GSH_HIP(hipFree(g->work));
    g->work = nullptr;
    g->work_bytes = 0;
  }
  g->work_gen++;
  if (hipMalloc(&g->work, bytes) != hipSuccess) {
    g->work = nullptr;
    return GSH_ERR_OUT_OF_MEMORY;
  }
  g->work_bytes = bytes;
  return GSH_OK;
}

int ensure_work2(gsh_grid_s* g, size_t bytes) {
  if (bytes <= g->work2_bytes) return GSH_OK;
  if (g->work2) {
    GSH_HIP(hipDeviceSynchronize());
    GSH_HIP(hipFree(g->work2));
    g->work2 = nullptr;
    g->work2_bytes = 0;
  }
  g->work_gen++;
  if (hipMalloc(&g->work2, bytes) != hipSuccess) {
    g->work2 = nullptr;
    return GSH_ERR_OUT_OF_MEMORY;
  }
  g->work2_bytes = bytes;
  return GSH_OK;
}

// Chunk offsets of the split coefficient layout: order m (slot m - m_first)
// owns ceil((lmax - l0 + 1) / 32) chunks of 32 degrees, l0 = max(|m|, |n|).
// Split-block offsets per slot: one block per 32-degree chunk, or (sym, the
// parity layout of upper index 0) two per 64-degree chunk.
int ensure_cso(gsh_grid_s* g, int lmax, int n, int real, int sym, long** dev, long* total) {
  const long key = (((long)lmax * 4096 + (n + 2048)) * 2 + real) * 2 + sym;
  auto it = g->cso.find(key);
  if (it != g->cso.end()) {
    *dev = it->second.first;
    *total = it->second.second;
    return GSH_OK;
  }
  const int nslot = real ? lmax + 1 : 2 * lmax + 1;
  const int m_first = real ? 0 : -lmax;
  std::vector<long> h(nslot + 1);
  long acc = 0;
  for (int sl = 0; sl < nslot; sl++) {
    h[sl] = acc;
    const int m = sl + m_first;
    const int l0 = std::max(std::abs(m), std::abs(n));
    acc += sym ? 2 * ((lmax - l0 + 64) / 64) : (lmax - l0 + 32) / 32;
  }
  h[nslot] = acc;
  long* d = nullptr;
  GSH_HIP(hipMalloc(&d, h.size() * sizeof(long)));
  GSH_HIP(hipMemcpy(d, h.data(), h.size() * sizeof(long), hipMemcpyHostToDevice));
  g->cso.emplace(key, std::make_pair(d, acc));
  *dev = d;
  *total = acc;
  return GSH_OK;
}

// Checkpoints of every order of upper index n (gsh::rec_checkpoints): per
// order, one row of ntheta states per pass k = 1 .. npass(L) - 1.
int ensure_ck(gsh_grid_s* g, NData* d, int n, hipStream_t s) {
  if (d->ck) return GSH_OK;
  const int L = g->L, nt = g->ntheta();
  std::vector<long> off(2 * L + 1);
  long rows = 0;
  for (int m = -L; m <= L; m++) {
    off[m + L] = rows;
    const int l0 = std::max(std::abs(m), std::abs(n));
    rows += std::max(0, (L - l0 + gsh::kCk) / gsh::kCk - 1);
  }
  double2* ck = nullptr;
  long* ckoff = nullptr;
  int* flag = nullptr;
  auto release = [&](int status) {
    if (ck) (void)hipFree(ck);
    if (ckoff) (void)hipFree(ckoff);
    if (flag) (void)hipFree(flag);
    return status;
  };
  if (hipMalloc(&ck, std::max(rows, 1L) * nt * sizeof(double2)) != hipSuccess ||
      hipMalloc(&ckoff, off.size() * sizeof(long)) != hipSuccess ||
      (d->sseed && hipMalloc(&flag, sizeof(int)) != hipSuccess))
    return release(GSH_ERR_OUT_OF_MEMORY);
  if (hipMemcpy(ckoff, off.data(), off.size() * sizeof(long), hipMemcpyHostToDevice) !=
          hipSuccess ||
      (flag && hipMemset(flag, 0, sizeof(int)) != hipSuccess))
    return release(GSH_ERR_DEVICE);
  LegParams p{};
  p.L = L;
  p.ntheta = nt;
  p.n = n;
  p.nabs = std::abs(n);
  p.cosv = g->d_cos;
  p.seeds = d->seeds;
  p.rec = d->rec;
  p.roff = d->roff;
  p.ck = ck;
  p.ckoff = ckoff;
  p.sseed = d->sseed;
  p.skexp = d->skexp;
  p.ck_flag = flag;
  if (gsh::launch_rec_checkpoints(p, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
    return release(GSH_ERR_DEVICE);
  int rose = 0;
  if (flag && hipMemcpy(&rose, flag, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
    return release(GSH_ERR_DEVICE);
  if (flag) (void)hipFree(flag);
  d->ck = ck;
  d->ckoff = ckoff;
  d->ck_restart = rose != 0;
  return GSH_OK;
}

constexpr int kMaxGridDegree = 2048;
// Degrees per pass of the split-BF16 analysis when neither a plan's tuning
// nor the caller chose: gsh::kWideAnalPass (128) or gsh::kPass (64).
constexpr int kDefaultAnalPass = gsh::kWideAnalPass;

int order_of_host(int mw, int real) { return real ? mw : ((mw & 1) ? -((mw + 1) >> 1) : (mw >> 1)); }

// Orders per band of an analysis work list (kAnalBand below).
constexpr int kAnalBand = 8;

// Work list of an analysis launch: items (mw, pass | group << 16) in bands
// of kAnalBand neighbouring order indices mw, each band ordered by group,
// pass, order; padded to a multiple of 8 with (~0u, 0).  glmax[g] bounds the
// degrees of group g (empty: lmax for every group).  Passes are kPass
// degrees, or kWideAnalPass for family 3 (the 128-degree split-BF16 kernel);
// family 5's items are the orders m >= 0 (mw = m) of the paired kernel.
// An XCD runs a contiguous eighth of the list, so the workgroups on it at
// one time hold a few passes of each order of a band: the passes of one order
// and group read the same input blocks (from L2 after the first), and the
// orders of a band write neighbouring 16-byte coefficients of each degree
// (the coefficient layout runs over m within a degree), which meet in L2 as
// whole lines rather than leave as partial-line writes.
int ensure_worklist(gsh_grid_s* g, int family, int lmax, int n, int real, int ngroups,
                    const std::vector<int>& glmax, WorkList* out) {
  auto key = std::make_tuple(family, lmax, n, real,
                             glmax.empty() ? std::vector<int>{-ngroups} : glmax);
  auto it = g->worklists.find(key);
  if (it != g->worklists.end()) {
    *out = it->second;
    return GSH_OK;
  }
  // Family 5 (orders m and -m in one item) lists the orders m >= 0 in turn.
  const int nslot = family == 5 ? lmax + 1 : real ? lmax + 1 : 2 * lmax + 1;
  const int plen = family == 3 ? gsh::kWideAnalPass : gsh::kPass;  // degrees per pass
  std::vector<uint2> items;
  for (int b0 = 0; b0 < nslot; b0 += kAnalBand)
    for (int grp = 0; grp < ngroups; grp++) {
      const int lw = glmax.empty() ? lmax : glmax[grp];
      for (int k = 0; k <= lw / plen; k++)
        for (int mw = b0; mw < std::min(nslot, b0 + kAnalBand); mw++) {
          const int m = family == 5 ? mw : order_of_host(mw, real);
          const int l0 = std::max(std::abs(m), std::abs(n));
          if (l0 <= lw && k < (lw - l0 + plen) / plen)
            items.push_back(make_uint2(static_cast<unsigned>(mw),
                                       static_cast<unsigned>(k) | (static_cast<unsigned>(grp) << 16)));
        }
    }
  while (items.size() % 8 != 0 || items.empty()) items.push_back(make_uint2(~0u, 0u));
  WorkList w;
  w.per = static_cast<int>(items.size() / 8);
  GSH_HIP(hipMalloc(&w.d, items.size() * sizeof(uint2)));
  GSH_HIP(hipMemcpy(w.d, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice));
  g->worklists.emplace(key, w);
  *out = w;
  return GSH_OK;
}

// Work list of a synthesis launch: every (order, column group, colatitude
// super-block) item.  An item of order m sums lmax - |m| + 1 degrees.  Whole
// units (order, column group) -- all their super-blocks, which read the same
// coefficients -- go to one XCD, heaviest first onto the least loaded XCD, so
// the eight XCDs finish together; each XCD's list runs heaviest first and is
// padded to the longest with empty items.
int ensure_synth_worklist(gsh_grid_s* g, int family, int lmax, int real, int ngroups, int ntsb,
                          WorkList* out) {
  auto key = std::make_tuple(family, lmax, 0, real, std::vector<int>{ngroups, ntsb});
  auto it = g->worklists.find(key);
  if (it != g->worklists.end()) {
    *out = it->second;
    return GSH_OK;
  }
  // Family 22 (orders m and -m in one item) lists the orders m >= 0, an
  // item of m > 0 weighing twice its degrees.
  const bool paired = family == 22;
  const int nslot = real || paired ? lmax + 1 : 2 * lmax + 1;
  std::vector<std::pair<int, int>> units;  // (weight, mw * ngroups + grp)
  units.reserve((size_t)nslot * ngroups);
  // mw is the kernels' order index: complex orders interleave as
  // 0, -1, 1, -2, 2, ... (order_of_host); paired items take mw = m.
  for (int mw = 0; mw < nslot; mw++)
    for (int grp = 0; grp < ngroups; grp++)
      units.emplace_back(paired ? (lmax - mw + 1) * (mw > 0 ? 2 : 1)
                                : lmax - std::abs(order_of_host(mw, real)) + 1,
                         mw * ngroups + grp);
  std::stable_sort(units.begin(), units.end(),
                   [](const auto& a, const auto& b) { return a.first > b.first; });
  std::vector<uint2> xcd[8];
  long load[8] = {0, 0, 0, 0, 0, 0, 0, 0};
  for (const auto& u : units) {
    const int x = static_cast<int>(std::min_element(load, load + 8) - load);
    load[x] += static_cast<long>(u.first) * ntsb;
    const int mw = u.second / ngroups, grp = u.second % ngroups;
    for (int t = 0; t < ntsb; t++)
      xcd[x].push_back(make_uint2(static_cast<unsigned>(mw),
                                  static_cast<unsigned>(t) | (static_cast<unsigned>(grp) << 16)));
  }
  size_t per = 1;
  for (const auto& v : xcd) per = std::max(per, v.size());
  std::vector<uint2> items;
  items.reserve(8 * per);
  for (auto& v : xcd) {
    v.resize(per, make_uint2(~0u, 0u));
    items.insert(items.end(), v.begin(), v.end());
  }
  WorkList w;
  w.per = static_cast<int>(items.size() / 8);
  GSH_HIP(hipMalloc(&w.d, items.size() * sizeof(uint2)));
  GSH_HIP(hipMemcpy(w.d, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice));
  g->worklists.emplace(key, w);
  *out = w;
  return GSH_OK;
}

// Work list of coef_split: items (group gs of the 8 slots 8 gs .. 8 gs + 7,
// chunk | tile << 16).  Groups go in bands of kSplitBand neighbours, and
// within a band the list runs tile, then chunk, then group: consecutive
// items -- which run on one XCD -- are neighbouring order groups at the same
// degrees, whose 128-byte pieces of the coefficient rows (8 orders of one
// degree and field, not line-aligned) share cache lines, and the next chunk
// of a group, whose window overlaps the previous one by 8 degrees, follows
// a band later.
int ensure_split_worklist(gsh_grid_s* g, int lmax, int n, int real, int ntile, int sym,
                          WorkList* out) {
  auto key = std::make_tuple(sym ? 32 : 31, lmax, n, real, std::vector<int>{ntile});
  auto it = g->worklists.find(key);
  if (it != g->worklists.end()) {
    *out = it->second;
    return GSH_OK;
  }
  const int nslot = real ? lmax + 1 : 2 * lmax + 1;
  const int m_first = real ? 0 : -lmax;
  const int ngs = (nslot + 7) / 8;
  std::vector<int> nch(ngs, 0);  // chunks of each group's order with the most
  for (int gs = 0; gs < ngs; gs++)
    for (int sl = 8 * gs; sl < std::min(8 * gs + 8, nslot); sl++) {
      const int l0 = std::max(std::abs(sl + m_first), std::abs(n));
      const int cw = sym ? 64 : 32;  // degrees per chunk
      nch[gs] = std::max(nch[gs], (lmax - l0 + cw) / cw);
    }
  constexpr int kSplitBand = 4;
  std::vector<uint2> items;
  for (int b0 = 0; b0 < ngs; b0 += kSplitBand) {
    const int b1 = std::min(b0 + kSplitBand, ngs);
    int cmax = 0;
    for (int gs = b0; gs < b1; gs++) cmax = std::max(cmax, nch[gs]);
    for (int t = 0; t < ntile; t++)
      for (int c = 0; c < cmax; c++)
        for (int gs = b0; gs < b1; gs++)
          if (c < nch[gs])
            items.push_back(make_uint2(static_cast<unsigned>(gs),
                                       static_cast<unsigned>(c) | (static_cast<unsigned>(t) << 16)));
  }
  while (items.size() % 8 != 0 || items.empty()) items.push_back(make_uint2(~0u, 0u));
  WorkList w;
  w.per = static_cast<int>(items.size() / 8);
  GSH_HIP(hipMalloc(&w.d, items.size() * sizeof(uint2)));
  GSH_HIP(hipMemcpy(w.d, items.data(), items.size() * sizeof(uint2), hipMemcpyHostToDevice));
  g->worklists.emplace(key, w);
  *out = w;
  return GSH_OK;
}

int check_call(const gsh_grid_s* g, int lmax, int n, int value_kind, int precision, int storage,
               int batch) {
  if (!g) return GSH_ERR_INVALID_ARGUMENT;
  if (value_kind != GSH_REAL && value_kind != GSH_COMPLEX) return GSH_ERR_INVALID_ARGUMENT;
  if (precision < GSH_FP64 || precision > GSH_BF16X6) return GSH_ERR_INVALID_ARGUMENT;
  if (storage != GSH_STORE_FP64 && storage != GSH_STORE_FP32) return GSH_ERR_INVALID_ARGUMENT;
  if (batch < 0) return GSH_ERR_INVALID_ARGUMENT;
  if (lmax < 0 || lmax > g->L) return GSH_ERR_DEGREE;
  if (n < g->min_upper() || n > g->nmax || std::abs(n) > lmax) return GSH_ERR_UPPER_INDEX;
  if (value_kind == GSH_COMPLEX && g->mrange != GSH_ALL) return GSH_ERR_VALUE_KIND;
  return GSH_OK;
}

LegParams base_params(gsh_grid_s* g, const NData* d, int lmax, int n, int real, int batch) {
  LegParams p{};
  p.L = g->L;
  p.ntheta = g->ntheta();
  p.lmax = lmax;
  p.n = n;
  p.nabs = std::abs(n);
  p.real = real;
  p.m_first = real ? 0 : -lmax;
  p.nslot = real ? lmax + 1 : 2 * lmax + 1;
  p.nfield = batch;
  p.ncol = 2 * batch;
  p.cosv = g->d_cos;
  p.seeds = d->seeds;
  p.rec = d->rec;
  p.roff = d->roff;
  p.F = g->work;
  return p;
}

FFTParams fft_params(gsh_grid_s* g, int lmax, int real, int batch) {
  FFTParams q{};
  q.N = g->nphi();
  q.M = g->M;
  q.chirp = g->d_chirp;
  q.hf = g->d_hf;
  q.hb = g->d_hb;
  q.L = g->L;
  q.ntheta = g->ntheta();
  q.lmax = lmax;
  q.real = real;
  q.m_first = real ? 0 : -lmax;
  q.nslot = real ? lmax + 1 : 2 * lmax + 1;
  q.nfield = batch;
  q.ncol = 2 * batch;
  q.tw = g->d_tw;
  q.twa = g->d_twa;
  q.wq = g->d_wq;
  q.F = g->work;
  return q;
}

bool is_pow2(int N) { return N >= 1 && (N & (N - 1)) == 0; }

// LDS transform length for ring length N: N itself when a power of two,
// else the Bluestein convolution length (smallest power of two >= 2N - 1).
int fft_length(int N) {
  if (is_pow2(N)) return N;
  int M = 1;
  while (M < 2 * N - 1) M <<= 1;
  return M;
}

// In-place radix-2 FFT in long double (host, set-up only), sign -1.
void fft_ld(std::vector<std::complex<long double>>& a) {
  const size_t n = a.size();
  for (size_t i = 1, j = 0; i < n; i++) {
    size_t bit = n >> 1;
    for (; j & bit; bit >>= 1) j ^= bit;
    j ^= bit;
    if (i < j) std::swap(a[i], a[j]);
  }
  for (size_t len = 2; len <= n; len <<= 1) {
    const long double ang = -2.0L * kPiL / len;
    for (size_t i = 0; i < n; i += len)
      for (size_t k = 0; k < len / 2; k++) {
        const std::complex<long double> w(std::cos(ang * k), std::sin(ang * k));
        const auto u = a[i + k], v = a[i + k + len / 2] * w;
        a[i + k] = u + v;
        a[i + k + len / 2] = u - v;
      }
  }
}

// Releases a plan's device resources and detaches it from its grid
// (defined with the plans below).
void plan_detach(gsh_plan_s* P);

}  // namespace

// ===========================================================================
// C ABI
// ===========================================================================
extern "C" {

const char* gsh_status_string(int status) {
  switch (status) {
    case GSH_OK: return "ok";
    case GSH_ERR_INVALID_ARGUMENT: return "invalid argument";
    case GSH_ERR_UPPER_INDEX: return "upper index n outside the grid's UpperIndices() or |n| > lMax";
    case GSH_ERR_DEGREE: return "degree lMax outside [0, grid lMax]";
    case GSH_ERR_VALUE_KIND: return "complex field on an MRange=NonNegative grid";
    case GSH_ERR_UNSUPPORTED: return "configuration not supported by this build";
    case GSH_ERR_DEVICE: return "HIP runtime error";
    case GSH_ERR_NO_DEVICE: return "no gfx950 device";
    case GSH_ERR_OUT_OF_MEMORY: return "out of device memory";
    case GSH_ERR_SCRATCH_LIMIT:
      return "batch x colatitudes exceed the static scratch of the allocation-free entry point";
    default: return "unknown status";
  }
}

int gsh_version(void) { return 1; }

int gsh_gl_nodes(int lmax, double* theta, double* weight) {
  if (lmax < 0 || !theta || !weight) return GSH_ERR_INVALID_ARGUMENT;
  std::vector<long double> x, w;
  gauss_legendre(lmax + 1, x, w);
  for (int i = 0; i <= lmax; i++) {
    const double xi = static_cast<double>(x[i]);
    theta[i] = std::acos(-xi);
    weight[i] = static_cast<double>(w[i]);
  }
  return GSH_OK;
}

long gsh_coefficient_size(int lmax, int n, int value_kind) {
  if (lmax < 0 || std::abs(n) > lmax) return -1;
  const long l = lmax, na = std::abs(n);
  if (value_kind == GSH_REAL) return (l + 1) * (l + 2) / 2 - na * (na + 1) / 2;
  return (l + 1) * (l + 1) - na * na;
}

int gsh_wigner_seed(int l, int n, double theta, double* value) {
  if (!value || l < 0 || std::abs(n) > l) return GSH_ERR_INVALID_ARGUMENT;
  SeedArgs a(theta);
  *value = min_order(l, n, a);
  return GSH_OK;
}

int gsh_grid_create(int lmax, int nmax, int mrange, int nrange, gsh_grid_t* grid) {
  if (!grid) return GSH_ERR_INVALID_ARGUMENT;
  *grid = nullptr;
  if (lmax < 0 || nmax < 0 || nmax > lmax) return GSH_ERR_INVALID_ARGUMENT;
  if (mrange != GSH_ALL && mrange != GSH_NONNEGATIVE) return GSH_ERR_INVALID_ARGUMENT;
  if (nrange != GSH_ALL && nrange != GSH_NONNEGATIVE && nrange != GSH_SINGLE)
    return GSH_ERR_INVALID_ARGUMENT;
  // Grids up to lMax = 2048 (the reference's constructor has no cap).  The
  // LDS ring FFT transforms up to 8192 points, so every ring length up to
  // 4096 runs, power of two or through Bluestein's convolution; past 2048 the
  // recurrence checkpoint table of one upper index, about (L + 1)^3 bytes,
  // grows from 8.6 GB (L = 2048) to 69 GB (L = 4096).
  if (lmax > kMaxGridDegree) return GSH_ERR_UNSUPPORTED;
  if (lmax > 0 && fft_length(2 * lmax) > 8192) return GSH_ERR_UNSUPPORTED;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return GSH_ERR_NO_DEVICE;
  int dev = 0;
  if (hipGetDevice(&dev) != hipSuccess) return GSH_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return GSH_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return GSH_ERR_NO_DEVICE;

  auto g = std::make_unique<gsh_grid_s>();
  g->L = lmax;
  g->nmax = nmax;
  g->mrange = mrange;
  g->nrange = nrange;
  g->device = dev;
  g->theta.resize(lmax + 1);
  g->weight.resize(lmax + 1);
  gsh_gl_nodes(lmax, g->theta.data(), g->weight.data());
  const int N = g->nphi();
  std::vector<double> cs(lmax + 1), wq(lmax + 1);
  // scaleFactor = 2 pi / nPhi (GaussLegendreGrid.h:144-145)
  const double scale = N > 0 ? 2.0 * 3.14159265358979323846 / static_cast<double>(N) : 0.0;
  for (int i = 0; i <= lmax; i++) {
    cs[i] = std::cos(g->theta[i]);
    wq[i] = g->weight[i] * scale;
  }
  const int M = N > 0 ? fft_length(N) : 1;
  g->M = M;
  std::vector<double2> tw(M);
  for (int k = 0; k < M; k++) {
    const long double a = -2.0L * kPiL * k / M;
    tw[k] = make_double2(static_cast<double>(std::cos(a)), static_cast<double>(std::sin(a)));
  }
  if (N > 0 && M != N) {
    // Bluestein: chirp w_k = exp(-i pi k^2 / N); forward filter conj(w),
    // backward filter w, both circularly extended to M and transformed.
    std::vector<std::complex<long double>> w(N), hf(M), hb(M);
    for (int k = 0; k < N; k++) {
      const long long kk = (static_cast<long long>(k) * k) % (2LL * N);
      const long double a = -kPiL * kk / N;
      w[k] = std::complex<long double>(std::cos(a), std::sin(a));
    }
    for (int k = 0; k < N; k++) {
      hf[k] = std::conj(w[k]);
      hb[k] = w[k];
      if (k > 0) {
        hf[M - k] = std::conj(w[k]);
        hb[M - k] = w[k];
      }
    }
    fft_ld(hf);
    fft_ld(hb);
    std::vector<double2> cw(N), chf(M), chb(M);
    for (int k = 0; k < N; k++)
      cw[k] = make_double2(static_cast<double>(w[k].real()), static_cast<double>(w[k].imag()));
    for (int k = 0; k < M; k++) {
      chf[k] = make_double2(static_cast<double>(hf[k].real() / M),
                            static_cast<double>(hf[k].imag() / M));
      chb[k] = make_double2(static_cast<double>(hb[k].real() / M),
                            static_cast<double>(hb[k].imag() / M));
    }
    GSH_HIP(hipMalloc(&g->d_chirp, N * sizeof(double2)));
    GSH_HIP(hipMalloc(&g->d_hf, M * sizeof(double2)));
    GSH_HIP(hipMalloc(&g->d_hb, M * sizeof(double2)));
    GSH_HIP(hipMemcpy(g->d_chirp, cw.data(), N * sizeof(double2), hipMemcpyHostToDevice));
    GSH_HIP(hipMemcpy(g->d_hf, chf.data(), M * sizeof(double2), hipMemcpyHostToDevice));
    GSH_HIP(hipMemcpy(g->d_hb, chb.data(), M * sizeof(double2), hipMemcpyHostToDevice));
  }
  if (N > 0 && gsh::wave_fft_supported(N)) {
    // twa[k1 * 64 + j] = exp(-2 pi i j k1 / N), k1 < N / 64, j < 64
    std::vector<double2> ta(N);
    for (int k1 = 0; k1 < N / 64; k1++)
      for (int j = 0; j < 64; j++) {
        const long double a = -2.0L * kPiL * ((j * k1) % N) / N;
        ta[k1 * 64 + j] =
            make_double2(static_cast<double>(std::cos(a)), static_cast<double>(std::sin(a)));
      }
    GSH_HIP(hipMalloc(&g->d_twa, ta.size() * sizeof(double2)));
    GSH_HIP(hipMemcpy(g->d_twa, ta.data(), ta.size() * sizeof(double2), hipMemcpyHostToDevice));
  }
  GSH_HIP(hipMalloc(&g->d_cos, cs.size() * sizeof(double)));
  GSH_HIP(hipMalloc(&g->d_wq, wq.size() * sizeof(double)));
  GSH_HIP(hipMalloc(&g->d_tw, tw.size() * sizeof(double2)));
  GSH_HIP(hipMemcpy(g->d_cos, cs.data(), cs.size() * sizeof(double), hipMemcpyHostToDevice));
  GSH_HIP(hipMemcpy(g->d_wq, wq.data(), wq.size() * sizeof(double), hipMemcpyHostToDevice));
  {
    const std::vector<double> one(lmax + 1, 1.0);
    GSH_HIP(hipMalloc(&g->d_one, one.size() * sizeof(double)));
    GSH_HIP(hipMemcpy(g->d_one, one.data(), one.size() * sizeof(double), hipMemcpyHostToDevice));
  }
  GSH_HIP(hipMemcpy(g->d_tw, tw.data(), tw.size() * sizeof(double2), hipMemcpyHostToDevice));
  // Address of this device's copy of the allocation-free entry point's
  // scratch array (looked up once per device, with the device current).
  GSH_HIP(gsh::noalloc_scratch_init(dev));
  *grid = g.release();
  return GSH_OK;
}

int gsh_grid_destroy(gsh_grid_t g) {
  if (!g) return GSH_ERR_INVALID_ARGUMENT;
  (void)hipSetDevice(g->device);
  (void)hipDeviceSynchronize();
  // Plans still alive lose their graphs and device arrays (which embed this
  // grid's buffers) and are detached: their later calls fail with
  // GSH_ERR_INVALID_ARGUMENT and gsh_plan_destroy frees only the host object.
  for (gsh_plan_s* P : g->plans) plan_detach(P);
  g->plans.clear();
  for (auto& kv : g->ndata) {
    (void)hipFree(kv.second.seeds);
    (void)hipFree(kv.second.rec);
    (void)hipFree(kv.second.roff);
    if (kv.second.ck) (void)hipFree(kv.second.ck);
    if (kv.second.ckoff) (void)hipFree(kv.second.ckoff);
    if (kv.second.sseed) (void)hipFree(kv.second.sseed);
    if (kv.second.skexp) (void)hipFree(kv.second.skexp);
  }
  for (auto& kv : g->worklists) (void)hipFree(kv.second.d);
  (void)hipFree(g->d_cos);
  (void)hipFree(g->d_wq);
  (void)hipFree(g->d_one);
  (void)hipFree(g->d_tw);
  (void)hipFree(g->d_twa);
  if (g->d_chirp) (void)hipFree(g->d_chirp);
  if (g->d_hf) (void)hipFree(g->d_hf);
  if (g->d_hb) (void)hipFree(g->d_hb);
  for (int k = 0; k < GSH_KERNEL_COUNT; k++)
    for (auto& pr : g->pending[k]) {
      (void)hipEventDestroy(pr.first);
      (void)hipEventDestroy(pr.second);
    }
  for (auto e : g->event_pool) (void)hipEventDestroy(e);
  if (g->work) (void)hipFree(g->work);
  if (g->work2) (void)hipFree(g->work2);
  for (auto& kv : g->cso) (void)hipFree(kv.second.first);
  if (g->stage) (void)hipFree(g->stage);
  delete g;
  return GSH_OK;
}

int gsh_grid_info(gsh_grid_t g, int* lmax, int* nmax, int* min_upper, int* ntheta, int* nphi) {
  if (!g) return GSH_ERR_INVALID_ARGUMENT;
  if (lmax) *lmax = g->L;
  if (nmax) *nmax = g->nmax;
  if (min_upper) *min_upper = g->min_upper();
  if (ntheta) *ntheta = g->ntheta();
  if (nphi) *nphi = g->nphi();
  return GSH_OK;
}

int gsh_grid_colatitudes(gsh_grid_t g, double* theta, double* weight) {
  if (!g) return GSH_ERR_INVALID_ARGUMENT;
  if (theta) std::copy(g->theta.begin(), g->theta.end(), theta);
  if (weight) std::copy(g->weight.begin(), g->weight.end(), weight);
  return GSH_OK;
}

int gsh_grid_prepare(gsh_grid_t g, int n) {
  if (!g) return GSH_ERR_INVALID_ARGUMENT;
  if (n < g->min_upper() || n > g->nmax) return GSH_ERR_UPPER_INDEX;
  std::lock_guard<std::mutex> lock(g->mu);
  NData* d;
  return ensure_ndata(g, n, &d);
}

namespace {

// Options of one launch group of a batch plan (defaults: a plain batched call).
struct RunOpts {
  const int* fmap = nullptr;     // device: column field -> caller's field index
  const int* f_lmax = nullptr;   // device: truncation degree per column field
  const int* cg_lmax = nullptr;  // device: max f_lmax per 16-column group
  std::vector<int> cg_lmax_host;  // the same on the host (work lists)
  int f32 = 0;                   // fields stored in FP32
  int synth_tiles = 0;           // FP64 synthesis: 16-column tiles per workgroup (0: auto)
  int fft = 0;                   // 0: auto, 1: wavefront-per-ring kernel, 2: LDS kernel
  int anal_pass = 0;             // split-BF16 analysis degrees per pass: 64, 128 (0: auto)
  const void* field2 = nullptr;  // analysis of a pointwise expression: second operand
  int op = 0;                    // gsh::ExprOp applied as the analysis rings load
  int src_cplx = 0;              // real analysis of Re / Im of a complex field
  double2 scalar = {0.0, 0.0};   // scalar operand of the expression
};

bool use_wave_fft(const FFTParams& q, const RunOpts& o) {
  return q.twa != nullptr && o.fft != 2;
}

// Synthesis of `batch` column fields with lmax >= 1 (no argument checks; the
// caller holds the grid lock and has prepared n).
int run_inverse(gsh_grid_s* g, NData* d, int lmax, int n, int real, int precision, int batch,
                const double* cin, long cstride, void* field, long fstride, hipStream_t s,
                const RunOpts& o) {
  int st;
  const int nslot = real ? lmax + 1 : 2 * lmax + 1;
  const size_t ncg16 = (2 * (size_t)batch + 15) / 16;
  if ((st = ensure_work(g, sizeof(double) * g->ntheta() * nslot * 16 * ncg16))) return st;
  LegParams p = base_params(g, d, lmax, n, real, batch);
  p.cin = cin;
  p.cstride = cstride;
  p.fmap = o.fmap;
  p.f_lmax = o.f_lmax;
  p.cg_lmax = o.cg_lmax;
  const int ntsb = (g->ntheta() + 255) / 256;
  FFTParams q = fft_params(g, lmax, real, batch);
  q.fout = field;
  q.fstride = fstride;
  q.f32 = o.f32;
  q.fmap = o.fmap;
  // The split-BF16 synthesis hands its sums to the wavefront ring FFT in
  // FP32 (half the bytes of F written and read; the products are FP32-
  // accurate already).  The LDS ring FFT reads FP64 bins.
  const bool f32bins = precision != GSH_FP64 && precision != GSH_FP32 && use_wave_fft(q, o);
  p.f32out = q.fa32 = f32bins ? 1 : 0;
  if (precision == GSH_FP64 || precision == GSH_FP32) {
    // Orders whose seeds underflow but whose values rise to significance
    // before L: the synthesis takes its state from the scaled checkpoints.
    if (d->sseed) {
      if ((st = ensure_ck(g, d, n, s))) return st;
      if (d->ck_restart) {
        p.ck = d->ck;
        p.ckoff = d->ckoff;
        p.ck_restart = 1;
      }
    }
    const int T = o.synth_tiles == 1 || o.synth_tiles == 2 ? o.synth_tiles : (p.ncol > 16 ? 2 : 1);
    WorkList w;
    if ((st = ensure_synth_worklist(g, 10 + T, lmax, real, (p.ncol + 16 * T - 1) / (16 * T), ntsb,
                                    &w)))
      return st;
    p.items = w.d;
    p.items_per = w.per;
    GSH_HIP(timed(g, GSH_KERNEL_LEGENDRE_SYNTHESIS, s, [&] {
      return gsh::launch_leg_synth_f64(p, s, T, precision == GSH_FP32);
    }));
  } else {
    // Upper index 0 is synthesized over the northern colatitudes (items of
    // 128 of them; each writes the mirror rows too), the coefficients in the
    // parity layout, and the equator of an odd ntheta as a tail row.
    // Otherwise up to kSynthTailRows colatitudes past whole super-blocks of
    // 256 go to the tail kernels rather than to items of their own.
    // Complex fields of other upper indices run orders m and -m together
    // (items of 128 colatitudes; order -m's rows mirrored, its coefficients
    // signed by coef_split), with up to kSynthTailRows colatitudes past whole
    // blocks of 128 as tail rows (order -m's then sit at the top rows).
    const int nt = g->ntheta();
    const int sym = n == 0 ? 1 : 0;
    const int pm = !sym && !real ? 1 : 0;
    const int blk = pm ? 128 : 256;  // colatitudes per item (sym: northern ones)
    const bool tail =
        sym ? nt % 2 == 1 : nt > blk && nt % blk != 0 && nt % blk <= gsh::kSynthTailRows;
    const int nblk = sym ? (nt / 2 + 127) / 128 : tail ? nt / blk : (nt + blk - 1) / blk;
    WorkList w;
    if ((st = ensure_synth_worklist(g, sym ? 21 : pm ? 22 : 20, lmax, real,
                                    ((p.ncol + 15) / 16 + 7) / 8, nblk, &w)))
      return st;
    p.sym = sym;
    p.pm = pm;
    p.items = w.d;
    p.items_per = w.per;
    const int ns = precision == GSH_BF16X6 ? 3 : 2;
    // The split-BF16 synthesis starts its recurrences at the checkpoints.
    if ((st = ensure_ck(g, d, n, s))) return st;
    p.ck = d->ck;
    p.ckoff = d->ckoff;
    long* cso = nullptr;
    long nchunks = 0;
    if ((st = ensure_cso(g, lmax, n, real, sym, &cso, &nchunks))) return st;
    // Column tiles padded to whole super-groups of 8 (zero columns), so the
    // synthesis reads every tile of its super-group without a bounds check.
    p.ntile = ((p.ncol + 15) / 16 + 7) / 8 * 8;
    p.cso = cso;
    const long nblocks = nchunks * p.ntile;
    // The tail rows' d values (32 per block and row) and partial sums (16 per
    // block, tile and row) follow the split coefficients in the same buffer.
    const int trows = !tail ? 0 : sym ? 1 : nt % blk;
    const size_t bs_bytes = (size_t)nblocks * ns * 64 * 16;
    const size_t dv_bytes = (size_t)trows * nchunks * 32 * sizeof(double);
    const size_t pt_bytes = (size_t)trows * nblocks * 16 * sizeof(double);
    if ((st = ensure_work2(g, bs_bytes + dv_bytes + pt_bytes))) return st;
    p.bs = static_cast<const uint4*>(g->work2);
    p.bs_out = static_cast<uint4*>(g->work2);
    if (trows) {
      char* w2 = static_cast<char*>(g->work2);
      p.tail_rows = trows;
      p.tail_t0 = sym ? nt / 2 : nt - trows;
      p.tail_dv = reinterpret_cast<const double*>(w2 + bs_bytes);
      p.tail_dstride = nchunks * 32;
      p.tail_part = reinterpret_cast<double*>(w2 + bs_bytes + dv_bytes);
      p.tail_pstride = nblocks * 16;
    }
    {
      WorkList ws;
      if ((st = ensure_split_worklist(g, lmax, n, real, p.ntile, sym, &ws))) return st;
      p.items = ws.d;
      p.items_per = ws.per;
    }
    GSH_HIP(timed(g, GSH_KERNEL_COEF_SPLIT, s,
                  [&] {
                    hipError_t e = trows ? gsh::launch_tail_dvals(p, s) : hipSuccess;
                    return e == hipSuccess ? gsh::launch_coef_split(p, ns, s) : e;
                  }));
    p.items = w.d;  // the synthesis work list again
    p.items_per = w.per;
    GSH_HIP(timed(g, GSH_KERNEL_LEGENDRE_SYNTHESIS, s, [&] {
      hipError_t e = gsh::launch_leg_synth_bf16(p, ns, s);
      if (e == hipSuccess && trows) e = gsh::launch_tail_reduce(p, s);
      return e;
    }));
  }
  GSH_HIP(timed(g, GSH_KERNEL_FFT_SYNTHESIS, s, [&] {
    return use_wave_fft(q, o) ? gsh::launch_fft_synth_wave(q, s) : gsh::launch_fft_synth(q, s);
  }));
  return GSH_OK;
}

// Analysis (adjoint = 0) or the adjoint of the synthesis (adjoint = 1) of
// `batch` column fields with lmax >= 1: the same ring FFT and Legendre
// contraction, the adjoint with unit colatitude weights, no 2 pi / N, no
// (L, L) zeroing, and for real fields the c2r weights (2 on bins
// 0 < k < N/2, real part only on bins 0 and N/2).
int run_forward(gsh_grid_s* g, NData* d, int lmax, int n, int real, int precision, int batch,
                const void* field, long fstride, double* cout, long cstride, int accumulate,
                int adjoint, hipStream_t s, const RunOpts& o) {
  int st;
  const int nslot = real ? lmax + 1 : 2 * lmax + 1;
  const int ncol = 2 * batch;
  FFTParams q = fft_params(g, lmax, real, batch);
  q.fin = field;
  q.fin2 = o.field2;
  q.op = o.op;
  q.src_cplx = o.src_cplx;
  q.scalar = o.scalar;
  q.fstride = fstride;
  q.f32 = o.f32;
  q.fmap = o.fmap;
  if (adjoint) {
    q.wq = g->d_one;
    q.adjoint = 1;
  }
  LegParams p = base_params(g, d, lmax, n, real, batch);
  p.cout = cout;
  p.cstride = cstride;
  p.accumulate = accumulate ? 1 : 0;
  p.zero_last = (!adjoint && !real) ? 1 : 0;
  p.fmap = o.fmap;
  p.f_lmax = o.f_lmax;
  p.cg_lmax = o.cg_lmax;
  // FA (FP32 in the split-BF16 modes), plus 8 quads of 64 entries: the
  // split-BF16 analysis reads whole 32-colatitude chunks, up to 8 quads past
  // the last block.  In the split-BF16 modes FA starts gsh::kFAFrontPad
  // floats into the buffer: the paired-order kernel's mirrored reads of the
  // last chunk reach up to 31 rows of 16 floats before a block.
  const bool mixed = precision == GSH_BF16X6 || precision == GSH_BF16X3;
  const size_t fa = (size_t)nslot * ((ncol + 15) / 16) * ((g->ntheta() + 3) / 4) * 64 + 8 * 64 +
                    (mixed ? gsh::kFAFrontPad : 0);
  if ((st = ensure_work(g, (mixed ? sizeof(float) : sizeof(double)) * fa))) return st;
  q.F = p.F = mixed ? reinterpret_cast<double*>(reinterpret_cast<float*>(g->work) + gsh::kFAFrontPad)
                    : g->work;
  q.fa32 = mixed ? 1 : 0;
  p.FA32 = mixed ? reinterpret_cast<const float*>(q.F) : nullptr;
  if ((st = ensure_ck(g, d, n, s))) return st;
  p.ck = d->ck;
  p.ckoff = d->ckoff;
  const int ncg16 = (ncol + 15) / 16;
  GSH_HIP(timed(g, GSH_KERNEL_FFT_ANALYSIS, s, [&] {
    return use_wave_fft(q, o) ? gsh::launch_fft_anal_wave(q, s) : gsh::launch_fft_anal(q, s);
  }));
  if (precision == GSH_FP64 || precision == GSH_FP32) {
    WorkList w;
    if ((st = ensure_worklist(g, 0, lmax, n, real, ncg16, o.cg_lmax_host, &w))) return st;
    p.items = w.d;
    p.items_per = w.per;
    GSH_HIP(timed(g, GSH_KERNEL_LEGENDRE_ANALYSIS, s, [&] {
      return gsh::launch_leg_anal_f64(p, s, precision == GSH_FP32);
    }));
    return GSH_OK;
  }
  // Upper index 0 runs over the northern colatitudes only (the
  // Gauss-Legendre nodes are symmetric about the equator and d^l_{m0} has
  // parity (-1)^(l+m) under theta -> pi - theta).  Complex fields of other
  // upper indices run orders m and -m together in 64-degree passes
  // (d^l_{-m,n}(theta) = (-1)^(l+n) d^l_{m,n}(pi - theta)) by default;
  // everything else defaults to 128-degree passes.  o.anal_pass (64 or 128,
  // 0 for the default) overrides.
  const bool sym = n == 0;
  const bool pm_ok = !real && !sym;
  const int apass = o.anal_pass ? o.anal_pass : pm_ok ? gsh::kPass : kDefaultAnalPass;
  const bool wide = apass == gsh::kWideAnalPass;
  const bool pm = pm_ok && !wide;
  {
    // 128-column super-groups of the split-BF16 kernel.
    const int nsg = (ncg16 + 7) / 8;
    std::vector<int> sg;
    for (int k = 0; k < nsg && !o.cg_lmax_host.empty(); k++) {
      int mx = 0;
      for (int c = 8 * k; c < std::min(8 * k + 8, ncg16); c++) mx = std::max(mx, o.cg_lmax_host[c]);
      sg.push_back(mx);
    }
    WorkList w;
    if ((st = ensure_worklist(g, pm ? 5 : wide ? 3 : 1, lmax, n, real, nsg, sg, &w))) return st;
    p.items = w.d;
    p.items_per = w.per;
  }
  const int ns = precision == GSH_BF16X6 ? 3 : 2;
  GSH_HIP(timed(g, GSH_KERNEL_LEGENDRE_ANALYSIS, s,
                [&] {
                  return gsh::launch_leg_anal_bf16(p, ns, s,
                                                   pm ? 4 : (wide ? 1 : 0) + (sym ? 2 : 0));
                }));
  return GSH_OK;
}

}  // namespace

int gsh_inverse(gsh_grid_t g, int lmax, int n, int value_kind, int precision, int storage,
                int batch, const void* coeffs, long coeff_stride, void* field, long field_stride,
                void* stream) {
  int st = check_call(g, lmax, n, value_kind, precision, storage, batch);
  if (st) return st;
  if (batch == 0) return GSH_OK;
  if (!coeffs || !field) return GSH_ERR_INVALID_ARGUMENT;
  if (storage != GSH_STORE_FP64) return GSH_ERR_UNSUPPORTED;
  const int real = value_kind == GSH_REAL;
  const long csize = gsh_coefficient_size(lmax, n, value_kind);
  const long fsize = static_cast<long>(g->ntheta()) * g->nphi();
  if (batch > 1 && (coeff_stride < csize || field_stride < fsize)) return GSH_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(g->mu);
  GSH_HIP(hipSetDevice(g->device));
  if (lmax == 0) {
    GSH_HIP(gsh::launch_lmax0(1, real, batch, static_cast<const double*>(coeffs), coeff_stride,
                              static_cast<double*>(field), field_stride, s));
    return GSH_OK;
  }
  NData* d;
  if ((st = ensure_ndata(g, n, &d))) return st;
  return run_inverse(g, d, lmax, n, real, precision, batch, static_cast<const double*>(coeffs),
                     coeff_stride, field, field_stride, s, RunOpts{});
}

static int forward_impl(gsh_grid_t g, int lmax, int n, int value_kind, int precision, int storage,
                        int batch, const void* field, long field_stride, void* coeffs,
                        long coeff_stride, int accumulate, void* stream, int adjoint) {
  int st = check_call(g, lmax, n, value_kind, precision, storage, batch);
  if (st) return st;
  if (batch == 0) return GSH_OK;
  if (!coeffs || !field) return GSH_ERR_INVALID_ARGUMENT;
  if (storage != GSH_STORE_FP64) return GSH_ERR_UNSUPPORTED;
  const int real = value_kind == GSH_REAL;
  const long csize = gsh_coefficient_size(lmax, n, value_kind);
  const long fsize = static_cast<long>(g->ntheta()) * g->nphi();
  if (batch > 1 && (coeff_stride < csize || field_stride < fsize)) return GSH_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(g->mu);
  GSH_HIP(hipSetDevice(g->device));
  if (lmax == 0) {
    GSH_HIP(gsh::launch_lmax0(adjoint ? 2 : 0, real, batch, static_cast<const double*>(field),
                              field_stride, static_cast<double*>(coeffs), coeff_stride, s));
    return GSH_OK;
  }
  NData* d;
  if ((st = ensure_ndata(g, n, &d))) return st;
  return run_forward(g, d, lmax, n, real, precision, batch, field, field_stride,
                     static_cast<double*>(coeffs), coeff_stride, accumulate, adjoint, s,
                     RunOpts{});
}

int gsh_forward(gsh_grid_t g, int lmax, int n, int value_kind, int precision, int storage,
                int batch, const void* field, long field_stride, void* coeffs, long coeff_stride,
                int accumulate, void* stream) {
  return forward_impl(g, lmax, n, value_kind, precision, storage, batch, field, field_stride,
                      coeffs, coeff_stride, accumulate, stream, 0);
}

// Integrate (ScalarFieldOverloads.h:288-301) and the L2 inner product
// (:303-316) with the grid's weights (GridBase.h:60-66): ring sums on the
// device, then each field's sum over the colatitudes in order.
namespace {
// Argument checks and launch parameters shared by gsh_integrate and
// gsh_integrate_noalloc (everything but the ring-sum buffer p.part).
int integrate_params(gsh_grid_t g, int value_kind, int batch, const void* field_a,
                     const void* field_b, long field_stride, double* out,
                     gsh::GridOpParams* p) {
  if (!g || batch < 0 || (value_kind != GSH_REAL && value_kind != GSH_COMPLEX))
    return GSH_ERR_INVALID_ARGUMENT;
  if (value_kind == GSH_COMPLEX && g->mrange != GSH_ALL) return GSH_ERR_VALUE_KIND;
  if (batch == 0) return GSH_OK;  // nothing to do (the callers return here too)
  const long fsize = static_cast<long>(g->ntheta()) * g->nphi();
  if (!field_a || !out || (batch > 1 && field_stride < fsize)) return GSH_ERR_INVALID_ARGUMENT;
  if (g->L == 0) return GSH_ERR_UNSUPPORTED;  // no longitudes
  *p = gsh::GridOpParams{};
  p->ntheta = g->ntheta();
  p->nphi = g->nphi();
  p->batch = batch;
  p->real = value_kind == GSH_REAL;
  p->fstride = batch > 1 ? field_stride : fsize;
  p->fa = field_a;
  p->fb = field_b;
  p->wq = g->d_wq;  // w_i 2 pi / nphi: the colatitude weight times the longitude weight
  p->dphi = 1.0;
  p->out = out;
  return GSH_OK;
}
}  // namespace

int gsh_integrate(gsh_grid_t g, int value_kind, int batch, const void* field_a,
                  const void* field_b, long field_stride, double* out, void* stream) {
  gsh::GridOpParams p;
  int st = integrate_params(g, value_kind, batch, field_a, field_b, field_stride, out, &p);
  if (st || batch == 0) return st;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(g->mu);
  GSH_HIP(hipSetDevice(g->device));
  if ((st = ensure_work(g, sizeof(double2) * (size_t)batch * g->ntheta()))) return st;
  p.part = reinterpret_cast<double2*>(g->work);
  GSH_HIP(gsh::launch_integrate(p, s));
  return GSH_OK;
}

// The same with the ring sums in the library's static scratch array: no
// allocation, no host synchronisation, two kernel launches on `stream`.
int gsh_integrate_noalloc(gsh_grid_t g, int value_kind, int batch, const void* field_a,
                          const void* field_b, long field_stride, double* out, void* stream) {
  gsh::GridOpParams p;
  int st = integrate_params(g, value_kind, batch, field_a, field_b, field_stride, out, &p);
  if (st || batch == 0) return st;
  if (static_cast<long>(batch) * g->ntheta() > gsh::kNoAllocRingSums)
    return GSH_ERR_SCRATCH_LIMIT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(g->mu);
  GSH_HIP(hipSetDevice(g->device));
  // Every entry the second kernel reads, part[f * ntheta + i] for f < batch
  // and i < ntheta, is written by the first, so nothing is cleared first.
  p.part = gsh::noalloc_scratch(g->device);
  if (!p.part) return GSH_ERR_DEVICE;  // set up by gsh_grid_create
  GSH_HIP(gsh::launch_integrate(p, s));
  return GSH_OK;
}

// Seeded coefficients generated on the device: the counter-based stream of
// gshtrans_amd.random_coefficients (the reference's generator, GridBase.h:
// 115-172, draws from std::random_device), with its masks for the grid's L.
int gsh_random_coefficients(gsh_grid_t g, int lmax, int n, int value_kind,
                            unsigned long long seed, double decay, int batch, void* coeffs,
                            long coeff_stride, void* stream) {
  int st = check_call(g, lmax, n, value_kind, GSH_FP64, GSH_STORE_FP64, batch);
  if (st) return st;
  if (batch == 0) return GSH_OK;
  const long csize = gsh_coefficient_size(lmax, n, value_kind);
  if (!coeffs || (batch > 1 && coeff_stride < csize)) return GSH_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(g->mu);
  GSH_HIP(hipSetDevice(g->device));
  gsh::RandomParams p{};
  p.lmax = lmax;
  p.grid_lmax = g->L;
  p.nabs = std::abs(n);
  p.real = value_kind == GSH_REAL;
  p.batch = batch;
  p.size = csize;
  p.cstride = batch > 1 ? coeff_stride : csize;
  p.seed = seed;
  p.decay = decay;
  p.out = static_cast<double*>(coeffs);
  GSH_HIP(gsh::launch_random_coeffs(p, s));
  return GSH_OK;
}

// Analysis of a pointwise expression of one or two fields, formed as the
// ring FFT loads its rings (the expression fields of
// CanonicalComponentFieldBinary.h:17-308 and CanonicalComponentFieldUnary.h:
// 17-357, built by CanonicalComponentFieldOverloads.h:20-266).  value_kind is
// the input fields' kind; Re and Im of complex fields are real fields.
int gsh_forward_expr(gsh_grid_t g, int lmax, int n, int value_kind, int op, int precision,
                     int batch, const void* field_a, const void* field_b, long field_stride,
                     double scalar_re, double scalar_im, void* coeffs, long coeff_stride,
                     int accumulate, void* stream) {
  int st = check_call(g, lmax, n, value_kind, precision, GSH_STORE_FP64, batch);
  if (st) return st;
  if (op < GSH_EXPR_PRODUCT || op > GSH_EXPR_SHIFT) return GSH_ERR_INVALID_ARGUMENT;
  if (batch == 0) return GSH_OK;
  if (!coeffs || !field_a || (gsh::expr_binary(op) && !field_b)) return GSH_ERR_INVALID_ARGUMENT;
  if (value_kind == GSH_REAL && scalar_im != 0.0) return GSH_ERR_INVALID_ARGUMENT;
  if (lmax == 0) return GSH_ERR_UNSUPPORTED;  // the reference's lMax = 0 branch reads one value
  const bool src_cplx = value_kind == GSH_COMPLEX && (op == GSH_EXPR_REAL || op == GSH_EXPR_IMAG);
  const int real = value_kind == GSH_REAL || src_cplx;
  const long csize = gsh_coefficient_size(lmax, n, real ? GSH_REAL : GSH_COMPLEX);
  const long fsize = static_cast<long>(g->ntheta()) * g->nphi();
  if (batch > 1 && (coeff_stride < csize || field_stride < fsize)) return GSH_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(g->mu);
  GSH_HIP(hipSetDevice(g->device));
  NData* d;
  if ((st = ensure_ndata(g, n, &d))) return st;
  RunOpts o;
  o.field2 = gsh::expr_binary(op) ? field_b : nullptr;
  o.op = op;
  o.src_cplx = src_cplx ? 1 : 0;
  o.scalar = make_double2(scalar_re, scalar_im);
  // Re / Im of complex fields: the wavefront FFT reads two complex samples
  // per real pair; the LDS kernels handle every form.
  return run_forward(g, d, lmax, n, real, precision, batch, field_a, field_stride,
                     static_cast<double*>(coeffs), coeff_stride, accumulate, 0, s, o);
}

// Analysis of the pointwise product of two fields (the product field of
// CanonicalComponentFieldBinary.h:203-253 and CanonicalComponentFieldOverloads.h:
// 207-233, upper index n = N0 + N1), the product formed as the ring FFT loads
// its rings.
int gsh_forward_product(gsh_grid_t g, int lmax, int n, int value_kind, int precision, int batch,
                        const void* field_a, const void* field_b, long field_stride,
                        void* coeffs, long coeff_stride, int accumulate, void* stream) {
  return gsh_forward_expr(g, lmax, n, value_kind, GSH_EXPR_PRODUCT, precision, batch, field_a,
                          field_b, field_stride, 0.0, 0.0, coeffs, coeff_stride, accumulate,
                          stream);
}

int gsh_timing_enable(gsh_grid_t g, int enable) {
  if (!g) return GSH_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(g->mu);
  for (int k = 0; k < GSH_KERNEL_COUNT; k++) {
    for (auto& pr : g->pending[k]) {
      (void)hipEventSynchronize(pr.second);
      g->event_pool.push_back(pr.first);
      g->event_pool.push_back(pr.second);
    }
    g->pending[k].clear();
    g->total_ms[k] = 0;
    g->launches[k] = 0;
  }
  g->timing = enable != 0;
  return GSH_OK;
}

int gsh_timing_read(gsh_grid_t g, double* total_ms, long* launches, int nkernels) {
  if (!g || nkernels < 0) return GSH_ERR_INVALID_ARGUMENT;
  std::lock_guard<std::mutex> lock(g->mu);
  for (int k = 0; k < GSH_KERNEL_COUNT; k++) {
    for (auto& pr : g->pending[k]) {
      GSH_HIP(hipEventSynchronize(pr.second));
      float ms = 0;
      GSH_HIP(hipEventElapsedTime(&ms, pr.first, pr.second));
      g->total_ms[k] += ms;
      g->launches[k] += 1;
      g->event_pool.push_back(pr.first);
      g->event_pool.push_back(pr.second);
    }
    g->pending[k].clear();
  }
  for (int k = 0; k < nkernels && k < GSH_KERNEL_COUNT; k++) {
    if (total_ms) total_ms[k] = g->total_ms[k];
    if (launches) launches[k] = g->launches[k];
  }
  return GSH_OK;
}

int gsh_adjoint_inverse(gsh_grid_t g, int lmax, int n, int value_kind, int precision, int storage,
                        int batch, const void* field, long field_stride, void* coeffs,
                        long coeff_stride, void* stream) {
  return forward_impl(g, lmax, n, value_kind, precision, storage, batch, field, field_stride,
                      coeffs, coeff_stride, 0, stream, 1);
}

static int stage_ensure(gsh_grid_s* g, size_t bytes) {
  if (bytes <= g->stage_bytes) return GSH_OK;
  if (g->stage) GSH_HIP(hipFree(g->stage));
  g->stage = nullptr;
  g->stage_bytes = 0;
  if (hipMalloc(&g->stage, bytes) != hipSuccess) return GSH_ERR_OUT_OF_MEMORY;
  g->stage_bytes = bytes;
  return GSH_OK;
}

int gsh_inverse_host(gsh_grid_t g, int lmax, int n, int value_kind, const double* coeffs,
                     double* field) {
  int st = check_call(g, lmax, n, value_kind, GSH_FP64, GSH_STORE_FP64, 1);
  if (st) return st;
  if (!coeffs || !field) return GSH_ERR_INVALID_ARGUMENT;
  const size_t cbytes = 16 * (size_t)gsh_coefficient_size(lmax, n, value_kind);
  const size_t fbytes = (value_kind == GSH_REAL ? 8 : 16) * (size_t)g->ntheta() * g->nphi();
  {
    std::lock_guard<std::mutex> lock(g->mu);
    GSH_HIP(hipSetDevice(g->device));
    if ((st = stage_ensure(g, cbytes + fbytes))) return st;
    GSH_HIP(hipMemcpy(g->stage, coeffs, cbytes, hipMemcpyHostToDevice));
    // lMax == 0 writes element 0 only; keep the rest of the caller's field.
    if (lmax == 0) GSH_HIP(hipMemcpy(reinterpret_cast<char*>(g->stage) + cbytes, field, fbytes,
                                     hipMemcpyHostToDevice));
  }
  char* dfield = reinterpret_cast<char*>(g->stage) + cbytes;
  st = gsh_inverse(g, lmax, n, value_kind, GSH_FP64, GSH_STORE_FP64, 1, g->stage, 0, dfield, 0,
                   nullptr);
  if (st) return st;
  GSH_HIP(hipMemcpy(field, dfield, fbytes, hipMemcpyDeviceToHost));
  return GSH_OK;
}

int gsh_forward_host(gsh_grid_t g, int lmax, int n, int value_kind, const double* field,
                     double* coeffs, int accumulate) {
  int st = check_call(g, lmax, n, value_kind, GSH_FP64, GSH_STORE_FP64, 1);
  if (st) return st;
  if (!coeffs || !field) return GSH_ERR_INVALID_ARGUMENT;
  const size_t cbytes = 16 * (size_t)gsh_coefficient_size(lmax, n, value_kind);
  const size_t fbytes = (value_kind == GSH_REAL ? 8 : 16) * (size_t)g->ntheta() * g->nphi();
  {
    std::lock_guard<std::mutex> lock(g->mu);
    GSH_HIP(hipSetDevice(g->device));
    if ((st = stage_ensure(g, cbytes + fbytes))) return st;
    GSH_HIP(hipMemcpy(g->stage, field, fbytes, hipMemcpyHostToDevice));
    GSH_HIP(hipMemcpy(reinterpret_cast<char*>(g->stage) + fbytes, coeffs, cbytes,
                      hipMemcpyHostToDevice));
  }
  char* dco = reinterpret_cast<char*>(g->stage) + fbytes;
  st = gsh_forward(g, lmax, n, value_kind, GSH_FP64, GSH_STORE_FP64, 1, g->stage, 0, dco, 0,
                   accumulate, nullptr);
  if (st) return st;
  GSH_HIP(hipMemcpy(coeffs, dco, cbytes, hipMemcpyDeviceToHost));
  return GSH_OK;
}

}  // extern "C"

// ---------------------------------------------------------------------------
// Batch plans (gsh_plan_*): many fields of uneven truncation degree and upper
// index.  Fields are grouped by upper index; within a group they are ordered
// by descending degree, so each 16-column group of a launch has a degree
// bound close to its fields' own (cg_lmax) and every column masks its own
// degrees (f_lmax).  The first execution in each direction times the
// candidate tile shapes and keeps the fastest; graph executions capture the
// whole launch sequence once per pointer set and replay it.
// ---------------------------------------------------------------------------
struct PlanGroup {
  int n = 0, lmax = 0, count = 0;
  std::vector<int> fmap, flmax, cglmax;
  int* d_fmap = nullptr;
  int* d_flmax = nullptr;
  int* d_cglmax = nullptr;
  NData* nd = nullptr;
  int synth_tiles = 0;  // FP64 synthesis column tiles (0: not tuned)
  int fft_inv = 0;      // 1: wavefront FFT, 2: LDS FFT (0: not tuned)
  int fft_fwd = 0;
  int anal_pass = 0;    // split-BF16 analysis degrees per pass, 64 or 128 (0: not tuned)
};

struct GraphKey {
  int dir;
  const void* in;
  long istride;
  const void* out;
  long ostride;
  bool operator<(const GraphKey& o) const {
    return std::tie(dir, in, istride, out, ostride) <
           std::tie(o.dir, o.in, o.istride, o.out, o.ostride);
  }
};

// A captured launch sequence.  It embeds the grid's work buffers and the
// plan's staging buffer as they were at capture, so it is replayed only while
// both generations are unchanged (a larger call on the grid, or a larger
// stride on an FP32 plan, reallocates them); otherwise it is re-captured.
struct PlanGraph {
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  unsigned long long work_gen = 0, stage_gen = 0;
  unsigned long long last_use = 0;
};

// Graphs kept per plan (least recently used evicted first).
constexpr size_t kPlanGraphCap = 8;

struct gsh_plan_s {
  gsh_grid_s* g = nullptr;  // null once the grid has been destroyed
  int count = 0, kind = 0, precision = 0, storage = 0;
  std::vector<int> lmax, n;
  long cmax = 0;  // largest coefficient count of a field
  std::vector<PlanGroup> groups;
  bool tuned[2] = {false, false};
  double* stage = nullptr;  // FP64 coefficients of FP32-storage plans
  size_t stage_bytes = 0;
  unsigned long long stage_gen = 0;  // incremented whenever stage is (re)allocated
  long* d_csize = nullptr;  // coefficient count per field
  hipStream_t cap = nullptr;
  std::map<GraphKey, PlanGraph> graphs;
  unsigned long long clock = 0;  // graph use counter (LRU)
};

namespace {

void drop_graph(PlanGraph& G) {
  if (G.exec) (void)hipGraphExecDestroy(G.exec);
  if (G.graph) (void)hipGraphDestroy(G.graph);
  G.exec = nullptr;
  G.graph = nullptr;
}

// Releases every device resource of a plan (its grid's lock held, device
// current); the plan object itself stays.
void plan_release(gsh_plan_s* P) {
  (void)hipDeviceSynchronize();
  for (auto& kv : P->graphs) drop_graph(kv.second);
  P->graphs.clear();
  for (PlanGroup& G : P->groups) {
    (void)hipFree(G.d_fmap);
    (void)hipFree(G.d_flmax);
    (void)hipFree(G.d_cglmax);
    G.d_fmap = G.d_flmax = G.d_cglmax = nullptr;
  }
  if (P->stage) (void)hipFree(P->stage);
  P->stage = nullptr;
  P->stage_bytes = 0;
  if (P->d_csize) (void)hipFree(P->d_csize);
  P->d_csize = nullptr;
  if (P->cap) (void)hipStreamDestroy(P->cap);
  P->cap = nullptr;
}

void plan_detach(gsh_plan_s* P) {
  plan_release(P);
  P->g = nullptr;
}

int plan_stage(gsh_plan_s* P, long cstride) {
  const size_t bytes = 16 * (size_t)P->count * cstride;
  if (bytes <= P->stage_bytes) return GSH_OK;
  if (P->stage) {
    GSH_HIP(hipDeviceSynchronize());
    GSH_HIP(hipFree(P->stage));
  }
  P->stage = nullptr;
  P->stage_bytes = 0;
  P->stage_gen++;
  if (hipMalloc(&P->stage, bytes) != hipSuccess) return GSH_ERR_OUT_OF_MEMORY;
  P->stage_bytes = bytes;
  return GSH_OK;
}

RunOpts group_opts(const gsh_plan_s* P, const PlanGroup& G, int dir) {
  RunOpts o;
  o.fmap = G.d_fmap;
  o.f_lmax = G.d_flmax;
  o.cg_lmax = G.d_cglmax;
  o.cg_lmax_host = G.cglmax;
  o.f32 = P->storage == GSH_STORE_FP32;
  o.synth_tiles = G.synth_tiles;
  o.fft = dir == 0 ? G.fft_inv : G.fft_fwd;
  o.anal_pass = dir == 0 ? 0 : G.anal_pass;
  return o;
}

// One group in one direction (dir 0 synthesis, 1 analysis).
int run_group(gsh_plan_s* P, const PlanGroup& G, int dir, const double* cin, double* cout,
              long cstride, const void* fin, void* fout, long fstride, hipStream_t s,
              const RunOpts& o) {
  const int real = P->kind == GSH_REAL;
  if (dir == 0)
    return run_inverse(P->g, G.nd, G.lmax, G.n, real, P->precision, G.count, cin, cstride, fout,
                       fstride, s, o);
  return run_forward(P->g, G.nd, G.lmax, G.n, real, P->precision, G.count, fin, fstride, cout,
                     cstride, 0, 0, s, o);
}

// The whole plan in one direction on stream s.
int run_plan(gsh_plan_s* P, int dir, const void* in, long istride, void* out, long ostride,
             hipStream_t s) {
  int st;
  const bool f32 = P->storage == GSH_STORE_FP32;
  if (dir == 0) {
    const double* cin = static_cast<const double*>(in);
    if (f32) {
      if ((st = plan_stage(P, istride))) return st;
      GSH_HIP(gsh::launch_widen_c32(in, P->stage, (long)P->count * istride, s));
      cin = P->stage;
    }
    for (const PlanGroup& G : P->groups)
      if ((st = run_group(P, G, 0, cin, nullptr, istride, nullptr, out, ostride, s,
                          group_opts(P, G, 0))))
        return st;
    return GSH_OK;
  }
  double* cout = static_cast<double*>(out);
  if (f32) {
    if ((st = plan_stage(P, ostride))) return st;
    cout = P->stage;
  }
  for (const PlanGroup& G : P->groups)
    if ((st = run_group(P, G, 1, nullptr, cout, ostride, in, nullptr, istride, s,
                        group_opts(P, G, 1))))
      return st;
  if (f32)
    GSH_HIP(gsh::launch_narrow_c64(P->stage, out, ostride, P->d_csize, P->cmax, P->count, s));
  return GSH_OK;
}

// Tuned choices by launch shape: the key names the grid degree, value kind,
// precision, storage, direction, upper index, field count and the fields'
// degrees (hashed, in the group's order), so a later plan of the same shape
// takes the choices without timing.  Kept for the life of the process and,
// when the environment variable GSH_TUNE_CACHE names a file, read from and
// appended to that file (one "key tiles fft pass" line per shape;
// unreadable lines are skipped).
std::mutex tune_mu;
// Choices per shape: synthesis tiles, FFT family, analysis pass length.
struct TuneChoice {
  int tiles = 0, fft = 0, pass = 0;
};
std::map<std::string, TuneChoice> tune_memo;
std::string tune_file_loaded;

std::string tune_key(const gsh_plan_s* P, const PlanGroup& G, int dir) {
  uint64_t h = 1469598103934665603ULL;  // FNV-1a over the degrees
  for (int l : G.flmax) {
    h ^= static_cast<uint64_t>(l) & 0xffffffffULL;
    h *= 1099511628211ULL;
  }
  std::ostringstream k;
  k << "gsh2:L" << P->g->L << ":k" << P->kind << ":p" << P->precision << ":s" << P->storage
    << ":d" << dir << ":n" << G.n << ":c" << G.count << ":h" << std::hex << h;
  return k.str();
}

// Caller holds tune_mu.
void tune_load_file() {
  const char* path = std::getenv("GSH_TUNE_CACHE");
  if (!path || !*path || tune_file_loaded == path) return;
  tune_file_loaded = path;
  std::ifstream f(path);
  std::string line;
  while (std::getline(f, line)) {
    std::istringstream is(line);
    std::string key;
    TuneChoice c{-1, -1, -1};
    if ((is >> key >> c.tiles >> c.fft >> c.pass) && key.rfind("gsh2:", 0) == 0 &&
        c.tiles >= 0 && c.tiles <= 2 && c.fft >= 1 && c.fft <= 2 &&
        (c.pass == 0 || c.pass == gsh::kPass || c.pass == gsh::kWideAnalPass))
      tune_memo[key] = c;
  }
}

bool tune_lookup(const std::string& key, TuneChoice* out) {
  std::lock_guard<std::mutex> lock(tune_mu);
  tune_load_file();
  auto it = tune_memo.find(key);
  if (it == tune_memo.end()) return false;
  *out = it->second;
  return true;
}

void tune_store(const std::string& key, const TuneChoice& v) {
  std::lock_guard<std::mutex> lock(tune_mu);
  tune_memo[key] = v;
  const char* path = std::getenv("GSH_TUNE_CACHE");
  if (!path || !*path) return;
  std::ofstream f(path, std::ios::app);
  if (f) f << key << " " << v.tiles << " " << v.fft << " " << v.pass << "\n";
}

// Autotuner: per group, run each candidate (FP64 synthesis tiles, FFT family,
// and for the split-BF16 analysis 64- or 128-degree passes) once to warm up
// and once timed with events, and keep the fastest; a shape tuned before
// (this process, or the GSH_TUNE_CACHE file) is not timed again.
int tune_plan(gsh_plan_s* P, int dir, const void* in, long istride, void* out, long ostride,
              hipStream_t s) {
  int st;
  const bool f32 = P->storage == GSH_STORE_FP32;
  const double* cin = static_cast<const double*>(in);
  double* cout = static_cast<double*>(out);
  if (f32) {
    if ((st = plan_stage(P, dir == 0 ? istride : ostride))) return st;
    if (dir == 0) GSH_HIP(gsh::launch_widen_c32(in, P->stage, (long)P->count * istride, s));
    cin = P->stage;
    cout = P->stage;
  }
  hipEvent_t e0, e1;
  GSH_HIP(hipEventCreate(&e0));
  GSH_HIP(hipEventCreate(&e1));
  const bool was_timing = P->g->timing;
  P->g->timing = false;
  for (PlanGroup& G : P->groups) {
    const std::string key = tune_key(P, G, dir);
    TuneChoice known;
    if (tune_lookup(key, &known)) {
      if (dir == 0) {
        G.synth_tiles = known.tiles;
        G.fft_inv = known.fft;
      } else {
        G.fft_fwd = known.fft;
        G.anal_pass = known.pass;
      }
      continue;
    }
    std::vector<TuneChoice> cands;
    const bool wave_ok = P->g->d_twa != nullptr;
    const bool fp = P->precision == GSH_FP64 || P->precision == GSH_FP32;
    const bool tiles2 = dir == 0 && fp && G.count > 8;
    const std::vector<int> passes =
        dir == 1 && !fp ? std::vector<int>{gsh::kPass, gsh::kWideAnalPass} : std::vector<int>{0};
    for (int t : tiles2 ? std::vector<int>{1, 2} : std::vector<int>{0})
      for (int f : wave_ok ? std::vector<int>{1, 2} : std::vector<int>{2})
        for (int pl : passes) cands.push_back(TuneChoice{t, f, pl});
    float best = 1e30f;
    TuneChoice pick = cands[0];
    for (const auto& c : cands) {
      RunOpts o = group_opts(P, G, dir);
      o.synth_tiles = c.tiles;
      o.fft = c.fft;
      o.anal_pass = c.pass;
      for (int rep = 0; rep < 2; rep++) {
        if (rep == 1) GSH_HIP(hipEventRecord(e0, s));
        if ((st = run_group(P, G, dir, cin, cout, dir == 0 ? istride : ostride, in, out,
                            dir == 0 ? ostride : istride, s, o))) {
          P->g->timing = was_timing;
          return st;
        }
      }
      GSH_HIP(hipEventRecord(e1, s));
      GSH_HIP(hipEventSynchronize(e1));
      float ms = 0;
      GSH_HIP(hipEventElapsedTime(&ms, e0, e1));
      if (ms < best) {
        best = ms;
        pick = c;
      }
    }
    if (dir == 0) {
      G.synth_tiles = pick.tiles;
      G.fft_inv = pick.fft;
    } else {
      G.fft_fwd = pick.fft;
      G.anal_pass = pick.pass;
    }
    tune_store(key, pick);
  }
  P->g->timing = was_timing;
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  P->tuned[dir] = true;
  return GSH_OK;
}

int plan_exec(gsh_plan_s* P, int dir, const void* in, long istride, void* out, long ostride,
              int graph, void* stream) {
  if (!P || !in || !out) return GSH_ERR_INVALID_ARGUMENT;
  gsh_grid_s* g = P->g;
  if (!g) return GSH_ERR_INVALID_ARGUMENT;  // its grid has been destroyed
  const long fsize = (long)g->ntheta() * g->nphi();
  const long cs = dir == 0 ? istride : ostride;
  const long fs = dir == 0 ? ostride : istride;
  if (cs < P->cmax || fs < fsize) return GSH_ERR_INVALID_ARGUMENT;
  hipStream_t s = static_cast<hipStream_t>(stream);
  std::lock_guard<std::mutex> lock(g->mu);
  GSH_HIP(hipSetDevice(g->device));
  int st;
  const GraphKey key{dir, in, istride, out, ostride};
  if (graph) {
    auto it = P->graphs.find(key);
    if (it != P->graphs.end()) {
      PlanGraph& G = it->second;
      if (G.work_gen == g->work_gen && G.stage_gen == P->stage_gen) {
        G.last_use = ++P->clock;
        GSH_HIP(hipGraphLaunch(G.exec, s));
        return GSH_OK;
      }
      // A buffer the graph embeds has been reallocated since its capture:
      // the graph is stale and is captured again below.
      GSH_HIP(hipStreamSynchronize(s));
      drop_graph(G);
      P->graphs.erase(it);
    }
  }
  if (!P->tuned[dir] && (st = tune_plan(P, dir, in, istride, out, ostride, s))) return st;
  // Eager run: the result, and every work buffer sized before any capture.
  if ((st = run_plan(P, dir, in, istride, out, ostride, s))) return st;
  if (!graph) return GSH_OK;
  if (!P->cap) GSH_HIP(hipStreamCreateWithFlags(&P->cap, hipStreamNonBlocking));
  GSH_HIP(hipStreamSynchronize(s));
  const unsigned long long wgen = g->work_gen, sgen = P->stage_gen;
  const bool was_timing = g->timing;
  g->timing = false;
  hipGraph_t gr = nullptr;
  GSH_HIP(hipStreamBeginCapture(P->cap, hipStreamCaptureModeThreadLocal));
  st = run_plan(P, dir, in, istride, out, ostride, P->cap);
  const hipError_t ec = hipStreamEndCapture(P->cap, &gr);
  g->timing = was_timing;
  if (st) {
    if (gr) (void)hipGraphDestroy(gr);
    return st;
  }
  if (ec != hipSuccess) return GSH_ERR_DEVICE;
  // The eager run sized every buffer, so the capture must not have
  // reallocated one; if it had, the graph would embed freed memory.
  if (g->work_gen != wgen || P->stage_gen != sgen) {
    (void)hipGraphDestroy(gr);
    return GSH_ERR_DEVICE;
  }
  hipGraphExec_t ex = nullptr;
  if (hipGraphInstantiate(&ex, gr, nullptr, nullptr, 0) != hipSuccess) {
    (void)hipGraphDestroy(gr);
    return GSH_ERR_DEVICE;
  }
  if (P->graphs.size() >= kPlanGraphCap) {
    auto lru = P->graphs.begin();
    for (auto jt = P->graphs.begin(); jt != P->graphs.end(); ++jt)
      if (jt->second.last_use < lru->second.last_use) lru = jt;
    drop_graph(lru->second);
    P->graphs.erase(lru);
  }
  PlanGraph G;
  G.graph = gr;
  G.exec = ex;
  G.work_gen = wgen;
  G.stage_gen = sgen;
  G.last_use = ++P->clock;
  P->graphs.emplace(key, G);
  return GSH_OK;
}

}  // namespace

namespace {

// Groups of a new plan by upper index, with their device arrays.
int plan_build(gsh_grid_s* g, gsh_plan_s* P, const std::map<int, std::vector<int>>& byn,
               const int* lmax, const std::vector<long>& csize) {
  int st;
  const int count = P->count;
  for (auto& kv : byn) {
    PlanGroup G;
    G.n = kv.first;
    G.fmap = kv.second;
    std::stable_sort(G.fmap.begin(), G.fmap.end(),
                     [&](int a, int b) { return lmax[a] > lmax[b]; });
    G.count = static_cast<int>(G.fmap.size());
    for (int f : G.fmap) G.flmax.push_back(lmax[f]);
    G.lmax = G.flmax[0];
    const int ncg = (2 * G.count + 15) / 16;
    std::vector<int> cg(ncg);
    for (int c = 0; c < ncg; c++) cg[c] = G.flmax[8 * c];  // descending: first is largest
    G.cglmax = cg;
    if ((st = ensure_ndata(g, G.n, &G.nd))) return st;
    // Held by the plan before its arrays are allocated, so a failure below
    // leaves them to plan_release.
    P->groups.push_back(std::move(G));
    PlanGroup& H = P->groups.back();
    GSH_HIP(hipMalloc(&H.d_fmap, H.count * sizeof(int)));
    GSH_HIP(hipMalloc(&H.d_flmax, H.count * sizeof(int)));
    GSH_HIP(hipMalloc(&H.d_cglmax, ncg * sizeof(int)));
    GSH_HIP(hipMemcpy(H.d_fmap, H.fmap.data(), H.count * sizeof(int), hipMemcpyHostToDevice));
    GSH_HIP(hipMemcpy(H.d_flmax, H.flmax.data(), H.count * sizeof(int), hipMemcpyHostToDevice));
    GSH_HIP(hipMemcpy(H.d_cglmax, H.cglmax.data(), ncg * sizeof(int), hipMemcpyHostToDevice));
  }
  GSH_HIP(hipMalloc(&P->d_csize, count * sizeof(long)));
  GSH_HIP(hipMemcpy(P->d_csize, csize.data(), count * sizeof(long), hipMemcpyHostToDevice));
  return GSH_OK;
}

}  // namespace

extern "C" {

int gsh_plan_create(gsh_grid_t g, int count, const int* lmax, const int* n, int value_kind,
                    int precision, int storage, gsh_plan_t* plan) {
  if (!g || !plan || count <= 0 || !lmax || !n) return GSH_ERR_INVALID_ARGUMENT;
  *plan = nullptr;
  int st;
  for (int i = 0; i < count; i++) {
    if (lmax[i] < 1) return GSH_ERR_DEGREE;
    if ((st = check_call(g, lmax[i], n[i], value_kind, precision, storage, 1))) return st;
  }
  auto P = std::make_unique<gsh_plan_s>();
  P->g = g;
  P->count = count;
  P->kind = value_kind;
  P->precision = precision;
  P->storage = storage;
  P->lmax.assign(lmax, lmax + count);
  P->n.assign(n, n + count);
  std::vector<long> csize(count);
  for (int i = 0; i < count; i++) {
    csize[i] = gsh_coefficient_size(lmax[i], n[i], value_kind);
    P->cmax = std::max(P->cmax, csize[i]);
  }
  std::map<int, std::vector<int>> byn;
  for (int i = 0; i < count; i++) byn[n[i]].push_back(i);
  std::lock_guard<std::mutex> lock(g->mu);
  GSH_HIP(hipSetDevice(g->device));
  // A failure part-way releases what the plan has allocated so far.
  st = plan_build(g, P.get(), byn, lmax, csize);
  if (st) {
    plan_release(P.get());
    return st;
  }
  g->plans.push_back(P.get());
  *plan = P.release();
  return GSH_OK;
}

int gsh_plan_destroy(gsh_plan_t P) {
  if (!P) return GSH_ERR_INVALID_ARGUMENT;
  if (gsh_grid_s* g = P->g) {
    std::lock_guard<std::mutex> lock(g->mu);
    (void)hipSetDevice(g->device);
    plan_release(P);
    g->plans.erase(std::remove(g->plans.begin(), g->plans.end(), P), g->plans.end());
  }
  delete P;
  return GSH_OK;
}

int gsh_plan_inverse(gsh_plan_t P, const void* coeffs, long coeff_stride, void* field,
                     long field_stride, int graph, void* stream) {
  return plan_exec(P, 0, coeffs, coeff_stride, field, field_stride, graph, stream);
}

int gsh_plan_forward(gsh_plan_t P, const void* field, long field_stride, void* coeffs,
                     long coeff_stride, int graph, void* stream) {
  return plan_exec(P, 1, field, field_stride, coeffs, coeff_stride, graph, stream);
}

int gsh_plan_analysis_pass(gsh_plan_t P, int* ngroups, int cap, int* pass) {
  if (!P || !ngroups || !P->g) return GSH_ERR_INVALID_ARGUMENT;
  *ngroups = static_cast<int>(P->groups.size());
  for (int i = 0; i < cap && i < *ngroups; i++)
    if (pass) pass[i] = P->groups[i].anal_pass;
  return GSH_OK;
}

int gsh_plan_info(gsh_plan_t P, int* ngroups, int cap, int* n, int* lmax, int* count,
                  int* synth_tiles, int* fft_inverse, int* fft_forward) {
  if (!P || !ngroups || !P->g) return GSH_ERR_INVALID_ARGUMENT;
  *ngroups = static_cast<int>(P->groups.size());
  for (int i = 0; i < cap && i < *ngroups; i++) {
    const PlanGroup& G = P->groups[i];
    if (n) n[i] = G.n;
    if (lmax) lmax[i] = G.lmax;
    if (count) count[i] = G.count;
    if (synth_tiles) synth_tiles[i] = G.synth_tiles;
    if (fft_inverse) fft_inverse[i] = G.fft_inv;
    if (fft_forward) fft_forward[i] = G.fft_fwd;
  }
  return GSH_OK;
}

// Device storage of the facade's Device* containers (the layouts of
// CanonicalComponentField.h:74-81, 119-124 and CanonicalComponentExpansion.h:
// 69-78 in device memory).
int gsh_device_alloc(gsh_grid_t g, size_t bytes, void** ptr) {
  if (!g || !ptr) return GSH_ERR_INVALID_ARGUMENT;
  *ptr = nullptr;
  if (bytes == 0) return GSH_OK;
  GSH_HIP(hipSetDevice(g->device));
  void* p = nullptr;
  if (hipMalloc(&p, bytes) != hipSuccess) return GSH_ERR_OUT_OF_MEMORY;
  if (hipMemset(p, 0, bytes) != hipSuccess) {
    (void)hipFree(p);
    return GSH_ERR_DEVICE;
  }
  *ptr = p;
  return GSH_OK;
}

int gsh_device_free(gsh_grid_t g, void* ptr) {
  if (!g) return GSH_ERR_INVALID_ARGUMENT;
  if (!ptr) return GSH_OK;
  GSH_HIP(hipSetDevice(g->device));
  GSH_HIP(hipFree(ptr));
  return GSH_OK;
}

int gsh_copy(gsh_grid_t g, void* dst, const void* src, size_t bytes) {
  if (!g || (bytes && (!dst || !src))) return GSH_ERR_INVALID_ARGUMENT;
  if (bytes == 0) return GSH_OK;
  GSH_HIP(hipSetDevice(g->device));
  GSH_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDefault));
  return GSH_OK;
}

int gsh_synchronize(gsh_grid_t g, void* stream) {
  if (!g) return GSH_ERR_INVALID_ARGUMENT;
  GSH_HIP(hipSetDevice(g->device));
  GSH_HIP(hipStreamSynchronize(static_cast<hipStream_t>(stream)));
  return GSH_OK;
}

// ---- Wigner d tables (Wigner.h:99-286) ----------------------------------

namespace {

// GSHIndices<MRange>(lmax, mmax, n).Size() (Indexing.h:104-150).
long gsh_block_size(long lmax, long mmax, long n, bool all) {
  const long mm = std::min(lmax, mmax), nabs = std::labs(n);
  auto offset = [&](long l) {
    if (all) {
      if (mm >= nabs)
        return l <= mm ? l * l - nabs * nabs
                       : (mm + 1) * (mm + 1) - nabs * nabs + (l - 1 - mm) * (2 * mm + 1);
      return (l - nabs) * (2 * mm + 1);
    }
    if (mm >= nabs)
      return l <= mm ? (l * (l + 1)) / 2 - (nabs * (nabs + 1)) / 2
                     : ((mm + 1) * (mm + 2)) / 2 - (nabs * (nabs + 1)) / 2 + (l - 1 - mm) * (mm + 1);
    return (l - nabs) * (mm + 1);
  };
  return offset(lmax) + (all ? 2 * std::min(lmax, mm) + 1 : std::min(lmax, mm) + 1);
}

int wigner_check(int lmax, int mmax, int nmin, int nmax, int ntheta, int mrange) {
  if (lmax < 0 || mmax < 0 || ntheta < 1 || nmin > nmax) return GSH_ERR_INVALID_ARGUMENT;
  if (mrange != GSH_ALL && mrange != GSH_NONNEGATIVE) return GSH_ERR_INVALID_ARGUMENT;
  if (std::max(std::abs(nmin), std::abs(nmax)) > lmax) return GSH_ERR_UPPER_INDEX;
  return GSH_OK;
}

// Offsets of the (n, theta_i) blocks in the order of Wigner::Indices
// (Wigner.h:240-250): ColumnMajor n outer, RowMajor theta outer.
std::vector<long> wigner_offsets(int lmax, int mmax, int nmin, int nmax, int ntheta, bool all,
                                 int storage, long* total) {
  const int nn = nmax - nmin + 1;
  std::vector<long> size(nn), off((size_t)nn * ntheta);
  for (int ni = 0; ni < nn; ni++) size[ni] = gsh_block_size(lmax, mmax, nmin + ni, all);
  long acc = 0;
  if (storage == GSH_COLUMN_MAJOR) {
    for (int ni = 0; ni < nn; ni++)
      for (int i = 0; i < ntheta; i++) {
        off[(size_t)ni * ntheta + i] = acc;
        acc += size[ni];
      }
  } else {
    for (int i = 0; i < ntheta; i++)
      for (int ni = 0; ni < nn; ni++) {
        off[(size_t)ni * ntheta + i] = acc;
        acc += size[ni];
      }
  }
  *total = acc;
  return off;
}

}  // namespace

long gsh_wigner_size(int lmax, int mmax, int nmin, int nmax, int ntheta, int mrange) {
  if (wigner_check(lmax, mmax, nmin, nmax, ntheta, mrange)) return -1;
  long total = 0;
  for (int n = nmin; n <= nmax; n++) total += gsh_block_size(lmax, mmax, n, mrange == GSH_ALL);
  return total * ntheta;
}

long gsh_wigner_offset(int lmax, int mmax, int nmin, int nmax, int ntheta, int mrange,
                       int storage, int n, int itheta) {
  if (wigner_check(lmax, mmax, nmin, nmax, ntheta, mrange)) return -1;
  if (storage != GSH_COLUMN_MAJOR && storage != GSH_ROW_MAJOR) return -1;
  if (n < nmin || n > nmax || itheta < 0 || itheta >= ntheta) return -1;
  long total = 0;
  const auto off = wigner_offsets(lmax, mmax, nmin, nmax, ntheta, mrange == GSH_ALL, storage,
                                  &total);
  return off[(size_t)(n - nmin) * ntheta + itheta];
}

int gsh_wigner_table(int lmax, int mmax, int nmin, int nmax, int ntheta, const double* theta,
                     int mrange, int norm, int storage, double* table, void* stream) {
  int st = wigner_check(lmax, mmax, nmin, nmax, ntheta, mrange);
  if (st) return st;
  if (!theta || !table || (norm != GSH_ORTHO && norm != GSH_FOURPI) ||
      (storage != GSH_COLUMN_MAJOR && storage != GSH_ROW_MAJOR))
    return GSH_ERR_INVALID_ARGUMENT;
  int ndev = 0, dev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return GSH_ERR_NO_DEVICE;
  if (hipGetDevice(&dev) != hipSuccess) return GSH_ERR_NO_DEVICE;
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, dev) != hipSuccess) return GSH_ERR_NO_DEVICE;
  if (std::strncmp(prop.gcnArchName, "gfx950", 6) != 0) return GSH_ERR_NO_DEVICE;
  const bool all = mrange == GSH_ALL;
  gsh::WignerParams p{};
  p.lmax = lmax;
  p.mmax = std::min(lmax, mmax);
  p.mlo = all ? -p.mmax : 0;
  p.nm = p.mmax - p.mlo + 1;
  p.nmin = nmin;
  p.nn = nmax - nmin + 1;
  p.ntheta = ntheta;
  p.all = all ? 1 : 0;
  p.fourpi = norm == GSH_FOURPI ? 1 : 0;
  long total = 0;
  const auto off = wigner_offsets(lmax, mmax, nmin, nmax, ntheta, all, storage, &total);
  // Seeds (Ortho, at l0 = max(|m|, |n|)) and recurrence coefficients, as for
  // the grid's tables (ensure_ndata), for the caller's angles.
  std::vector<double> seeds((size_t)p.nn * ntheta * p.nm), cosv(ntheta), inv(lmax + 1);
  parallel_for((long)p.nn * ntheta, [&](long k) {
    const int ni = static_cast<int>(k / ntheta), i = static_cast<int>(k % ntheta);
    const long n = nmin + ni;
    SeedArgs a(theta[i]);
    for (int mi = 0; mi < p.nm; mi++) {
      const long m = p.mlo + mi;
      seeds[(size_t)k * p.nm + mi] = ortho_scale(std::max(std::labs(m), std::labs(n)),
                                                 seed_value(m, n, a));
    }
  });
  std::vector<double4> rec((size_t)p.nn * (lmax + 1) * p.nm, make_double4(0.0, 0.0, 0.0, 0.0));
  parallel_for((long)p.nn * p.nm, [&](long k) {
    const int ni = static_cast<int>(k / p.nm), mi = static_cast<int>(k % p.nm);
    const long n = nmin + ni, m = p.mlo + mi;
    const long l0 = std::max(std::labs(m), std::labs(n));
    for (long l = l0 + 1; l <= lmax; l++)
      rec[((size_t)ni * (lmax + 1) + l) * p.nm + mi] = rec_coef(l, m, n);
  });
  for (int i = 0; i < ntheta; i++) cosv[i] = std::cos(theta[i]);
  for (int l = 0; l <= lmax; l++) inv[l] = 1.0 / ortho_scale(l, 1.0);
  hipStream_t s = static_cast<hipStream_t>(stream);
  void* buf[5] = {nullptr, nullptr, nullptr, nullptr, nullptr};
  const size_t bytes[5] = {cosv.size() * sizeof(double), seeds.size() * sizeof(double),
                           rec.size() * sizeof(double4), inv.size() * sizeof(double),
                           off.size() * sizeof(long)};
  const void* src[5] = {cosv.data(), seeds.data(), rec.data(), inv.data(), off.data()};
  st = GSH_OK;
  for (int k = 0; k < 5 && st == GSH_OK; k++) {
    if (hipMalloc(&buf[k], bytes[k]) != hipSuccess) st = GSH_ERR_OUT_OF_MEMORY;
    else if (hipMemcpy(buf[k], src[k], bytes[k], hipMemcpyHostToDevice) != hipSuccess)
      st = GSH_ERR_DEVICE;
  }
  if (st == GSH_OK) {
    p.cosv = static_cast<const double*>(buf[0]);
    p.seeds = static_cast<const double*>(buf[1]);
    p.rec = static_cast<const double4*>(buf[2]);
    p.inv_ortho = static_cast<const double*>(buf[3]);
    p.blockoff = static_cast<const long*>(buf[4]);
    p.out = table;
    if (gsh::launch_wigner_table(p, s) != hipSuccess || hipStreamSynchronize(s) != hipSuccess)
      st = GSH_ERR_DEVICE;
  }
  for (auto b : buf)
    if (b) (void)hipFree(b);
  return st;
}

int gsh_wigner_table_host(int lmax, int mmax, int nmin, int nmax, int ntheta,
                          const double* theta, int mrange, int norm, int storage, double* table) {
  const long size = gsh_wigner_size(lmax, mmax, nmin, nmax, ntheta, mrange);
  if (size < 0) return wigner_check(lmax, mmax, nmin, nmax, ntheta, mrange);
  if (!table) return GSH_ERR_INVALID_ARGUMENT;
  void* d = nullptr;
  if (hipMalloc(&d, size * sizeof(double)) != hipSuccess) return GSH_ERR_OUT_OF_MEMORY;
  int st = gsh_wigner_table(lmax, mmax, nmin, nmax, ntheta, theta, mrange, norm, storage,
                            static_cast<double*>(d), nullptr);
  if (st == GSH_OK && hipMemcpy(table, d, size * sizeof(double), hipMemcpyDeviceToHost) != hipSuccess)
    st = GSH_ERR_DEVICE;
  (void)hipFree(d);
  return st;
}

}  // extern "C"
