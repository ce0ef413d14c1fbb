// gsh_internal.h -- shared declarations between the host layer (gsh_capi.cpp)
// and the HIP kernels (gsh_legendre.hip, gsh_fft.hip).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>

namespace gsh {

// Parameters of one Legendre-contraction launch (synthesis, analysis or
// adjoint) for a batch of fields sharing (lmax, n, value kind).
//
// Work layout in HBM ("F"): F[theta][slot][col] doubles, where slot indexes
// the order m (slot = m - m_first; m_first = -lmax for complex fields, 0 for
// real fields) and col = 2*field + part (part 0 = real, 1 = imaginary).
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
  // per-grid / per-n tables
  const double* cosv;   // cos(theta_t), t < ntheta
  const double* seeds;  // Ortho-normalised d^{l0}_{m n}(theta_t) at [(m+L)*ntheta + t]
  const double* RA;     // recurrence: e_l = (RA*cos - RB) e_{l-1} - RC e_{l-2}
  const double* RB;
  const double* RC;
  const long* roff;     // [m + L]: index of (m, l0) in RA/RB/RC
  int accumulate;       // analysis: coefficients += result
  int zero_last;        // analysis: zero the (L, L) coefficient (complex, lmax == L)
  int n_tsb;            // synthesis: theta super-blocks of 256
  int n_cg;             // column groups
  int nblk;             // analysis: theta blocks of 64
};

// Parameters of one ring-FFT launch.
struct FFTParams {
  int N;            // ring length 2L (power of two)
  int L, ntheta, lmax, nslot, ncol, nfield, real, m_first;
  int rings;        // rings per workgroup (rings * N == 4096 or N == 4096)
  int nfg;          // field groups
  const double2* tw;  // tw[k] = exp(-2 pi i k / N)
  const double* wq;   // analysis weights w_t * 2 pi / N
  double* F;
  const void* fin;    // analysis input field
  void* fout;         // synthesis output field
  long fstride;       // elements between fields
};

hipError_t launch_leg_synth_f64(const LegParams& p, hipStream_t s);
hipError_t launch_leg_anal_f64(const LegParams& p, hipStream_t s);
hipError_t launch_fft_synth(const FFTParams& p, hipStream_t s);
hipError_t launch_fft_anal(const FFTParams& p, hipStream_t s);
hipError_t launch_lmax0(int inverse, int real, int nfield, const double* in, long istride,
                        double* out, long ostride, int accumulate, hipStream_t s);

}  // namespace gsh
