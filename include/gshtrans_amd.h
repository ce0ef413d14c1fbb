/* gshtrans_amd.h -- C ABI of the MI355X (gfx950) GSH transform library.
 *
 * One shared library, gshtrans_amd/libgshtrans_hip.so, holds the HIP kernels
 * and this plain-C entry layer: no C++ or torch types cross it, only
 * integers, doubles, raw pointers and an opaque stream handle (hipStream_t
 * passed as void*; NULL means the default stream).
 *
 * Each entry point replaces one piece of the reference's header-only C++ API
 * (paths relative to the reference tree):
 *
 *   gsh_grid_create        GaussLegendreGrid(int lMax, int nMax, Flag)
 *                          GSHTrans/src/GaussLegendreGrid.h:38-73
 *   gsh_grid_colatitudes   GaussLegendreGrid::CoLatitudes / CoLatitudeWeights
 *                          GaussLegendreGrid.h:89-92
 *   gsh_inverse            GaussLegendreGrid::InverseTransformation
 *                          GaussLegendreGrid.h:222-321 (synthesis)
 *   gsh_forward            GaussLegendreGrid::ForwardTransformation
 *                          GaussLegendreGrid.h:108-217 (analysis)
 *   gsh_inverse_host /     the same two calls on host ranges, as the reference
 *   gsh_forward_host       takes them (used by the C++ facade in include/GSHTrans)
 *   gsh_coefficient_size   GridBase::{Real,Complex}CoefficientSize
 *                          GSHTrans/src/GridBase.h:79-113 via
 *                          GSHIndices::Size, GSHTrans/src/Indexing.h:148-150
 *   gsh_gl_nodes           GaussQuad::GaussQuadrature + Transform(acos(-x))
 *                          as called at GaussLegendreGrid.h:47-49
 *   gsh_wigner_seed        WignerDetails::WignerMinOrder, Wigner.h:52-80
 *   gsh_adjoint_inverse    no reference counterpart: the transpose of the
 *                          synthesis, used for least-squares gradients
 *   gsh_plan_*             no reference counterpart: batches of fields with
 *                          uneven lMax and upper index in one plan (graphs,
 *                          autotuned tiles, FFT family and analysis passes)
 *   gsh_timing_enable /    no reference counterpart (the reference has no
 *   gsh_timing_read        timing code): HIP events around each kernel
 *                          launch, on the launch stream, for benchmarks
 *   gsh_integrate          Integrate / L2InnerProduct, ScalarFieldOverloads.h:
 *                          288-316, with GridBase::Weights (GridBase.h:60-66)
 *   gsh_integrate_noalloc  no reference counterpart: gsh_integrate for callers
 *                          that cannot allocate (static scratch, see below)
 *   gsh_random_coefficients GridBase::Random*Coefficient (GridBase.h:115-172),
 *                          seeded, on the device
 *   gsh_forward_product    ForwardTransformation of the product field of
 *                          CanonicalComponentFieldBinary.h:203-253
 *                          (CanonicalComponentFieldOverloads.h:207-233)
 *   gsh_forward_expr       ForwardTransformation of the expression fields of
 *                          CanonicalComponentFieldOverloads.h:20-266 (sums,
 *                          differences, quotients, conj / real / imag,
 *                          negation, scalar multiples and shifts, products)
 *   gsh_wigner_table       Wigner<...>(lMax, mMax, nMax, theta) table build,
 *   (_host, _size,         Wigner.h:99-286 (ComputeAll :305-314, offsets
 *    _offset)              :116-132, 280-286), computed on the GPU
 *   gsh_device_alloc /     the storage of CanonicalComponentField.h:74-81,
 *   gsh_device_free /      119-124 and CanonicalComponentExpansion.h:69-78
 *   gsh_copy /             in device memory (the C++ facade's Device*
 *   gsh_synchronize        containers), so transforms take no per-call
 *                          host staging; no reference counterpart
 *
 * Layouts are the reference's:
 *   coefficients  GSHIndices<All>(lMax, lMax, n) for complex fields and
 *                 GSHIndices<NonNegative>(lMax, lMax, n) for real fields
 *                 (Indexing.h:104-146), interleaved complex (re, im)
 *   field         row-major (iTheta, iPhi), (L+1) x 2L points
 *                 (CanonicalComponentField.h:122-124, GaussLegendreGrid.h:94-103)
 * A batch is `batch` such arrays spaced by a stride counted in elements of
 * the array's own type (complex elements for coefficients and complex
 * fields, doubles or floats for real fields).
 *
 * Streams: the work of a call is queued on its stream.  Calls on one grid
 * handle share the handle's work buffers; calls on different streams must be
 * ordered by the caller (or use separate handles).
 *
 * Errors: every entry returns a gsh_status; the reference only asserts
 * (GaussLegendreGrid.h:41-44, 126-134, 240-248), which the C++ facade
 * restores by asserting on non-zero status.
 */
#ifndef GSHTRANS_AMD_H
#define GSHTRANS_AMD_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct gsh_grid_s* gsh_grid_t;

enum gsh_status {
  GSH_OK = 0,
  GSH_ERR_INVALID_ARGUMENT = 1, /* null pointer, negative size, bad enum   */
  GSH_ERR_UPPER_INDEX = 2,      /* n not in UpperIndices() or |n| > lMax   */
  GSH_ERR_DEGREE = 3,           /* lMax > grid lMax or lMax < 0            */
  GSH_ERR_VALUE_KIND = 4,       /* complex field on an MRange=NonNegative grid */
  GSH_ERR_UNSUPPORTED = 5,      /* precision/size combination not built    */
  GSH_ERR_DEVICE = 6,           /* HIP runtime error                       */
  GSH_ERR_NO_DEVICE = 7,        /* no gfx950 device visible                */
  GSH_ERR_OUT_OF_MEMORY = 8,
  GSH_ERR_SCRATCH_LIMIT = 9     /* call larger than the static scratch of an
                                   allocation-free entry point            */
};

/* MRange / NRange tags (GSHTrans/src/Concepts.h:24-27). */
enum gsh_range { GSH_ALL = 0, GSH_NONNEGATIVE = 1, GSH_SINGLE = 2 };

/* Value type of the field (Concepts.h:53-54). */
enum gsh_value_kind { GSH_REAL = 0, GSH_COMPLEX = 1 };

/* Arithmetic of the Legendre contraction.  The Wigner recurrence, its seeds,
 * the quadrature weights and the FFT arithmetic are FP64 in every mode; the
 * split-BF16 modes hand the order bins between the Legendre sums and the
 * wavefront ring FFT (ring lengths 512, 1024, 2048) over in FP32. */
enum gsh_precision {
  GSH_FP64 = 0,   /* FP64 MFMA, the reference's arithmetic                 */
  GSH_FP32 = 1,   /* FP32 MFMA on FP32-rounded operands                     */
  GSH_BF16X3 = 2, /* 2-way BF16 split, 3 products, FP32 accumulation        */
  GSH_BF16X6 = 3  /* 3-way BF16 split, 6 products, FP32 accumulation with
                     periodic FP64 flush                                     */
};

/* Storage type of a batched call's arrays. */
enum gsh_storage { GSH_STORE_FP64 = 0, GSH_STORE_FP32 = 1 };

const char* gsh_status_string(int status);
int gsh_version(void);

/* ---- host-only helpers (no device needed) ------------------------------ */
int gsh_gl_nodes(int lmax, double* theta, double* weight);
long gsh_coefficient_size(int lmax, int n, int value_kind);
int gsh_wigner_seed(int l, int n, double theta, double* value);

/* ---- grid -------------------------------------------------------------- */
int gsh_grid_create(int lmax, int nmax, int mrange, int nrange, gsh_grid_t* grid);
int gsh_grid_destroy(gsh_grid_t grid);
int gsh_grid_info(gsh_grid_t grid, int* lmax, int* nmax, int* min_upper_index,
                  int* ntheta, int* nphi);
int gsh_grid_colatitudes(gsh_grid_t grid, double* theta, double* weight);
/* Build the per-upper-index device data (seeds, recurrence coefficients)
 * ahead of the first transform; transforms build it on demand otherwise. */
int gsh_grid_prepare(gsh_grid_t grid, int n);

/* ---- transforms on device memory -------------------------------------- */
/* Synthesis: field = S_n(coeffs) for `batch` fields. */
int gsh_inverse(gsh_grid_t grid, int lmax, int n, int value_kind, int precision,
                int storage, int batch, const void* coeffs, long coeff_stride,
                void* field, long field_stride, void* stream);
/* Analysis: coeffs (+)= A_n(field).  accumulate != 0 adds into coeffs as the
 * reference does (GaussLegendreGrid.h:192); 0 overwrites. */
int gsh_forward(gsh_grid_t grid, int lmax, int n, int value_kind, int precision,
                int storage, int batch, const void* field, long field_stride,
                void* coeffs, long coeff_stride, int accumulate, void* stream);
/* Analysis of the pointwise product field_a * field_b (coeffs (+)= A_n(a b)),
 * the product formed as the ring FFT loads its rings; n is the product's
 * upper index N0 + N1 (CanonicalComponentFieldBinary.h:203-253).  FP64
 * storage; lmax >= 1. */
int gsh_forward_product(gsh_grid_t grid, int lmax, int n, int value_kind, int precision,
                        int batch, const void* field_a, const void* field_b, long field_stride,
                        void* coeffs, long coeff_stride, int accumulate, void* stream);
/* Analysis of a pointwise expression of the fields a (and b), formed as the
 * ring FFT loads its rings, so no expression field is written: coeffs (+)=
 * A_n(e) for the expression e of CanonicalComponentFieldOverloads.h:20-266
 * (classes in CanonicalComponentFieldUnary.h / CanonicalComponentFieldBinary.h).
 * value_kind is the kind of a and b; e is real when they are, or for
 * GSH_EXPR_REAL / GSH_EXPR_IMAG (then coeffs take the NonNegative layout),
 * complex otherwise.  n is e's upper index: N0 + N1 for products, N0 for the
 * rest (the reference's types require N0 == N1 for sums and differences, a
 * divisor of upper index 0, and N == 0 for adding a scalar).  The scalar
 * (scalar_re + i scalar_im; real for real fields) is the operand of
 * GSH_EXPR_SCALE and GSH_EXPR_SHIFT.  FP64 storage; lmax >= 1. */
enum gsh_field_expr {
  GSH_EXPR_PRODUCT = 1,    /* a * b   CanonicalComponentFieldMultiply */
  GSH_EXPR_SUM = 2,        /* a + b   CanonicalComponentFieldAdd      */
  GSH_EXPR_DIFFERENCE = 3, /* a - b   CanonicalComponentFieldSubtract */
  GSH_EXPR_QUOTIENT = 4,   /* a / b   CanonicalComponentFieldDivide   */
  GSH_EXPR_NEGATE = 5,     /* -a      operator-, Overloads.h:23-32    */
  GSH_EXPR_CONJ = 6,       /* conj(a) CanonicalComponentFieldConj     */
  GSH_EXPR_REAL = 7,       /* real(a) CanonicalComponentFieldReal     */
  GSH_EXPR_IMAG = 8,       /* imag(a) CanonicalComponentFieldImag     */
  GSH_EXPR_SCALE = 9,      /* s * a   UnaryWithScalar, multiplies     */
  GSH_EXPR_SHIFT = 10      /* a + s   UnaryWithScalar, plus           */
};
int gsh_forward_expr(gsh_grid_t grid, int lmax, int n, int value_kind, int op, int precision,
                     int batch, const void* field_a, const void* field_b, long field_stride,
                     double scalar_re, double scalar_im, void* coeffs, long coeff_stride,
                     int accumulate, void* stream);
/* Integrate (ScalarFieldOverloads.h:288-301) of `batch` fields, or with
 * field_b the L2 inner product Integrate(conj(a) b) (:303-316), with the
 * grid's quadrature weights w_i 2 pi / nphi (GridBase.h:60-66): out[f] is
 * complex (interleaved; imaginary part 0 for real fields), device memory. */
int gsh_integrate(gsh_grid_t grid, int value_kind, int batch, const void* field_a,
                  const void* field_b, long field_stride, double* out, void* stream);
/* gsh_integrate without any allocation.  No counterpart in the reference:
 * it is for callers that cannot allocate device memory (code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc).  The ring
 * sums go to a static device array of the library (4 MiB per device, held
 * until the process ends) instead of the grid's work buffer, so batch x
 * (lMax + 1) may not exceed 262144 (GSH_ERR_SCRATCH_LIMIT otherwise; e.g.
 * 255 fields at lMax = 1024).  The call makes no allocation, copy or host
 * synchronisation: it queues two kernels on `stream`.  Every call on one GPU
 * shares that GPU's array, so calls that overlap on one GPU (other streams,
 * other grids, other threads) race; order them. */
int gsh_integrate_noalloc(gsh_grid_t grid, int value_kind, int batch, const void* field_a,
                          const void* field_b, long field_stride, double* out, void* stream);
/* Seeded coefficients generated on the device (device memory): the stream
 * of gshtrans_amd.random_coefficients, scaled by (l+1)^-decay, with the
 * masks of GridBase.h:115-172 for the grid's lMax. */
int gsh_random_coefficients(gsh_grid_t grid, int lmax, int n, int value_kind,
                            unsigned long long seed, double decay, int batch, void* coeffs,
                            long coeff_stride, void* stream);
/* Adjoint of the synthesis: coeffs = S_n^H(field) (no quadrature weights). */
int gsh_adjoint_inverse(gsh_grid_t grid, int lmax, int n, int value_kind,
                        int precision, int storage, int batch, const void* field,
                        long field_stride, void* coeffs, long coeff_stride,
                        void* stream);

/* ---- batch plans: many fields of uneven truncation and upper index ---- */
/* No reference counterpart (the reference transforms one field per call,
 * GaussLegendreGrid.h:108-321).  A plan fixes `count` fields on one grid:
 * field i has truncation degree lmax[i] (1 <= lmax[i] <= grid lMax) and
 * upper index n[i]; all share one value kind, precision and storage.  The
 * plan groups the fields by upper index, orders each group by degree so that
 * every 16-column group of a launch carries its own degree bound, and on its
 * first execution in each direction times the candidate tile shapes
 * (synthesis column tiles, ring-FFT kernel family) and keeps the fastest.
 *
 * Arrays: coefficients of field i at coeffs + i*coeff_stride (complex
 * elements of the storage type; GSHIndices(lmax[i], lmax[i], n[i]) layout,
 * coeff_stride >= the largest coefficient count), field i at
 * field + i*field_stride (elements of the storage type).
 *
 * graph != 0 captures the whole launch sequence for these pointers into a
 * HIP graph on first use (after one eager run) and replays that graph on
 * later calls with the same pointers and strides. */
typedef struct gsh_plan_s* gsh_plan_t;
int gsh_plan_create(gsh_grid_t grid, int count, const int* lmax, const int* n,
                    int value_kind, int precision, int storage, gsh_plan_t* plan);
int gsh_plan_destroy(gsh_plan_t plan);
int gsh_plan_inverse(gsh_plan_t plan, const void* coeffs, long coeff_stride, void* field,
                     long field_stride, int graph, void* stream);
/* Analysis; overwrites the coefficients (no accumulation). */
int gsh_plan_forward(gsh_plan_t plan, const void* field, long field_stride, void* coeffs,
                     long coeff_stride, int graph, void* stream);
/* Launch groups and the tuned choices: per group g < *ngroups (up to cap),
 * n, lmax, field count, synthesis column tiles (0 = not tuned yet), and the
 * FFT family of synthesis / analysis (1 wavefront-per-ring, 2 LDS). */
int gsh_plan_info(gsh_plan_t plan, int* ngroups, int cap, int* n, int* lmax, int* count,
                  int* synth_tiles, int* fft_inverse, int* fft_forward);
/* The tuned degrees per pass of each group's split-BF16 analysis (64 or
 * 128; 0 for FP64 / FP32 plans and before the first analysis; for complex
 * fields of upper index n != 0, 64 is the kernel that runs orders m and -m
 * together, whose -m sums agree with the 128-degree kernel's to the
 * split-BF16 error, not bit for bit).  Tuned
 * choices are kept per launch shape for the life of the process and, when
 * the environment variable GSH_TUNE_CACHE names a file, read from and
 * appended to it, so later plans of a shape skip the timing. */
int gsh_plan_analysis_pass(gsh_plan_t plan, int* ngroups, int cap, int* pass);

/* ---- per-kernel timing ------------------------------------------------- */
/* In the split-BF16 synthesis of a grid whose colatitude count is one or two
 * past a multiple of 256, the LEGENDRE_SYNTHESIS pair also spans the kernel
 * that sums those last rows, and the COEF_SPLIT pair the kernel that writes
 * their recurrence values. */
enum gsh_kernel {
  GSH_KERNEL_LEGENDRE_SYNTHESIS = 0,
  GSH_KERNEL_FFT_SYNTHESIS = 1,
  GSH_KERNEL_FFT_ANALYSIS = 2,
  GSH_KERNEL_LEGENDRE_ANALYSIS = 3,
  GSH_KERNEL_COEF_SPLIT = 4,      /* mixed precision: coefficient split   */
  GSH_KERNEL_BIN_SPLIT = 5,       /* not launched (always 0): the mixed
                                     analysis reads FP32 bins directly     */
  GSH_KERNEL_COUNT = 6
};
/* enable != 0 records a HIP event pair around every kernel launch of this
 * grid (on the launch stream) and clears the accumulated totals. */
int gsh_timing_enable(gsh_grid_t grid, int enable);
/* Waits for the recorded events and returns, per kernel id, the summed
 * duration in milliseconds and the number of launches since enabling. */
int gsh_timing_read(gsh_grid_t grid, double* total_ms, long* launches, int nkernels);

/* ---- device memory on the grid's GPU ---------------------------------- */
/* Zero-filled device allocation of `bytes` (0 gives NULL). */
int gsh_device_alloc(gsh_grid_t grid, size_t bytes, void** ptr);
int gsh_device_free(gsh_grid_t grid, void* ptr);
/* Synchronous copy of `bytes` between any two of host and device memory. */
int gsh_copy(gsh_grid_t grid, void* dst, const void* src, size_t bytes);
/* Waits for the work queued on `stream` (NULL: the default stream). */
int gsh_synchronize(gsh_grid_t grid, void* stream);

/* ---- Wigner d tables ---------------------------------------------------- */
/* The table of Wigner<double, Norm, MRange, NRange, AngleRange, Storage>(
 * lmax, mmax, nmax, theta) (Wigner.h:99-286) for upper indices nmin..nmax
 * (NRange All: -nmax..nmax, NonNegative: 0..nmax, Single: nmax..nmax) and
 * ntheta angles: one GSHIndices<MRange>(lmax, mmax, n) block of d^l_{mn}
 * (Ortho: times sqrt(2l+1) / (2 sqrt(pi)), Wigner.h:475-485) per
 * (n, theta_i), blocks in ColumnMajor (n outer) or RowMajor (theta outer)
 * order (Wigner.h:240-250, 280-286).  mrange is GSH_ALL or GSH_NONNEGATIVE. */
enum gsh_wigner_norm { GSH_ORTHO = 0, GSH_FOURPI = 1 };
enum gsh_wigner_storage { GSH_COLUMN_MAJOR = 0, GSH_ROW_MAJOR = 1 };
/* Number of doubles of the table (-1 for invalid arguments). */
long gsh_wigner_size(int lmax, int mmax, int nmin, int nmax, int ntheta, int mrange);
/* First entry of block (n, itheta) (-1 for invalid arguments). */
long gsh_wigner_offset(int lmax, int mmax, int nmin, int nmax, int ntheta, int mrange,
                       int storage, int n, int itheta);
/* Fills `table` (device memory of the current device, gsh_wigner_size
 * doubles; theta in host memory) and waits for the GPU. */
int gsh_wigner_table(int lmax, int mmax, int nmin, int nmax, int ntheta, const double* theta,
                     int mrange, int norm, int storage, double* table, void* stream);
/* The same into host memory. */
int gsh_wigner_table_host(int lmax, int mmax, int nmin, int nmax, int ntheta,
                          const double* theta, int mrange, int norm, int storage,
                          double* table);

/* ---- transforms on host memory (staged through device buffers) -------- */
int gsh_inverse_host(gsh_grid_t grid, int lmax, int n, int value_kind,
                     const double* coeffs, double* field);
int gsh_forward_host(gsh_grid_t grid, int lmax, int n, int value_kind,
                     const double* field, double* coeffs, int accumulate);

#ifdef __cplusplus
}
#endif

#endif /* GSHTRANS_AMD_H */
