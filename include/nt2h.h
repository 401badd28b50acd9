/*
 * nt2h.h — C ABI of the MI355X (gfx950) architecture site for NT2's table
 * expression path.  Plain pointers and sizes only; every entry point returns
 * NT2H_OK (0) or a negative NT2H_E* code and never throws.  Device pointers
 * are owned by the caller unless they come from nt2h_malloc.  Calls are
 * asynchronous on the given HIP stream (NULL = the default stream) unless
 * stated otherwise; nt2h_last_error() returns a per-thread message.
 *
 * Reference interfaces replaced (paths relative to the reference tree):
 *  - transform_(out, in, range):
 *      modules/arch/shared_memory/include/nt2/core/functions/shared_memory/transform.hpp:26-36
 *  - inner_fold_(out, in, neutral, bop, uop):
 *      modules/arch/shared_memory/include/nt2/core/functions/shared_memory/inner_fold.hpp:24-38
 *  - the site tag and its device buffer:
 *      modules/arch/cuda/include/nt2/sdk/cuda/cuda.hpp:16-24,
 *      modules/arch/cuda/include/nt2/sdk/cuda/memory/buffer.hpp:36-242
 * The C++ front end (nt2_amd/include/nt2/table.hpp) lowers arbitrary
 * expressions at compile time; the entry points below are the fixed
 * workloads and runtime services behind it.
 */
#ifndef NT2H_H
#define NT2H_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define NT2H_OK 0
#define NT2H_EINVAL (-1)     /* bad argument (null pointer, size, op code) */
#define NT2H_EHIP (-2)       /* HIP runtime error */
#define NT2H_ENOMEM (-3)     /* device allocation failed */
#define NT2H_EBUDGET (-4)    /* request larger than the server's memory budget */
#define NT2H_EAGAIN (-5)     /* not ready yet (poll) */
#define NT2H_ECANCELLED (-6) /* request cancelled */
#define NT2H_ESTATE (-7)     /* object in the wrong state */

/* evaluation policy of floating-point kernels */
#define NT2H_EXACT 0 /* NT2's arithmetic: fma = a*b+c with two roundings; bit-identical to the CPU evaluation */
#define NT2H_FAST 1  /* NT2's fma call sites fused in hardware; within each function's ULP bound */
#define NT2H_ACCURATE 2 /* Black-Scholes only: f32 tables with an FP64 internal evaluation, f64
                           tables with double-double tails, both with this library's own
                           approximations; within 4 ulp of the exact price */

/* elementwise operations for nt2h_map_* (NT2 functors exp_, log_, sqrt_, erfc_,
 * normcdf_, fastnormcdf_, cos_, sin_, abs_, exp2_ ... atan_, tan_, asin_, acos_, erf_,
 * cbrt_, asinh_, acosh_, atanh_, cot_, sec_, csc_, sinpi_ ... cotpi_, sind_ ... cotd_) */
#define NT2H_OP_EXP 0
#define NT2H_OP_LOG 1
#define NT2H_OP_SQRT 2
#define NT2H_OP_ERFC 3
#define NT2H_OP_NORMCDF 4
#define NT2H_OP_FASTNORMCDF 5 /* single precision only, as in NT2 */
#define NT2H_OP_COS 6
#define NT2H_OP_SIN 7
#define NT2H_OP_ABS 8
#define NT2H_OP_EXP2 9
#define NT2H_OP_EXP10 10
#define NT2H_OP_LOG2 11
#define NT2H_OP_LOG10 12
#define NT2H_OP_FLOOR 13
#define NT2H_OP_CEIL 14
#define NT2H_OP_TRUNC 15
#define NT2H_OP_ROUND 16      /* half away from zero */
#define NT2H_OP_ROUND2EVEN 17 /* half to even */
#define NT2H_OP_SIGN 18
#define NT2H_OP_EXPM1 19
#define NT2H_OP_LOG1P 20
#define NT2H_OP_TANH 21
#define NT2H_OP_SINH 22
#define NT2H_OP_COSH 23
#define NT2H_OP_ATAN 24
#define NT2H_OP_TAN 25
#define NT2H_OP_ASIN 26
#define NT2H_OP_ACOS 27
#define NT2H_OP_ERF 28   /* f64: a pack's tail takes NT2's scalar form (+1 beyond |x| = 6) */
#define NT2H_OP_CBRT 29
#define NT2H_OP_ASINH 30
#define NT2H_OP_ACOSH 31
#define NT2H_OP_ATANH 32
#define NT2H_OP_COT 33   /* radian, per-pack range reduction as cos/sin/tan */
#define NT2H_OP_SEC 34   /* 1/cos */
#define NT2H_OP_CSC 35   /* 1/sin */
#define NT2H_OP_SINPI 36 /* sin(pi x); pi and degree forms reduce lane by lane */
#define NT2H_OP_COSPI 37
#define NT2H_OP_TANPI 38
#define NT2H_OP_COTPI 39
#define NT2H_OP_SIND 40  /* sin of x degrees */
#define NT2H_OP_COSD 41
#define NT2H_OP_TAND 42
#define NT2H_OP_COTD 43
/* Compositions, NT2's definitions over the kernels above (NT2's scalar form
 * for the tail where the inner function or the definition has one): */
#define NT2H_OP_COTH 44   /* hyperbolic/functions/simd/common/coth.hpp:35-63 */
#define NT2H_OP_SECH 45   /* 1/cosh */
#define NT2H_OP_CSCH 46   /* 1/sinh */
#define NT2H_OP_ACOTH 47  /* atanh(1/x) */
#define NT2H_OP_ASECH 48  /* acosh(1/x) */
#define NT2H_OP_ACSCH 49  /* asinh(1/x) */
#define NT2H_OP_ACOT 50   /* atan(1/x) */
#define NT2H_OP_ASEC 51   /* pi/2 - acsc(x); scalar forms: trigonometric/functions/scalar/asec.hpp:26-67 */
#define NT2H_OP_ACSC 52   /* asin(1/x) */
#define NT2H_OP_ACOTD 53  /* degrees: 90 - atand(|x|), sign of x */
#define NT2H_OP_ASECD 54
#define NT2H_OP_ACSCD 55
#define NT2H_OP_ACOSD 56
#define NT2H_OP_ACOSPI 57 /* units of pi: acos(x)/pi */
#define NT2H_OP_ASIND 58
#define NT2H_OP_ASINPI 59
#define NT2H_OP_ATAND 60
#define NT2H_OP_ATANPI 61
#define NT2H_OP_INDEG 62  /* radians to degrees */
#define NT2H_OP_INRAD 63  /* degrees to radians */
#define NT2H_OP_SECD 64   /* 1/cosd, NaN at its zeros */
#define NT2H_OP_CSCD 65   /* cscpi(x/180) */
#define NT2H_OP_SECPI 66  /* 1/cospi, NaN at half-integers */
#define NT2H_OP_CSCPI 67  /* 1/sinpi, NaN at nonzero integers */
#define NT2H_OP_SINC 68   /* sin(x)/x, radian: per-pack range reduction */
#define NT2H_OP_SINCPI 69 /* sinpi(x)/(pi x) */
#define NT2H_OP_SQRT1PM1 70  /* sqrt(1 + x) - 1 */
#define NT2H_OP_EXPRECNEG 71 /* exp(-1/x) */
#define NT2H_OP_EXPRECNEGC 72 /* -expm1(-1/x) */
#define NT2H_OP_REALLOG 73   /* log */
#define NT2H_OP_REALSQRT 74  /* sqrt */
#define NT2H_OP_EXPX2 75     /* exp(x^2): exponential/functions/simd/common/expx2.hpp:37-61 */
#define NT2H_OP_ERFINV 76    /* euler/functions/simd/common/erfinv.hpp:32-108 */
#define NT2H_OP_ERFCINV 77   /* euler/functions/simd/common/erfcinv.hpp:32-107 */
#define NT2H_OP_SINHC 78     /* sinh(x)/x: hyperbolic/functions/simd/common/sinhc.hpp:41-80 */

/* Black-Scholes formula variants */
#define NT2H_BS_TEXTBOOK 0 /* d1 = (log(S/X)+(r+v^2/2)T)/(v sqrt T), call and put, erfc-based normcdf */
#define NT2H_BS_DEMO 1     /* demo/blackscholes/nt2/blackandscholes_nt2.cpp:55-59: call only, fastnormcdf */

/* ---- runtime ----------------------------------------------------------- */
int nt2h_abi_version(void);
const char* nt2h_last_error(void);
/* device properties: compute units, HBM bytes; name may be NULL */
int nt2h_device_info(int* cus, size_t* hbm_bytes, char* name, size_t name_len);

/* Device storage for tables: the HIP counterpart of cuda_buffer
 * (modules/arch/cuda/include/nt2/sdk/cuda/memory/buffer.hpp:36-242). */
int nt2h_malloc(void** dptr, size_t bytes);
int nt2h_free(void* dptr);
#define NT2H_H2D 1
#define NT2H_D2H 2
#define NT2H_D2D 3
int nt2h_memcpy(void* dst, const void* src, size_t bytes, int kind, void* stream);
int nt2h_stream_create(void** stream);
int nt2h_stream_destroy(void* stream);
int nt2h_stream_sync(void* stream);

/* ---- seeded synthetic tables (counter-based SplitMix64) ---------------- */
/* out[i] = (float)(lo + (hi-lo)*u_i), u_i = (splitmix64(seed, i) >> 11) * 2^-53 */
int nt2h_fill_uniform_f32(float* out, size_t n, double lo, double hi, uint64_t seed, void* stream);
int nt2h_fill_uniform_f64(double* out, size_t n, double lo, double hi, uint64_t seed, void* stream);
/* standard normal via Box-Muller in double, rounded to float */
int nt2h_fill_normal_f32(float* out, size_t n, uint64_t seed, void* stream);
/* sigma-delta test sequence: nframes frames of npix = h*w uint8 pixels
 * (column-major, frame k at frames + k*npix): seeded background U{16..200},
 * nrect moving bright rectangles, uniform noise +-noise, saturated. */
int nt2h_fill_frames_u8(uint8_t* frames, int h, int w, int nframes, int nrect, int noise,
                        uint64_t seed, void* stream);

/* ---- transform_: elementwise, NT2 pack semantics ----------------------- */
/* y = op(x).  Packs of 8 floats / 4 doubles from index 0 as on NT2's AVX site,
 * the last n%8 (n%4) elements with NT2's scalar kernels (trigonometric range
 * reduction is chosen per pack, as in NT2). */
int nt2h_map_f32(int op, size_t n, const float* x, float* y, int policy, void* stream);
int nt2h_map_f64(int op, size_t n, const double* x, double* y, int policy, void* stream);
/* out = f(x, y) elementwise for the binary functions below, NT2's pack
 * schedule as nt2h_map_*: replaces transform_ over a binary node such as
 * atan2(a, b) (modules/core/trigonometric/include/nt2/trigonometric/functions/
 * simd/common/atan2.hpp:30-54, scalar/atan2.hpp:29-45 for the pack tail) or
 * pow(a, b) (modules/core/exponential/include/nt2/exponential/functions/
 * simd/common/pow.hpp:38-46, pow_abs.hpp:65-133; scalar/pow.hpp:30-36,
 * pow_abs.hpp:51-111 for the pack tail). */
#define NT2H_OP2_ATAN2 0
#define NT2H_OP2_POW 1
#define NT2H_OP2_ATAN2D 2 /* atan2 in degrees: trigonometric/functions/generic/atan2d.hpp:16-27 */
#define NT2H_OP2_LOGSPACE_ADD 3 /* log(exp(x) + exp(y)): exponential/functions/scalar/logspace_add.hpp:25-42 */
#define NT2H_OP2_LOGSPACE_SUB 4 /* log(exp(x) - exp(y)): exponential/functions/scalar/logspace_sub.hpp:20-32 */
int nt2h_map2_f32(int op, size_t n, const float* x, const float* y, float* out, int policy, void* stream);
int nt2h_map2_f64(int op, size_t n, const double* x, const double* y, double* out, int policy, void* stream);
/* [s, c] = sincos(x): modules/core/trigonometric/include/nt2/trigonometric/functions/generic/sincos.hpp:50-64 */
int nt2h_sincos_f32(size_t n, const float* x, float* s, float* c, int policy, void* stream);
int nt2h_sincos_f64(size_t n, const double* x, double* s, double* c, int policy, void* stream);
/* dst = src with 16-byte streaming loads/stores; the HBM bandwidth reference
 * (grid <= 0 picks one vector per thread).  No counterpart in the reference. */
int nt2h_stream_copy(const void* src, void* dst, size_t bytes, int grid, void* stream);
/* Y = a*X + Y: demo/saxpy/nt2/saxpy_nt2.cpp:39-42 */
int nt2h_saxpy_f32(size_t n, float a, const float* x, float* y, void* stream);

/* ---- inner_fold_: sum / globalsum / dot -------------------------------- */
/* *out = sum(x) or sum(x.*y) (demo/sdot/nt2/sdot_nt2.cpp:35-38), written to
 * device memory as double.  Products are rounded to float as in the NT2
 * expression X*Y; partial sums are compensated FP32 (each addition's rounding
 * error recovered with TwoSum and accumulated per lane, error-free pairwise
 * combine across the wave), blocks are combined in FP64 in a fixed order, so
 * results are deterministic. */
int nt2h_sum_f32(size_t n, const float* x, double* out, void* stream);
int nt2h_dot_f32(size_t n, const float* x, const float* y, double* out, void* stream);
/* nt2h_dot_f32 for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops): no counterpart in the reference.
 * Same kernels and bits as nt2h_dot_f32; the block partials live in a static
 * device array of 256 doubles per device, looked up on the first call on a
 * device, and no hipMalloc, hipFree or memset is made.  The fold never needs
 * more than those 256 partials, whatever n (NT2H_EINVAL otherwise).  Calls
 * that overlap on one device share the array and race: serialize them. */
int nt2h_dot_f32_noalloc(size_t n, const float* x, const float* y, double* out, void* stream);

/* ---- sum(a, dim) and cumsum(a, dim) over an N-D table -------------------- */
/* x: device table of extent ext[0..3] (host array; column-major, trailing
 * sizes 1 for fewer dimensions); dim >= 1 (beyond the last dimension: a
 * dimension of size 1).  sum writes the table of extent ext with ext[dim-1]
 * replaced by 1; cumsum writes a table of extent ext (out may equal x).
 * policy NT2H_EXACT: NT2's order in the table's precision, bit-identical to
 * NT2's CPU result (inner fold: pack accumulators, horizontal sum, scalar
 * tail; outer fold and scans: sequential along the dimension); NT2H_FAST:
 * parallel chunked forms with compensated / FP64 accumulation, each output
 * rounded once.
 * sum replaces NT2's reduction run_assign_ (modules/core/reduction/include/nt2/dsl/functions/reduction/run.hpp:30-93)
 * with inner_fold_ (modules/core/reduction/include/nt2/core/functions/simd/inner_fold.hpp:44-67)
 * and outer_fold_ (modules/core/reduction/include/nt2/core/functions/simd/outer_fold.hpp:46-76);
 * cumsum replaces inner_scan_ / outer_scan_
 * (modules/core/swar/include/nt2/core/functions/scalar/inner_scan.hpp:36-55, scalar/outer_scan.hpp:34-55). */
int nt2h_sum_dim_f32(const float* x, const size_t* ext, int dim, float* out, int policy, void* stream);
int nt2h_sum_dim_f64(const double* x, const size_t* ext, int dim, double* out, int policy, void* stream);
int nt2h_cumsum_dim_f32(const float* x, const size_t* ext, int dim, float* out, int policy, void* stream);
int nt2h_cumsum_dim_f64(const double* x, const size_t* ext, int dim, double* out, int policy, void* stream);
/* cumprod(a, dim), cummin(a, dim), cummax(a, dim): arguments as for cumsum;
 * each line runs NT2's sequential step bop(summary, x) from the neutral (1,
 * +Inf, -Inf) under both policies, bit-identical to NT2's CPU result.  min
 * and max are NT2's (s < x ? s : x), (s > x ? s : x): a NaN element yields NaN
 * at its position only.  They replace inner_scan_ / outer_scan_ for the
 * cumulative_ functors cumprod_, cummin_, cummax_
 * (modules/boost/simd/base/include/boost/simd/swar/functions/cumprod.hpp:27,
 * cummin.hpp:26, cummax.hpp:27). */
int nt2h_cumprod_dim_f32(const float* x, const size_t* ext, int dim, float* out, int policy, void* stream);
int nt2h_cumprod_dim_f64(const double* x, const size_t* ext, int dim, double* out, int policy, void* stream);
int nt2h_cummin_dim_f32(const float* x, const size_t* ext, int dim, float* out, int policy, void* stream);
int nt2h_cummin_dim_f64(const double* x, const size_t* ext, int dim, double* out, int policy, void* stream);
int nt2h_cummax_dim_f32(const float* x, const size_t* ext, int dim, float* out, int policy, void* stream);
int nt2h_cummax_dim_f64(const double* x, const size_t* ext, int dim, double* out, int policy, void* stream);

/* ---- image demos ------------------------------------------------------ */
/* grey = 0.3 r + 0.59 g + 0.11 b over n pixels (planar float tables):
 * replaces the statement of demo/rgb2grey/nt2/rgb2grey_nt2.cpp:42-45. */
int nt2h_rgb2grey_f32(size_t n, const float* r, const float* g, const float* b, float* grey, int policy,
                      void* stream);
/* uint8 planes: grey = min(255, (77 r + 151 g + 28 b + 128) >> 8), the same
 * weights in 8.8 fixed point, rounded and saturated (no counterpart in the
 * reference, whose demo is float only: an integer variant of
 * demo/rgb2grey/nt2/rgb2grey_nt2.cpp:42-45). */
int nt2h_rgb2grey_u8(size_t n, const uint8_t* r, const uint8_t* g, const uint8_t* b, uint8_t* grey,
                     void* stream);
/* y = 0.299 r + 0.587 g + 0.114 b; u = 0.492 (b - y); v = 0.877 (r - y), one
 * pass: replaces rgb2yuv_work, demo/rgb2yuv/nt2/rgb2yuv_nt2.cpp:28-35. */
int nt2h_rgb2yuv_f32(size_t n, const float* r, const float* g, const float* b, float* y, float* u, float* v,
                     int policy, void* stream);
/* out[p] = hi for p = n - 1, else p*step + lo with step = (hi - lo)/(n - 1):
 * replaces linspace(lo, hi, n),
 * modules/core/generative/include/nt2/core/functions/details/linspace.hpp:22-48. */
int nt2h_linspace_f32(size_t n, float lo, float hi, float* out, int policy, void* stream);
int nt2h_linspace_f64(size_t n, double lo, double hi, double* out, int policy, void* stream);
/* out = arrayfun(step, a, b): escape counts of z = z^2 + (a + ib) from z = 0,
 * iterated per pack of 8 until no lane of the pack has |z|^2 < 4 or max_iter
 * iterations (256 in the demo); replaces demo/mandelbrot/nt2/mandelbrot_nt2.cpp:35-70
 * and :100-103.  max_iter must be at least 1. */
int nt2h_mandelbrot_f32(size_t n, const float* a, const float* b, int32_t* out, int max_iter, int policy,
                        void* stream);

/* ---- fused Black-Scholes ---------------------------------------------- */
/* call/put from tables S, X, T and scalars r, v; one pass, one kernel.
 * demo mode writes call only (put may be NULL). */
int nt2h_blackscholes_f32(size_t n, const float* S, const float* X, const float* T, float r,
                          float v, float* call, float* put, int mode, int policy, void* stream);
/* workgroups of the accurate-policy kernel's grid-stride launch (default 8192);
 * a tuning hook with no counterpart in the reference */
int nt2h_bs_set_grid(int workgroups);
/* pixels per thread of nt2h_sigmadelta_u8's main kernel: 16 (default), 8 or
 * 4; results do not depend on it (no counterpart in the reference: a tuning
 * hook) */
int nt2h_sigmadelta_set_config(int pixels_per_thread);
/* f64 tables: policy NT2H_EXACT / NT2H_FAST run NT2's double-precision
 * evaluation (demo/blackscholes/nt2/blackandscholes_nt2.cpp:53-60 with the
 * erfc-based normcdf, modules/core/statistics/include/nt2/statistics/
 * functions/generic/normcdf.hpp:28-35); NT2H_ACCURATE prices every option
 * within 4 ulp of the exact price (S, X, T, v > 0). */
int nt2h_blackscholes_f64(size_t n, const double* S, const double* X, const double* T,
                          double r, double v, double* call, double* put, int policy,
                          void* stream);

/* ---- sigma-delta motion detection (uint8, bit-exact) ------------------- */
/* Runs frames 1..nframes-1 over state bkg/var/label (demo/sigmadelta/nt2/
 * sigmadelta_nt2.cpp:37-58, driver :100-106) in one fused kernel that keeps
 * the state in registers; N is the variance factor (mul = adds applied N-1
 * times).  frame_counts (optional, nframes entries) receives the number of
 * label==1 pixels after each frame (entry 0 is zeroed). */
int nt2h_sigmadelta_u8(size_t npix, int nframes, const uint8_t* frames, size_t frame_stride,
                       uint8_t* bkg, uint8_t* var, uint8_t* label, int N,
                       uint32_t* frame_counts, void* stream);
/* One frame step per launch (the plain NT2 schedule, for comparison). */
int nt2h_sigmadelta_step_u8(size_t npix, const uint8_t* frame, uint8_t* bkg, uint8_t* var,
                            uint8_t* label, int N, void* stream);

/* ---- batched expression z = exp(-x.*x/2).*cos(y) + sqrt(abs(x)) ------- */
int nt2h_expr4_f32(size_t n, const float* x, const float* y, float* z, int policy, void* stream);
/* *out = sum(z) for one table, one launch (the plain schedule) */
int nt2h_expr4_sum_f32(size_t n, const float* x, const float* y, double* out, int policy,
                       void* stream);

/* ---- batched evaluation: many uneven tables, one persistent launch ------ */
/* No counterpart in the reference (NT2 runs one statement per transform_ /
 * inner_fold_ call); batches the expression above over `count` tables.
 * x, y, lens: host arrays of device pointers and lengths; the tables must
 * stay allocated while the batch is used. */
typedef struct nt2h_batch nt2h_batch;
int nt2h_batch_create(nt2h_batch** batch, int count, const float* const* x, const float* const* y,
                      const size_t* lens, int policy);
/* tile = block*ept elements (ept a multiple of 8), wg_per_cu persistent workgroups per CU */
int nt2h_batch_set_config(nt2h_batch* batch, int block, int ept, int wg_per_cu);
int nt2h_batch_get_config(nt2h_batch* batch, int* block, int* ept, int* wg_per_cu);
/* sums: device array of `count` doubles; replays the captured hipGraph on stream */
int nt2h_batch_run(nt2h_batch* batch, double* sums, void* stream);
/* the plain schedule: one nt2h_expr4_sum_f32 launch per table */
int nt2h_batch_run_plain(nt2h_batch* batch, double* sums, void* stream);
/* times the candidate configurations and keeps the fastest (synchronous) */
int nt2h_batch_autotune(nt2h_batch* batch, double* sums, double* best_ms);
int nt2h_batch_destroy(nt2h_batch* batch);

/* ---- serving: Black-Scholes requests with continuous batching ---------- */
/* No counterpart in the reference (NT2 evaluates synchronously on the calling
 * thread).  Requests are admitted in arrival order against a device-memory
 * budget (input + output tables), batched into shared launches while up to
 * max_inflight launches run concurrently, and completed asynchronously. */
#define NT2H_F32 0
#define NT2H_F64 1
#define NT2H_BACKEND_HIP 0  /* device memory, HIP streams */
#define NT2H_BACKEND_HOST 1 /* host memory and CPU evaluation: scheduler tests without a GPU */
typedef struct nt2h_server nt2h_server;
typedef struct {
  size_t budget_bytes; /* bytes of admitted requests' tables (5*n*sizeof(T) each) */
  int max_inflight;    /* concurrent batched launches (streams), 1..16 */
  int max_batch;       /* requests per launch; 1 with max_inflight 1 = plain schedule */
  int policy;          /* NT2H_EXACT, NT2H_FAST, or NT2H_ACCURATE (f32 and f64 requests) */
  int backend;         /* NT2H_BACKEND_HIP or NT2H_BACKEND_HOST */
  int fail_alloc_every; /* test hook: fail every k-th output allocation (0 = never) */
  int fail_launch_every; /* test hook: report every k-th launched batch as failed (0 = never) */
} nt2h_server_config;
typedef struct {
  int dtype;           /* NT2H_F32 or NT2H_F64 */
  size_t n;            /* options */
  const void* S;       /* resident input tables (device memory, caller-owned) */
  const void* X;
  const void* T;
  double r, v;         /* rounded to float for f32 requests */
  void* call;          /* NULL: the server allocates call/put within the budget */
  void* put;
  int64_t t_arrival_ns; /* steady-clock arrival stamp; 0 = time of submission */
} nt2h_bs_request;
typedef struct {
  uint64_t completed, failed, cancelled, options, launches;
  double mean_batch, p50_us, p99_us, mean_us, max_us, span_s, options_per_s;
  size_t used_bytes, peak_bytes;
} nt2h_server_stats;
int nt2h_server_create(const nt2h_server_config* cfg, nt2h_server** server);
/* NT2H_EINVAL for a malformed request, NT2H_EBUDGET if it can never fit */
int nt2h_server_submit(nt2h_server* server, const nt2h_bs_request* req, uint64_t* id);
/* NT2H_OK when done, NT2H_EAGAIN while queued or running, else the failure code */
int nt2h_server_poll(nt2h_server* server, uint64_t id);
int nt2h_server_wait(nt2h_server* server, uint64_t id, double timeout_s);
int nt2h_server_result(nt2h_server* server, uint64_t id, void** call, void** put,
                       int64_t* latency_ns);
int nt2h_server_cancel(nt2h_server* server, uint64_t id); /* before launch only */
int nt2h_server_release(nt2h_server* server, uint64_t id);
int nt2h_server_drain(nt2h_server* server, double timeout_s);
int nt2h_server_get_stats(nt2h_server* server, nt2h_server_stats* stats);
int nt2h_server_reset_stats(nt2h_server* server);
/* submit items[k] at t0 + arrival_s[k] (busy-wait pacing), release each on completion */
int nt2h_server_run_trace(nt2h_server* server, int count, const nt2h_bs_request* items,
                          const double* arrival_s, double timeout_s);
int nt2h_server_destroy(nt2h_server* server);

#ifdef __cplusplus
}
#endif
#endif /* NT2H_H */
