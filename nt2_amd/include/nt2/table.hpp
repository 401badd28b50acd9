// nt2::table<T> and its expression DSL on the HIP architecture site (gfx950).
//
// User code written against NT2's table API compiles against this header with
// hipcc: each statement `A = expression;`, `tie(A, B) = tie(e1, e2);`,
// `globalsum(expression)`, `dot(a, b)` is lowered at compile time to ONE fused
// kernel that evaluates the whole expression tree per 128-bit pack, keeping
// every intermediate in registers.
//
// Semantics kept from NT2 (paths relative to the reference tree):
//  * table<T>: owning contiguous storage, column-major, 1-based operator()
//    (modules/core/container/table/include/nt2/core/container/table/table.hpp:26-121,
//    modules/core/sdk/include/nt2/core/settings/index.hpp:29, storage_order.hpp:22);
//  * assignment resizes the left-hand side to the right-hand side's extent and
//    evaluates elementwise, in place when the LHS appears on the RHS
//    (modules/core/base/include/nt2/dsl/functions/elementwise/run.hpp:21-36);
//  * transform over packs of the AVX width from index 0, scalar tail
//    (modules/core/base/include/nt2/core/functions/simd/transform.hpp:40-57), so
//    pack-dependent kernels (the trigonometric range reduction) give NT2's bits;
//  * tie reads every right-hand side at an index before writing any output
//    (modules/core/container/table/include/nt2/core/functions/expr/tie.hpp:40-121);
//  * globalsum/dot fold the whole table (modules/core/reduction/include/nt2/
//    core/functions/globalsum.hpp:24-88, reduction/functions/expr/dot.hpp:22-29),
//    here with compensated FP32 partials and an FP64 cross-block combine;
//    integer tables sum in the element type, wrapping, as NT2's sum does;
//  * sum(a[, dim]) and cumsum(a[, dim]) reduce or scan along a dimension
//    (default: the first non-singleton one) in NT2's order under the exact
//    policy (modules/core/reduction/include/nt2/dsl/functions/reduction/run.hpp:30-93),
//    in the parallel fast order under the fast policy (nt2/hip/dimwise.hpp);
//    cumprod, cummin and cummax scan in NT2's order under both policies.
//
// Storage is mirrored: a table keeps a host copy and a device copy with
// validity flags, so host element access (setup loops) and device evaluation
// can be mixed; the device copy is the HIP counterpart of cuda_buffer
// (modules/arch/cuda/include/nt2/sdk/cuda/memory/buffer.hpp:36-242).
//
// Floating-point policy: NT2_HIP_POLICY (default nt2::hip::exact_policy: NT2's
// two-rounding fma, bit-identical to the CPU evaluation; nt2::hip::fast_policy
// fuses NT2's fma call sites).
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <tuple>
#include <utility>
#include <vector>

#include "nt2/hip/dimwise.hpp"
#include "nt2/hip/math.hpp"
#include "nt2/hip/math_compose.hpp"

#ifndef NT2_HIP_POLICY
#define NT2_HIP_POLICY nt2::hip::exact_policy
#endif

namespace nt2 {

// ---- architecture site tags ------------------------------------------------
// NT2 stacks sites (modules/arch/cuda/include/nt2/sdk/cuda/cuda.hpp:16-24);
// this header's evaluator is the hip_ site over the host site.
namespace tag {
struct cpu_ {};
template <class Site> struct hip_ : Site { typedef Site parent; };
}  // namespace tag
typedef tag::hip_<tag::cpu_> default_site;

namespace hip { namespace fe {

inline void check(hipError_t e, const char* what) {
  if (e != hipSuccess) throw std::runtime_error(std::string("nt2 hip site: ") + what + ": " + hipGetErrorString(e));
}

// pack width per element type: NT2's AVX cardinal for f32/f64, one 128-bit vector for bytes
template <class T> struct pack_width { static constexpr int value = 16 / (int)sizeof(T) * (sizeof(T) >= 4 ? 2 : 1); };
template <> struct pack_width<float> { static constexpr int value = 8; };
template <> struct pack_width<double> { static constexpr int value = 4; };
template <> struct pack_width<uint8_t> { static constexpr int value = 16; };
template <> struct pack_width<bool> { static constexpr int value = 16; };

}}  // namespace hip::fe

// ---- extents ----------------------------------------------------------------
// up to NT2_MAX_DIMENSIONS = 4 dimensions, trailing sizes 1
struct of_size_ {
  size_t d[4];
  size_t numel() const { return d[0] * d[1] * d[2] * d[3]; }
  bool operator==(const of_size_& o) const {
    return d[0] == o.d[0] && d[1] == o.d[1] && d[2] == o.d[2] && d[3] == o.d[3];
  }
};
// extent of an expression: none for a scalar
struct ext_t {
  bool has;
  of_size_ s;
};
inline ext_t merge_ext(const ext_t& x, const ext_t& y) {
  if (x.has && y.has && !(x.s == y.s)) throw std::invalid_argument("nt2: non-conformant table sizes");
  return x.has ? x : y;
}
namespace meta { template <class T> struct as_ {}; }
inline of_size_ of_size(size_t n) { return of_size_{{n, 1, 1, 1}}; }
inline of_size_ of_size(size_t h, size_t w) { return of_size_{{h, w, 1, 1}}; }
inline of_size_ of_size(size_t a, size_t b, size_t c) { return of_size_{{a, b, c, 1}}; }
inline of_size_ of_size(size_t a, size_t b, size_t c, size_t e) { return of_size_{{a, b, c, e}}; }
// first dimension of size other than 1 (1 when there is none)
inline int firstnonsingleton(const of_size_& s) {
  for (int k = 0; k < 4; ++k)
    if (s.d[k] != 1) return k + 1;
  return 1;
}

// table<T> is table<T, void>; table<T, S> takes NT2's settings S (below)
template <class T, class S = void> class table;

// ---- expression node types ----------------------------------------------------
namespace expr {

struct is_expr_base {};

// device views: trivially copyable, passed to the fused kernel by value
template <class T> struct term_v { const T* p; };
template <class T> struct scal_v { T v; };

template <class T> struct terminal : is_expr_base {
  typedef T value_type;
  const table<T>* t;
  ext_t ext() const;
  void prepare() const;
  term_v<T> view() const;
};
template <class T> struct scalar : is_expr_base {
  typedef T value_type;
  T v;
  ext_t ext() const { return ext_t{false, of_size_{{0, 1, 1, 1}}}; }
  void prepare() const {}
  scal_v<T> view() const { return scal_v<T>{v}; }
};

template <class Tag, class A> struct node1 : is_expr_base {
  typedef typename Tag::template result<typename A::value_type>::type value_type;
  A a;
  ext_t ext() const { return a.ext(); }
  void prepare() const { a.prepare(); }
  struct view_t { decltype(std::declval<A>().view()) a; };
  view_t view() const { return view_t{a.view()}; }
};
template <class Tag, class A, class B> struct node2 : is_expr_base {
  typedef typename Tag::template result<typename A::value_type, typename B::value_type>::type value_type;
  A a;
  B b;
  ext_t ext() const { return merge_ext(a.ext(), b.ext()); }
  void prepare() const { a.prepare(); b.prepare(); }
  struct view_t { decltype(std::declval<A>().view()) a; decltype(std::declval<B>().view()) b; };
  view_t view() const { return view_t{a.view(), b.view()}; }
};
template <class Tag, class A, class B, class C> struct node3 : is_expr_base {
  typedef typename Tag::template result<typename A::value_type, typename B::value_type,
                                        typename C::value_type>::type value_type;
  A a;
  B b;
  C c;
  ext_t ext() const { return merge_ext(merge_ext(a.ext(), b.ext()), c.ext()); }
  void prepare() const { a.prepare(); b.prepare(); c.prepare(); }
  struct view_t {
    decltype(std::declval<A>().view()) a;
    decltype(std::declval<B>().view()) b;
    decltype(std::declval<C>().view()) c;
  };
  view_t view() const { return view_t{a.view(), b.view(), c.view()}; }
};

}  // namespace expr

// ---- result-type helpers and tags ---------------------------------------------
namespace tagfn {
template <class X, class Y> struct common { typedef typename std::conditional<sizeof(X) >= sizeof(Y), X, Y>::type type; };
template <class X> struct common<X, bool> { typedef X type; };
template <class Y> struct common<bool, Y> { typedef Y type; };

struct same1 { template <class A> struct result { typedef A type; }; };
// a unary function whose pack form and scalar (tail) form may differ
struct same1p : same1 {};
struct same2 { template <class A, class B> struct result { typedef typename common<A, B>::type type; }; };
struct logical2 { template <class A, class B> struct result { typedef bool type; }; };
struct sel3 { template <class C, class A, class B> struct result { typedef typename common<A, B>::type type; }; };
struct same3 { template <class A, class B, class C> struct result { typedef typename common<typename common<A, B>::type, C>::type type; }; };
struct selinc2 { template <class C, class A> struct result { typedef A type; }; };
struct izeo1 { template <class C> struct result { typedef uint8_t type; }; };
// pow: the base's type for an integral exponent (generic/pow.hpp:110-168)
struct pow2 {
  template <class A, class B> struct result {
    typedef typename std::conditional<std::is_integral<B>::value, A, typename common<A, B>::type>::type type;
  };
};

#define NT2H_TAG(name, base) struct name : base {};
NT2H_TAG(plus_, same2) NT2H_TAG(minus_, same2) NT2H_TAG(multiplies_, same2) NT2H_TAG(divides_, same2)
NT2H_TAG(max_, same2) NT2H_TAG(min_, same2) NT2H_TAG(adds_, same2) NT2H_TAG(subs_, same2)
NT2H_TAG(lt_, logical2) NT2H_TAG(gt_, logical2) NT2H_TAG(le_, logical2) NT2H_TAG(ge_, logical2)
NT2H_TAG(eq_, logical2) NT2H_TAG(neq_, logical2) NT2H_TAG(land_, logical2) NT2H_TAG(lor_, logical2)
NT2H_TAG(neg_, same1) NT2H_TAG(exp_, same1) NT2H_TAG(log_, same1) NT2H_TAG(sqrt_, same1)
NT2H_TAG(erfc_, same1) NT2H_TAG(normcdf_mag_, same1) NT2H_TAG(normcdf_join_, same2) NT2H_TAG(fastnormcdf_, same1) NT2H_TAG(cos_, same1)
NT2H_TAG(sin_, same1) NT2H_TAG(abs_, same1) NT2H_TAG(sqr_, same1) NT2H_TAG(rec_, same1)
NT2H_TAG(exp2_, same1) NT2H_TAG(exp10_, same1) NT2H_TAG(log2_, same1) NT2H_TAG(log10_, same1)
NT2H_TAG(floor_, same1) NT2H_TAG(ceil_, same1) NT2H_TAG(trunc_, same1) NT2H_TAG(round_, same1)
NT2H_TAG(round2even_, same1) NT2H_TAG(sign_, same1) NT2H_TAG(expm1_, same1) NT2H_TAG(log1p_, same1)
NT2H_TAG(tanh_, same1) NT2H_TAG(sinh_, same1) NT2H_TAG(cosh_, same1) NT2H_TAG(atan_, same1)
NT2H_TAG(atan2_, same2) NT2H_TAG(tan_, same1) NT2H_TAG(asin_, same1) NT2H_TAG(acos_, same1)
NT2H_TAG(erf_, same1) NT2H_TAG(cbrt_, same1) NT2H_TAG(asinh_, same1) NT2H_TAG(acosh_, same1)
NT2H_TAG(atanh_, same1) NT2H_TAG(cot_, same1) NT2H_TAG(sec_, same1) NT2H_TAG(csc_, same1)
NT2H_TAG(sinpi_, same1) NT2H_TAG(cospi_, same1) NT2H_TAG(tanpi_, same1) NT2H_TAG(cotpi_, same1)
NT2H_TAG(sind_, same1) NT2H_TAG(cosd_, same1) NT2H_TAG(tand_, same1) NT2H_TAG(cotd_, same1)
NT2H_TAG(coth_, same1) NT2H_TAG(sech_, same1) NT2H_TAG(csch_, same1) NT2H_TAG(acoth_, same1)
NT2H_TAG(asech_, same1p) NT2H_TAG(acsch_, same1p) NT2H_TAG(acot_, same1p) NT2H_TAG(asec_, same1p)
NT2H_TAG(acsc_, same1p) NT2H_TAG(acotd_, same1p) NT2H_TAG(asecd_, same1p) NT2H_TAG(acscd_, same1p)
NT2H_TAG(acosd_, same1p) NT2H_TAG(acospi_, same1p) NT2H_TAG(asind_, same1p) NT2H_TAG(asinpi_, same1p)
NT2H_TAG(atand_, same1p) NT2H_TAG(atanpi_, same1p) NT2H_TAG(indeg_, same1) NT2H_TAG(inrad_, same1)
NT2H_TAG(secd_, same1) NT2H_TAG(cscd_, same1) NT2H_TAG(secpi_, same1) NT2H_TAG(cscpi_, same1)
NT2H_TAG(sinc_, same1p) NT2H_TAG(sincpi_, same1p) NT2H_TAG(atan2d_, same2)
NT2H_TAG(sqrt1pm1_, same1) NT2H_TAG(exprecneg_, same1) NT2H_TAG(exprecnegc_, same1) NT2H_TAG(reallog_, same1)
NT2H_TAG(realsqrt_, same1) NT2H_TAG(expx2_, same1p) NT2H_TAG(erfinv_, same1) NT2H_TAG(erfcinv_, same1) NT2H_TAG(sinhc_, same1) NT2H_TAG(logspace_add_, same2) NT2H_TAG(logspace_sub_, same2)
NT2H_TAG(pow_, pow2)
NT2H_TAG(fma_, same3) NT2H_TAG(fnms_, same3) NT2H_TAG(if_else_, sel3)
NT2H_TAG(selinc_, selinc2) NT2H_TAG(seldec_, selinc2) NT2H_TAG(if_zero_else_one_, izeo1)
#undef NT2H_TAG
}  // namespace tagfn

namespace expr {
// A rough VALU cost of a node's own operation (arithmetic 1, a division or
// square root 4, a transcendental function 20)
template <class Tag> struct tag_cost {
  static constexpr int value =
      std::is_same<Tag, tagfn::sqrt_>::value || std::is_same<Tag, tagfn::rec_>::value ||
              std::is_same<Tag, tagfn::divides_>::value
          ? 4
          : (std::is_same<Tag, tagfn::neg_>::value || std::is_same<Tag, tagfn::abs_>::value ||
                     std::is_same<Tag, tagfn::sqr_>::value || std::is_same<Tag, tagfn::sign_>::value ||
                     std::is_same<Tag, tagfn::floor_>::value || std::is_same<Tag, tagfn::ceil_>::value ||
                     std::is_same<Tag, tagfn::trunc_>::value || std::is_same<Tag, tagfn::round_>::value ||
                     std::is_same<Tag, tagfn::round2even_>::value ||
                     std::is_same<Tag, tagfn::if_zero_else_one_>::value || std::is_same<Tag, tagfn::plus_>::value ||
                     std::is_same<Tag, tagfn::minus_>::value || std::is_same<Tag, tagfn::multiplies_>::value ||
                     std::is_same<Tag, tagfn::max_>::value || std::is_same<Tag, tagfn::min_>::value ||
                     std::is_same<Tag, tagfn::adds_>::value || std::is_same<Tag, tagfn::subs_>::value ||
                     std::is_base_of<tagfn::logical2, Tag>::value || std::is_same<Tag, tagfn::fma_>::value ||
                     std::is_same<Tag, tagfn::fnms_>::value || std::is_same<Tag, tagfn::if_else_>::value ||
                     std::is_same<Tag, tagfn::selinc_>::value || std::is_same<Tag, tagfn::seldec_>::value ||
                     std::is_same<Tag, tagfn::normcdf_join_>::value
                 ? 1
                 : 20);
};
}  // namespace expr

// ---- device evaluation of packs -------------------------------------------------
namespace hip { namespace fe {
namespace dm = nt2::hip::dm;
typedef NT2_HIP_POLICY P;

template <int G, class T> struct pk { T v[G]; };

// A (1 or 2): i is a pack start and the terminal 16-byte aligned, so the pack
// is read with 16-byte loads, nontemporal for A == 2 (a cheap statement whose
// terminals are each read once; a costly one may read a terminal several
// times, which nontemporal loads would not share)
template <class T, int G, int A = 0>
__device__ __forceinline__ void load(const expr::term_v<T>& t, size_t i, T (&r)[G]) {
  if constexpr (A && std::is_arithmetic<T>::value && !std::is_same<T, bool>::value &&
                (G * sizeof(T)) % 16 == 0) {
    constexpr int V = 16 / (int)sizeof(T);
    typedef T vt __attribute__((ext_vector_type(V)));
    const vt* q = reinterpret_cast<const vt*>(t.p + i);
#pragma unroll
    for (int k = 0; k < G / V; ++k) {
      vt w;
      if constexpr (A == 2) w = __builtin_nontemporal_load(q + k);  // streamed once
      else w = q[k];
#pragma unroll
      for (int l = 0; l < V; ++l) r[k * V + l] = w[l];
    }
  } else {
#pragma unroll
    for (int l = 0; l < G; ++l) r[l] = t.p[i + l];
  }
}
template <class T, int G, int A = 0>
__device__ __forceinline__ void load(const expr::scal_v<T>& s, size_t, T (&r)[G]) {
#pragma unroll
  for (int l = 0; l < G; ++l) r[l] = s.v;
}

template <class T> struct is_f { static constexpr bool value = std::is_floating_point<T>::value; };

// Functions with no integer form here fail to compile on integer tables
// instead of producing a value.
template <class T> __device__ __forceinline__ T no_integer_form() {
  static_assert(sizeof(T) == 0, "this function is defined for floating-point tables only");
  return T();
}

template <class Tag> struct apply1;
#define NT2H_UN(TAG, EXPR_F, EXPR_D, EXPR_I)                                         \
  template <> struct apply1<tagfn::TAG> {                                            \
    template <int G, class T> __device__ __forceinline__ static void run(T (&a)[G], T (&r)[G]) { \
      _Pragma("clang fp contract(off)")                                              \
      _Pragma("unroll") for (int l = 0; l < G; ++l) {                                \
        const T x = a[l];                                                            \
        if constexpr (std::is_same<T, float>::value) { r[l] = (EXPR_F); }          \
        else if constexpr (std::is_same<T, double>::value) { r[l] = (EXPR_D); }    \
        else { r[l] = (T)(EXPR_I); }                                                 \
      }                                                                              \
    }                                                                                \
  };
NT2H_UN(neg_, -x, -x, -x)
NT2H_UN(exp_, dm::exp_f<P>(x), dm::exp_d<P>(x), no_integer_form<T>())
NT2H_UN(log_, dm::log_f<P>(x), dm::log_d<P>(x), no_integer_form<T>())
NT2H_UN(sqrt_, __builtin_sqrtf(x), __builtin_sqrt(x), no_integer_form<T>())
NT2H_UN(erfc_, dm::erfc_f<P>(x), (G == 1 ? dm::erfc_d<P, true>(x) : dm::erfc_d<P, false>(x)), no_integer_form<T>())
// normcdf(x) is built as normcdf_join(x, normcdf_mag(|x|)) (the builder below)
NT2H_UN(normcdf_mag_, dm::normcdf_mag_f<P>(x), dm::normcdf_mag_d<P>(x), no_integer_form<T>())
NT2H_UN(fastnormcdf_, dm::fastnormcdf_f<P>(x), (double)dm::fastnormcdf_f<P>((float)x), no_integer_form<T>())
NT2H_UN(abs_, __builtin_fabsf(x), __builtin_fabs(x), x)
NT2H_UN(sqr_, x * x, x * x, x * x)
NT2H_UN(rec_, 1.0f / x, 1.0 / x, no_integer_form<T>())
// exp2, exp10, log2, log10: NT2's expa / logarithm kernels (math.hpp)
NT2H_UN(exp2_, dm::exp2_f<P>(x), dm::exp2_d<P>(x), no_integer_form<T>())
NT2H_UN(exp10_, dm::exp10_f<P>(x), dm::exp10_d<P>(x), no_integer_form<T>())
NT2H_UN(log2_, dm::log2_f<P>(x), dm::log2_d<P>(x), no_integer_form<T>())
NT2H_UN(log10_, dm::log10_f<P>(x), dm::log10_d<P>(x), no_integer_form<T>())
// rounding is the identity on integers (boost/simd/arithmetic/functions/scalar/
// round.hpp:26-34 and the floor/ceil/trunc/round2even integer forms)
NT2H_UN(floor_, __builtin_floorf(x), __builtin_floor(x), x)
NT2H_UN(ceil_, __builtin_ceilf(x), __builtin_ceil(x), x)
NT2H_UN(trunc_, __builtin_truncf(x), __builtin_trunc(x), x)
NT2H_UN(round_, dm::round_nt2(x), dm::round_nt2(x), x)
NT2H_UN(round2even_, __builtin_rintf(x), __builtin_rint(x), x)
NT2H_UN(sign_, dm::sign_nt2(x), dm::sign_nt2(x), (x > T(0)) - (x < T(0)))
NT2H_UN(expm1_, dm::expm1_f<P>(x), dm::expm1_d<P>(x), no_integer_form<T>())
NT2H_UN(log1p_, dm::log1p_f<P>(x), dm::log1p_d<P>(x), no_integer_form<T>())
NT2H_UN(tanh_, dm::tanh_f<P>(x), dm::tanh_d<P>(x), no_integer_form<T>())
NT2H_UN(sinh_, dm::sinh_f<P>(x), dm::sinh_d<P>(x), no_integer_form<T>())
NT2H_UN(cosh_, dm::cosh_f<P>(x), dm::cosh_d<P>(x), no_integer_form<T>())
// atan: the pack form for packs, NT2's scalar form for the tail (G == 1)
NT2H_UN(atan_, (G == 1 ? dm::atan_f<P, true>(x) : dm::atan_f<P, false>(x)),
        (G == 1 ? dm::atan_d<P, true>(x) : dm::atan_d<P, false>(x)), no_integer_form<T>())
// asin, acos: one f32 form; f64 packs and tails differ (G == 1 is the tail)
NT2H_UN(asin_, dm::asin_f<P>(x), (G == 1 ? dm::asin_d<P, true>(x) : dm::asin_d<P, false>(x)), no_integer_form<T>())
NT2H_UN(acos_, dm::acos_f<P>(x), (G == 1 ? dm::acos_d<P, true>(x) : dm::acos_d<P, false>(x)), no_integer_form<T>())
// erf: one f32 form, f64 tails differ; cbrt, atanh: one form; asinh, acosh:
// tails differ in both precisions
NT2H_UN(erf_, dm::erf_f<P>(x), (G == 1 ? dm::erf_d<P, true>(x) : dm::erf_d<P, false>(x)), no_integer_form<T>())
NT2H_UN(cbrt_, dm::cbrt_f<P>(x), dm::cbrt_d<P>(x), no_integer_form<T>())
NT2H_UN(asinh_, (G == 1 ? dm::asinh_f<P, true>(x) : dm::asinh_f<P, false>(x)),
        (G == 1 ? dm::asinh_d<P, true>(x) : dm::asinh_d<P, false>(x)), no_integer_form<T>())
NT2H_UN(acosh_, (G == 1 ? dm::acosh_f<P, true>(x) : dm::acosh_f<P, false>(x)),
        (G == 1 ? dm::acosh_d<P, true>(x) : dm::acosh_d<P, false>(x)), no_integer_form<T>())
NT2H_UN(atanh_, dm::atanh_f<P>(x), dm::atanh_d<P>(x), no_integer_form<T>())
// trigonometric functions of pi x and of x degrees: lane-wise reductions, one form
#define NT2H_UNIT(TAG, U, K) \
  NT2H_UN(TAG, (dm::trigu_f<P, dm::U>(dm::K, x)), (dm::trigu_d<P, dm::U>(dm::K, x)), no_integer_form<T>())
NT2H_UNIT(sinpi_, u_pi, k_sin) NT2H_UNIT(cospi_, u_pi, k_cos) NT2H_UNIT(tanpi_, u_pi, k_tan)
NT2H_UNIT(cotpi_, u_pi, k_cot) NT2H_UNIT(sind_, u_deg, k_sin) NT2H_UNIT(cosd_, u_deg, k_cos)
NT2H_UNIT(tand_, u_deg, k_tan) NT2H_UNIT(cotd_, u_deg, k_cot)
#undef NT2H_UNIT
// the compositions (nt2/hip/math_compose.hpp); G == 1 is NT2's scalar form
#define NT2H_COMP(TAG, CALL)                                                                   \
  NT2H_UN(TAG, (G == 1 ? dm::CALL<P, true>(x) : dm::CALL<P, false>(x)),                        \
          (G == 1 ? dm::CALL<P, true>(x) : dm::CALL<P, false>(x)), no_integer_form<T>())
NT2H_COMP(asech_, asech_t) NT2H_COMP(acsch_, acsch_t) NT2H_COMP(acot_, acot_t) NT2H_COMP(asec_, asec_t)
NT2H_COMP(acsc_, acsc_t) NT2H_COMP(acotd_, acotd_t) NT2H_COMP(asecd_, asecd_t) NT2H_COMP(acscd_, acscd_t)
NT2H_COMP(acosd_, acosd_t) NT2H_COMP(acospi_, acospi_t) NT2H_COMP(asind_, asind_t) NT2H_COMP(asinpi_, asinpi_t)
NT2H_COMP(atand_, atand_t) NT2H_COMP(atanpi_, atanpi_t) NT2H_COMP(sincpi_, sincpi_t)
#undef NT2H_COMP
NT2H_UN(coth_, dm::coth_t<P>(x), dm::coth_t<P>(x), no_integer_form<T>())
NT2H_UN(sech_, dm::sech_t<P>(x), dm::sech_t<P>(x), no_integer_form<T>())
NT2H_UN(csch_, dm::csch_t<P>(x), dm::csch_t<P>(x), no_integer_form<T>())
NT2H_UN(acoth_, dm::acoth_t<P>(x), dm::acoth_t<P>(x), no_integer_form<T>())
NT2H_UN(indeg_, dm::indeg(x), dm::indeg(x), no_integer_form<T>())
NT2H_UN(inrad_, dm::inrad(x), dm::inrad(x), no_integer_form<T>())
NT2H_UN(secd_, dm::secd_t<P>(x), dm::secd_t<P>(x), no_integer_form<T>())
NT2H_UN(cscd_, dm::cscd_t<P>(x), dm::cscd_t<P>(x), no_integer_form<T>())
NT2H_UN(secpi_, dm::secpi_t<P>(x), dm::secpi_t<P>(x), no_integer_form<T>())
NT2H_UN(cscpi_, dm::cscpi_t<P>(x), dm::cscpi_t<P>(x), no_integer_form<T>())
NT2H_UN(sqrt1pm1_, dm::sqrt1pm1(x), dm::sqrt1pm1(x), no_integer_form<T>())
NT2H_UN(erfinv_, dm::erfinv_t<P>(x), dm::erfinv_t<P>(x), no_integer_form<T>())
NT2H_UN(erfcinv_, dm::erfcinv_t<P>(x), dm::erfcinv_t<P>(x), no_integer_form<T>())
NT2H_UN(sinhc_, dm::sinhc_t<P>(x), dm::sinhc_t<P>(x), no_integer_form<T>())
NT2H_UN(exprecneg_, dm::exprecneg_t<P>(x), dm::exprecneg_t<P>(x), no_integer_form<T>())
NT2H_UN(exprecnegc_, dm::exprecnegc_t<P>(x), dm::exprecnegc_t<P>(x), no_integer_form<T>())
NT2H_UN(reallog_, dm::log_t<P>(x), dm::log_t<P>(x), no_integer_form<T>())
NT2H_UN(realsqrt_, dm::sqrtv(x), dm::sqrtv(x), no_integer_form<T>())
NT2H_UN(expx2_, (G == 1 ? dm::expx2_t<P, true>(x) : dm::expx2_t<P, false>(x)),
        (G == 1 ? dm::expx2_t<P, true>(x) : dm::expx2_t<P, false>(x)), no_integer_form<T>())
#undef NT2H_UN
template <> struct apply1<tagfn::sinc_> {
  template <int G, class T> __device__ __forceinline__ static void run(T (&a)[G], T (&r)[G]) {
    static_assert(std::is_floating_point<T>::value, "sinc is defined for floating-point tables only");
    dm::sinc_group<P, G>(a, r);
  }
};

template <> struct apply1<tagfn::cos_> {
  template <int G, class T> __device__ __forceinline__ static void run(T (&a)[G], T (&r)[G]) {
    if constexpr (std::is_same<T, float>::value) dm::trig_group_f<P, G>(dm::k_cos, a, r, nullptr);
    else dm::trig_group_d<P, G>(dm::k_cos, a, r, nullptr);
  }
};
template <> struct apply1<tagfn::tan_> {
  template <int G, class T> __device__ __forceinline__ static void run(T (&a)[G], T (&r)[G]) {
    if constexpr (std::is_same<T, float>::value) dm::trig_group_f<P, G>(dm::k_tan, a, r, nullptr);
    else dm::trig_group_d<P, G>(dm::k_tan, a, r, nullptr);
  }
};
template <> struct apply1<tagfn::sin_> {
  template <int G, class T> __device__ __forceinline__ static void run(T (&a)[G], T (&r)[G]) {
    if constexpr (std::is_same<T, float>::value) dm::trig_group_f<P, G>(dm::k_sin, a, r, nullptr);
    else dm::trig_group_d<P, G>(dm::k_sin, a, r, nullptr);
  }
};
// cot (radian, per-pack reduction), sec = 1/cos, csc = 1/sin
#define NT2H_RADIAN(TAG, K)                                                                          \
  template <> struct apply1<tagfn::TAG> {                                                            \
    template <int G, class T> __device__ __forceinline__ static void run(T (&a)[G], T (&r)[G]) {    \
      if constexpr (std::is_same<T, float>::value) dm::trig_group_f<P, G>(dm::K, a, r, nullptr);     \
      else dm::trig_group_d<P, G>(dm::K, a, r, nullptr);                                             \
    }                                                                                                \
  };
NT2H_RADIAN(cot_, k_cot) NT2H_RADIAN(sec_, k_sec) NT2H_RADIAN(csc_, k_csc)
#undef NT2H_RADIAN

template <class Tag> struct apply2;
#define NT2H_BIN(TAG, EXPR)                                                           \
  template <> struct apply2<tagfn::TAG> {                                             \
    template <class R, class T, class U>                                              \
    __device__ __forceinline__ static R one(T x, U y) {                               \
      _Pragma("clang fp contract(off)")                                               \
      return (R)(EXPR);                                                               \
    }                                                                                 \
  };
NT2H_BIN(plus_, x + y)
NT2H_BIN(minus_, x - y)
NT2H_BIN(multiplies_, x * y)
NT2H_BIN(divides_, x / y)
NT2H_BIN(max_, x > y ? x : y)
NT2H_BIN(min_, x < y ? x : y)
NT2H_BIN(lt_, x < y)
NT2H_BIN(gt_, x > y)
NT2H_BIN(le_, x <= y)
NT2H_BIN(ge_, x >= y)
NT2H_BIN(eq_, x == y)
NT2H_BIN(neq_, x != y)
// logical_and / logical_or: a nonzero (or NaN) operand is true, as NT2's is_nez
NT2H_BIN(land_, (x != 0) && (y != 0))
NT2H_BIN(lor_, (x != 0) || (y != 0))
#undef NT2H_BIN
template <> struct apply2<tagfn::adds_> {
  template <class R, class T, class U> __device__ __forceinline__ static R one(T x, U y) {
    if constexpr (std::is_same<R, uint8_t>::value) return (R)dm::adds_u8(x, y);
    else return (R)(x + y);
  }
};
template <> struct apply2<tagfn::subs_> {
  template <class R, class T, class U> __device__ __forceinline__ static R one(T x, U y) {
    if constexpr (std::is_same<R, uint8_t>::value) return (R)dm::subs_u8(x, y);
    else return (R)(x - y);
  }
};
template <> struct apply2<tagfn::selinc_> {  // selinc(c, a) = c ? a+1 : a (wrapping)
  template <class R, class C, class A> __device__ __forceinline__ static R one(C c, A a) {
    return c ? (R)(a + 1) : (R)a;
  }
};
template <> struct apply2<tagfn::seldec_> {
  template <class R, class C, class A> __device__ __forceinline__ static R one(C c, A a) {
    return c ? (R)(a - 1) : (R)a;
  }
};

}}  // namespace hip::fe

// ---- user-facing table ------------------------------------------------------
template <class T>
class table<T, void> {
 public:
  typedef T value_type;
  table() : ext_{{0, 1, 1, 1}} {}
  explicit table(of_size_ s) : ext_{{0, 1, 1, 1}} { resize(s); }
  table(const table& o) : ext_{{0, 1, 1, 1}} { *this = o; }
  template <class E, class = typename std::enable_if<std::is_base_of<expr::is_expr_base, E>::value>::type>
  table(const E& e) : ext_{{0, 1, 1, 1}} { assign(e); }
  ~table() { release(); }

  table& operator=(const table& o) {
    if (this == &o) return *this;
    resize(o.ext_);
    if (o.dev_valid_ && !o.host_valid_) {
      ensure_dev_alloc();
      hip::fe::check(hipMemcpy(dev_, o.dev_, bytes(), hipMemcpyDeviceToDevice), "copy");
      dev_valid_ = true;
      host_valid_ = false;
    } else {
      o.sync_host();
      host_ = o.host_;
      host_valid_ = true;
      dev_valid_ = false;
    }
    return *this;
  }
  template <class E>
  typename std::enable_if<std::is_base_of<expr::is_expr_base, E>::value, table&>::type
  operator=(const E& e) {
    assign(e);
    return *this;
  }

  void resize(of_size_ s) {
    if (s == ext_) return;
    if (s.numel() != numel()) {
      host_.assign(s.numel(), T());
      if (dev_) { hip::fe::check(hipFree(dev_), "free"); dev_ = nullptr; }
      host_valid_ = true;
      dev_valid_ = false;
    }
    ext_ = s;
  }
  size_t numel() const { return ext_.numel(); }
  of_size_ extent() const { return ext_; }
  size_t size(int dim) const { return dim >= 1 && dim <= 4 ? ext_.d[dim - 1] : 1; }

  // 1-based, column-major host access (synchronizes with the device copy)
  T& operator()(size_t i) {
    sync_host();
    dev_valid_ = false;
    return host_[i - 1];
  }
  T operator()(size_t i) const {
    sync_host();
    return host_[i - 1];
  }
  T& operator()(size_t i, size_t j) { return (*this)((j - 1) * ext_.d[0] + i); }
  T operator()(size_t i, size_t j) const { return (*this)((j - 1) * ext_.d[0] + i); }
  T& operator()(size_t i, size_t j, size_t k) { return (*this)(((k - 1) * ext_.d[1] + (j - 1)) * ext_.d[0] + i); }
  T operator()(size_t i, size_t j, size_t k) const {
    return (*this)(((k - 1) * ext_.d[1] + (j - 1)) * ext_.d[0] + i);
  }
  // the value of a one-element table (a reduction to a scalar: T s = sum(x))
  operator T() const {
    if (numel() != 1) throw std::invalid_argument("nt2: table of more than one element used as a scalar");
    return (*this)(1);
  }
  // same elements, new extent of the same number of elements
  void reshape_to(of_size_ s) {
    if (s.numel() != numel()) throw std::invalid_argument("nt2: reshape changes the number of elements");
    ext_ = s;
  }

  // raw access
  const T* device_data() const { sync_dev(); return dev_; }
  // device buffer for a kernel that overwrites every element: the host copy
  // is not uploaded first (an in-place statement has already made this
  // table's device copy current as one of its own terminals)
  T* device_data_for_write() {
    ensure_dev_alloc();
    host_valid_ = false;
    dev_valid_ = true;
    return dev_;
  }
  const T* data() const { sync_host(); return host_.data(); }

  void sync_host() const {
    if (!host_valid_ && dev_valid_) {
      host_.resize(numel());
      hip::fe::check(hipMemcpy(host_.data(), dev_, bytes(), hipMemcpyDeviceToHost), "download");
      host_valid_ = true;
    }
  }
  void sync_dev() const {
    if (numel() == 0) return;
    ensure_dev_alloc();
    if (!dev_valid_) {
      hip::fe::check(hipMemcpy(dev_, host_.data(), bytes(), hipMemcpyHostToDevice), "upload");
      dev_valid_ = true;
    }
  }

  template <class E> void assign(const E& e);

 private:
  size_t bytes() const { return numel() * sizeof(T); }
  void ensure_dev_alloc() const {
    if (!dev_ && numel()) hip::fe::check(hipMalloc((void**)&dev_, bytes()), "alloc");
  }
  void release() {
    if (dev_) (void)hipFree(dev_);
    dev_ = nullptr;
  }
  of_size_ ext_;
  mutable std::vector<T> host_;
  mutable T* dev_ = nullptr;
  mutable bool host_valid_ = true, dev_valid_ = false;
};

// ---- terminals ------------------------------------------------------------------
namespace expr {
template <class T> ext_t terminal<T>::ext() const { return ext_t{true, t->extent()}; }
template <class T> void terminal<T>::prepare() const { t->sync_dev(); }
template <class T> term_v<T> terminal<T>::view() const { return term_v<T>{t->device_data()}; }

template <class X> struct wrap { typedef X type; static const X& get(const X& x) { return x; } };
template <class T, class S> struct wrap<table<T, S>> {
  typedef terminal<T> type;
  static terminal<T> get(const table<T>& t) { terminal<T> r; r.t = &t; return r; }
};
template <class X, class Hint, bool = std::is_arithmetic<X>::value> struct as_expr : wrap<X> {};
template <class X, class Hint> struct as_expr<X, Hint, true> {
  typedef scalar<Hint> type;
  static scalar<Hint> get(X x) { scalar<Hint> s; s.v = (Hint)x; return s; }
};
template <class X> struct is_dsl {
  static constexpr bool value = std::is_base_of<is_expr_base, X>::value;
};
template <class T, class S> struct is_dsl<table<T, S>> { static constexpr bool value = true; };
template <class X> struct vt { typedef typename X::value_type type; };
template <class X, bool = std::is_arithmetic<X>::value> struct vt_or { typedef typename X::value_type type; };
template <class X> struct vt_or<X, true> { typedef X type; };
}  // namespace expr

// ---- table settings -----------------------------------------------------------------
// NT2's table<T, S> takes one option or settings(options...)
// (modules/core/sdk/include/nt2/core/settings/forward/storage_order.hpp:17-33,
// index.hpp:34-57, sharing.hpp:17-18, shape.hpp:32-49, alignment.hpp:16-23,
// storage_duration.hpp:16-17; _1D.._4D in core/utility/of_size/predef.hpp:25-29
// are of_size_ types with every size set at run time).  Here the element
// storage is always column-major, owned, aligned and dense; options that ask
// for exactly that are accepted, the index base and the number of dimensions
// are honoured, and the others stop the compilation with a message.
struct settings {};
struct column_major_ {};
typedef column_major_ matlab_order_;
typedef column_major_ fortran_order_;
struct row_major_ {};
typedef row_major_ C_order_;
template <int B> struct index_ {};
typedef index_<0> C_index_;
typedef index_<1> matlab_index_;
template <int N> struct dims_ {};  // NT2's _1D ... _4D: at most N dimensions, sizes set at run time
typedef dims_<1> _1D;
typedef dims_<2> _2D;
typedef dims_<3> _3D;
typedef dims_<4> _4D;
struct automatic_ {};
struct dynamic_ {};
struct owned_ {};
struct shared_ {};
struct aligned_ {};
struct unaligned_ {};
struct rectangular_ {};
struct conventional_ {};
struct packed_ {};
struct upper_triangular_ {};
struct lower_triangular_ {};
struct diagonal_ {};
struct symmetric_ {};
struct positive_definite_ {};

namespace detail {
// what one option sets: the index base, the number of dimensions, or nothing
template <class O> struct option {
  static constexpr bool known = false;
  static constexpr int base = 0, nbase = 0, dims = 0, ndims = 0;
};
template <int B> struct option<index_<B>> {
  static constexpr bool known = true;
  static constexpr int base = B, nbase = 1, dims = 0, ndims = 0;
};
template <int N> struct option<dims_<N>> {
  static constexpr bool known = true;
  static constexpr int base = 0, nbase = 0, dims = N, ndims = 1;
};
#define NT2H_NEUTRAL(O)                                           \
  template <> struct option<O> {                                  \
    static constexpr bool known = true;                           \
    static constexpr int base = 0, nbase = 0, dims = 0, ndims = 0; \
  };
NT2H_NEUTRAL(column_major_) NT2H_NEUTRAL(automatic_) NT2H_NEUTRAL(dynamic_) NT2H_NEUTRAL(owned_)
NT2H_NEUTRAL(aligned_) NT2H_NEUTRAL(unaligned_) NT2H_NEUTRAL(rectangular_) NT2H_NEUTRAL(conventional_)
#undef NT2H_NEUTRAL
template <class... O> struct options_of {
  static constexpr bool known = (option<O>::known && ... && true);
  static constexpr int nbase = (option<O>::nbase + ... + 0), ndims = (option<O>::ndims + ... + 0);
  static constexpr long base = nbase ? (option<O>::base + ... + 0) : 1;  // NT2's default: matlab_index_
  static constexpr int dims = ndims ? (option<O>::dims + ... + 0) : 4;   // NT2_MAX_DIMENSIONS
};
template <class S> struct parse : options_of<S> {};
template <class... O> struct parse<settings(O...)> : options_of<O...> {};
}  // namespace detail

template <class T, class S>
class table : public table<T, void> {
  typedef table<T, void> base_t;
  typedef detail::parse<S> cfg;
  static_assert(cfg::known, "nt2 table settings: only column-major, owned, dense storage (with an index "
                             "base and a number of dimensions) is provided by this implementation");
  static_assert(cfg::nbase <= 1 && cfg::ndims <= 1, "nt2 table settings: one index base and one number of dimensions");
  static_assert(cfg::dims >= 1 && cfg::dims <= 4, "nt2 table settings: 1 to 4 dimensions");

 public:
  static constexpr long index_base = cfg::base;
  static constexpr int max_dims = cfg::dims;
  typedef T value_type;

  table() = default;
  explicit table(of_size_ s) : base_t() { resize(s); }
  table(const table& o) : base_t(o) {}
  table(const base_t& o) : base_t() { *this = o; }
  template <class E, class = typename std::enable_if<std::is_base_of<expr::is_expr_base, E>::value>::type>
  table(const E& e) : base_t() { *this = e; }

  table& operator=(const table& o) {
    base_t::operator=(o);
    return *this;
  }
  table& operator=(const base_t& o) {
    check_dims(o.extent());
    base_t::operator=(o);
    return *this;
  }
  template <class E>
  typename std::enable_if<std::is_base_of<expr::is_expr_base, E>::value, table&>::type operator=(const E& e) {
    const ext_t x = expr::wrap<E>::get(e).ext();
    if (x.has) check_dims(x.s);
    base_t::operator=(e);
    return *this;
  }
  void resize(of_size_ s) {
    check_dims(s);
    base_t::resize(s);
  }
  // index_base-based, column-major host access
  T& operator()(size_t i) { return base_t::operator()(i + 1 - index_base); }
  T operator()(size_t i) const { return base_t::operator()(i + 1 - index_base); }
  T& operator()(size_t i, size_t j) { return base_t::operator()(i + 1 - index_base, j + 1 - index_base); }
  T operator()(size_t i, size_t j) const { return base_t::operator()(i + 1 - index_base, j + 1 - index_base); }
  T& operator()(size_t i, size_t j, size_t k) {
    return base_t::operator()(i + 1 - index_base, j + 1 - index_base, k + 1 - index_base);
  }
  T operator()(size_t i, size_t j, size_t k) const {
    return base_t::operator()(i + 1 - index_base, j + 1 - index_base, k + 1 - index_base);
  }

 private:
  static void check_dims(const of_size_& s) {
    for (int d = max_dims; d < 4; ++d)
      if (s.d[d] != 1) throw std::invalid_argument("nt2: extent has more dimensions than the table's settings allow");
  }
};

// ---- builders ---------------------------------------------------------------------
#define NT2H_ENABLE1(A) typename std::enable_if<expr::is_dsl<A>::value, int>::type = 0
#define NT2H_ENABLE2(A, B) typename std::enable_if<expr::is_dsl<A>::value || expr::is_dsl<B>::value, int>::type = 0

template <class Tag, class A>
expr::node1<Tag, typename expr::wrap<A>::type> make1(const A& a) {
  expr::node1<Tag, typename expr::wrap<A>::type> n;
  n.a = expr::wrap<A>::get(a);
  return n;
}
template <class Tag, class A, class B>
auto make2(const A& a, const B& b) {
  typedef typename std::conditional<expr::is_dsl<A>::value, typename expr::vt_or<typename expr::wrap<A>::type>::type,
                                    typename expr::vt_or<typename expr::wrap<B>::type>::type>::type hint;
  typedef typename expr::as_expr<A, hint>::type EA;
  typedef typename expr::as_expr<B, hint>::type EB;
  expr::node2<Tag, EA, EB> n;
  n.a = expr::as_expr<A, hint>::get(a);
  n.b = expr::as_expr<B, hint>::get(b);
  return n;
}
template <class Tag, class A, class B, class C>
auto make3(const A& a, const B& b, const C& c) {
  // the branches' type; if_else(c, 1.0f, 0.0f) takes its scalars' type, not the condition's
  typedef typename std::conditional<
      expr::is_dsl<B>::value, typename expr::vt_or<typename expr::wrap<B>::type>::type,
      typename std::conditional<
          expr::is_dsl<C>::value, typename expr::vt_or<typename expr::wrap<C>::type>::type,
          typename std::conditional<std::is_same<Tag, tagfn::if_else_>::value && std::is_arithmetic<B>::value, B,
                                    typename expr::vt_or<typename expr::wrap<A>::type>::type>::type>::type>::type hint;
  typedef typename expr::as_expr<A, hint>::type EA;
  typedef typename expr::as_expr<B, hint>::type EB;
  typedef typename expr::as_expr<C, hint>::type EC;
  expr::node3<Tag, EA, EB, EC> n;
  n.a = expr::as_expr<A, hint>::get(a);
  n.b = expr::as_expr<B, hint>::get(b);
  n.c = expr::as_expr<C, hint>::get(c);
  return n;
}

#define NT2H_OP2(op, Tag) \
  template <class A, class B, NT2H_ENABLE2(A, B)> auto operator op(const A& a, const B& b) { return make2<tagfn::Tag>(a, b); }
NT2H_OP2(+, plus_) NT2H_OP2(-, minus_) NT2H_OP2(*, multiplies_) NT2H_OP2(/, divides_)
NT2H_OP2(<, lt_) NT2H_OP2(>, gt_) NT2H_OP2(<=, le_) NT2H_OP2(>=, ge_) NT2H_OP2(==, eq_) NT2H_OP2(!=, neq_)
NT2H_OP2(&&, land_) NT2H_OP2(||, lor_)
#undef NT2H_OP2
#define NT2H_FN2(name, Tag) \
  template <class A, class B, NT2H_ENABLE2(A, B)> auto name(const A& a, const B& b) { return make2<tagfn::Tag>(a, b); }
NT2H_FN2(logical_and, land_) NT2H_FN2(logical_or, lor_)
#undef NT2H_FN2
template <class A, NT2H_ENABLE1(A)> auto operator-(const A& a) { return make1<tagfn::neg_>(a); }

#define NT2H_FN1(name, Tag) \
  template <class A, NT2H_ENABLE1(A)> auto name(const A& a) { return make1<tagfn::Tag>(a); }
NT2H_FN1(exp, exp_) NT2H_FN1(log, log_) NT2H_FN1(sqrt, sqrt_) NT2H_FN1(erfc, erfc_)
NT2H_FN1(fastnormcdf, fastnormcdf_) NT2H_FN1(cos, cos_)
NT2H_FN1(sin, sin_) NT2H_FN1(abs, abs_) NT2H_FN1(sqr, sqr_) NT2H_FN1(rec, rec_)
NT2H_FN1(exp2, exp2_) NT2H_FN1(exp10, exp10_) NT2H_FN1(log2, log2_) NT2H_FN1(log10, log10_)
NT2H_FN1(floor, floor_) NT2H_FN1(ceil, ceil_) NT2H_FN1(trunc, trunc_) NT2H_FN1(round, round_)
NT2H_FN1(round2even, round2even_) NT2H_FN1(sign, sign_)
NT2H_FN1(expm1, expm1_) NT2H_FN1(log1p, log1p_) NT2H_FN1(tanh, tanh_) NT2H_FN1(sinh, sinh_)
NT2H_FN1(cosh, cosh_) NT2H_FN1(atan, atan_) NT2H_FN1(tan, tan_) NT2H_FN1(asin, asin_) NT2H_FN1(acos, acos_)
NT2H_FN1(erf, erf_) NT2H_FN1(cbrt, cbrt_) NT2H_FN1(asinh, asinh_) NT2H_FN1(acosh, acosh_) NT2H_FN1(atanh, atanh_)
NT2H_FN1(cot, cot_) NT2H_FN1(sec, sec_) NT2H_FN1(csc, csc_) NT2H_FN1(sinpi, sinpi_) NT2H_FN1(cospi, cospi_)
NT2H_FN1(tanpi, tanpi_) NT2H_FN1(cotpi, cotpi_) NT2H_FN1(sind, sind_) NT2H_FN1(cosd, cosd_) NT2H_FN1(tand, tand_)
NT2H_FN1(cotd, cotd_)
NT2H_FN1(coth, coth_) NT2H_FN1(sech, sech_) NT2H_FN1(csch, csch_) NT2H_FN1(acoth, acoth_)
NT2H_FN1(asech, asech_) NT2H_FN1(acsch, acsch_) NT2H_FN1(acot, acot_) NT2H_FN1(asec, asec_)
NT2H_FN1(acsc, acsc_) NT2H_FN1(acotd, acotd_) NT2H_FN1(asecd, asecd_) NT2H_FN1(acscd, acscd_)
NT2H_FN1(acosd, acosd_) NT2H_FN1(acospi, acospi_) NT2H_FN1(asind, asind_) NT2H_FN1(asinpi, asinpi_)
NT2H_FN1(atand, atand_) NT2H_FN1(atanpi, atanpi_) NT2H_FN1(indeg, indeg_) NT2H_FN1(inrad, inrad_)
NT2H_FN1(secd, secd_) NT2H_FN1(cscd, cscd_) NT2H_FN1(secpi, secpi_) NT2H_FN1(cscpi, cscpi_)
NT2H_FN1(sinc, sinc_) NT2H_FN1(sincpi, sincpi_)
NT2H_FN1(sqrt1pm1, sqrt1pm1_) NT2H_FN1(exprecneg, exprecneg_) NT2H_FN1(exprecnegc, exprecnegc_)
NT2H_FN1(reallog, reallog_) NT2H_FN1(realsqrt, realsqrt_) NT2H_FN1(expx2, expx2_)
NT2H_FN1(erfinv, erfinv_) NT2H_FN1(erfcinv, erfcinv_) NT2H_FN1(sinhc, sinhc_)
NT2H_FN1(if_zero_else_one, if_zero_else_one_)
#undef NT2H_FN1
#define NT2H_FN2(name, Tag) \
  template <class A, class B, NT2H_ENABLE2(A, B)> auto name(const A& a, const B& b) { return make2<tagfn::Tag>(a, b); }
NT2H_FN2(max, max_) NT2H_FN2(min, min_) NT2H_FN2(adds, adds_) NT2H_FN2(subs, subs_)
NT2H_FN2(selinc, selinc_) NT2H_FN2(seldec, seldec_) NT2H_FN2(atan2, atan2_)
NT2H_FN2(atan2d, atan2d_) NT2H_FN2(logspace_add, logspace_add_) NT2H_FN2(logspace_sub, logspace_sub_)
#undef NT2H_FN2
// pow(a, b): Cephes' pow for floating-point exponents; an integral exponent
// (a scalar, kept in its own type, or an integer table) takes NT2's repeated
// squaring instead, 1/p only for a negative signed exponent
// (modules/core/exponential/include/nt2/exponential/functions/generic/pow.hpp:110-168).
template <class A, class B, NT2H_ENABLE2(A, B)> auto pow(const A& a, const B& b) {
  static_assert(!std::is_same<B, bool>::value, "pow: a logical exponent");
  if constexpr (std::is_integral<B>::value && expr::is_dsl<A>::value) {
    typedef typename expr::wrap<A>::type EA;
    expr::node2<tagfn::pow_, EA, expr::scalar<B>> n;
    n.a = expr::wrap<A>::get(a);
    n.b.v = b;
    return n;
  } else {
    return make2<tagfn::pow_>(a, b);
  }
}
// normcdf(a) = Half*erfc(-Sqrt_2o_2*a)
// (modules/core/statistics/include/nt2/statistics/functions/generic/normcdf.hpp:33),
// built as normcdf_join(a, normcdf_mag(abs(b))) with b = a stripped of a
// negation: |-b| = |b| bit for bit, so normcdf(b) and normcdf(-b) in one
// statement hold the same normcdf_mag node, which is evaluated once.
namespace expr {
template <class X> struct strip_neg {
  typedef X type;
  static const X& get(const X& x) { return x; }
};
template <class A> struct strip_neg<node1<tagfn::neg_, A>> {
  typedef A type;
  static const A& get(const node1<tagfn::neg_, A>& x) { return x.a; }
};
}  // namespace expr
template <class A, NT2H_ENABLE1(A)> auto normcdf(const A& a) {
  typedef typename expr::wrap<A>::type EA;
  const EA x = expr::wrap<A>::get(a);
  return make2<tagfn::normcdf_join_>(x, make1<tagfn::normcdf_mag_>(make1<tagfn::abs_>(expr::strip_neg<EA>::get(x))));
}
// fma(a,b,c) = a*b+c and fnms(a,b,c) = c-a*b, NT2's two-rounding forms
template <class A, class B, class C,
          typename std::enable_if<expr::is_dsl<A>::value || expr::is_dsl<B>::value || expr::is_dsl<C>::value, int>::type = 0>
auto fma(const A& a, const B& b, const C& c) { return make3<tagfn::fma_>(a, b, c); }
template <class A, class B, class C,
          typename std::enable_if<expr::is_dsl<A>::value || expr::is_dsl<B>::value || expr::is_dsl<C>::value, int>::type = 0>
auto fnms(const A& a, const B& b, const C& c) { return make3<tagfn::fnms_>(a, b, c); }
template <class Cd, class A, class B,
          typename std::enable_if<expr::is_dsl<Cd>::value || expr::is_dsl<A>::value || expr::is_dsl<B>::value, int>::type = 0>
auto if_else(const Cd& c, const A& a, const B& b) { return make3<tagfn::if_else_>(c, a, b); }
// scalar fma for host-side constants (NT2's scalar fma, two roundings)
inline float fma(float a, float b, float c) { volatile float p = a * b; return p + c; }
inline double fma(double a, double b, double c) { volatile double p = a * b; return p + c; }
inline float sqr(float a) { return a * a; }
inline double sqr(double a) { return a * a; }

// ---- device evaluation -------------------------------------------------------------
namespace hip { namespace fe {

template <int G, int A = 0, class T>
__device__ __forceinline__ void ev(const expr::term_v<T>& v, size_t i, T (&r)[G]) { load<T, G, A>(v, i, r); }
template <int G, int A = 0, class T>
__device__ __forceinline__ void ev(const expr::scal_v<T>& v, size_t i, T (&r)[G]) { load<T, G, A>(v, i, r); }

template <class E> struct vview;  // value type of a view
template <class T> struct vview<expr::term_v<T>> { typedef T type; };
template <class T> struct vview<expr::scal_v<T>> { typedef T type; };

// node views carry their tag and child views through their host type
template <class Tag, class VA> struct n1v { VA a; };
template <class Tag, class VA, class VB> struct n2v { VA a; VB b; };
template <class Tag, class VA, class VB, class VC> struct n3v { VA a; VB b; VC c; };

}}  // namespace hip::fe

namespace expr {
// host node -> device view (tagged)
template <class X> struct dview;
template <class T> struct dview<terminal<T>> {
  typedef term_v<T> type;
  static type get(const terminal<T>& t) { return t.view(); }
};
template <class T> struct dview<scalar<T>> {
  typedef scal_v<T> type;
  static type get(const scalar<T>& s) { return s.view(); }
};
template <class Tag, class A> struct dview<node1<Tag, A>> {
  typedef hip::fe::n1v<Tag, typename dview<A>::type> type;
  static type get(const node1<Tag, A>& n) { return type{dview<A>::get(n.a)}; }
};
template <class Tag, class A, class B> struct dview<node2<Tag, A, B>> {
  typedef hip::fe::n2v<Tag, typename dview<A>::type, typename dview<B>::type> type;
  static type get(const node2<Tag, A, B>& n) { return type{dview<A>::get(n.a), dview<B>::get(n.b)}; }
};
template <class Tag, class A, class B, class C> struct dview<node3<Tag, A, B, C>> {
  typedef hip::fe::n3v<Tag, typename dview<A>::type, typename dview<B>::type, typename dview<C>::type> type;
  static type get(const node3<Tag, A, B, C>& n) {
    return type{dview<A>::get(n.a), dview<B>::get(n.b), dview<C>::get(n.c)};
  }
};
}  // namespace expr

namespace hip { namespace fe {

template <class Tag, class VA> struct vview<n1v<Tag, VA>> {
  typedef typename Tag::template result<typename vview<VA>::type>::type type;
};
template <class Tag, class VA, class VB> struct vview<n2v<Tag, VA, VB>> {
  typedef typename Tag::template result<typename vview<VA>::type, typename vview<VB>::type>::type type;
};
template <class Tag, class VA, class VB, class VC> struct vview<n3v<Tag, VA, VB, VC>> {
  typedef typename Tag::template result<typename vview<VA>::type, typename vview<VB>::type,
                                        typename vview<VC>::type>::type type;
};

// ---- common subexpressions ----------------------------------------------------------
// A statement evaluates each costly node view (a division, square root or
// function) once per element group even when the user wrote it several times
// (the textbook Black-Scholes tie statement holds d1, sqrt(T), log(S/X) and
// exp(-rT) two to six times each; NT2 evaluates every occurrence).  Views are
// numbered in pre-order.  Before a launch the host tests whether every
// occurrence of each costly node type has the view of that type's first
// occurrence (same tables, same scalars, bit for bit).  If so, the kernel
// instantiated for that case takes every later occurrence's value from the
// first one, a choice made at compile time: the compiler keeps one live value
// per repeated node and no test runs in the kernel.  Otherwise every
// occurrence is evaluated, as NT2 does (tests in the kernel would keep every
// candidate value and every occurrence's operands live, and their registers
// cost more than the evaluations they save).  Either way the values are those
// NT2's evaluation gives, bit for bit.
template <class A, class B> struct vpair { A a; B b; };  // the two right-hand sides of a tie

template <class V> struct vsize { static constexpr int value = 1; };
template <class Tag, class VA> struct vsize<n1v<Tag, VA>> { static constexpr int value = 1 + vsize<VA>::value; };
template <class Tag, class VA, class VB> struct vsize<n2v<Tag, VA, VB>> {
  static constexpr int value = 1 + vsize<VA>::value + vsize<VB>::value;
};
template <class Tag, class VA, class VB, class VC> struct vsize<n3v<Tag, VA, VB, VC>> {
  static constexpr int value = 1 + vsize<VA>::value + vsize<VB>::value + vsize<VC>::value;
};
template <class VA, class VB> struct vsize<vpair<VA, VB>> {
  static constexpr int value = 1 + vsize<VA>::value + vsize<VB>::value;
};

template <class V> struct is_node { static constexpr bool value = false; };
template <class Tag, class VA> struct is_node<n1v<Tag, VA>> { static constexpr bool value = true; };
template <class Tag, class VA, class VB> struct is_node<n2v<Tag, VA, VB>> { static constexpr bool value = true; };
template <class Tag, class VA, class VB, class VC> struct is_node<n3v<Tag, VA, VB, VC>> {
  static constexpr bool value = true;
};

template <class V> struct nchild { static constexpr int value = 0; };
template <class Tag, class VA> struct nchild<n1v<Tag, VA>> { static constexpr int value = 1; };
template <class Tag, class VA, class VB> struct nchild<n2v<Tag, VA, VB>> { static constexpr int value = 2; };
template <class Tag, class VA, class VB, class VC> struct nchild<n3v<Tag, VA, VB, VC>> {
  static constexpr int value = 3;
};
template <class VA, class VB> struct nchild<vpair<VA, VB>> { static constexpr int value = 2; };

// the view at pre-order index K
template <int K, class V>
__host__ __device__ __forceinline__ const auto& vat(const V& v) {
  if constexpr (K == 0) {
    return v;
  } else {
    constexpr int sa = vsize<decltype(v.a)>::value;
    if constexpr (K - 1 < sa) {
      return vat<K - 1>(v.a);
    } else if constexpr (nchild<V>::value == 2) {
      return vat<K - 1 - sa>(v.b);
    } else {
      constexpr int sb = vsize<decltype(v.b)>::value;
      if constexpr (K - 1 - sa < sb) return vat<K - 1 - sa>(v.b);
      else return vat<K - 1 - sa - sb>(v.c);
    }
  }
}
template <class V, int K> struct vat_t { typedef std::decay_t<decltype(vat<K>(std::declval<const V&>()))> type; };

// occurrences of type X in root R at pre-order indices in [K0, K1)
template <class R, class X, int K0, int K1>
constexpr int occurrences() {
  if constexpr (K0 >= K1) return 0;
  else return (std::is_same<typename vat_t<R, K0>::type, X>::value ? 1 : 0) + occurrences<R, X, K0 + 1, K1>();
}

// equality of two views of one type: the same tables and scalars
template <class T> inline bool veq(const expr::term_v<T>& x, const expr::term_v<T>& y) {
  return x.p == y.p;
}
template <class T> inline bool veq(const expr::scal_v<T>& x, const expr::scal_v<T>& y) {
  if constexpr (std::is_floating_point<T>::value) {
    typedef typename std::conditional<sizeof(T) == 4, uint32_t, uint64_t>::type U;
    return __builtin_bit_cast(U, x.v) == __builtin_bit_cast(U, y.v);
  } else {
    return x.v == y.v;
  }
}
template <class Tag, class VA> inline bool veq(const n1v<Tag, VA>& x, const n1v<Tag, VA>& y) {
  return veq(x.a, y.a);
}
template <class Tag, class VA, class VB>
inline bool veq(const n2v<Tag, VA, VB>& x, const n2v<Tag, VA, VB>& y) {
  return veq(x.a, y.a) && veq(x.b, y.b);
}
template <class Tag, class VA, class VB, class VC>
inline bool veq(const n3v<Tag, VA, VB, VC>& x, const n3v<Tag, VA, VB, VC>& y) {
  return veq(x.a, y.a) && veq(x.b, y.b) && veq(x.c, y.c);
}
template <class X> inline bool veq(const X&, const X&) { return false; }  // other views: never shared

// pre-order index of the first occurrence of type X in root R (from K)
template <class R, class X, int K = 0>
constexpr int first_of() {
  if constexpr (K >= vsize<R>::value) return -1;
  else if constexpr (std::is_same<typename vat_t<R, K>::type, X>::value) return K;
  else return first_of<R, X, K + 1>();
}
// node types worth sharing: a division, square root or function at the top
// (an arithmetic node is evaluated at every occurrence, from shared operands)
template <class V> struct shareable { static constexpr bool value = false; };
template <class Tag, class VA> struct shareable<n1v<Tag, VA>> {
  static constexpr bool value = expr::tag_cost<Tag>::value > 1;
};
template <class Tag, class VA, class VB> struct shareable<n2v<Tag, VA, VB>> {
  static constexpr bool value = expr::tag_cost<Tag>::value > 1;
};
template <class Tag, class VA, class VB, class VC> struct shareable<n3v<Tag, VA, VB, VC>> {
  static constexpr bool value = expr::tag_cost<Tag>::value > 1;
};
// some shareable node type occurs more than once in R
template <class R, int K = 0>
constexpr bool has_repeats() {
  if constexpr (K >= vsize<R>::value) return false;
  else {
    typedef typename vat_t<R, K>::type X;
    if constexpr (shareable<X>::value && first_of<R, X>() < K) return true;
    else return has_repeats<R, K + 1>();
  }
}
// the host test before a launch: every occurrence of each shareable node type
// has the view of that type's first occurrence
template <class R, int K = 0>
bool shared_by_type(const R& root) {
  if constexpr (K >= vsize<R>::value) {
    return true;
  } else {
    typedef typename vat_t<R, K>::type X;
    if constexpr (shareable<X>::value && first_of<R, X>() < K) {
      if (!veq(vat<K>(root), vat<first_of<R, X>()>(root))) return false;
    }
    return shared_by_type<R, K + 1>(root);
  }
}

// per-index value slots of one element group (only node indices hold values)
template <class R, int G, int K, bool N = is_node<typename vat_t<R, K>::type>::value> struct slot_of {
  typedef char type;
};
template <class R, int G, int K> struct slot_of<R, G, K, true> {
  typedef pk<G, typename vview<typename vat_t<R, K>::type>::type> type;
};
template <class R, int G, class Seq> struct slots;
template <class R, int G, int... Ps> struct slots<R, G, std::integer_sequence<int, Ps...>> {
  typedef std::tuple<typename slot_of<R, G, Ps>::type...> type;
};
template <class R, int G, bool TS> struct cse_ctx {
  typedef R root_t;
  static constexpr bool by_type = TS;  // shared_by_type holds for the launch
  typename slots<R, G, std::make_integer_sequence<int, vsize<R>::value>>::type s;
};

template <int G, int A, int K, class C, class V, class R>
__device__ __forceinline__ void evm(C& c, const V& v, size_t i, R (&r)[G]);

template <int G, int A, int K, class C, class Tag, class VA, class R>
__device__ __forceinline__ void evn(C& c, const n1v<Tag, VA>& v, size_t i, R (&r)[G]) {
  typedef typename vview<VA>::type TA;
  TA a[G];
  evm<G, A, K + 1>(c, v.a, i, a);
  if constexpr (std::is_same<Tag, tagfn::if_zero_else_one_>::value) {
#pragma unroll
    for (int l = 0; l < G; ++l) r[l] = a[l] ? (R)0 : (R)1;
  } else {
    apply1<Tag>::template run<G, TA>(a, r);
  }
}

template <int G, int A, int K, class C, class Tag, class VA, class VB, class R>
__device__ __forceinline__ void evn(C& c, const n2v<Tag, VA, VB>& v, size_t i, R (&r)[G]) {
  typedef typename vview<VA>::type TA;
  typedef typename vview<VB>::type TB;
  TA a[G];
  TB b[G];
  evm<G, A, K + 1>(c, v.a, i, a);
  evm<G, A, K + 1 + vsize<VA>::value>(c, v.b, i, b);
  if constexpr (std::is_same<Tag, tagfn::pow_>::value && std::is_integral<TB>::value) {
    // generic/pow.hpp:110-132 and :154-168: |b| by repeated squaring, 1/p for b < 0
    static_assert(std::is_floating_point<R>::value, "pow is defined for floating-point bases only");
#pragma unroll
    for (int l = 0; l < G; ++l) {
      _Pragma("clang fp contract(off)")
      R base = (R)a[l], p = R(1);
      typedef typename std::make_unsigned<TB>::type UB;
      bool neg = false;
      if constexpr (std::is_signed<TB>::value) neg = b[l] < 0;
      UB e = neg ? (UB)(UB(0) - (UB)b[l]) : (UB)b[l];
      while (e) {
        if (e & 1u) p = p * base;
        e >>= 1;
        base = base * base;
      }
      r[l] = neg ? R(1) / p : p;
    }
  } else if constexpr (std::is_same<Tag, tagfn::pow_>::value) {
    static_assert(std::is_floating_point<R>::value, "pow is defined for floating-point tables only");
#pragma unroll
    for (int l = 0; l < G; ++l) {
      if constexpr (std::is_same<R, float>::value)
        r[l] = G == 1 ? dm::pow_f<P, true>((float)a[l], (float)b[l]) : dm::pow_f<P, false>((float)a[l], (float)b[l]);
      else
        r[l] = G == 1 ? dm::pow_d<P, true>((double)a[l], (double)b[l]) : dm::pow_d<P, false>((double)a[l], (double)b[l]);
    }
  } else if constexpr (std::is_same<Tag, tagfn::normcdf_join_>::value) {
    // the f64 scalar form (a pack's tail) is erfc's own, which ignores b
#pragma unroll
    for (int l = 0; l < G; ++l) {
      if constexpr (std::is_same<R, float>::value) r[l] = dm::normcdf_join_f((float)a[l], (float)b[l]);
      else if constexpr (std::is_same<R, double>::value)
        r[l] = G == 1 ? dm::normcdf_d<P, true>((double)a[l]) : dm::normcdf_join_d((double)a[l], (double)b[l]);
      else r[l] = no_integer_form<R>();
    }
  } else if constexpr (std::is_same<Tag, tagfn::logspace_add_>::value || std::is_same<Tag, tagfn::logspace_sub_>::value) {
    static_assert(std::is_floating_point<R>::value, "logspace_add / logspace_sub are defined for floating-point tables only");
#pragma unroll
    for (int l = 0; l < G; ++l)
      r[l] = std::is_same<Tag, tagfn::logspace_add_>::value ? dm::logspace_add_t<P>((R)a[l], (R)b[l])
                                                            : dm::logspace_sub_t<P>((R)a[l], (R)b[l]);
  } else if constexpr (std::is_same<Tag, tagfn::atan2d_>::value) {
    static_assert(std::is_floating_point<R>::value, "atan2d is defined for floating-point tables only");
#pragma unroll
    for (int l = 0; l < G; ++l)
      r[l] = G == 1 ? dm::atan2d_t<P, true>((R)a[l], (R)b[l]) : dm::atan2d_t<P, false>((R)a[l], (R)b[l]);
  } else if constexpr (std::is_same<Tag, tagfn::atan2_>::value) {
    static_assert(std::is_floating_point<R>::value, "atan2 is defined for floating-point tables only");
#pragma unroll
    for (int l = 0; l < G; ++l) {
      if constexpr (std::is_same<R, float>::value)
        r[l] = G == 1 ? dm::atan2_f<P, true>((float)a[l], (float)b[l]) : dm::atan2_f<P, false>((float)a[l], (float)b[l]);
      else
        r[l] = G == 1 ? dm::atan2_d<P, true>((double)a[l], (double)b[l]) : dm::atan2_d<P, false>((double)a[l], (double)b[l]);
    }
  } else {
#pragma unroll
    for (int l = 0; l < G; ++l) r[l] = apply2<Tag>::template one<R>(a[l], b[l]);
  }
}

template <int G, int A, int K, class C, class Tag, class VA, class VB, class VC, class R>
__device__ __forceinline__ void evn(C& c, const n3v<Tag, VA, VB, VC>& v, size_t i, R (&r)[G]) {
  _Pragma("clang fp contract(off)")
  typedef typename vview<VA>::type TA;
  typedef typename vview<VB>::type TB;
  typedef typename vview<VC>::type TC;
  TA a[G];
  TB b[G];
  TC cc[G];
  evm<G, A, K + 1>(c, v.a, i, a);
  evm<G, A, K + 1 + vsize<VA>::value>(c, v.b, i, b);
  evm<G, A, K + 1 + vsize<VA>::value + vsize<VB>::value>(c, v.c, i, cc);
#pragma unroll
  for (int l = 0; l < G; ++l) {
    if constexpr (std::is_same<Tag, tagfn::if_else_>::value) r[l] = a[l] ? (R)b[l] : (R)cc[l];
    else if constexpr (std::is_same<Tag, tagfn::fma_>::value) r[l] = dm::fma<P>((R)a[l], (R)b[l], (R)cc[l]);
    else r[l] = dm::fnms<P>((R)a[l], (R)b[l], (R)cc[l]);
  }
}

template <int G, int A, int K, class C, class V, class R>
__device__ __forceinline__ void evm(C& c, const V& v, size_t i, R (&r)[G]) {
  if constexpr (!is_node<V>::value) {
    ev<G, A>(v, i, r);
  } else {
    typedef typename C::root_t Root;
    constexpr bool sh = C::by_type && shareable<V>::value;
    constexpr int before = sh ? occurrences<Root, V, 0, K>() : 0;
    constexpr int after = sh ? occurrences<Root, V, K + 1, vsize<Root>::value>() : 0;
    if constexpr (before > 0) {
      const auto& q = std::get<first_of<Root, V>()>(c.s);
#pragma unroll
      for (int l = 0; l < G; ++l) r[l] = q.v[l];
    } else {
      evn<G, A, K>(c, v, i, r);
    }
    if constexpr (before == 0 && after > 0) {
      auto& q = std::get<K>(c.s);
#pragma unroll
      for (int l = 0; l < G; ++l) q.v[l] = r[l];
    }
  }
}

// a node view evaluated with its common subexpressions shared
// a node view evaluated on its own (nested evaluations: expand_to's operand)
template <int G, int A = 0, class Tag, class VA, class R>
__device__ __forceinline__ void ev(const n1v<Tag, VA>& v, size_t i, R (&r)[G]) {
  cse_ctx<n1v<Tag, VA>, G, false> c{};
  evm<G, A, 0>(c, v, i, r);
}
template <int G, int A = 0, class Tag, class VA, class VB, class R>
__device__ __forceinline__ void ev(const n2v<Tag, VA, VB>& v, size_t i, R (&r)[G]) {
  cse_ctx<n2v<Tag, VA, VB>, G, false> c{};
  evm<G, A, 0>(c, v, i, r);
}
template <int G, int A = 0, class Tag, class VA, class VB, class VC, class R>
__device__ __forceinline__ void ev(const n3v<Tag, VA, VB, VC>& v, size_t i, R (&r)[G]) {
  cse_ctx<n3v<Tag, VA, VB, VC>, G, false> c{};
  evm<G, A, 0>(c, v, i, r);
}
// a statement's root, its repeats shared when TS (shared_by_type held)
template <int G, int A, bool TS, class V, class R>
__device__ __forceinline__ void ev_root(const V& v, size_t i, R (&r)[G]) {
  if constexpr (is_node<V>::value) {
    cse_ctx<V, G, TS> c{};
    evm<G, A, 0>(c, v, i, r);
  } else {
    ev<G, A>(v, i, r);
  }
}
// the two right-hand sides of a tie, sharing their common subexpressions
template <int G, int A, bool TS, class V1, class V2, class T>
__device__ __forceinline__ void ev2(const vpair<V1, V2>& root, size_t i, T (&r1)[G], T (&r2)[G]) {
  cse_ctx<vpair<V1, V2>, G, TS> c{};
  evm<G, A, 1>(c, root.a, i, r1);
  evm<G, A, 1 + vsize<V1>::value>(c, root.b, i, r2);
}

// the pack's W results stored with 16-byte nontemporal stores (A) or one by one
template <int A, class T, int W>
__device__ __forceinline__ void store_pack(T* __restrict__ out, size_t i, const T (&r)[W]) {
  if constexpr (A && std::is_arithmetic<T>::value && !std::is_same<T, bool>::value &&
                (W * sizeof(T)) % 16 == 0) {
    constexpr int V = 16 / (int)sizeof(T);
    typedef T vt __attribute__((ext_vector_type(V)));
#pragma unroll
    for (int k = 0; k < W / V; ++k) {
      vt w;
#pragma unroll
      for (int l = 0; l < V; ++l) w[l] = r[k * V + l];
      __builtin_nontemporal_store(w, reinterpret_cast<vt*>(out + i) + k);
    }
  } else {
#pragma unroll
    for (int l = 0; l < W; ++l) out[i + l] = r[l];
  }
}

// Elements a thread evaluates at once for an expression in which no function
// decides per pack (expr::pack_free): a quarter (f32) or half (f64) pack,
// which keeps the registers of a long expression at half the occupancy cost
// and gives the same bits, every lane's result being its own.  Otherwise the
// whole pack of W.
template <class T> struct sub_width { static constexpr int value = pack_width<T>::value; };
template <> struct sub_width<float> { static constexpr int value = 4; };
template <> struct sub_width<double> { static constexpr int value = 2; };

// one output: packs of W from index 0, then the scalar tail; a grid-stride
// loop over work items (SG elements of a pack for t < nsub, then the tail
// elements one by one); A: every terminal and the output are 16-byte aligned
// U work items per thread and iteration are evaluated before any of them is
// stored (an in-place statement reads only its own indices, so the loads of
// later items may go ahead of the earlier items' stores)
template <class T, class V, int A, int SG, int U, bool TS>
__global__ __launch_bounds__(256) void transform_k(T* __restrict__ out, V v, size_t n) {
  constexpr int W = pack_width<T>::value;
  static_assert(W % SG == 0, "a sub-pack divides the pack");
  const size_t npacks = n / W;
  const size_t nsub = npacks * (W / SG);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  for (size_t t0 = g; t0 < nsub; t0 += U * stride) {
    T r[U][SG];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (t0 + u * stride < nsub) ev_root<SG, A, TS>(v, (t0 + u * stride) * SG, r[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (t0 + u * stride < nsub) store_pack<A>(out, (t0 + u * stride) * SG, r[u]);
  }
  for (size_t t = g; t < n - npacks * W; t += stride) {  // the scalar tail
    const size_t i = npacks * W + t;
    T r[1];
    ev_root<1, 0, TS>(v, i, r);
    out[i] = r[0];
  }
}

// tie: every right-hand side of an index is evaluated before any output of
// that index is written
template <class T, class V1, class V2, int A, int SG, bool TS>
__global__ __launch_bounds__(256) void transform2_k(T* __restrict__ o1, T* __restrict__ o2, vpair<V1, V2> v,
                                                    size_t n) {
  constexpr int W = pack_width<T>::value;
  static_assert(W % SG == 0, "a sub-pack divides the pack");
  const size_t npacks = n / W;
  const size_t nsub = npacks * (W / SG);
  const size_t items = nsub + (n - npacks * W);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  for (size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x; t < items; t += stride) {
    if (t < nsub) {
      T r1[SG], r2[SG];
      ev2<SG, A, TS>(v, t * SG, r1, r2);
      store_pack<A>(o1, t * SG, r1);
      store_pack<A>(o2, t * SG, r2);
    } else {
      const size_t i = npacks * W + (t - nsub);
      T r1[1], r2[1];
      ev2<1, 0, TS>(v, i, r1, r2);
      o1[i] = r1[0];
      o2[i] = r2[0];
    }
  }
}

// fold: compensated FP32/FP64 partials per lane, error-free wave merge,
// per-block FP64 partials combined in block order by a second launch
template <class T>
__device__ __forceinline__ void acc_add(T& s, T& c, T v) {
  _Pragma("clang fp contract(off)")
  const T t = s + v;
  const T bp = t - s;
  const T e = (s - (t - bp)) + (v - bp);
  s = t;
  c = c + e;
}

// fixed-order sum of one double per thread of a 256-thread block (valid in thread 0)
__device__ __forceinline__ double block_sum256(double d) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) d += __shfl_xor(d, off, 64);
  __shared__ double w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = d;
  __syncthreads();
  return ((w[0] + w[1]) + w[2]) + w[3];
}

template <class T, class V, int A, int SG, int U, bool TS>
__global__ __launch_bounds__(256) void fold_k(V v, size_t n, double* __restrict__ partial) {
  constexpr int W = pack_width<T>::value;
  static_assert(W % SG == 0, "a sub-pack divides the pack");
  const size_t npacks = n / W;
  const size_t nsub = npacks * (W / SG);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  T s = 0, c = 0;
  for (size_t t0 = g; t0 < nsub; t0 += U * stride) {
    T r[U][SG];
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (t0 + u * stride < nsub) ev_root<SG, A, TS>(v, (t0 + u * stride) * SG, r[u]);
#pragma unroll
    for (int u = 0; u < U; ++u)
      if (t0 + u * stride < nsub)
#pragma unroll
        for (int l = 0; l < SG; ++l) acc_add(s, c, r[u][l]);
  }
  for (size_t t = g; t < n - npacks * W; t += stride) {  // the scalar tail
    T r[1];
    ev_root<1, 0, TS>(v, npacks * W + t, r);
    acc_add(s, c, r[0]);
  }
  const double b = block_sum256((double)s + (double)c);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}

static __global__ __launch_bounds__(256) void fold_final_k(const double* __restrict__ partial, int nb,
                                                          double* __restrict__ out) {
  double d = 0.0;
  for (int k = threadIdx.x; k < nb; k += 256) d += partial[k];
  const double t = block_sum256(d);
  if (threadIdx.x == 0) *out = t;
}

// integer tables: NT2 sums in the element type, wrapping; the sum modulo 2^64
// (in any order) reduced to the type's width is that sum
__device__ __forceinline__ uint64_t block_sum256_u64(uint64_t d) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) d += __shfl_xor(d, off, 64);
  __shared__ uint64_t w[4];
  if ((threadIdx.x & 63) == 0) w[threadIdx.x >> 6] = d;
  __syncthreads();
  return ((w[0] + w[1]) + w[2]) + w[3];
}
template <class T, class V, int A, int SG, bool TS>
__global__ __launch_bounds__(256) void fold_int_k(V v, size_t n, uint64_t* __restrict__ partial) {
  constexpr int W = pack_width<T>::value;
  const size_t npacks = n / W;
  const size_t nsub = npacks * (W / SG);
  const size_t stride = (size_t)gridDim.x * blockDim.x;
  const size_t g = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
  uint64_t s = 0;
  for (size_t t = g; t < nsub; t += stride) {
    T r[SG];
    ev_root<SG, A, TS>(v, t * SG, r);
#pragma unroll
    for (int l = 0; l < SG; ++l) s += (uint64_t)(long long)r[l];
  }
  for (size_t t = g; t < n - npacks * W; t += stride) {
    T r[1];
    ev_root<1, 0, TS>(v, npacks * W + t, r);
    s += (uint64_t)(long long)r[0];
  }
  const uint64_t b = block_sum256_u64(s);
  if (threadIdx.x == 0) partial[blockIdx.x] = b;
}
static __global__ __launch_bounds__(256) void fold_final_int_k(const uint64_t* __restrict__ partial, int nb,
                                                              uint64_t* __restrict__ out) {
  uint64_t d = 0;
  for (int k = threadIdx.x; k < nb; k += 256) d += partial[k];
  const uint64_t t = block_sum256_u64(d);
  if (threadIdx.x == 0) *out = t;
}

// workgroups of a grid-stride launch over `items` work items (at most 32 per
// compute unit of a 256-CU MI355X)
inline unsigned grid(size_t items) {
  size_t g = (items + 255) / 256;
  if (g > 8192) g = 8192;
  return (unsigned)(g ? g : 1);
}

// f(integral_constant<int, A>, bool_constant<TS>) with A = AL when al holds,
// else 0, and TS = shared_by_type(root); the TS kernel is instantiated only
// for statements that repeat a shareable node type
template <int AL, class R, class F>
void dispatch(bool al, const R& root, F&& f) {
  typedef std::integral_constant<int, AL> a1;
  typedef std::integral_constant<int, 0> a0;
  if constexpr (has_repeats<R>()) {
    const bool ts = shared_by_type(root);
    if (al && ts) f(a1{}, std::true_type{});
    else if (al) f(a1{}, std::false_type{});
    else if (ts) f(a0{}, std::true_type{});
    else f(a0{}, std::false_type{});
  } else {
    if (al) f(a1{}, std::false_type{});
    else f(a0{}, std::false_type{});
  }
}

inline unsigned grid_full(size_t items) {
  size_t g = (items + 255) / 256;
  if (g > 0x7fffffffu) g = 0x7fffffffu;
  return (unsigned)(g ? g : 1);
}

// every terminal of a view 16-byte aligned (the output is checked by the caller)
template <class T> inline bool aligned16(const expr::term_v<T>& v) { return ((uintptr_t)v.p & 15u) == 0; }
template <class T> inline bool aligned16(const expr::scal_v<T>&) { return true; }
template <class Tag, class VA> inline bool aligned16(const n1v<Tag, VA>& v) { return aligned16(v.a); }
template <class Tag, class VA, class VB> inline bool aligned16(const n2v<Tag, VA, VB>& v) {
  return aligned16(v.a) && aligned16(v.b);
}
template <class Tag, class VA, class VB, class VC> inline bool aligned16(const n3v<Tag, VA, VB, VC>& v) {
  return aligned16(v.a) && aligned16(v.b) && aligned16(v.c);
}
template <class V> inline bool aligned16(const V&) { return false; }  // other views: element loads

// a pinned host double per thread for fold results (a copy to pageable memory
// goes through a staging buffer), kept until the process ends
inline double* fold_result_host() {
  thread_local double* p = nullptr;
  if (!p) check(hipHostMalloc((void**)&p, sizeof(double), 0), "fold result buffer");
  return p;
}

// the fold's partials: one device buffer per device, grown on demand, kept
// until the process ends
inline double* fold_scratch(size_t nd) {
  static std::mutex mu;
  static std::vector<std::pair<double*, size_t>> per_dev;
  int dev = 0;
  check(hipGetDevice(&dev), "hipGetDevice");
  std::lock_guard<std::mutex> lk(mu);
  if ((int)per_dev.size() <= dev) per_dev.resize(dev + 1, {nullptr, 0});
  auto& e = per_dev[dev];
  if (e.second < nd) {
    if (e.first) check(hipFree(e.first), "fold scratch free");
    e.first = nullptr;
    check(hipMalloc(&e.first, sizeof(double) * nd), "fold scratch alloc");
    e.second = nd;
  }
  return e.first;
}

}}  // namespace hip::fe

namespace expr {
// false when a function of the expression decides per pack (NT2's
// trigonometric range reduction) or has a separate scalar form for a pack's
// tail (atan, asinh, acosh; the f64 erfc, normcdf, asin, acos and erf)
template <class X> struct pack_free { static constexpr bool value = true; };
template <class Tag, class A> struct pack_free<node1<Tag, A>> {
  static constexpr bool value = pack_free<A>::value && !std::is_same<Tag, tagfn::cos_>::value &&
                                !std::is_same<Tag, tagfn::sin_>::value && !std::is_same<Tag, tagfn::tan_>::value &&
                                !std::is_same<Tag, tagfn::cot_>::value && !std::is_same<Tag, tagfn::sec_>::value &&
                                !std::is_same<Tag, tagfn::csc_>::value && !std::is_base_of<tagfn::same1p, Tag>::value &&
                                !std::is_same<Tag, tagfn::atan_>::value && !std::is_same<Tag, tagfn::asinh_>::value &&
                                !std::is_same<Tag, tagfn::acosh_>::value &&
                                !((std::is_same<Tag, tagfn::erfc_>::value ||
                                   std::is_same<Tag, tagfn::asin_>::value || std::is_same<Tag, tagfn::acos_>::value ||
                                   std::is_same<Tag, tagfn::erf_>::value) &&
                                  std::is_same<typename A::value_type, double>::value);
};
template <class Tag, class A, class B> struct pack_free<node2<Tag, A, B>> {
  static constexpr bool value = pack_free<A>::value && pack_free<B>::value && !std::is_same<Tag, tagfn::atan2_>::value &&
                                !std::is_same<Tag, tagfn::atan2d_>::value &&
                                !std::is_same<Tag, tagfn::pow_>::value &&
                                !(std::is_same<Tag, tagfn::normcdf_join_>::value &&
                                  std::is_same<typename A::value_type, double>::value);
};
template <class Tag, class A, class B, class C> struct pack_free<node3<Tag, A, B, C>> {
  static constexpr bool value = pack_free<A>::value && pack_free<B>::value && pack_free<C>::value;
};

// A rough VALU cost of one element (the sum of its nodes' tag_cost), which
// picks the lowering's work shape:
// cheap expressions are memory bound and get several packs in flight per
// thread; costly ones get fewer elements per thread.  Views other than nodes,
// terminals and scalars count as costly.
template <class X> struct cost { static constexpr int value = 20; };
template <class T> struct cost<terminal<T>> { static constexpr int value = 0; };
template <class T> struct cost<scalar<T>> { static constexpr int value = 0; };
template <class Tag, class A> struct cost<node1<Tag, A>> { static constexpr int value = cost<A>::value + tag_cost<Tag>::value; };
template <class Tag, class A, class B> struct cost<node2<Tag, A, B>> {
  static constexpr int value = cost<A>::value + cost<B>::value + tag_cost<Tag>::value;
};
template <class Tag, class A, class B, class C> struct cost<node3<Tag, A, B, C>> {
  static constexpr int value = cost<A>::value + cost<B>::value + cost<C>::value + tag_cost<Tag>::value;
};
}  // namespace expr

namespace hip { namespace fe {
// elements per work item (SG) and work items in flight per thread (U) for an
// expression type
template <class T, class EX> struct shape {
  static constexpr bool cheap = expr::cost<EX>::value < 8;
  static constexpr int SG = expr::pack_free<EX>::value ? sub_width<T>::value : pack_width<T>::value;
  static constexpr int U = 1;
  // a cheap statement streams: one work item per thread over a full grid (the
  // shape of the library's own saxpy); a costly one loops over a grid of at
  // most 32 workgroups per compute unit
  static constexpr bool full_grid = cheap;
};
}}  // namespace hip::fe

// ---- lowering entry points -------------------------------------------------------
// What an NT2 site override calls: transform_(out, in, range) over the whole
// table (range = {0, numel}) and inner_fold_ of a whole expression.  Each
// expression type instantiates its own fused kernel; the terminals' device
// copies must be current (expr.prepare()).
namespace hip {
template <class T, class EX>
void transform(T* out, const EX& x, size_t n, hipStream_t s = nullptr) {
  if (n == 0) return;
  typedef fe::shape<T, EX> sh;
  auto v = expr::dview<EX>::get(x);
  const size_t items = (n / sh::SG + sh::U - 1) / sh::U;
  const dim3 g(sh::full_grid ? fe::grid_full(items) : fe::grid(items)), b(256);
  constexpr int AL = sh::cheap ? 2 : 1;
  typedef decltype(v) V;
  fe::dispatch<AL>(fe::aligned16(v) && ((uintptr_t)out & 15u) == 0, v, [&](auto a, auto ts) {
    hipLaunchKernelGGL((fe::transform_k<T, V, decltype(a)::value, sh::SG, sh::U, decltype(ts)::value>), g, b, 0, s,
                       out, v, n);
  });
  fe::check(hipGetLastError(), "transform launch");
}

// sum of an integer expression's n elements in T (wrapping as NT2's sum in T
// does), returned on the host
template <class T, class EX>
T fold_int(const EX& x, size_t n, hipStream_t s = nullptr) {
  typedef fe::shape<T, EX> sh;
  const size_t items = n / sh::SG;
  const unsigned nb = (unsigned)std::min<size_t>(1024, (items + 255) / 256 ? (items + 255) / 256 : 1);
  uint64_t* buf = reinterpret_cast<uint64_t*>(fe::fold_scratch(nb + 1));  // 8-byte slots either way
  auto v = expr::dview<EX>::get(x);
  typedef decltype(v) V;
  fe::dispatch<0>(false, v, [&](auto a, auto ts) {
    hipLaunchKernelGGL((fe::fold_int_k<T, V, decltype(a)::value, sh::SG, decltype(ts)::value>), dim3(nb), dim3(256),
                       0, s, v, n, buf);
  });
  fe::check(hipGetLastError(), "fold launch");
  hipLaunchKernelGGL(fe::fold_final_int_k, dim3(1), dim3(256), 0, s, buf, (int)nb, buf + nb);
  double* hr = fe::fold_result_host();
  fe::check(hipMemcpyAsync(hr, buf + nb, sizeof(uint64_t), hipMemcpyDeviceToHost, s), "fold result");
  fe::check(hipStreamSynchronize(s), "fold sync");
  uint64_t u;
  std::memcpy(&u, hr, sizeof u);
  return (T)(typename std::make_unsigned<T>::type)u;
}

// FP64 value of the compensated fold of x's n elements (synchronises s: the
// value is returned on the host, as NT2's globalsum returns it)
template <class T, class EX>
double fold(const EX& x, size_t n, hipStream_t s = nullptr) {
  typedef fe::shape<T, EX> sh;
  const size_t items = (n / sh::SG + sh::U - 1) / sh::U;
  const unsigned nb = (unsigned)std::min<size_t>(1024, (items + 255) / 256 ? (items + 255) / 256 : 1);
  double* buf = fe::fold_scratch(nb + 1);
  auto v = expr::dview<EX>::get(x);
  constexpr int AL = sh::cheap ? 2 : 1;
  typedef decltype(v) V;
  fe::dispatch<AL>(fe::aligned16(v), v, [&](auto a, auto ts) {
    hipLaunchKernelGGL((fe::fold_k<T, V, decltype(a)::value, sh::SG, sh::U, decltype(ts)::value>), dim3(nb),
                       dim3(256), 0, s, v, n, buf);
  });
  fe::check(hipGetLastError(), "fold launch");
  hipLaunchKernelGGL(fe::fold_final_k, dim3(1), dim3(256), 0, s, buf, (int)nb, buf + nb);
  double* hr = fe::fold_result_host();
  fe::check(hipMemcpyAsync(hr, buf + nb, sizeof(double), hipMemcpyDeviceToHost, s), "fold result");
  fe::check(hipStreamSynchronize(s), "fold sync");
  return *hr;
}
}  // namespace hip

template <class T>
template <class E>
void table<T>::assign(const E& e) {
  typedef typename expr::wrap<E>::type EX;
  const EX& x = expr::wrap<E>::get(e);
  const ext_t ex = x.ext();
  if (!ex.has) throw std::invalid_argument("nt2: assignment of a scalar expression to a table");
  const of_size_ s = ex.s;
  x.prepare();  // terminals' device copies are current (this table included, for in-place)
  resize(s);
  T* out = device_data_for_write();
  hip::transform(out, x, numel());
}

// globalsum(expr): fold of the whole expression, returned on the host
template <class E, NT2H_ENABLE1(E)>
typename expr::vt<typename expr::wrap<E>::type>::type globalsum(const E& e) {
  typedef typename expr::wrap<E>::type EX;
  typedef typename expr::vt<EX>::type T;
  const EX& x = expr::wrap<E>::get(e);
  const ext_t ex = x.ext();
  const size_t n = ex.has ? ex.s.numel() : 1;
  static_assert(!std::is_same<T, bool>::value, "globalsum of a logical expression: convert it to a number first");
  x.prepare();
  if constexpr (std::is_integral<T>::value) return hip::fold_int<T>(x, n);
  else return (T)hip::fold<T>(x, n);
}
// ---- reductions and scans along a dimension ------------------------------------------
namespace hip { namespace fe {
// an expression as a source of the dimension-wise kernels
template <class T, class V, bool PF> struct expr_src {
  static constexpr bool pack_free = PF;
  V v;
  template <int G> __device__ __forceinline__ void load(size_t i, T (&r)[G]) const { ev<G>(v, i, r); }
};
}}  // namespace hip::fe
namespace hip {
// sum (scan false) or cumsum (scan true) of x along dim into a new table;
// Op other than dw::op_sum: that scan (cumprod, cummin, cummax) in NT2's
// order under every policy
template <class T, class Op = dw::op_sum, class EX>
table<T> dimwise(const EX& x, int dim, bool scan, hipStream_t s = nullptr) {
  static_assert(std::is_same<T, float>::value || std::is_same<T, double>::value,
                "sum/cumsum/cumprod/cummin/cummax along a dimension: float or double tables");
  if (dim < 1) throw std::invalid_argument("nt2: dimension must be >= 1");
  const ext_t ex = x.ext();
  if (!ex.has) throw std::invalid_argument("nt2: reduction of a scalar expression");
  x.prepare();
  of_size_ os = ex.s;
  if (!scan && dim <= 4) os.d[dim - 1] = 1;
  table<T> r(os);
  if (r.numel() == 0) return r;
  const size_t e4[4] = {ex.s.d[0], ex.s.d[1], ex.s.d[2], ex.s.d[3]};
  const dw::lines L = dw::lines_of(e4, dim);
  constexpr bool summing = std::is_same<Op, dw::op_sum>::value;
  const bool fast = NT2_HIP_POLICY::fused && summing;
  double* scratch = nullptr;
  if (const size_t need = dw::scratch_doubles(L, fast, scan)) fe::check(hipMalloc(&scratch, need * sizeof(double)), "scratch");
  auto v = expr::dview<EX>::get(x);
  const fe::expr_src<T, decltype(v), expr::pack_free<EX>::value> src{v};
  T* out = r.device_data_for_write();
  if (!summing) dw::launch_scan_ordered<T, Op>(src, L, out, s);
  else if (scan) dw::launch_scan<T>(src, L, fast, scratch, out, s);
  else dw::launch_sum<T>(src, L, fast, scratch, out, s);
  fe::check(hipGetLastError(), "dimwise launch");
  if (scratch) {
    fe::check(hipStreamSynchronize(s), "dimwise sync");
    fe::check(hipFree(scratch), "scratch free");
  }
  return r;
}
}  // namespace hip

// sum(a, dim): NT2's reduction along dim (a table of a's extent with dim of size 1)
template <class E, NT2H_ENABLE1(E)>
auto sum(const E& e, int dim) {
  typedef typename expr::wrap<E>::type EX;
  typedef typename expr::vt<EX>::type T;
  return hip::dimwise<T>(expr::wrap<E>::get(e), dim, false);
}
// sum(a): along the first non-singleton dimension; a one-element result
// converts to a scalar (T s = sum(x) for a vector x)
template <class E, NT2H_ENABLE1(E)>
auto sum(const E& e) {
  typedef typename expr::wrap<E>::type EX;
  const EX& x = expr::wrap<E>::get(e);
  const ext_t ex = x.ext();
  return sum(e, ex.has ? firstnonsingleton(ex.s) : 1);
}
// cumsum(a, dim) and cumsum(a): running sums along a dimension
// (modules/core/swar/include/nt2/swar/functions/container/cumsum.hpp)
template <class E, NT2H_ENABLE1(E)>
auto cumsum(const E& e, int dim) {
  typedef typename expr::wrap<E>::type EX;
  typedef typename expr::vt<EX>::type T;
  return hip::dimwise<T>(expr::wrap<E>::get(e), dim, true);
}
template <class E, NT2H_ENABLE1(E)>
auto cumsum(const E& e) {
  typedef typename expr::wrap<E>::type EX;
  const EX& x = expr::wrap<E>::get(e);
  const ext_t ex = x.ext();
  return cumsum(e, ex.has ? firstnonsingleton(ex.s) : 1);
}
// cumprod, cummin, cummax (modules/core/swar/include/nt2/swar/functions/
// container/cumprod.hpp, cummin.hpp, cummax.hpp): running product / minimum /
// maximum along a dimension, the first non-singleton one by default
#define NT2H_CUMOP(name, Op)                                                      \
  template <class E, NT2H_ENABLE1(E)>                                            \
  auto name(const E& e, int dim) {                                               \
    typedef typename expr::wrap<E>::type EX;                                     \
    typedef typename expr::vt<EX>::type T;                                       \
    return hip::dimwise<T, hip::dw::Op>(expr::wrap<E>::get(e), dim, true);      \
  }                                                                              \
  template <class E, NT2H_ENABLE1(E)>                                            \
  auto name(const E& e) {                                                        \
    typedef typename expr::wrap<E>::type EX;                                     \
    const EX& x = expr::wrap<E>::get(e);                                         \
    const ext_t ex = x.ext();                                                    \
    return name(e, ex.has ? firstnonsingleton(ex.s) : 1);                        \
  }
NT2H_CUMOP(cumprod, op_prod)
NT2H_CUMOP(cummin, op_min)
NT2H_CUMOP(cummax, op_max)
#undef NT2H_CUMOP

// ---- reshape and colon ---------------------------------------------------------------
// reshape(a, sizes): the elements of a (evaluated if an expression) with a new
// extent of the same number of elements
template <class E, NT2H_ENABLE1(E)>
auto reshape(const E& e, of_size_ s) {
  typedef typename expr::wrap<E>::type EX;
  typedef typename expr::vt<EX>::type T;
  table<T> r;
  r = expr::wrap<E>::get(e);
  r.reshape_to(s);
  return r;
}
template <class E, NT2H_ENABLE1(E)> auto reshape(const E& e, size_t a, size_t b) { return reshape(e, of_size(a, b)); }
template <class E, NT2H_ENABLE1(E)> auto reshape(const E& e, size_t a, size_t b, size_t c) {
  return reshape(e, of_size(a, b, c));
}
template <class E, NT2H_ENABLE1(E)> auto reshape(const E& e, size_t a, size_t b, size_t c, size_t d) {
  return reshape(e, of_size(a, b, c, d));
}
// _(a, b): the row vector a, a+1, ..., b (NT2's colon)
struct colon_ {
  template <class T> table<T> operator()(T a, T b) const {
    const size_t n = b >= a ? (size_t)(b - a) + 1 : 0;
    table<T> r(of_size(1, n));
    for (size_t i = 0; i < n; ++i) r(i + 1) = (T)(a + (T)i);
    return r;
  }
};
static const colon_ _ = {};
// dot(a, b) = sum(a .* conj(b)): modules/core/reduction/include/nt2/reduction/functions/expr/dot.hpp:22-29
template <class A, class B, NT2H_ENABLE2(A, B)>
auto dot(const A& a, const B& b) { return globalsum(a * b); }

// ---- tie ---------------------------------------------------------------------------
template <class E1, class E2> struct tie2_rhs { E1 e1; E2 e2; };
template <class T> struct tie2_lhs {
  table<T>& a;
  table<T>& b;
  template <class E1, class E2> tie2_lhs& operator=(const tie2_rhs<E1, E2>& r) {
    const ext_t x = merge_ext(r.e1.ext(), r.e2.ext());
    if (!x.has) throw std::invalid_argument("nt2: tie of scalar expressions");
    r.e1.prepare();
    r.e2.prepare();
    a.resize(x.s);
    b.resize(x.s);
    T* o1 = a.device_data_for_write();
    T* o2 = b.device_data_for_write();
    const size_t n = x.s.numel();
    constexpr int W = hip::fe::pack_width<T>::value;
    auto v1 = expr::dview<E1>::get(r.e1);
    auto v2 = expr::dview<E2>::get(r.e2);
    constexpr int SG = expr::pack_free<E1>::value && expr::pack_free<E2>::value ? hip::fe::sub_width<T>::value : W;
    const dim3 g(hip::fe::grid(n / SG + n % W)), blk(256);
    typedef decltype(v1) V1;
    typedef decltype(v2) V2;
    const hip::fe::vpair<V1, V2> root{v1, v2};
    hip::fe::dispatch<1>(hip::fe::aligned16(v1) && hip::fe::aligned16(v2) && (((uintptr_t)o1 | (uintptr_t)o2) & 15u) == 0,
                         root, [&](auto al, auto ts) {
                           hipLaunchKernelGGL((hip::fe::transform2_k<T, V1, V2, decltype(al)::value, SG,
                                                                     decltype(ts)::value>),
                                              g, blk, 0, 0, o1, o2, root, n);
                         });
    hip::fe::check(hipGetLastError(), "tie launch");
    return *this;
  }
};
template <class T> tie2_lhs<T> tie(table<T>& a, table<T>& b) { return tie2_lhs<T>{a, b}; }
template <class E1, class E2>
tie2_rhs<typename expr::wrap<E1>::type, typename expr::wrap<E2>::type> tie(const E1& e1, const E2& e2) {
  return tie2_rhs<typename expr::wrap<E1>::type, typename expr::wrap<E2>::type>{
      expr::wrap<E1>::get(e1), expr::wrap<E2>::get(e2)};
}

// ---- packs for arrayfun -------------------------------------------------------------
// arrayfun(f, a, b) calls f once per pack of NT2's width (8 floats) and once
// per element of the tail with nt2::pack<T, 1>, as NT2's transform calls a
// functor on native packs and then on scalars.  A functor decides per pack
// with any()/all(); its operator() must be callable on the device
// (__host__ __device__).  Pack arithmetic rounds every operation (no
// contraction), as on AVX without FMA.
template <class T, int N> struct pack {
  T v[N];
  __host__ __device__ pack() {}
  __host__ __device__ pack(T x) {
    for (int l = 0; l < N; ++l) v[l] = x;
  }
  __host__ __device__ T& operator[](int l) { return v[l]; }
  __host__ __device__ const T& operator[](int l) const { return v[l]; }
};
#define NT2H_PACK_OP(OP, R)                                                                      \
  template <class T, int N>                                                                      \
  __host__ __device__ pack<R, N> operator OP(const pack<T, N>& a, const pack<T, N>& b) {         \
    _Pragma("clang fp contract(off)") pack<R, N> r;                                              \
    for (int l = 0; l < N; ++l) r.v[l] = a.v[l] OP b.v[l];                                       \
    return r;                                                                                    \
  }                                                                                              \
  template <class T, int N, class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type> \
  __host__ __device__ pack<R, N> operator OP(const pack<T, N>& a, S b) {                         \
    return a OP pack<T, N>((T)b);                                                                \
  }                                                                                              \
  template <class T, int N, class S, class = typename std::enable_if<std::is_arithmetic<S>::value>::type> \
  __host__ __device__ pack<R, N> operator OP(S a, const pack<T, N>& b) {                         \
    return pack<T, N>((T)a) OP b;                                                                \
  }
NT2H_PACK_OP(+, T) NT2H_PACK_OP(-, T) NT2H_PACK_OP(*, T) NT2H_PACK_OP(/, T)
NT2H_PACK_OP(<, bool) NT2H_PACK_OP(>, bool) NT2H_PACK_OP(<=, bool) NT2H_PACK_OP(>=, bool)
NT2H_PACK_OP(==, bool) NT2H_PACK_OP(!=, bool)
#undef NT2H_PACK_OP
template <class T, int N> __host__ __device__ pack<T, N> operator-(const pack<T, N>& a) {
  pack<T, N> r;
  for (int l = 0; l < N; ++l) r.v[l] = -a.v[l];
  return r;
}
// any / all lanes of a mask (modules/boost/simd/base/include/boost/simd/reduction/functions/any.hpp)
template <int N> __host__ __device__ bool any(const pack<bool, N>& m) {
  bool r = false;
  for (int l = 0; l < N; ++l) r = r || m.v[l];
  return r;
}
template <int N> __host__ __device__ bool all(const pack<bool, N>& m) {
  bool r = true;
  for (int l = 0; l < N; ++l) r = r && m.v[l];
  return r;
}
__host__ __device__ inline bool any(bool m) { return m; }
__host__ __device__ inline bool all(bool m) { return m; }
// seladd(c, a, b) = c ? a + b : a, lane by lane
// (modules/boost/simd/base/include/boost/simd/boolean/functions/seladd.hpp)
template <class T, int N>
__host__ __device__ pack<T, N> seladd(const pack<bool, N>& c, const pack<T, N>& a, const pack<T, N>& b) {
  pack<T, N> r;
  for (int l = 0; l < N; ++l) r.v[l] = c.v[l] ? (T)(a.v[l] + b.v[l]) : a.v[l];
  return r;
}
template <class T> __host__ __device__ T seladd(bool c, T a, T b) { return c ? (T)(a + b) : a; }
template <class T> __host__ __device__ T Zero() { return T(0); }
template <class T> __host__ __device__ T One() { return T(1); }
namespace meta {
template <class T> struct scalar_of { typedef T type; };
template <class T, int N> struct scalar_of<pack<T, N>> { typedef T type; };
template <class T> struct as_integer {
  typedef typename std::conditional<sizeof(T) == 8, int64_t,
          typename std::conditional<sizeof(T) == 4, int32_t,
          typename std::conditional<sizeof(T) == 2, int16_t, int8_t>::type>::type>::type type;
};
template <class T, int N> struct as_integer<pack<T, N>> { typedef pack<typename as_integer<T>::type, N> type; };
template <class T> struct as_logical { typedef bool type; };
template <class T, int N> struct as_logical<pack<T, N>> { typedef pack<bool, N> type; };
}  // namespace meta

// ---- linspace, expand_to, colvect, arrayfun ---------------------------------------------
namespace hip { namespace fe {
template <class T> struct lin_v { size_t n; T lo, step, hi; };
template <class VA> struct bcast_v { VA a; size_t to[4]; size_t from[4]; };
template <class F, class VA, class VB> struct afun2_v { F f; VA a; VB b; };
template <class T> struct vview<lin_v<T>> { typedef T type; };
template <class VA> struct vview<bcast_v<VA>> { typedef typename vview<VA>::type type; };
template <class F, class VA, class VB> struct vview<afun2_v<F, VA, VB>> {
  typedef typename meta::scalar_of<decltype(std::declval<const F&>()(
      std::declval<const nt2::pack<typename vview<VA>::type, 8>&>(),
      std::declval<const nt2::pack<typename vview<VB>::type, 8>&>()))>::type type;
};

// linspace: hi at the last index, p*step + lo elsewhere
// (modules/core/generative/include/nt2/core/functions/details/linspace.hpp:33-45)
template <int G, int A = 0, class T>
__device__ __forceinline__ void ev(const lin_v<T>& v, size_t i, T (&r)[G]) {
#pragma unroll
  for (int l = 0; l < G; ++l) {
    const size_t p = i + l;
    r[l] = (p == v.n - 1) ? v.hi : dm::fma<P>((T)p, v.step, v.lo);
  }
}
// expand_to: each output index reads the source at its coordinates, 0 along
// the source's singleton dimensions (one element at a time)
template <int G, int A = 0, class VA, class R>
__device__ __forceinline__ void ev(const bcast_v<VA>& v, size_t i, R (&r)[G]) {
#pragma unroll
  for (int l = 0; l < G; ++l) {
    size_t k = i + l, src = 0, stride = 1;
#pragma unroll
    for (int d = 0; d < 4; ++d) {
      const size_t c = k % v.to[d];
      k /= v.to[d];
      src += (v.from[d] == 1 ? 0 : c) * stride;
      stride *= v.from[d];
    }
    R t[1];
    ev<1>(v.a, src, t);
    r[l] = t[0];
  }
}
// arrayfun: the functor sees the whole pack
template <int G, int A = 0, class F, class VA, class VB, class R>
__device__ __forceinline__ void ev(const afun2_v<F, VA, VB>& v, size_t i, R (&r)[G]) {
  nt2::pack<typename vview<VA>::type, G> a;
  nt2::pack<typename vview<VB>::type, G> b;
  ev<G, A>(v.a, i, a.v);
  ev<G, A>(v.b, i, b.v);
  const auto res = v.f(a, b);
#pragma unroll
  for (int l = 0; l < G; ++l) r[l] = (R)res[l];
}
}}  // namespace hip::fe

namespace expr {
template <class T> struct linspace_node : is_expr_base {
  typedef T value_type;
  size_t n;
  T lo, hi;
  ext_t ext() const { return ext_t{true, of_size(1, n)}; }
  void prepare() const {}
};
template <class A> struct expand_node : is_expr_base {
  typedef typename A::value_type value_type;
  A a;
  of_size_ to;
  ext_t ext() const { return ext_t{true, to}; }
  void prepare() const { a.prepare(); }
};
template <class F, class A, class B> struct arrayfun_node : is_expr_base {
  typedef typename hip::fe::vview<hip::fe::afun2_v<F, typename dview<A>::type, typename dview<B>::type>>::type
      value_type;
  F f;
  A a;
  B b;
  ext_t ext() const { return merge_ext(a.ext(), b.ext()); }
  void prepare() const { a.prepare(); b.prepare(); }
};
template <class T> struct dview<linspace_node<T>> {
  typedef hip::fe::lin_v<T> type;
  static type get(const linspace_node<T>& x) {
    return type{x.n, x.lo, (T)((x.hi - x.lo) / (T)(x.n - 1)), x.hi};
  }
};
template <class A> struct dview<expand_node<A>> {
  typedef hip::fe::bcast_v<typename dview<A>::type> type;
  static type get(const expand_node<A>& x) {
    type v;
    v.a = dview<A>::get(x.a);
    const ext_t e = x.a.ext();
    for (int d = 0; d < 4; ++d) {
      v.to[d] = x.to.d[d];
      v.from[d] = e.has ? e.s.d[d] : 1;
    }
    return v;
  }
};
template <class F, class A, class B> struct dview<arrayfun_node<F, A, B>> {
  typedef hip::fe::afun2_v<F, typename dview<A>::type, typename dview<B>::type> type;
  static type get(const arrayfun_node<F, A, B>& x) { return type{x.f, dview<A>::get(x.a), dview<B>::get(x.b)}; }
};
template <class A> struct pack_free<expand_node<A>> { static constexpr bool value = pack_free<A>::value; };
template <class F, class A, class B> struct pack_free<arrayfun_node<F, A, B>> { static constexpr bool value = false; };
}  // namespace expr

// linspace(a, b, n): n points from a to b in a 1 x n row (100 by default)
template <class T> expr::linspace_node<T> linspace(T a, T b, size_t n = 100) {
  expr::linspace_node<T> x;
  x.n = n;
  x.lo = a;
  x.hi = b;
  return x;
}
// expand_to(a, sizes): a broadcast along its singleton dimensions to the new
// extent (every other dimension must match)
template <class E, NT2H_ENABLE1(E)>
auto expand_to(const E& e, of_size_ s) {
  typedef typename expr::wrap<E>::type EX;
  expr::expand_node<EX> x;
  x.a = expr::wrap<E>::get(e);
  x.to = s;
  const ext_t from = x.a.ext();
  for (int d = 0; d < 4 && from.has; ++d)
    if (from.s.d[d] != 1 && from.s.d[d] != s.d[d])
      throw std::invalid_argument("nt2: expand_to of a non-singleton dimension to another size");
  return x;
}
template <class E, NT2H_ENABLE1(E)> auto expand_to(const E& e, size_t h, size_t w) { return expand_to(e, of_size(h, w)); }
// colvect(a): the elements of a as one column
template <class E, NT2H_ENABLE1(E)>
auto colvect(const E& e) {
  typedef typename expr::wrap<E>::type EX;
  const ext_t x = expr::wrap<E>::get(e).ext();
  return reshape(e, of_size(x.has ? x.s.numel() : 1));
}
// arrayfun(f, a, b): f applied to a and b pack by pack (tail element by element)
template <class F, class A, class B, NT2H_ENABLE2(A, B)>
auto arrayfun(const F& f, const A& a, const B& b) {
  typedef typename expr::wrap<A>::type EA;
  typedef typename expr::wrap<B>::type EB;
  expr::arrayfun_node<F, EA, EB> x;
  x.f = f;
  x.a = expr::wrap<A>::get(a);
  x.b = expr::wrap<B>::get(b);
  return x;
}

// ---- generators used by the demos ----------------------------------------------------
template <class T> table<T> ones(of_size_ s, meta::as_<T>) {
  table<T> t(s);
  for (size_t i = 1; i <= s.numel(); ++i) t(i) = T(1);
  return t;
}
template <class T> table<T> zeros(of_size_ s, meta::as_<T>) { return table<T>(s); }

}  // namespace nt2
