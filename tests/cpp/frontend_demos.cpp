// NT2 demo statements compiled unchanged against the HIP site (nt2/table.hpp)
// and run on the GPU; every result is written to <outdir>/<name>.bin so the
// test can compare it with the oracle.  The statement bodies are those of
// demo/saxpy/nt2/saxpy_nt2.cpp:39-42, demo/sdot/nt2/sdot_nt2.cpp:35-38,
// demo/blackscholes/nt2/blackandscholes_nt2.cpp:55-59,
// demo/sigmadelta/nt2/sigmadelta_nt2.cpp:37-58, demo/rgb2grey/nt2/rgb2grey_nt2.cpp:44,
// demo/rgb2yuv/nt2/rgb2yuv_nt2.cpp:30-35 and demo/mandelbrot/nt2/mandelbrot_nt2.cpp:86-88,
// :102 (NT2's table API).
#include <nt2/table.hpp>

#include <cstdio>
#include <cstdlib>
#include <string>
#include <vector>

using namespace nt2;

static std::string outdir;

template <class T> static void dump(const char* name, const table<T>& t) {
  FILE* f = std::fopen((outdir + "/" + name + ".bin").c_str(), "wb");
  std::fwrite(t.data(), sizeof(T), t.numel(), f);
  std::fclose(f);
}
template <class T> static void dump_scalar(const char* name, T v) {
  FILE* f = std::fopen((outdir + "/" + name + ".bin").c_str(), "wb");
  std::fwrite(&v, sizeof(T), 1, f);
  std::fclose(f);
}

// simple seeded inputs (the test regenerates them with numpy)
static uint64_t lcg(uint64_t& s) { s = s * 6364136223846793005ull + 1442695040888963407ull; return s >> 11; }
static float unif(uint64_t& s, float lo, float hi) { return lo + (hi - lo) * (float)((double)lcg(s) * 0x1.0p-53); }

template <class Pixel>
void do_work(Pixel& bkg, const Pixel& fr, Pixel& var, Pixel& diff_img, Pixel& mul_img, Pixel& out) {
  typename Pixel::value_type zero = 0;
  bkg = nt2::selinc(bkg < fr, nt2::seldec(bkg > fr, bkg));
  diff_img = nt2::max(bkg, fr) - nt2::min(bkg, fr);
  mul_img = nt2::adds(nt2::adds(diff_img, diff_img), diff_img);
  var = nt2::if_else(diff_img != zero, nt2::selinc(var < mul_img, nt2::seldec(var > mul_img, var)), var);
  out = nt2::if_zero_else_one(diff_img < var);
}

// Escape-time count of z = z^2 + c from z = 0, one call per pack: a lane
// counts the iterations whose |z|^2 (before the update) is below 4, and the
// pack goes on while any lane still counts, at most `limit` times.
struct escape_count {
  int limit;
  template <class V>
  __host__ __device__ typename nt2::meta::as_integer<V>::type operator()(const V& cr, const V& ci) const {
    typedef typename nt2::meta::as_integer<V>::type count_t;
    typedef typename nt2::meta::scalar_of<V>::type real_t;
    count_t count = nt2::Zero<count_t>();
    V zr = nt2::Zero<V>(), zi = nt2::Zero<V>();
    int k = 0;
    bool again;
    do {
      const V rr = zr * zr, ii = zi * zi, ri = real_t(2) * zr * zi;
      const auto inside = (rr + ii) < real_t(4);
      zr = rr - ii + cr;
      zi = ri + ci;
      count = nt2::seladd(inside, count, nt2::One<count_t>());
      again = nt2::any(inside);
    } while (again && ++k < limit);
    return count;
  }
};

int main(int argc, char** argv) {
  outdir = argc > 1 ? argv[1] : ".";
  const std::size_t size = argc > 2 ? std::strtoul(argv[2], nullptr, 10) : 100003;
  uint64_t seed = 12345;

  // ---- saxpy / sdot
  table<float> X(of_size(size)), Y(of_size(size));
  for (std::size_t i = 1; i <= size; ++i) { X(i) = unif(seed, -1, 1); Y(i) = unif(seed, -1, 1); }
  dump("x", X);
  dump("y", Y);
  float alpha = 2.5f;
  Y = alpha * X + Y;
  dump("saxpy", Y);
  float res = globalsum(X * Y);
  dump_scalar("sdot", res);
  dump_scalar("dot", nt2::dot(X, Y));

  // ---- Black-Scholes, the demo's four statements
  table<float> Sa(of_size(size)), Xa(of_size(size)), Ta(of_size(size)), ra(of_size(size)),
      va(of_size(size)), R(of_size(size));
  for (std::size_t i = 1; i <= size; ++i) {
    Sa(i) = unif(seed, 5, 30);
    Xa(i) = unif(seed, 1, 100);
    Ta(i) = unif(seed, 0.25f, 10);
    ra(i) = 0.02f;
    va(i) = 0.30f;
  }
  dump("S", Sa);
  dump("X", Xa);
  dump("T", Ta);
  {
    typedef float T;
    nt2::table<T> da = nt2::sqrt(Ta);
    nt2::table<T> d1 = nt2::log(Sa / Xa) + (nt2::fma(nt2::sqr(va), T(0.5), ra) * Ta) / (va * da);
    nt2::table<T> d2 = nt2::fnms(va, da, d1);
    R = nt2::fnms(Xa * nt2::exp(-ra * Ta), nt2::fastnormcdf(d2), Sa * nt2::fastnormcdf(d1));
  }
  dump("bs_demo", R);

  // ---- one fused statement: call and put with erfc-based normcdf, via tie
  table<float> call, put;
  {
    const float r = 0.02f, v = 0.30f;
    table<float> vsT = v * nt2::sqrt(Ta);
    table<float> d1 = (nt2::log(Sa / Xa) + nt2::fma(nt2::sqr(v), 0.5f, r) * Ta) / vsT;
    table<float> d2 = d1 - vsT;
    table<float> e = Xa * nt2::exp(-r * Ta);
    nt2::tie(call, put) = nt2::tie(Sa * nt2::normcdf(d1) - e * nt2::normcdf(d2),
                                   e * nt2::normcdf(-d2) - Sa * nt2::normcdf(-d1));
  }
  dump("bs_call", call);
  dump("bs_put", put);

  // ---- trigonometric expression with packs (config-4 expression)
  table<float> xn(of_size(size)), yn(of_size(size)), z(of_size(size));
  for (std::size_t i = 1; i <= size; ++i) { xn(i) = unif(seed, -3, 3); yn(i) = unif(seed, -1e4, 1e4); }
  dump("xn", xn);
  dump("yn", yn);
  z = nt2::exp(-xn * xn / 2.0f) * nt2::cos(yn) + nt2::sqrt(nt2::abs(xn));
  dump("expr4", z);
  dump_scalar("expr4_sum", nt2::globalsum(nt2::exp(-xn * xn / 2.0f) * nt2::cos(yn) + nt2::sqrt(nt2::abs(xn))));

  // ---- sigma-delta: 8 frames of 37x29 pixels
  const std::size_t h = 37, w = 29, nf = 8;
  std::vector<table<uint8_t>> frames(nf);
  for (std::size_t k = 0; k < nf; ++k) {
    frames[k].resize(of_size(h, w));
    for (std::size_t j = 1; j <= w; ++j)
      for (std::size_t i = 1; i <= h; ++i) frames[k](i, j) = (uint8_t)(lcg(seed) & 0xff);
  }
  {
    table<uint8_t> all(of_size(h * w * nf));
    for (std::size_t k = 0; k < nf; ++k)
      for (std::size_t p = 1; p <= h * w; ++p) all(k * h * w + p) = frames[k](p);
    dump("frames", all);
  }
  table<uint8_t> image_variance = ones(of_size(h, w), meta::as_<uint8_t>());
  table<uint8_t> background_img(of_size(h, w)), diff_img, mul_img;
  table<uint8_t> binary_label = zeros(of_size(h, w), meta::as_<uint8_t>());
  for (std::size_t i = 1; i <= h; ++i)
    for (std::size_t j = 1; j <= w; ++j) background_img(i, j) = frames[0](i, j);
  for (std::size_t k = 1; k < nf; ++k)
    do_work(background_img, frames[k], image_variance, diff_img, mul_img, binary_label);
  dump("sd_bkg", background_img);
  dump("sd_var", image_variance);
  dump("sd_label", binary_label);
  dump_scalar("frame1_sum", nt2::globalsum(frames[1]));  // uint8: NT2 sums in the element type, wrapping

  // ---- sum / cumsum along dimensions (NT2's unit tests
  // modules/core/reduction/unit/table/sum.cpp and modules/core/swar/unit/table/cumsum.cpp)
  table<double> a = reshape(_(1.0, 30.0), 3, 5, 2);
  table<double> b = cumsum(a);
  dump("cumsum_inner", b);
  b = cumsum(a, 2);
  dump("cumsum_outer", b);
  table<double> r = sum(a);
  dump("sum_a", r);
  r = sum(a, 2);
  dump("sum_a2", r);
  r = sum(a, 3);
  dump("sum_a3", r);
  r = sum(a, 4);
  dump("sum_a4", r);
  const std::size_t mh = 256, mw = 129;
  table<float> M(of_size(mh, mw));
  for (std::size_t i = 1; i <= mh * mw; ++i) M(i) = unif(seed, -1e4f, 1e4f);
  dump("M", M);
  table<float> s1 = sum(M, 1), s2 = sum(M, 2), c1 = cumsum(M), c2 = cumsum(M, 2);
  dump("M_sum1", s1);
  dump("M_sum2", s2);
  dump("M_cumsum1", c1);
  dump("M_cumsum2", c2);
  s1 = sum(nt2::cos(M), 1);  // cos decides its range reduction per pack
  s2 = sum(nt2::cos(M), 2);
  dump("M_sumcos1", s1);
  dump("M_sumcos2", s2);
  s1 = sum(M * 2.0f + 1.0f);  // first non-singleton dimension: 1
  dump("M_sumexpr", s1);
  c1 = cumprod(M / 1.0e4f + 1.0f, 1);
  dump("M_cumprod1", c1);
  c2 = cummin(M, 2);
  dump("M_cummin2", c2);
  c1 = cummax(nt2::cos(M));  // scans evaluate each element in NT2's scalar form
  dump("M_cummaxcos1", c1);
  table<float> v = M(1) * ones(of_size(1000), meta::as_<float>());
  const float vs = sum(v);  // a one-element result used as a scalar
  dump_scalar("v_sum", vs);

  // ---- exponentials, logarithms, rounding and sign in one statement each
  table<float> ew(of_size(size));
  ew = nt2::round(xn * 10.0f) + nt2::log2(nt2::abs(yn) + 1.0f) - nt2::exp10(xn / 8.0f);
  dump("ew_f32", ew);
  ew = nt2::round2even(yn / 16.0f) * nt2::sign(xn) + nt2::exp2(nt2::floor(xn) + nt2::ceil(-xn) / 4.0f);
  dump("ew2_f32", ew);
  ew = nt2::tanh(xn) * nt2::expm1(xn / 4.0f) + nt2::log1p(nt2::abs(yn));
  dump("ew3_f32", ew);
  table<float> at(of_size(size - 3));  // a ragged tail of NT2's scalar forms
  for (std::size_t i = 1; i <= size - 3; ++i) at(i) = xn(i);
  table<float> bt(of_size(size - 3));
  for (std::size_t i = 1; i <= size - 3; ++i) bt(i) = yn(i) / 1000.0f;
  table<float> at2;
  at2 = nt2::atan(at * 3.0f) + nt2::atan2(at, bt);
  dump("at_a", at);
  dump("at_b", bt);
  dump("atan_f32", at2);
  at2 = nt2::tan(at) * 0.5f - nt2::tan(bt * 100.0f);
  dump("tan_f32", at2);
  at2 = nt2::asin(at / 3.0f) * nt2::acos(bt / 10.0f);
  dump("asin_f32", at2);
  at2 = nt2::pow(nt2::abs(at) + 0.5f, bt / 4.0f) + nt2::pow(at, nt2::round(bt));
  dump("pow_f32", at2);
  at2 = nt2::pow(at, 3) + nt2::pow(bt, -2);  // integral exponents: repeated squaring
  dump("powi_f32", at2);
  // unsigned and 64-bit exponents keep their type: 3e9 squares, no 1/p
  at2 = nt2::pow(at, 2u) + nt2::pow(bt * 0.001f + 1.0f, 3000000000u) - nt2::pow(at, (long long)-3);
  dump("powu_f32", at2);
  at2 = nt2::erf(at) + nt2::cbrt(bt * 50.0f) - nt2::asinh(bt * at * 100.0f);
  dump("erf_f32", at2);
  at2 = nt2::acosh(nt2::abs(bt) * 1.0e6f + 1.0f) * nt2::atanh(at / 3.0f);
  dump("acosh_f32", at2);
  at2 = nt2::cot(at) + nt2::sec(bt) * nt2::csc(at * 7.0f);
  dump("cot_f32", at2);
  at2 = nt2::sinpi(at) - nt2::cospi(bt) + nt2::tanpi(at / 7.0f) * nt2::cotpi(bt / 3.0f);
  dump("sinpi_f32", at2);
  // repeated subexpressions: shared when each repeated node type has one view,
  // otherwise every occurrence is evaluated; the values are the same either way
  at2 = nt2::exp(at) * nt2::exp(at) + nt2::sqrt(nt2::exp(at)) - nt2::cos(bt * 2.0f) / nt2::cos(bt * 2.0f);
  dump("rep_same", at2);
  at2 = nt2::exp(at) * nt2::exp(bt) + nt2::cos(bt * 2.0f) - nt2::cos(bt * 3.0f) + nt2::exp(at);
  dump("rep_mixed", at2);
  {
    table<float> p1, p2;
    nt2::tie(p1, p2) = nt2::tie(nt2::normcdf(at) + nt2::normcdf(bt), nt2::normcdf(-at) - nt2::normcdf(-bt));
    dump("rep_tie1", p1);
    dump("rep_tie2", p2);
    nt2::tie(p1, p2) = nt2::tie(nt2::normcdf(at) * 2.0f, nt2::normcdf(-at));
    dump("rep_tie3", p1);
    dump("rep_tie4", p2);
  }
  dump_scalar("rep_sum_mixed", nt2::globalsum(nt2::exp(at) * nt2::exp(at) + nt2::exp(bt)));
  dump_scalar("rep_sum_same", nt2::globalsum(nt2::exp(at) * nt2::exp(at)));
  at2 = nt2::sind(at * 100.0f) - nt2::cosd(bt * 50.0f) + nt2::tand(at * 30.0f) * nt2::cotd(bt * 40.0f);
  dump("sind_f32", at2);
  // the compositions: reciprocal hyperbolic, inverse trigonometric of 1/x and in
  // degree / pi units, reciprocal trigonometric of degree / pi arguments, sinc
  at2 = nt2::coth(at) + nt2::sech(bt) * nt2::csch(at * 3.0f) - nt2::acoth(bt * 4.0f);
  dump("coth_f32", at2);
  at2 = nt2::asech(nt2::abs(at) / 4.0f) + nt2::acsch(bt) - nt2::acot(at) * nt2::asec(bt * 3.0f);
  dump("asech_f32", at2);
  at2 = nt2::acsc(at * 2.0f) + nt2::acotd(bt) - nt2::asecd(at * 5.0f) + nt2::acscd(bt * 4.0f);
  dump("acsc_f32", at2);
  at2 = nt2::acosd(at / 4.0f) - nt2::acospi(bt / 4.0f) + nt2::asind(at / 3.0f) * nt2::asinpi(bt / 5.0f);
  dump("acosd_f32", at2);
  at2 = nt2::atand(at) + nt2::atanpi(bt) + nt2::indeg(at) - nt2::inrad(bt * 100.0f);
  dump("atand_f32", at2);
  at2 = nt2::secd(at * 90.0f) + nt2::cscd(bt * 45.0f) - nt2::secpi(at) * nt2::cscpi(bt / 3.0f) + nt2::sinc(at * 4.0f) +
        nt2::sincpi(bt);
  dump("secd_f32", at2);
  at2 = nt2::atan2d(at, bt);
  dump("atan2d_f32", at2);
  at2 = nt2::sqrt1pm1(at) + nt2::exprecneg(bt) - nt2::exprecnegc(at * 2.0f) + nt2::reallog(nt2::abs(bt)) *
        nt2::realsqrt(nt2::abs(at)) + nt2::expx2(bt / 2.0f);
  dump("sqrt1pm1_f32", at2);
  at2 = nt2::logspace_add(at, bt) - nt2::logspace_sub(at * 2.0f, at * 2.0f - nt2::abs(bt) - 0.5f);
  dump("logspace_f32", at2);
  at2 = if_else((at > 0.0f) && (bt < 1.0f), at, bt) + if_else(logical_or(at < -1.0f, bt > 2.0f), 1.0f, 0.0f);
  dump("logical_f32", at2);
  table<double> dm(of_size(mh * mw)), dr;
  for (std::size_t i = 1; i <= mh * mw; ++i) dm(i) = (double)M(i) / 100.0;
  dr = nt2::trunc(dm) * nt2::exp2(nt2::ceil(dm / 50.0)) + nt2::sign(dm) * nt2::log10(nt2::abs(dm) + 1.0);
  dump("ew_f64", dr);
  dr = nt2::sinh(dm / 20.0) - nt2::cosh(dm / 30.0);
  dump("ew2_f64", dr);
  dr = nt2::tan(dm / 7.0) + nt2::atan(dm);
  dump("tan_f64", dr);
  table<double> dq(of_size(1001)), dq2;  // a ragged tail of NT2's scalar forms
  for (std::size_t i = 1; i <= 1001; ++i) dq(i) = dm(i) / 100.0;
  dq2 = nt2::asin(dq) - nt2::acos(dq * 0.9);
  dump("dq", dq);
  dump("asin_f64", dq2);
  dq2 = nt2::pow(dq + 2.0, dq * 300.0);
  dump("pow_f64", dq2);
  dq2 = nt2::pow(dq * 10.0, 7);
  dump("powi_f64", dq2);
  dq2 = nt2::erf(dq * 8.0) + nt2::cbrt(dq * 1e5) - nt2::asinh(dq * 1e9);
  dump("erf_f64", dq2);
  dq2 = nt2::acosh(nt2::abs(dq) * 1e17 + 1.0) * nt2::atanh(dq * 0.999);
  dump("acosh_f64", dq2);
  dq2 = nt2::cot(dq * 30.0) + nt2::sec(dq * 9.0) * nt2::csc(dq * 7.0);
  dump("cot_f64", dq2);
  dq2 = nt2::sinpi(dq * 5.0) - nt2::cospi(dq * 3.0) + nt2::tanpi(dq / 0.7) * nt2::cotpi(dq * 1.3);
  dump("sinpi_f64", dq2);
  dq2 = nt2::sind(dq * 500.0) - nt2::cosd(dq * 300.0) + nt2::tand(dq * 200.0) * nt2::cotd(dq * 170.0);
  dump("sind_f64", dq2);
  dq2 = nt2::coth(dq * 3.0) - nt2::asech(nt2::abs(dq)) + nt2::acot(dq) * nt2::asec(dq * 30.0);
  dump("coth_f64", dq2);
  dq2 = nt2::acotd(dq * 7.0) + nt2::acscd(dq * 20.0) - nt2::asind(dq * 0.9) + nt2::atanpi(dq * 9.0);
  dump("acotd_f64", dq2);
  dq2 = nt2::secd(dq * 900.0) - nt2::cscpi(dq * 6.0) + nt2::sinc(dq * 40.0) * nt2::sincpi(dq * 5.0) +
        nt2::atan2d(dq, dq * 0.5 - 0.1);
  dump("secd_f64", dq2);

  // ---- rgb2grey and rgb2yuv on a 96 x 131 image: the demos' ramp and random planes
  const std::size_t ih = 96, iw = 131;
  table<float> pr(of_size(ih, iw)), pg(of_size(ih, iw)), pb(of_size(ih, iw)), grey;
  for (std::size_t i = 1; i <= ih * iw; ++i) pr(i) = pg(i) = pb(i) = float(i - 1);
  grey = float(0.3) * pr + float(0.59) * pg + float(0.11) * pb;
  dump("grey_ramp", grey);
  for (std::size_t i = 1; i <= ih * iw; ++i) {
    pr(i) = unif(seed, 0, 255);
    pg(i) = unif(seed, 0, 255);
    pb(i) = unif(seed, 0, 255);
  }
  dump("pr", pr);
  dump("pg", pg);
  dump("pb", pb);
  grey = float(0.3) * pr + float(0.59) * pg + float(0.11) * pb;
  dump("grey", grey);
  table<float> yy, uu, vv;
  yy = float(0.299f) * pr + float(0.587f) * pg + float(0.114f) * pb;
  uu = float(0.492f) * (pb - yy);
  vv = float(0.877f) * (pr - yy);
  dump("yuv_y", yy);
  dump("yuv_u", uu);
  dump("yuv_v", vv);

  // ---- mandelbrot: a 203 x 203 grid of [-1.5, 1.5]^2 (203^2 is not a multiple of 8)
  const std::size_t mn = 203;
  table<float> MA, MB;
  MA = expand_to(linspace(-1.5f, 1.5f, mn), mn, mn);
  MB = expand_to(colvect(linspace(-1.5f, 1.5f, mn)), mn, mn);
  table<int> MC;
  MC = arrayfun(escape_count{256}, MA, MB);
  dump("mandel_a", MA);
  dump("mandel_b", MB);
  dump("mandel", MC);
  dump_scalar("mandel_sum", nt2::globalsum(MC * 1000003));  // int32, wrapping past 2^31
  table<float> lin = linspace(0.0f, 1.0f);
  dump("linspace100", lin);

  // ---- device storage: host and device copies stay coherent across statements
  {
    const std::size_t m = 100003;
    table<float> h(of_size(m));
    for (std::size_t i = 1; i <= m; ++i) h(i) = float(i);   // host writes
    table<float> d2 = h * 2.0f;                               // uploads h, result on the device
    table<float> probe(of_size(8));
    probe(1) = d2(m);                                         // downloads d2: 2m
    h(7) = -1.0f;                                             // host write after h went to the device
    table<float> d3 = h + 0.0f;                               // must see -1 at 7 (re-upload)
    probe(2) = d3(7);
    probe(3) = d3(m);
    table<float> cp = d3;                                     // device-to-device copy
    cp(1) = 42.0f;                                            // the copy's host write
    cp = cp + 1.0f;                                           // in place: 43 at 1
    probe(4) = cp(1);
    probe(5) = d3(1);                                         // the source is unchanged: 1
    d3 = d3 * 3.0f;                                           // in place on the device, host stale
    probe(6) = d3(m);                                         // 3m
    h.resize(of_size(2 * m));                                 // new size: zeros
    table<float> d4 = h + 1.0f;
    probe(7) = globalsum(d4);                                 // 2m
    probe(8) = d4(2 * m);
    dump("coherence", probe);
    dump("coherence_cp", cp);
  }
  // ---- table settings: index base and number of dimensions
  {
    table<float, settings(C_index_, _2D)> c0(of_size(3, 4));
    for (std::size_t j = 0; j < 4; ++j)
      for (std::size_t i = 0; i < 3; ++i) c0(i, j) = float(10 * i + j);  // 0-based
    table<float, _2D> m1 = c0 * 2.0f;                                     // 1-based
    table<double, matlab_index_> a6 = reshape(_(1.0, 6.0), 2, 3);
    table<double, settings(matlab_order_, matlab_index_, _2D)> d1 = cumsum(a6);
    table<float> probe(of_size(7));
    probe(1) = m1(1, 1);
    probe(2) = m1(3, 4);
    probe(3) = c0(2, 3);
    probe(4) = float(d1(2, 3));
    bool threw = false;
    try { table<float, _2D> bad(of_size(2, 2, 2)); } catch (const std::invalid_argument&) { threw = true; }
    probe(5) = threw ? 1.0f : 0.0f;
    threw = false;
    try { table<float, _1D> v = c0 + 1.0f; } catch (const std::invalid_argument&) { threw = true; }
    probe(6) = threw ? 1.0f : 0.0f;
    table<float, _1D> v1 = reshape(_(1.0f, 5.0f), of_size(5));  // a column: _(a, b) is a row
    probe(7) = globalsum(v1 * v1);
    dump("settings", probe);
    dump("settings_m1", m1);
  }
  std::printf("frontend demos done: %zu elements\n", size);
  return 0;
}
