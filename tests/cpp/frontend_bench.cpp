// Front-end statements at 2^26 elements against their C ABI twins: NT2 user
// code (nt2/table.hpp, lowered at compile time to one kernel per statement)
// timed beside the library's fixed kernels on the same device tables.
//   Y = alpha * X + Y            vs nt2h_saxpy_f32      (demo/saxpy/nt2/saxpy_nt2.cpp:39-42)
//   s = globalsum(X * Y)         vs nt2h_dot_f32        (demo/sdot/nt2/sdot_nt2.cpp:35-38)
//   tie(call, put) = tie(...)    vs nt2h_blackscholes_f32, exact policy (one fused statement)
//   the saxpy then sdot step     vs nt2h_saxpy_f32 then nt2h_dot_f32
// Prints one JSON object.  Usage: frontend_bench [log2 n] [reps]
#include <nt2/table.hpp>

#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "nt2h.h"

using namespace nt2;

static void ck(int rc, const char* what) {
  if (rc != NT2H_OK) {
    std::fprintf(stderr, "%s: %s\n", what, nt2h_last_error());
    std::exit(1);
  }
}

// 50 untimed launches first: the GPU clock ramps up over the first tens of
// launches after an idle period or host transfers
template <class F> static double time_ms(F&& f, int reps) {
  for (int i = 0; i < 50; ++i) f();
  hipDeviceSynchronize();
  hipEvent_t a, b;
  hipEventCreate(&a);
  hipEventCreate(&b);
  hipEventRecord(a, nullptr);
  for (int i = 0; i < reps; ++i) f();
  hipEventRecord(b, nullptr);
  hipEventSynchronize(b);
  float ms = 0;
  hipEventElapsedTime(&ms, a, b);
  hipEventDestroy(a);
  hipEventDestroy(b);
  return ms / reps;
}

template <class T> static bool same_bits(const T* a, const T* b, size_t n) {
  return std::memcmp(a, b, n * sizeof(T)) == 0;
}

int main(int argc, char** argv) {
  const int lg = argc > 1 ? std::atoi(argv[1]) : 26;
  const int reps = argc > 2 ? std::atoi(argv[2]) : 20;
  const size_t n = size_t(1) << lg;
  const float alpha = 2.5f, r = 0.02f, v = 0.30f;

  table<float> X(of_size(n)), Y(of_size(n)), Y2(of_size(n));
  ck(nt2h_fill_uniform_f32(X.device_data_for_write(), n, -1.0, 1.0, 1, nullptr), "fill");
  ck(nt2h_fill_uniform_f32(Y.device_data_for_write(), n, -1.0, 1.0, 11, nullptr), "fill");
  ck(nt2h_memcpy(Y2.device_data_for_write(), Y.device_data(), n * 4, NT2H_D2D, nullptr), "copy");

  // saxpy: one in-place statement vs the fixed kernel
  const double fe_saxpy = time_ms([&] { Y = alpha * X + Y; }, reps);
  const double abi_saxpy =
      time_ms([&] { ck(nt2h_saxpy_f32(n, alpha, X.device_data(), Y2.device_data_for_write(), nullptr), "saxpy"); },
              reps);
  // both tables took the same 50 + reps updates
  const bool saxpy_equal = same_bits(Y.data(), Y2.data(), n);

  // dot: globalsum returns on the host (one synchronisation per statement)
  float fe_s = 0.0f;
  const double fe_dot = time_ms([&] { fe_s = globalsum(X * Y); }, reps);
  double* d_s = nullptr;
  ck(nt2h_malloc((void**)&d_s, sizeof(double)), "malloc");
  const double abi_dot = time_ms([&] { ck(nt2h_dot_f32(n, X.device_data(), Y.device_data(), d_s, nullptr), "dot"); }, reps);
  double abi_s = 0.0;
  ck(nt2h_memcpy(&abi_s, d_s, sizeof(double), NT2H_D2H, nullptr), "copy");
  const bool dot_close = std::fabs((double)fe_s - abi_s) <= 1e-6 * std::fabs(abi_s) + 1e-6;
  const double abi_dot_sync = time_ms([&] {
    ck(nt2h_dot_f32(n, X.device_data(), Y.device_data(), d_s, nullptr), "dot");
    double h = 0.0;
    ck(nt2h_memcpy(&h, d_s, sizeof(double), NT2H_D2H, nullptr), "copy");
  }, reps);

  // the saxpy then sdot step (the two demo statements back to back): the C
  // ABI's three launches, and the front end's two statements (globalsum returns
  // its value to the host, as the sdot demo does)
  const double abi_step = time_ms([&] {
    ck(nt2h_saxpy_f32(n, alpha, X.device_data(), Y2.device_data_for_write(), nullptr), "saxpy");
    ck(nt2h_dot_f32(n, X.device_data(), Y2.device_data(), d_s, nullptr), "dot");
  }, reps);
  const double fe_step = time_ms([&] {
    Y = alpha * X + Y;
    fe_s = globalsum(X * Y);
  }, reps);

  // Black-Scholes: call and put in one tie statement vs the fixed exact kernel
  table<float> S(of_size(n)), K(of_size(n)), T(of_size(n)), call, put, c2(of_size(n)), p2(of_size(n));
  ck(nt2h_fill_uniform_f32(S.device_data_for_write(), n, 5.0, 30.0, 33, nullptr), "fill");
  ck(nt2h_fill_uniform_f32(K.device_data_for_write(), n, 1.0, 100.0, 34, nullptr), "fill");
  ck(nt2h_fill_uniform_f32(T.device_data_for_write(), n, 0.25, 10.0, 35, nullptr), "fill");
  const double fe_bs = time_ms([&] {
    nt2::tie(call, put) = nt2::tie(
        S * nt2::normcdf((nt2::log(S / K) + nt2::fma(nt2::sqr(v), 0.5f, r) * T) / (v * nt2::sqrt(T))) -
            (K * nt2::exp(-r * T)) *
                nt2::normcdf((nt2::log(S / K) + nt2::fma(nt2::sqr(v), 0.5f, r) * T) / (v * nt2::sqrt(T)) -
                             v * nt2::sqrt(T)),
        (K * nt2::exp(-r * T)) *
                nt2::normcdf(-((nt2::log(S / K) + nt2::fma(nt2::sqr(v), 0.5f, r) * T) / (v * nt2::sqrt(T)) -
                               v * nt2::sqrt(T))) -
            S * nt2::normcdf(-((nt2::log(S / K) + nt2::fma(nt2::sqr(v), 0.5f, r) * T) / (v * nt2::sqrt(T)))));
  }, reps);
  const double abi_bs = time_ms([&] {
    ck(nt2h_blackscholes_f32(n, S.device_data(), K.device_data(), T.device_data(), r, v, c2.device_data_for_write(),
                             p2.device_data_for_write(), NT2H_BS_TEXTBOOK, NT2H_EXACT, nullptr), "bs");
  }, reps);
  const bool bs_equal = same_bits(call.data(), c2.data(), n) && same_bits(put.data(), p2.data(), n);
  nt2h_free(d_s);

  auto gbs = [&](double bytes_per, double ms) { return bytes_per * (double)n / (ms * 1e-3) / 1e9; };
  std::printf("{\"n\": %zu, \"reps\": %d, "
              "\"saxpy\": {\"frontend_ms\": %.5f, \"cabi_ms\": %.5f, \"frontend_gbs\": %.1f, \"cabi_gbs\": %.1f, "
              "\"frontend_over_cabi\": %.4f, \"bitwise_equal\": %s}, "
              "\"dot\": {\"frontend_ms\": %.5f, \"cabi_ms\": %.5f, \"cabi_with_result_copy_ms\": %.5f, "
              "\"frontend_gbs\": %.1f, \"cabi_gbs\": %.1f, \"frontend_over_cabi_with_copy\": %.4f, \"close\": %s}, "
              "\"saxpy_sdot_step\": {\"frontend_ms\": %.5f, \"cabi_ms\": %.5f}, "
              "\"blackscholes_tie\": {\"frontend_ms\": %.5f, \"cabi_exact_ms\": %.5f, \"frontend_gbs\": %.1f, "
              "\"cabi_gbs\": %.1f, \"frontend_over_cabi\": %.4f, \"bitwise_equal\": %s}}\n",
              n, reps, fe_saxpy, abi_saxpy, gbs(12, fe_saxpy), gbs(12, abi_saxpy), fe_saxpy / abi_saxpy,
              saxpy_equal ? "true" : "false", fe_dot, abi_dot, abi_dot_sync, gbs(8, fe_dot), gbs(8, abi_dot),
              fe_dot / abi_dot_sync, dot_close ? "true" : "false", fe_step, abi_step, fe_bs, abi_bs, gbs(20, fe_bs), gbs(20, abi_bs),
              fe_bs / abi_bs, bs_equal ? "true" : "false");
  return 0;
}
