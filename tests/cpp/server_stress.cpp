// Concurrent clients of the serving scheduler on its host back end (no GPU):
// four threads submit f32 and f64 requests, poll, wait, cancel and release
// while a fifth reads the statistics, with forced allocation and launch
// failures.  Exit status 0 when every request ends (done, failed or
// cancelled) and the budget returns to zero.  tests/test_server_host.py runs
// it; tests/test_sanitizers.py builds it under ThreadSanitizer and
// AddressSanitizer.  Usage: server_stress
#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>
#include "nt2h.h"

int main() {
  nt2h_server_config cfg{};
  cfg.budget_bytes = size_t(64) << 20;
  cfg.max_inflight = 4;
  cfg.max_batch = 8;
  cfg.policy = NT2H_ACCURATE;
  cfg.backend = NT2H_BACKEND_HOST;
  cfg.fail_alloc_every = 7;
  cfg.fail_launch_every = 5;
  nt2h_server* s = nullptr;
  if (nt2h_server_create(&cfg, &s) != NT2H_OK) { std::printf("create failed: %s\n", nt2h_last_error()); return 1; }
  const int nthr = 4, per = 200;
  std::vector<std::vector<float>> f32(3, std::vector<float>(4096));
  std::vector<std::vector<double>> f64(3, std::vector<double>(4096));
  for (int i = 0; i < 4096; ++i) {
    f32[0][i] = 5 + i % 25; f32[1][i] = 1 + i % 99; f32[2][i] = 0.25f + (i % 39) * 0.25f;
    f64[0][i] = 5 + i % 25; f64[1][i] = 1 + i % 99; f64[2][i] = 0.25 + (i % 39) * 0.25;
  }
  std::atomic<int> done{0}, failed{0}, cancelled{0};
  std::vector<std::thread> th;
  for (int t = 0; t < nthr; ++t)
    th.emplace_back([&, t] {
      std::vector<uint64_t> ids;
      for (int k = 0; k < per; ++k) {
        nt2h_bs_request r{};
        const bool d = (k + t) % 5 == 0;
        r.dtype = d ? NT2H_F64 : NT2H_F32;
        r.n = 1 + (size_t)((k * 37 + t * 11) % 4096);
        r.S = d ? (const void*)f64[0].data() : f32[0].data();
        r.X = d ? (const void*)f64[1].data() : f32[1].data();
        r.T = d ? (const void*)f64[2].data() : f32[2].data();
        r.r = 0.02; r.v = 0.3;
        uint64_t id = 0;
        if (nt2h_server_submit(s, &r, &id) != NT2H_OK) continue;
        if (k % 9 == 0 && nt2h_server_cancel(s, id) == NT2H_OK) { ++cancelled; continue; }
        ids.push_back(id);
        if (ids.size() >= 8) {
          for (uint64_t q : ids) {
            nt2h_server_poll(s, q);
            const int rc = nt2h_server_wait(s, q, 10.0);
            if (rc == NT2H_OK) ++done; else ++failed;
            nt2h_server_release(s, q);
          }
          ids.clear();
        }
      }
      for (uint64_t q : ids) {
        const int rc = nt2h_server_wait(s, q, 10.0);
        if (rc == NT2H_OK) ++done; else ++failed;
        nt2h_server_release(s, q);
      }
    });
  std::thread stats([&] {
    for (int i = 0; i < 50; ++i) { nt2h_server_stats st; nt2h_server_get_stats(s, &st); std::this_thread::yield(); }
  });
  for (auto& x : th) x.join();
  stats.join();
  nt2h_server_drain(s, 10.0);
  nt2h_server_stats st;
  nt2h_server_get_stats(s, &st);
  std::printf("{\"done\": %d, \"failed\": %d, \"cancelled\": %d, \"server_completed\": %llu, "
              "\"server_failed\": %llu, \"used_bytes\": %zu}\n", done.load(), failed.load(), cancelled.load(),
              (unsigned long long)st.completed, (unsigned long long)st.failed, st.used_bytes);
  const bool ok = st.used_bytes == 0 && (uint64_t)done.load() == st.completed && (uint64_t)failed.load() == st.failed &&
                  done.load() + failed.load() + cancelled.load() > 0;
  nt2h_server_destroy(s);
  return ok ? 0 : 2;
}
