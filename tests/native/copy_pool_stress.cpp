// Stress test of the host copy pool (goofy_amd/csrc/copy_pool.h): several
// threads call run() concurrently; every item must run exactly once.  Built
// with -fsanitize=thread by tests/test_native_host.py (host code only).
#include "copy_pool.h"
#include <cstdio>
#include <cstring>
#include <vector>
int main() {
  using goofy_impl::CopyPool;
  std::vector<std::thread> callers;
  std::atomic<long> bad{0};
  for (int t = 0; t < 6; ++t) callers.emplace_back([t, &bad] {
    for (int it = 0; it < 400; ++it) {
      const int n = 1 + (it * 7 + t) % 23;
      std::vector<int> hit(n, 0);
      CopyPool::get().run(n, [&](int i) { hit[i] += 1; });
      for (int i = 0; i < n; ++i) if (hit[i] != 1) bad++;
    }
  });
  for (auto& c : callers) c.join();
  printf("pool workers=%d mismatches=%ld\n", CopyPool::get().workers(), bad.load());
  return bad != 0;
}
