// Host-only checks of dida::FitAlloc (the serving loop's region allocator),
// built with -fsanitize=address,undefined by tests/test_sanitize.py.
// Random alloc/release sequences are checked against a byte map of the
// region: blocks never overlap, lie inside the region, and every release
// coalesces back to a single free block once everything is returned.
#include <cstdio>
#include <cstdlib>
#include <random>
#include <vector>

#include "../../dida-thhpca_amd/csrc/letkf_pool.h"

using dida::FitAlloc;

static int fail(const char* what, long long a, long long b) {
  std::printf("FAIL %s (%lld, %lld)\n", what, a, b);
  return 1;
}

int main() {
  const size_t G = FitAlloc::kGranule;
  for (int seed = 1; seed <= 20; ++seed) {
    std::mt19937_64 rng(seed);
    FitAlloc fa;
    const size_t size = G * (64 + rng() % 4096);
    fa.reset(size + rng() % G);                    // unaligned sizes round down
    if (fa.size() != size) return fail("size", (long long)fa.size(), (long long)size);
    std::vector<unsigned char> owner(size / G, 0);
    struct Blk { size_t off, n; };
    std::vector<Blk> live;
    for (int step = 0; step < 4000; ++step) {
      const bool do_alloc = live.empty() || rng() % 3 != 0;
      if (do_alloc) {
        const size_t n = 1 + rng() % (size / 8);
        const size_t off = fa.alloc(n);
        if (off == FitAlloc::npos) {
          if (fa.largest() >= FitAlloc::round(n)) return fail("refused though it fits", n, fa.largest());
          continue;
        }
        if (off % G || off + FitAlloc::round(n) > size) return fail("outside", off, n);
        for (size_t g = off / G; g < (off + FitAlloc::round(n)) / G; ++g) {
          if (owner[g]) return fail("overlap", off, n);
          owner[g] = 1;
        }
        live.push_back({off, n});
      } else {
        const size_t i = rng() % live.size();
        const Blk b = live[i];
        if (!fa.release(b.off, b.n)) return fail("release refused", b.off, b.n);
        for (size_t g = b.off / G; g < (b.off + FitAlloc::round(b.n)) / G; ++g) owner[g] = 0;
        live[i] = live.back();
        live.pop_back();
        // a double release and a release outside the region are refused
        if (fa.release(b.off, b.n)) return fail("double release accepted", b.off, b.n);
      }
      size_t used = 0;
      for (const auto& b : live) used += FitAlloc::round(b.n);
      if (used != fa.used()) return fail("used", used, fa.used());
    }
    if (fa.release(size, G)) return fail("release past the end accepted", size, G);
    for (const auto& b : live)
      if (!fa.release(b.off, b.n)) return fail("final release", b.off, b.n);
    if (fa.used() != 0 || fa.n_free_blocks() != 1 || fa.largest() != size)
      return fail("not coalesced", fa.n_free_blocks(), fa.largest());
  }
  std::printf("pool_check ok\n");
  return 0;
}
