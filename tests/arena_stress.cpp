// Host-thread stress of the per-net arena allocator (arena_take in
// hvm-64_amd/csrc/hvm64_rules.h, built for the host): threads take chunks of
// random sizes from the arenas of a few nets, which grow by extents from a
// pool, with an overflow region for takes that arrive while an extent is
// being installed.  Checks that no two chunks overlap, that every chunk
// outside the overflow region lies in granules the granule map gives to the
// chunk's net, that the pool and the overflow region stay within their
// bounds, and that every take is served.  Prints one summary line; exits 1
// on any violation.
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

#include "hvm64_rules.h"

using namespace hvm;

int main(int argc, char **argv) {
  const u32 threads = argc > 1 ? (u32)atoi(argv[1]) : 8;
  const u32 iters = argc > 2 ? (u32)atoi(argv[2]) : 100000;
  const u32 nets = 5, gs = 6;  // granules of 64 nodes: many installs
  const u64 gmap_len = 1 << 22;
  std::vector<u32> gmap(gmap_len, NO_NET);
  std::vector<u64> word(nets), base0(nets);
  std::vector<u32> len(nets);
  u64 cur = 1;  // granule 0: the fixed cells
  for (u32 i = 0; i < nets; i++) {
    base0[i] = cur;
    len[i] = 1 + 7 * i;  // initial extents of 1 to 29 granules
    word[i] = cur << 32;
    for (u32 k = 0; k < len[i]; k++) gmap[cur + k] = i;
    cur += len[i];
  }
  const u64 pool_start = cur, pool_end = gmap_len - (gmap_len >> 4);
  u64 pool_next = pool_start, ov_next = pool_end << gs;
  const u64 ov_start = ov_next, ov_end = gmap_len << gs;
  const u64 max_refill = 64;
  std::vector<std::vector<ArenaChunk>> got(threads);
  std::vector<std::vector<u32>> got_net(threads);
  std::vector<u32> errors(threads, 0);
  std::vector<std::thread> pool;
  for (u32 t = 0; t < threads; t++)
    pool.emplace_back([&, t] {
      u64 x = 0x9E3779B97F4A7C15ull * (t + 1);
      for (u32 i = 0; i < iters; i++) {
        x ^= x << 13;
        x ^= x >> 7;
        x ^= x << 17;
        const u32 net = (u32)(x % nets), n = 1 + (u32)((x >> 8) % 64);
        const ArenaChunk c = arena_take(&word[net], base0[net], len[net], max_refill, gs, &pool_next, pool_end,
                                        gmap.data(), net, &ov_next, ov_end, n);
        if (c.err || c.end - c.pos != n) {
          errors[t]++;
          continue;
        }
        got[t].push_back(c);
        got_net[t].push_back(net);
      }
    });
  for (auto &th : pool) th.join();
  std::vector<std::pair<u64, u64>> all;
  u64 wrong_net = 0, outside = 0, nodes = 0, overflow_nodes = 0;
  for (u32 t = 0; t < threads; t++)
    for (size_t i = 0; i < got[t].size(); i++) {
      const ArenaChunk &c = got[t][i];
      all.push_back({c.pos, c.end});
      nodes += c.end - c.pos;
      if (c.pos >= ov_start) {
        overflow_nodes += c.end - c.pos;
        if (c.end > ov_end) outside++;
        continue;
      }
      if (c.pos < (pool_start << gs) && c.pos < (1ull << gs)) outside++;  // the fixed cells' granule
      for (u64 nd = c.pos; nd < c.end; nd++)
        if (gmap[nd >> gs] != got_net[t][i]) {
          wrong_net++;
          break;
        }
    }
  std::sort(all.begin(), all.end());
  u64 overlaps = 0;
  for (size_t i = 1; i < all.size(); i++)
    if (all[i].first < all[i - 1].second) overlaps++;
  u64 errs = 0;
  for (u32 e : errors) errs += e;
  if (pool_next > pool_end) outside++;
  printf("takes %zu nodes %llu overlaps %llu wrong_net %llu outside %llu errors %llu pool_granules %llu overflow_nodes %llu\n",
         all.size(), (unsigned long long)nodes, (unsigned long long)overlaps, (unsigned long long)wrong_net,
         (unsigned long long)outside, (unsigned long long)errs, (unsigned long long)(pool_next - pool_start),
         (unsigned long long)overflow_nodes);
  return overlaps || wrong_net || outside || errs ? 1 : 0;
}
