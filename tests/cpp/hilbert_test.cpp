// hbtx::hilbert3_10 (hbt-herons_amd/csrc/hilbert.h) over every cell of the
// 1024^3 grid: a bijection onto [0, 2^30), index 0 at the origin cell, and
// consecutive indices on cells that share a face (the defining property of a
// Hilbert curve), checked over the first 2^24 indices (a 256^3 octant).
#include "hilbert.h"

#include <atomic>
#include <cstdio>
#include <cstdlib>
#include <thread>
#include <vector>

static int coord(uint64_t m, int a)
{
  int v = 0;
  for (int b = 0; b < 10; b++)
    v |= (int)((m >> (3 * b + a)) & 1u) << b;
  return v;
}

int main()
{
  const uint64_t n = 1ull << 30, nh = 1ull << 24;
  std::vector<std::atomic<uint64_t>> seen(n / 64);
  for (auto &w : seen)
    w.store(0ull);
  std::vector<uint32_t> inv(nh, 0xffffffffu);
  std::atomic<int> bad(0);
  const int nthr = 8;
  std::vector<std::thread> pool;
  for (int t = 0; t < nthr; t++)
    pool.emplace_back([&, t]() {
      for (uint64_t m = n / nthr * t; m < n / nthr * (t + 1); m++)
      {
        const uint32_t h = hbtx::hilbert3_10(m);
        const uint64_t bit = 1ull << (h & 63);
        if (h >= n || (seen[h >> 6].fetch_or(bit) & bit))
        {
          bad.store(1);
          return;
        }
        if (h < nh)
          inv[h] = (uint32_t)m;
      }
    });
  for (auto &th : pool)
    th.join();
  if (bad.load())
  {
    printf("not a bijection\n");
    return 1;
  }
  if (inv[0] != 0)
  {
    printf("index 0 is not the origin cell\n");
    return 1;
  }
  for (uint64_t h = 1; h < nh; h++)
  {
    int d = 0;
    for (int a = 0; a < 3; a++)
      d += std::abs(coord(inv[h], a) - coord(inv[h - 1], a));
    if (d != 1)
    {
      printf("indices %llu and %llu are not face neighbours\n", (unsigned long long)(h - 1),
             (unsigned long long)h);
      return 1;
    }
  }
  printf("ok\n");
  return 0;
}
