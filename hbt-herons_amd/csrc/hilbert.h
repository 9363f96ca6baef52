// Hilbert-curve index of an octree cell (the warm round's target order).
// Plain C++ over integers, so that the host test driver compiles it as well.
#pragma once
#include <cstdint>
#ifndef HBTX_HD
#define HBTX_HD
#endif

namespace hbtx
{

// Index along the 3-D Hilbert curve of the 1024^3 cell whose 30-bit octree key
// (x, y, z bits interleaved from bit 0, top level first) is m: Skilling's
// transposition of the axes (J. Skilling, "Programming the Hilbert curve", AIP
// Conf. Proc. 707, 381 (2004)), its Gray-code step, then the bits interleaved
// again.
HBTX_HD uint32_t hilbert3_10(uint64_t m)
{
  uint32_t X[3] = {0u, 0u, 0u};
  for (int b = 0; b < 10; b++)
    for (int a = 0; a < 3; a++)
      X[a] |= (uint32_t)((m >> (3 * b + a)) & 1u) << b;
  for (uint32_t Q = 1u << 9; Q > 1u; Q >>= 1)
  {
    const uint32_t P = Q - 1u;
    for (int i = 0; i < 3; i++)
    {
      if (X[i] & Q)
        X[0] ^= P;
      else
      {
        const uint32_t t = (X[0] ^ X[i]) & P;
        X[0] ^= t;
        X[i] ^= t;
      }
    }
  }
  X[1] ^= X[0];
  X[2] ^= X[1];
  uint32_t t = 0u;
  for (uint32_t Q = 1u << 9; Q > 1u; Q >>= 1)
    if (X[2] & Q)
      t ^= Q - 1u;
  uint32_t h = 0u;
  for (int b = 0; b < 10; b++)
    for (int a = 0; a < 3; a++)
      h |= (((X[a] ^ t) >> b) & 1u) << (3 * b + (2 - a));
  return h;
}

} // namespace hbtx
