// Active-slot tiles: per round, the slot ranges of the subhaloes that still
// iterate, cut into tiles of SLOT_TILE slots (one block per tile).  Per-slot
// passes (source marking, partition, compaction) and their prefix sums run over
// these tiles only, so a round costs in proportion to the particles of the
// subhaloes it advances, not to the whole batch.
#pragma once
#include "common.h"
#include "scan.h"
#include "unbind.h"

namespace hbtx
{
#define SLOT_TILE 1024
#define SLOT_THREADS 256
#define SLOT_ITEMS (SLOT_TILE / SLOT_THREADS)

__global__ void k_slot_ntiles(SubArrays S, int32_t *ntile)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  const int n = S.off[s + 1] - S.off[s];
  ntile[s] = (S.mode[s] != MODE_IDLE) ? (n + SLOT_TILE - 1) / SLOT_TILE : 0;
}

// tile = (subhalo, first slot, end slot, index of the subhalo's first tile)
__global__ void k_slot_tiles(SubArrays S, const int32_t *toff, int4 *tiles)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub || S.mode[s] == MODE_IDLE)
    return;
  const int b = S.off[s], e = S.off[s + 1];
  const int t0 = toff[s];
  for (int k = 0; b + k * SLOT_TILE < e; k++)
    tiles[t0 + k] = make_int4(s, b + k * SLOT_TILE, min(e, b + (k + 1) * SLOT_TILE), t0);
}

// Pass 1: per tile sum of the flags (tsum[t] = 0 for t in [ntiles, maxt)).
template <typename T>
__global__ __launch_bounds__(SLOT_THREADS) void k_tile_sum(const int4 *__restrict__ tiles, const int32_t *ntiles_p,
                                                          int32_t maxt, const T *__restrict__ flag,
                                                          T *__restrict__ tsum)
{
  __shared__ T lds[SLOT_THREADS / HBTX_WAVE];
  const int nt = *ntiles_p;
  for (int t = blockIdx.x; t < maxt; t += gridDim.x)
  {
    if (t >= nt)
    {
      if (threadIdx.x == 0)
        tsum[t] = 0;
      continue;
    }
    const int4 tl = tiles[t];
    T v = 0;
    for (int q = tl.y + threadIdx.x; q < tl.z; q += SLOT_THREADS)
      v += flag[q];
    T tot;
    block_excl_scan(v, lds, &tot);
    if (threadIdx.x == 0)
      tsum[t] = tot;
  }
}

// Pass 3 (pass 2 is the exclusive scan of tsum into tpre): per-slot global
// exclusive prefix out[q] = tpre[t] + prefix inside the tile, and per subhalo the
// prefix at its first slot (sub_first) and its total (sub_total).  Thread i owns
// slots 4i..4i+3 of the tile.
template <typename T>
__global__ __launch_bounds__(SLOT_THREADS) void k_tile_down(SubArrays S, const int4 *__restrict__ tiles,
                                                           const int32_t *ntiles_p, const T *__restrict__ flag,
                                                           const T *__restrict__ tpre, T *__restrict__ out,
                                                           T *__restrict__ sub_first, T *__restrict__ sub_total)
{
  __shared__ T lds[SLOT_THREADS / HBTX_WAVE];
  TILE_LOOP(tiles, ntiles_p)
  {
    const int4 tl = tiles[t_];
    const int q0 = tl.y + SLOT_ITEMS * threadIdx.x;
    T v[SLOT_ITEMS];
    T sum = 0;
#pragma unroll
    for (int k = 0; k < SLOT_ITEMS; k++)
    {
      v[k] = (q0 + k < tl.z) ? flag[q0 + k] : (T)0;
      sum += v[k];
    }
    T tot;
    T ex = block_excl_scan(sum, lds, &tot) + tpre[t_];
#pragma unroll
    for (int k = 0; k < SLOT_ITEMS; k++)
    {
      if (q0 + k < tl.z)
        out[q0 + k] = ex;
      ex += v[k];
    }
    if (threadIdx.x == 0)
    {
      const int s = tl.x;
      if (tl.w == t_)
        sub_first[s] = tpre[t_];
      if (tl.z == S.off[s + 1])
        sub_total[s] = tpre[t_] + tot - tpre[tl.w];
    }
  }
}

} // namespace hbtx
