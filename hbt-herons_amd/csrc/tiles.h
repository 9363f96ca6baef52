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

// The flags of one tile, SLOT_ITEMS consecutive slots a thread, and their
// exclusive prefix across the block (pass 2, the exclusive scan of the tile totals
// into tpre, gives the tile's base).
template <typename T>
HBTX_DEV T tile_flags_prefix(const int4 &tl, const T *__restrict__ flag, T v[SLOT_ITEMS], T *lds, T *tot)
{
  const int q0 = tl.y + SLOT_ITEMS * threadIdx.x;
  T sum = 0;
#pragma unroll
  for (int k = 0; k < SLOT_ITEMS; k++)
  {
    v[k] = (q0 + k < tl.z) ? flag[q0 + k] : (T)0;
    sum += v[k];
  }
  return block_excl_scan(sum, lds, tot);
}

// Pass 3 of the source compaction: every flagged slot to its place in the dense
// source list (the tile's base + the prefix inside the tile), and per subhalo the
// position of its first source.
__global__ __launch_bounds__(SLOT_THREADS) void k_scatter_sources(SubArrays S, const int4 *tiles,
                                                                  const int32_t *ntiles_p, const int32_t *flag,
                                                                  const int32_t *tpre, int32_t *src,
                                                                  int32_t *src_sub, int32_t *sub_first)
{
  __shared__ int32_t lds[SLOT_THREADS / HBTX_WAVE];
  TILE_LOOP(tiles, ntiles_p)
  {
    const int4 tl = tiles[t_];
    int32_t v[SLOT_ITEMS], tot;
    int32_t ex = tile_flags_prefix(tl, flag, v, lds, &tot) + tpre[t_];
    const int q0 = tl.y + SLOT_ITEMS * threadIdx.x;
#pragma unroll
    for (int k = 0; k < SLOT_ITEMS; k++)
    {
      if (v[k])
      {
        src[ex] = q0 + k;
        src_sub[ex] = tl.x;
      }
      ex += v[k];
    }
    if (threadIdx.x == 0 && tl.w == t_)
      sub_first[tl.x] = tpre[t_];
  }
}

// RemoveUnboundParticles (:35-44): stable compaction of the bound candidates of
// each subhalo (rank = prefix of the kept flags from the subhalo's first tile);
// removed particles are tagged with the round.  Per subhalo the totals (kept in
// the low word, tracers in the high word) go to sub_total.
__global__ __launch_bounds__(SLOT_THREADS) void k_partition_scatter(SubArrays S, PartArrays Q, const int4 *tiles,
                                                                    const int32_t *ntiles_p, const int32_t *round_p,
                                                                    const int64_t *flag, const int64_t *tpre,
                                                                    int64_t *sub_total)
{
  __shared__ int64_t lds[SLOT_THREADS / HBTX_WAVE];
  const int round = *round_p;
  TILE_LOOP(tiles, ntiles_p)
  {
    const int4 tl = tiles[t_];
    const int s = tl.x;
    const int end = min(tl.z, S.off[s] + S.ntgt[s]);
    const int64_t first = tpre[tl.w];
    int64_t v[SLOT_ITEMS], tot;
    int64_t ex = tile_flags_prefix(tl, flag, v, lds, &tot) + tpre[t_];
    const int q0 = tl.y + SLOT_ITEMS * threadIdx.x;
#pragma unroll
    for (int k = 0; k < SLOT_ITEMS; k++)
    {
      const int q = q0 + k;
      if (q < end)
      {
        const int p = Q.cur[q];
        if (v[k] & 0xffffffffll)
          Q.cur2[S.off[s] + (int)((ex - first) & 0xffffffffll)] = p;
        else
          Q.state[p] = round;
      }
      ex += v[k];
    }
    if (threadIdx.x == 0 && tl.z == S.off[s + 1])
      sub_total[s] = tpre[t_] + tot - first;
  }
}

} // namespace hbtx
