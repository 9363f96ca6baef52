// Segmented sort of (uint64 key, int32 value) pairs.
//  * segments of <= SORT_WAVE_MAX elements: one wave each, bitonic sort in lanes;
//  * segments of <= SORT_SMALL_MAX elements: one workgroup each, the same bitonic
//    network in LDS;
//  * larger segments: tiled LSD radix sort (8- or 9-bit digits, whichever needs
//    fewer passes) where every tile of SORT_TILE elements ranks its keys stably
//    with wave ballots (match of the digit bits) and LDS per-digit counters, so
//    all large segments advance together in one launch per pass.
#pragma once
#include "common.h"
#include "scan.h"

namespace hbtx
{
#ifndef SORT_THREADS
#define SORT_THREADS 256
#endif
#define SORT_SMALL_MAX 2048
#define SORT_WAVE_MAX HBTX_WAVE // segments of at most one wave: k_sort_wave
#define SORT_WAVE4_MAX (4 * HBTX_WAVE) // segments of at most four per lane: k_sort_wave4
#ifndef SORT_TILE
#define SORT_TILE 2048
#endif
#define RADIX_MAX_BINS 512 // (9-bit digits)

__global__ __launch_bounds__(SORT_THREADS) void k_sort_small(uint64_t *__restrict__ keys, int32_t *__restrict__ vals,
                                                            const int32_t *__restrict__ seg_begin,
                                                            const int32_t *__restrict__ seg_len, int32_t nseg)
{
  __shared__ uint64_t sk[SORT_SMALL_MAX];
  __shared__ int32_t sv[SORT_SMALL_MAX];
  const int tid = threadIdx.x;
  for (int s = blockIdx.x; s < nseg; s += gridDim.x)
  {
    const int len = seg_len[s];
    if (len <= SORT_WAVE4_MAX || len > SORT_SMALL_MAX)
      continue; // (segments of up to four per lane: k_sort_wave, k_sort_wave4)
    const int b = seg_begin[s];
    int P = 1;
    while (P < len)
      P <<= 1;
    for (int i = tid; i < P; i += SORT_THREADS)
    {
      sk[i] = (i < len) ? keys[b + i] : ~0ull;
      sv[i] = (i < len) ? vals[b + i] : 0;
    }
    __syncthreads();
    for (int k = 2; k <= P; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1)
      {
        for (int i = tid; i < P; i += SORT_THREADS)
        {
          const int ixj = i ^ j;
          if (ixj > i)
          {
            const bool up = (i & k) == 0;
            const uint64_t a = sk[i], c = sk[ixj];
            if ((a > c) == up)
            {
              sk[i] = c;
              sk[ixj] = a;
              const int32_t t = sv[i];
              sv[i] = sv[ixj];
              sv[ixj] = t;
            }
          }
        }
        __syncthreads();
      }
    for (int i = tid; i < len; i += SORT_THREADS)
    {
      keys[b + i] = sk[i];
      vals[b + i] = sv[i];
    }
    __syncthreads();
  }
}

// Segments of at most one wave (the common case for small subhaloes): one wave per
// segment, element i in lane i, the same bitonic network as k_sort_small (so the
// same permutation, ties included) through lane shuffles, with no barrier.
__global__ __launch_bounds__(SORT_THREADS) void k_sort_wave(uint64_t *__restrict__ keys, int32_t *__restrict__ vals,
                                                           const int32_t *__restrict__ seg_begin,
                                                           const int32_t *__restrict__ seg_len, int32_t nseg)
{
  const int lane = lane_id();
  const int wpb = SORT_THREADS / HBTX_WAVE;
  for (int s = uniform(blockIdx.x * wpb + (int)(threadIdx.x / HBTX_WAVE)); s < nseg; s += gridDim.x * wpb)
  {
    const int len = seg_len[s];
    if (len <= 1 || len > SORT_WAVE_MAX)
      continue;
    const int b = seg_begin[s];
    int P = 1;
    while (P < len)
      P <<= 1;
    uint64_t key = (lane < len) ? keys[b + lane] : ~0ull;
    int32_t val = (lane < len) ? vals[b + lane] : 0;
    for (int k = 2; k <= P; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1)
      {
        const uint64_t ok = __shfl_xor(key, j, HBTX_WAVE);
        const int32_t ov = __shfl_xor(val, j, HBTX_WAVE);
        const bool low = (lane & j) == 0;     // this lane holds the pair's first element
        const bool up = (lane & k) == 0;      // the pair's direction (the same for both lanes)
        const uint64_t a = low ? key : ok, c = low ? ok : key;
        const bool swap = (a > c) == up;
        if (swap)
        {
          key = ok;
          val = ov;
        }
      }
    if (lane < len)
    {
      keys[b + lane] = key;
      vals[b + lane] = val;
    }
  }
}

// Segments of SORT_WAVE_MAX + 1 .. SORT_WAVE4_MAX elements: one wave each, element
// lane + 64 r in register r, padded to 128 or 256 like k_sort_small, and the same
// bitonic network (so the same permutation, ties included): pairs less than 64
// apart meet through lane shuffles, pairs 64 or more apart in a lane's registers.
// No LDS and no barrier.
__global__ __launch_bounds__(SORT_THREADS) void k_sort_wave4(uint64_t *__restrict__ keys, int32_t *__restrict__ vals,
                                                            const int32_t *__restrict__ seg_begin,
                                                            const int32_t *__restrict__ seg_len, int32_t nseg)
{
  const int lane = lane_id();
  const int wpb = SORT_THREADS / HBTX_WAVE;
  for (int s = uniform(blockIdx.x * wpb + (int)(threadIdx.x / HBTX_WAVE)); s < nseg; s += gridDim.x * wpb)
  {
    const int len = seg_len[s];
    if (len <= SORT_WAVE_MAX || len > SORT_WAVE4_MAX)
      continue;
    const int b = seg_begin[s];
    const int P = (len > 2 * HBTX_WAVE) ? 4 * HBTX_WAVE : 2 * HBTX_WAVE;
    const int R = P / HBTX_WAVE; // registers in use: 2 or 4
    uint64_t key[4];
    int32_t val[4];
#pragma unroll
    for (int r = 0; r < 4; r++)
    {
      const int e = lane + HBTX_WAVE * r;
      key[r] = (r < R && e < len) ? keys[b + e] : ~0ull;
      val[r] = (r < R && e < len) ? vals[b + e] : 0;
    }
    for (int k = 2; k <= P; k <<= 1)
      for (int j = k >> 1; j > 0; j >>= 1)
      {
        if (j < HBTX_WAVE)
        {
#pragma unroll
          for (int r = 0; r < 4; r++)
          {
            if (r >= R)
              continue;
            const int e = lane + HBTX_WAVE * r;
            const uint64_t ok = __shfl_xor(key[r], j, HBTX_WAVE);
            const int32_t ov = __shfl_xor(val[r], j, HBTX_WAVE);
            const bool low = (lane & j) == 0; // this lane holds the pair's first element
            const bool up = (e & k) == 0;     // the pair's direction (the same for both lanes)
            const uint64_t a = low ? key[r] : ok, c = low ? ok : key[r];
            if ((a > c) == up)
            {
              key[r] = ok;
              val[r] = ov;
            }
          }
        }
        else
        {
          const int jr = j / HBTX_WAVE; // the pair's registers differ by jr (1 or 2)
#pragma unroll
          for (int r = 0; r < 4; r++)
          {
            const int r2 = r ^ jr;
            if (r >= R || r2 < r)
              continue; // (r2 < R: pairs stay inside the padded length)
            const int e = lane + HBTX_WAVE * r;
            const bool up = (e & k) == 0;
            if ((key[r] > key[r2]) == up)
            {
              const uint64_t tk = key[r];
              key[r] = key[r2];
              key[r2] = tk;
              const int32_t tv = val[r];
              val[r] = val[r2];
              val[r2] = tv;
            }
          }
        }
      }
#pragma unroll
    for (int r = 0; r < 4; r++)
    {
      const int e = lane + HBTX_WAVE * r;
      if (r < R && e < len)
      {
        keys[b + e] = key[r];
        vals[b + e] = val[r];
      }
    }
  }
}

// Number of radix tiles of every segment (0 for segments sorted in LDS).
__global__ void k_sort_ntiles(const int32_t *__restrict__ seg_len, int32_t nseg, int32_t *__restrict__ ntile)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nseg)
  {
    const int len = seg_len[s];
    ntile[s] = (len > SORT_SMALL_MAX) ? (len + SORT_TILE - 1) / SORT_TILE : 0;
  }
}

// tiles[t] = (segment, begin, end, 0)
__global__ void k_sort_tiles(const int32_t *__restrict__ seg_begin, const int32_t *__restrict__ seg_len,
                             const int32_t *__restrict__ tile_off, int32_t nseg, int4 *__restrict__ tiles)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nseg)
    return;
  const int len = seg_len[s];
  if (len <= SORT_SMALL_MAX)
    return;
  const int b = seg_begin[s], e = b + len;
  int t = tile_off[s];
  for (int x = b; x < e; x += SORT_TILE, t++)
    tiles[t] = make_int4(s, x, min(e, x + SORT_TILE), 0);
}

template <int RB>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_hist(const uint64_t *__restrict__ kin,
                                                            const int4 *__restrict__ tiles,
                                                            const int32_t *__restrict__ ntiles_p, int shift,
                                                            int32_t *__restrict__ ghist)
{
  constexpr int BINS = 1 << RB;
  __shared__ int32_t hist[BINS];
  const int nt = *ntiles_p;
  const int tid = threadIdx.x;
  for (int t = blockIdx.x; t < nt; t += gridDim.x)
  {
    for (int b = tid; b < BINS; b += SORT_THREADS)
      hist[b] = 0;
    __syncthreads();
    const int4 tl = tiles[t];
    for (int i = tl.y + tid; i < tl.z; i += SORT_THREADS)
      atomicAdd(&hist[(int)((kin[i] >> shift) & (BINS - 1))], 1);
    __syncthreads();
    for (int b = tid; b < BINS; b += SORT_THREADS)
      ghist[(int64_t)t * BINS + b] = hist[b];
    __syncthreads();
  }
}

// Tile histograms -> global scatter bases, in two steps so that no block walks a
// large segment's whole column of tile counts:
//  * k_radix_chunk_scan: every chunk of SORT_CHUNK consecutive tiles of a segment
//    turns its counts into exclusive prefixes within the chunk (in place) and
//    stores the chunk's totals at its head tile in gchunk;
//  * k_radix_seg_scan: per large segment, the chunk totals become exclusive
//    prefixes over the segment's chunks plus the digit's base in the segment
//    (segment start + the counts of the smaller digits).
// A tile's base for digit d is then ghist[t][d] + gchunk[head(t)][d].  Both are
// driven by the radix tiles, so launches with no large segment cost nothing;
// thread i owns digits [PER i, PER i + PER).
#define SORT_CHUNK 16
template <int RB>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_chunk_scan(const int4 *__restrict__ tiles,
                                                                  const int32_t *__restrict__ ntiles_p,
                                                                  const int32_t *__restrict__ tile_off,
                                                                  const int32_t *__restrict__ ntile,
                                                                  int32_t *__restrict__ ghist,
                                                                  int32_t *__restrict__ gchunk)
{
  constexpr int BINS = 1 << RB, PER = BINS / SORT_THREADS;
  const int d0 = threadIdx.x * PER;
  const int nt_all = *ntiles_p;
  for (int t = blockIdx.x; t < nt_all; t += gridDim.x)
  {
    const int s = tiles[t].x;
    const int t0 = tile_off[s], te = t0 + ntile[s];
    if ((t - t0) % SORT_CHUNK != 0)
      continue; // not a chunk's head tile
    const int ce = min(te, t + SORT_CHUNK);
#pragma unroll
    for (int j = 0; j < PER; j++)
    {
      const int d = d0 + j;
      int c[SORT_CHUNK];
#pragma unroll
      for (int k = 0; k < SORT_CHUNK; k++)
        c[k] = (t + k < ce) ? ghist[(int64_t)(t + k) * BINS + d] : 0;
      int run = 0;
#pragma unroll
      for (int k = 0; k < SORT_CHUNK; k++)
      {
        if (t + k < ce)
          ghist[(int64_t)(t + k) * BINS + d] = run;
        run += c[k];
      }
      gchunk[(int64_t)t * BINS + d] = run;
    }
  }
}

template <int RB>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_seg_scan(const int32_t *__restrict__ seg_begin,
                                                                const int4 *__restrict__ tiles,
                                                                const int32_t *__restrict__ ntiles_p,
                                                                const int32_t *__restrict__ tile_off,
                                                                const int32_t *__restrict__ ntile,
                                                                int32_t *__restrict__ gchunk)
{
  constexpr int BINS = 1 << RB, PER = BINS / SORT_THREADS;
  __shared__ int32_t lds[SORT_THREADS / HBTX_WAVE];
  const int d0 = threadIdx.x * PER;
  const int nt_all = *ntiles_p;
  for (int t = blockIdx.x; t < nt_all; t += gridDim.x)
  {
    const int4 tl = tiles[t];
    const int s = tl.x;
    if (tl.y != seg_begin[s])
      continue; // not the segment's first tile
    const int t0 = tile_off[s], te = t0 + ntile[s];
    int run[PER];
#pragma unroll
    for (int j = 0; j < PER; j++)
    {
      const int d = d0 + j;
      run[j] = 0;
      for (int u = t0; u < te; u += SORT_CHUNK)
      {
        const int c = gchunk[(int64_t)u * BINS + d];
        gchunk[(int64_t)u * BINS + d] = run[j];
        run[j] += c;
      }
    }
    int mine = 0;
#pragma unroll
    for (int j = 0; j < PER; j++)
      mine += run[j];
    int tot;
    int base = block_excl_scan(mine, lds, &tot) + seg_begin[s];
#pragma unroll
    for (int j = 0; j < PER; j++)
    {
      const int d = d0 + j;
      for (int u = t0; u < te; u += SORT_CHUNK)
        gchunk[(int64_t)u * BINS + d] += base;
      base += run[j];
    }
    __syncthreads();
  }
}

// One pass of the tiled LSD radix sort.  Wave w ranks elements [1024 w, 1024 w +
// 1024) of the tile in order (ballot match of the digit bits, wave-private digit
// counters in LDS), the tile is then reordered by digit in LDS, and threads write
// it out in that order, so consecutive threads store to consecutive addresses of
// each digit's run.
#define SCAT_ITEMS (SORT_TILE / SORT_THREADS)
template <int RB>
__global__ __launch_bounds__(SORT_THREADS) void k_radix_scatter(const uint64_t *__restrict__ kin,
                                                               const int32_t *__restrict__ vin,
                                                               uint64_t *__restrict__ kout,
                                                               int32_t *__restrict__ vout,
                                                               const int4 *__restrict__ tiles,
                                                               const int32_t *__restrict__ ntiles_p, int shift,
                                                               const int32_t *__restrict__ ghist,
                                                               const int32_t *__restrict__ tile_off,
                                                               const int32_t *__restrict__ gchunk)
{
  constexpr int BINS = 1 << RB, PER = BINS / SORT_THREADS;
  __shared__ uint64_t sk[SORT_TILE];
  __shared__ int32_t sv[SORT_TILE];
  __shared__ int32_t wcnt[SORT_THREADS / HBTX_WAVE][BINS];
  __shared__ int32_t dbase[BINS];
  __shared__ int32_t scan_lds[SORT_THREADS / HBTX_WAVE];
  const int nt = *ntiles_p;
  const int tid = threadIdx.x;
  const int w = tid / HBTX_WAVE;
  const int lane = lane_id();
  const int per_wave = SCAT_ITEMS * HBTX_WAVE;
  // Workgroups are dispatched round-robin over the 8 XCDs.  With a grid that is a
  // multiple of 8, XCD x takes the x-th eighth of the tiles, so the short digit
  // runs that neighbouring tiles write next to each other meet in one L2 instead
  // of leaving partial lines in eight of them.
  const bool by_xcd = (gridDim.x & 7) == 0;
  const int xcd = by_xcd ? (int)(blockIdx.x & 7) : 0;
  const int step = by_xcd ? (int)(gridDim.x >> 3) : (int)gridDim.x;
  const int per_xcd = by_xcd ? (nt + 7) / 8 : nt;
  const int t_lo = xcd * per_xcd, t_hi = min(nt, t_lo + per_xcd);
  for (int t = t_lo + (by_xcd ? (int)(blockIdx.x >> 3) : (int)blockIdx.x); t < t_hi; t += step)
  {
    const int4 tl = tiles[t];
    const int n = tl.z - tl.y;
#pragma unroll
    for (int k = 0; k < SORT_THREADS / HBTX_WAVE; k++)
#pragma unroll
      for (int j = 0; j < PER; j++)
        wcnt[k][tid * PER + j] = 0;
    __syncthreads();
    uint64_t key[SCAT_ITEMS];
    int32_t val[SCAT_ITEMS];
    int32_t rank[SCAT_ITEMS];
#pragma unroll
    for (int r = 0; r < SCAT_ITEMS; r++)
    {
      const int li = w * per_wave + r * HBTX_WAVE + lane;
      const bool valid = li < n;
      key[r] = valid ? kin[tl.y + li] : 0ull;
      val[r] = valid ? vin[tl.y + li] : 0;
      const int d = (int)((key[r] >> shift) & (BINS - 1));
      uint64_t m = ballot(valid);
#pragma unroll
      for (int bit = 0; bit < RB; bit++)
      {
        const bool on = (d >> bit) & 1;
        const uint64_t bb = ballot(valid && on);
        m &= on ? bb : ~bb;
      }
      const int base = valid ? wcnt[w][d] : 0;
      rank[r] = base + __popcll(m & lanemask_lt());
      // the highest lane of each digit group advances the wave's counter
      if (valid && (m >> lane) == 1ull)
        wcnt[w][d] = base + __popcll(m);
    }
    __syncthreads();
    {
      // digits PER tid ..: offsets of the waves' runs and the first slot of the digit in the tile
      int tot[PER];
      int mine = 0;
#pragma unroll
      for (int j = 0; j < PER; j++)
      {
        const int d = tid * PER + j;
        tot[j] = 0;
#pragma unroll
        for (int k = 0; k < SORT_THREADS / HBTX_WAVE; k++)
        {
          const int c = wcnt[k][d];
          wcnt[k][d] = tot[j];
          tot[j] += c;
        }
        mine += tot[j];
      }
      int all;
      int b = block_excl_scan(mine, scan_lds, &all);
#pragma unroll
      for (int j = 0; j < PER; j++)
      {
        dbase[tid * PER + j] = b;
        b += tot[j];
      }
    }
    __syncthreads();
#pragma unroll
    for (int r = 0; r < SCAT_ITEMS; r++)
    {
      const int li = w * per_wave + r * HBTX_WAVE + lane;
      if (li < n)
      {
        const int d = (int)((key[r] >> shift) & (BINS - 1));
        const int pos = dbase[d] + wcnt[w][d] + rank[r];
        sk[pos] = key[r];
        sv[pos] = val[r];
      }
    }
    __syncthreads();
    const int32_t *off = ghist + (int64_t)t * BINS;
    const int t0 = tile_off[tl.x];
    const int32_t *coff = gchunk + (int64_t)(t0 + ((t - t0) / SORT_CHUNK) * SORT_CHUNK) * BINS;
    for (int i = tid; i < n; i += SORT_THREADS)
    {
      const uint64_t k = sk[i];
      const int d = (int)((k >> shift) & (BINS - 1));
      const int dst = off[d] + coff[d] + (i - dbase[d]);
      kout[dst] = k;
      vout[dst] = sv[i];
    }
    __syncthreads();
  }
}

} // namespace hbtx
