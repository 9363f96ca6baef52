// Global exclusive prefix sum (three-pass: block reduce, scan of block sums,
// block scan + offset).  out[i] = sum(in[0..i)), out[n] = total.  In-place safe.
#pragma once
#include "common.h"

namespace hbtx
{
#define SCAN_THREADS 256
#define SCAN_ITEMS 16
#define SCAN_TILE (SCAN_THREADS * SCAN_ITEMS)

template <typename T>
HBTX_DEV T wave_incl_scan(T v)
{
  const int l = lane_id();
#pragma unroll
  for (int d = 1; d < HBTX_WAVE; d <<= 1)
  {
    T o = __shfl_up(v, d, HBTX_WAVE);
    if (l >= d)
      v += o;
  }
  return v;
}

// Exclusive scan of one value per thread across a 256-thread block; returns the
// thread's exclusive prefix and writes the block total to *total.
template <typename T>
HBTX_DEV T block_excl_scan(T v, T *lds_wave, T *total)
{
  const int tid = threadIdx.x;
  const int w = tid / HBTX_WAVE;
  T incl = wave_incl_scan(v);
  if (lane_id() == HBTX_WAVE - 1)
    lds_wave[w] = incl;
  __syncthreads();
  T wbase = 0, tot = 0;
#pragma unroll
  for (int k = 0; k < SCAN_THREADS / HBTX_WAVE; k++)
  {
    T x = lds_wave[k];
    if (k < w)
      wbase += x;
    tot += x;
  }
  __syncthreads();
  *total = tot;
  return wbase + incl - v;
}

// Per-tile totals with coalesced loads (thread t reads elements t, t+256, ...).
// n_dev (may be null): the length is min(n, *n_dev), read on the device.
template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_reduce(const T *__restrict__ in, int64_t n,
                                                             const int32_t *n_dev, T *__restrict__ partials)
{
  __shared__ T lds[SCAN_THREADS / HBTX_WAVE];
  if (n_dev)
    n = min(n, (int64_t)*n_dev);
  if ((int64_t)blockIdx.x * SCAN_TILE >= n)
  {
    if (threadIdx.x == 0)
      partials[blockIdx.x] = 0;
    return;
  }
  const int64_t base = (int64_t)blockIdx.x * SCAN_TILE + threadIdx.x;
  T s = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++)
  {
    const int64_t i = base + (int64_t)k * SCAN_THREADS;
    if (i < n)
      s += in[i];
  }
  T tot;
  block_excl_scan(s, lds, &tot);
  if (threadIdx.x == 0)
    partials[blockIdx.x] = tot;
}

// Single block: exclusive scan of partials[0..nb) in place.
template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_partials(T *__restrict__ partials, int64_t nb)
{
  __shared__ T lds[SCAN_THREADS / HBTX_WAVE];
  T carry = 0;
  for (int64_t c = 0; c < nb; c += SCAN_TILE)
  {
    const int64_t base = c + (int64_t)threadIdx.x * SCAN_ITEMS;
    T v[SCAN_ITEMS];
    T s = 0;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
    {
      int64_t i = base + k;
      v[k] = (i < nb) ? partials[i] : (T)0;
      s += v[k];
    }
    T tot;
    T ex = block_excl_scan(s, lds, &tot) + carry;
#pragma unroll
    for (int k = 0; k < SCAN_ITEMS; k++)
    {
      int64_t i = base + k;
      if (i < nb)
        partials[i] = ex;
      ex += v[k];
    }
    carry += tot;
    __syncthreads();
  }
}

// Tile scan: each wave owns 1024 consecutive elements, read in 16 coalesced
// passes of 64 and scanned in registers; wave totals are combined through LDS.
template <typename T>
__global__ __launch_bounds__(SCAN_THREADS) void k_scan_down(const T *in, T *out, int64_t n, const int32_t *n_dev,
                                                           const T *__restrict__ partials)
{
  __shared__ T lds[SCAN_THREADS / HBTX_WAVE];
  if (n_dev)
    n = min(n, (int64_t)*n_dev);
  const int64_t bb = (int64_t)blockIdx.x * SCAN_TILE;
  // the block holding index n (or the last block) writes the grand total at out[n]
  const bool writes_total = n >= bb && (n < bb + SCAN_TILE || blockIdx.x == gridDim.x - 1);
  if (bb >= n && !writes_total)
    return;
  const int w = threadIdx.x / HBTX_WAVE;
  const int lane = lane_id();
  const int64_t wbase = (int64_t)blockIdx.x * SCAN_TILE + (int64_t)w * (HBTX_WAVE * SCAN_ITEMS);
  T v[SCAN_ITEMS];
  T run = 0;
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++)
  {
    const int64_t i = wbase + (int64_t)k * HBTX_WAVE + lane;
    const T x = (i < n) ? in[i] : (T)0;
    const T incl = wave_incl_scan(x);
    v[k] = run + incl - x; // exclusive prefix inside the wave's span
    run += __shfl(incl, HBTX_WAVE - 1, HBTX_WAVE);
  }
  if (lane == 0)
    lds[w] = run;
  __syncthreads();
  T off = partials[blockIdx.x];
  for (int k = 0; k < w; k++)
    off += lds[k];
#pragma unroll
  for (int k = 0; k < SCAN_ITEMS; k++)
  {
    const int64_t i = wbase + (int64_t)k * HBTX_WAVE + lane;
    if (i < n)
      out[i] = off + v[k];
  }
  if (writes_total && threadIdx.x == 0)
  {
    T tot = partials[blockIdx.x];
    for (int k = 0; k < SCAN_THREADS / HBTX_WAVE; k++)
      tot += lds[k];
    out[n] = tot;
  }
}

} // namespace hbtx
