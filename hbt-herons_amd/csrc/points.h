// Potential at arbitrary points (GravityTree_t::EvaluatePotential on a batch of
// targets, src/gravity_tree.cpp:82-166) and the exact direct sum.
#pragma once
#include "walk.h"

namespace hbtx
{

struct PointArgs
{
  const float4 *tgt;       // target position + self mass (targetMass)
  const int32_t *tgt_order;// walk order of targets (spatially sorted) or null
  int32_t ntgt;
  int32_t eb, ee;          // walk-record unit range
  int32_t treemode;
  const uint2 *wt;         // walk records
  float *phi;              // specific potential, scattered to the target's input index, or null
  double *phi_d;           // optional double output
  int32_t *counter;
  unsigned long long *stats;
  int32_t *overflow;       // set to WALK_STACK_OVERFLOW if the walk stack was full, or null
  DevParams P;
};

template <bool PERIODIC, bool STATS>
__global__ __launch_bounds__(WALK_WG_THREADS, 8) void k_points_potential(PointArgs A)
{
  const int lane = lane_id();
  __shared__ WalkLds wls[HBTX_WALK_WG_WAVES];
  WalkLds &wl = wls[threadIdx.x / HBTX_WAVE];
  unsigned long long st_steps = 0, st_visits = 0, st_accepts = 0, st_pairs = 0, st_near = 0;
  const int nitems = (A.ntgt + HBTX_WAVE - 1) / HBTX_WAVE;
  for (;;)
  {
    int it = 0;
    if (lane == 0)
      it = atomicAdd(A.counter, 1);
    it = uniform(__shfl(it, 0, HBTX_WAVE));
    if (it >= nitems)
      break;
    const int t = it * HBTX_WAVE + lane;
    const bool valid = t < A.ntgt;
    const int ti = valid ? (A.tgt_order ? A.tgt_order[t] : t) : 0;
    float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid)
      x = A.tgt[ti];
    Kahan acc;
    acc.init();
    WalkCounts wc = {0u, 0u, 0u, 0u};
    if (A.ee > A.eb)
    {
      if (A.treemode)
      {
        tree_walk_stack<PERIODIC, STATS>(A.wt, A.eb, A.ee, valid, x.x, x.y, x.z, A.P, acc, wc, wl, A.overflow);
      }
      else
      {
        direct_sum<PERIODIC>(A.wt, A.eb, A.ee, x.x, x.y, x.z, A.P, acc, wc);
        if (STATS)
          st_pairs += (unsigned long long)__popcll(wave_mask(valid)) * (unsigned long long)((A.ee - A.eb) / 2);
      }
    }
    if (STATS)
    {
      st_steps += wc.steps;
      st_visits += wc.visits;
      st_accepts += wc.accepts;
      st_near += wc.nnear;
    }
    if (valid)
    {
      const float self = x.w / A.P.eps;
      const double phi = ((double)self + (double)acc.s - (double)acc.c) * A.P.G_over_a;
      if (A.phi)
        A.phi[ti] = (float)phi;
      if (A.phi_d)
        A.phi_d[ti] = phi;
    }
  }
  if (STATS && A.stats && lane == 0)
  {
    atomicAdd(A.stats + STAT_TREE_STEPS, st_steps);
    atomicAdd(A.stats + STAT_TREE_VISITS, st_visits);
    atomicAdd(A.stats + STAT_TREE_ACCEPT, st_accepts);
    atomicAdd(A.stats + STAT_DIRECT_PAIRS, st_pairs);
    atomicAdd(A.stats + STAT_NEAR, st_near);
  }
}

// keys of external target points in a tree's cube (walk order for coherence)
__global__ void k_point_keys(const float4 *tgt, int32_t n, const double *cube, uint64_t *keys, int32_t *vals)
{
  GRID_STRIDE(i, n)
  {
    const float4 x = tgt[i];
    keys[i] = octree_key(x.x, x.y, x.z, cube[0], cube[1], cube[2], cube[3]);
    vals[i] = (int32_t)i;
  }
}

// Caller-owned device arrays packed into the direct sum's records {x, y, z, m}
// (the layout the tree build emits for a direct-sum source set) and into the
// walk's target points {x, y, z, target mass}.
__global__ void k_pack_records(const float *pos, const float *mass, int32_t n, uint2 *wt)
{
  GRID_STRIDE(i, n)
  {
    wt[2 * i] = make_uint2(__float_as_uint(pos[3 * i]), __float_as_uint(pos[3 * i + 1]));
    wt[2 * i + 1] = make_uint2(__float_as_uint(pos[3 * i + 2]), __float_as_uint(mass[i]));
  }
}

__global__ void k_pack_points(const float *pos, const float *mass, int32_t n, float4 *out)
{
  GRID_STRIDE(i, n)
  out[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], mass ? mass[i] : 0.f);
}

} // namespace hbtx
