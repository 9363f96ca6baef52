// Potential + binding-energy kernel (one wave = 64 targets of one subhalo).
//
// Tree mode: wavefront-cooperative walk of the depth-first walk records (16-byte
// particles, 24-byte cells).  The wave visits records in preorder at a
// wave-uniform byte offset (scalar loads).  Each
// lane applies the reference's own opening test to its own target
// (open if len^2 > r^2 theta^2 with r to the cell's centre of mass,
// src/gravity_tree.cpp:139, stored as r^2 < crit2 = len^2/theta^2) and keeps a
// private "skip" bound: after accepting an entry it ignores that entry's
// subtree.  The wave descends (index+1) if any active lane opened the entry,
// otherwise it jumps to the entry's skip.  Every lane therefore sees exactly the
// entries its own stackless walk would see, in the same order, while lanes on
// spatially adjacent targets share the traversal.
//
// Direct mode: the entry list holds particles only and every pair is summed.
//
// The fused epilogue forms the binding energy of the full iteration
// (src/subhalo_unbind.cpp:602-616, src/gravity_tree.cpp:168-178) or the
// correction update (src/subhalo_unbind.cpp:577-589).
#pragma once
#include "common.h"

namespace hbtx
{

enum
{
  STAT_TREE_STEPS = 0,    // wave steps of tree walks
  STAT_TREE_VISITS = 1,   // active lane-entry visits (opening test evaluated)
  STAT_TREE_ACCEPT = 2,   // accepted lane interactions (monopole or particle)
  STAT_DIRECT_PAIRS = 3,  // lane-pair evaluations in direct mode
  STAT_TARGETS = 4,       // targets processed
  STAT_ITEMS = 5,         // work items processed
  STAT_NEAR = 6,          // wave steps that evaluated the softening kernel
  NSTATS = 16
};

struct WalkArgs
{
  const int2 *items;        // (subhalo, first target) per work item
  const int32_t *nitems;    // number of items (device)
  int32_t *counter;         // work-stealing counters, one per item chunk, 16 ints apart (zeroed before launch)
  int32_t nchunks;          // item list split into this many contiguous chunks (1 = one shared counter)
  const int32_t *cur;       // slot-layout target order
  const int32_t *off;       // per subhalo slot offset
  const int32_t *ntgt;
  const int32_t *mode;
  const int32_t *treemode;
  const int32_t *ent_begin; // per subhalo: first walk-record unit of its tree
  const int32_t *ent_end;   // per subhalo: end unit
  const uint2 *wt;          // walk records (8-byte units, two spare units past the end)
  const float4 *posm;
  const float4 *velu;
  const float *selfm;       // target self mass (sources of a full iteration), else 0
  const float *ref;         // per subhalo frame [6]
  const float *oldref;      // per subhalo previous frame [6]
  float *energy;
  float *esl;               // the same energies by slot (off[s] + target index)
  float *eref;              // refinement energies (MODE_REFINE targets)
  float *phi_out;           // optional: specific potential per particle (float), may be null
  unsigned long long *stats;
  DevParams P;
};

HBTX_DEV uint64_t wave_mask(bool p)
{
  return __builtin_amdgcn_ballot_w64(p);
}

template <bool PERIODIC>
HBTX_DEV float sep2(float ex, float ey, float ez, float px, float py, float pz, const DevParams &P)
{
  float dx, dy, dz;
  if (PERIODIC)
  {
    dx = nearest_diff_f(ex, px, P.box, P.box_half);
    dy = nearest_diff_f(ey, py, P.box, P.box_half);
    dz = nearest_diff_f(ez, pz, P.box, P.box_half);
  }
  else
  {
    dx = ex - px;
    dy = ey - py;
    dz = ez - pz;
  }
  return dx * dx + dy * dy + dz * dz;
}

// Per-wave work counters of one walk (uniform values, kept only in the stats
// instantiation; the walk is deterministic, so a stats run counts exactly the
// work of a timed run).
struct WalkCounts
{
  uint32_t steps, visits, accepts, nnear;
};

// Walk the records of units [ub, ue) for this lane's target, accumulating the
// pair terms (-m/r outside h, the spline inside): plain FP32 sums over blocks of
// 16 steps, Kahan-compensated across blocks.  The wave position is a byte offset
// from the tree's first record (an SGPR, the scalar loads' offset); cells carry
// their skip in the same units.
template <bool PERIODIC, bool STATS>
HBTX_DEV void tree_walk(const uint2 *wt, int ub, int ue, bool valid, float px, float py, float pz,
                        const DevParams &P, Kahan &acc, WalkCounts &wc)
{
  const uint2 *tree = wt + ub;
  const int32_t span = 8 * (ue - ub);
  int32_t off = 0;
  int32_t myskip = valid ? 0 : 0x7fffffff;
  // (offsets stay below 2^31: k_sub_entries leaves a tree over WALK_MAX_SPAN_UNITS units empty
  // and reports it)
  for (;;)
  {
    float blk = 0.f;
#pragma unroll
    for (int k = 0; k < 16; k++)
    {
      const WRec e = const_load_wrec(tree, (uint32_t)off);
      const bool active = off >= myskip;
      const float r2 = sep2<PERIODIC>(e.x, e.y, e.z, px, py, pz, P);
      const bool open = r2 < e.crit2;
      const bool accept = active && !open;
      // lane masks straight from the compares (no round trip through a VGPR)
      const bool anyopen = (__builtin_amdgcn_sicmp(off, myskip, 39 /* signed >= */) &
                            __builtin_amdgcn_fcmpf(r2, e.crit2, 4 /* ordered < */)) != 0ull;
      const float rinv = __builtin_amdgcn_rsqf(r2);
      float term = -e.m * rinv;
      if (wave_mask(accept && r2 < P.h2))
      {
        if (STATS)
          wc.nnear++;
        const float tn = spline_near(r2, rinv, e.m, P.hinv);
        term = (r2 < P.h2) ? tn : term;
      }
      blk += accept ? term : 0.f;
      myskip = accept ? e.rskip : myskip;
      if (STATS)
      {
        wc.steps++;
        wc.visits += __popcll(wave_mask(active));
        wc.accepts += __popcll(wave_mask(accept));
      }
      // records of a subtree follow it, so rskip > off; max() keeps the walk finite regardless
      // (only a cell is ever opened: the next record then follows its 24 bytes)
      const int32_t next = off + (e.cell ? 24 : 16);
      off = uniform(anyopen ? next : max(e.rskip, next));
      if (off >= span)
      {
        acc.add(blk);
        return;
      }
    }
    acc.add(blk);
  }
}

template <bool PERIODIC>
HBTX_DEV float direct_term(const uint2 *wt, int i, float px, float py, float pz, const DevParams &P, bool &near)
{
  const float4 e = const_load_f4u(wt + 2 * i);
  const float r2 = sep2<PERIODIC>(e.x, e.y, e.z, px, py, pz, P);
  const float rinv = __builtin_amdgcn_rsqf(r2);
  float term = -e.w * rinv;
  const bool nr = r2 < P.h2;
  if (wave_mask(nr))
  {
    near = true;
    const float tn = spline_near(r2, rinv, e.w, P.hinv);
    term = nr ? tn : term;
  }
  return term;
}

// Direct mode: the records are particles only (two units each).
template <bool PERIODIC>
HBTX_DEV void direct_sum(const uint2 *wt, int ub, int ue, float px, float py, float pz, const DevParams &P,
                         Kahan &acc, WalkCounts &wc)
{
  const uint2 *ent = wt + ub;
  const int ee = (ue - ub) / 2;
  int i = 0;
  for (; i + 4 <= ee; i += 4)
  {
    bool nr = false;
    float part = direct_term<PERIODIC>(ent, i, px, py, pz, P, nr);
    part += direct_term<PERIODIC>(ent, i + 1, px, py, pz, P, nr);
    part += direct_term<PERIODIC>(ent, i + 2, px, py, pz, P, nr);
    part += direct_term<PERIODIC>(ent, i + 3, px, py, pz, P, nr);
    wc.nnear += nr;
    acc.add(part);
  }
  for (; i < ee; i++)
  {
    bool nr = false;
    acc.add(direct_term<PERIODIC>(ent, i, px, py, pz, P, nr));
  }
}

// Snapshot_t::RelativeVelocity in float (src/snapshot.h:135-148)
HBTX_DEV void rel_vel(const DevParams &P, float x0, float x1, float x2, float v0, float v1, float v2,
                      const float *rp, const float *rv, float dv[3])
{
  const float ha = P.hz * P.a;
  float dx[3] = {x0 - rp[0], x1 - rp[1], x2 - rp[2]};
  const float v[3] = {v0, v1, v2};
#pragma unroll
  for (int j = 0; j < 3; j++)
  {
    if (P.periodic)
      dx[j] = nearest_f(dx[j], P.box, P.box_half);
    dv[j] = v[j] - rv[j];
    dv[j] += ha * dx[j];
  }
}

template <bool PERIODIC, bool STATS>
__global__ __launch_bounds__(64) void k_walk_energy(WalkArgs A)
{
  const DevParams &P = A.P;
  const int lane = lane_id();
  unsigned long long st_steps = 0, st_visits = 0, st_accepts = 0, st_pairs = 0, st_targets = 0, st_items = 0,
                     st_near = 0;
  // XCD-aware work distribution: workgroups are dispatched round-robin over the
  // XCDs, so workgroup b runs on XCD b % 8.  The item list (targets in key order,
  // largest subhalo first) is cut into nchunks contiguous chunks; a wave drains the
  // chunk of its XCD first, so neighbouring targets - which walk the same part of
  // the tree - share that XCD's L2, then helps with the other chunks.
  const int nit = *A.nitems;
  const int nch = A.nchunks;
  int chunk = blockIdx.x % nch;
  int tried = 0;
  for (;;)
  {
    int it = -1;
    while (tried < nch)
    {
      const int cb = (int)((int64_t)nit * chunk / nch), ce = (int)((int64_t)nit * (chunk + 1) / nch);
      int got = 0;
      if (lane == 0)
        got = atomicAdd(A.counter + 16 * chunk, 1);
      got = uniform(__shfl(got, 0, HBTX_WAVE));
      if (cb + got < ce)
      {
        it = cb + got;
        break;
      }
      tried++;
      chunk = (chunk + 1 == nch) ? 0 : chunk + 1;
    }
    if (it < 0)
      break;
    const int2 item = A.items[it];
    const int s = uniform(item.x);
    const int t = item.y + lane;
    const int nt = A.ntgt[s];
    const bool valid = t < nt;
    const int q = A.off[s] + t; // the target's slot
    const int p = valid ? A.cur[q] : 0;
    float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid)
      x = A.posm[p];
    const int eb = uniform(A.ent_begin[s]), ee = uniform(A.ent_end[s]); // walk-record units
    Kahan acc;
    acc.init();
    WalkCounts wc = {0u, 0u, 0u, 0u};
    if (eb < ee)
    {
      if (A.treemode[s])
        tree_walk<PERIODIC, STATS>(A.wt, eb, ee, valid, x.x, x.y, x.z, P, acc, wc);
      else
      {
        direct_sum<PERIODIC>(A.wt, eb, ee, x.x, x.y, x.z, P, acc, wc);
        if (STATS)
          st_pairs += (unsigned long long)__popcll(wave_mask(valid)) * (unsigned long long)((ee - eb) / 2);
      }
    }
    if (STATS)
    {
      st_steps += wc.steps;
      st_visits += wc.visits;
      st_accepts += wc.accepts;
      st_near += wc.nnear;
      st_targets += __popcll(wave_mask(valid));
      st_items += 1;
    }
    if (!valid)
      continue;
    const int md = A.mode[s];
    const float self = (md == MODE_FULL || md == MODE_REFINE || md == MODE_ADD) ? (A.selfm[p] / P.eps) : 0.f;
    const double phi = ((double)self + (double)acc.s - (double)acc.c) * P.G_over_a;
    if (A.phi_out)
      A.phi_out[p] = (float)phi;
    const float4 v = A.velu[p];
    const float *rf = A.ref + 6 * s;
    if (md == MODE_FULL || md == MODE_REFINE)
    {
      float dv[3];
      rel_vel(P, x.x, x.y, x.z, v.x, v.y, v.z, rf, rf + 3, dv);
      const float k2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
      float E = (float)((double)k2 * 0.5 + phi);
      if (md == MODE_REFINE)
        A.eref[p] = E; // RefineBindingEnergyOrder: BindingEnergy with the true mass, no thermal term
      else
      {
        if (P.thermal)
          E += v.w;
        A.energy[p] = E;
        A.esl[q] = E;
      }
    }
    else if (md == MODE_ADD)
    {
      // potential of the bound particles from outside the proposal
      const float E = (float)((double)A.energy[p] + phi);
      A.energy[p] = E;
      A.esl[q] = E;
    }
    else
    {
      const float *of = A.oldref + 6 * s;
      float dV[3], ov[3];
      rel_vel(P, of[0], of[1], of[2], of[3], of[4], of[5], rf, rf + 3, dV);
      const float dK = (float)(0.5 * (double)(dV[0] * dV[0] + dV[1] * dV[1] + dV[2] * dV[2]));
      rel_vel(P, x.x, x.y, x.z, v.x, v.y, v.z, of, of + 3, ov);
      const float tkin = (ov[0] * dV[0] + ov[1] * dV[1] + ov[2] * dV[2]) + dK;
      const float E = (float)((double)A.energy[p] + ((double)tkin - phi));
      A.energy[p] = E;
      A.esl[q] = E;
    }
  }
  if (STATS && A.stats && lane == 0)
  {
    atomicAdd(A.stats + STAT_TREE_STEPS, st_steps);
    atomicAdd(A.stats + STAT_TREE_VISITS, st_visits);
    atomicAdd(A.stats + STAT_TREE_ACCEPT, st_accepts);
    atomicAdd(A.stats + STAT_DIRECT_PAIRS, st_pairs);
    atomicAdd(A.stats + STAT_TARGETS, st_targets);
    atomicAdd(A.stats + STAT_ITEMS, st_items);
    atomicAdd(A.stats + STAT_NEAR, st_near);
  }
}

} // namespace hbtx
