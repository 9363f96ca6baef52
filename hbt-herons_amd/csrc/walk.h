// Potential + binding-energy kernel (one wave = 64 targets of one subhalo).
//
// Tree mode: wavefront-cooperative walk of the depth-first entry array.  The
// wave visits entries in preorder at a wave-uniform index (scalar loads).  Each
// lane applies the reference's own opening test to its own target
// (open if len^2 > r^2 theta^2 with r to the cell's centre of mass,
// src/gravity_tree.cpp:139) and keeps a private "skip" bound: after accepting
// an entry it ignores that entry's subtree.  The wave descends (index+1) if
// any active lane opened the entry, otherwise it jumps to the entry's skip.
// Every lane therefore sees exactly the entries its own stackless walk would
// see, in the same order, while lanes on spatially adjacent targets share the
// traversal.
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
  STAT_NEAR = 6,          // accepted interactions inside the softening kernel
  NSTATS = 16
};

struct WalkArgs
{
  const int2 *items;        // (subhalo, first target) per work item
  const int32_t *nitems;    // number of items (device)
  int32_t *counter;         // work-stealing counter (zeroed before launch)
  const int32_t *cur;       // slot-layout target order
  const int32_t *off;       // per subhalo slot offset
  const int32_t *ntgt;
  const int32_t *mode;
  const int32_t *treemode;
  const int32_t *ent_begin;
  const int32_t *ent_end;
  const float4 *__restrict__ ent_pm;
  const int2 *__restrict__ ent_aux;
  const float4 *posm;
  const float4 *velu;
  const float *selfm;       // target self mass (sources of a full iteration), else 0
  const float *ref;         // per subhalo frame [6]
  const float *oldref;      // per subhalo previous frame [6]
  float *energy;
  float *phi_out;           // optional: specific potential per particle (float), may be null
  unsigned long long *stats;
  DevParams P;
};

// Far term -m/r and the spline kernel inside h, for the accepting lanes.
HBTX_DEV float pair_term(float r2, float m, bool acc, const DevParams &P, unsigned long long &nnear)
{
  float term = -m * __builtin_amdgcn_rsqf(r2);
  const bool nr = acc && (r2 < P.h2);
  const uint64_t bn = ballot(nr);
  if (bn)
  {
    nnear += __popcll(bn);
    const float tn = spline_near(r2, m, P.hinv);
    term = nr ? tn : term;
  }
  return acc ? term : 0.f;
}

HBTX_DEV void tree_walk(const WalkArgs &A, int eb, int ee, bool valid, float px, float py, float pz, Kahan &acc,
                        unsigned long long st[NSTATS])
{
  const DevParams &P = A.P;
  int i = eb;
  int myskip = valid ? eb : 0x7fffffff;
  unsigned long long steps = 0, visits = 0, accepts = 0, nnear = 0;
  while (i < ee)
  {
    const float4 e = const_load(A.ent_pm, i);
    const int2 a = const_load(A.ent_aux, i);
    const bool active = i >= myskip;
    float dx = e.x - px, dy = e.y - py, dz = e.z - pz;
    if (P.periodic)
    {
      dx = nearest_f(dx, P.box, P.box_half);
      dy = nearest_f(dy, P.box, P.box_half);
      dz = nearest_f(dz, P.box, P.box_half);
    }
    const float r2 = dx * dx + dy * dy + dz * dz;
    const bool open = __int_as_float(a.x) > r2 * P.theta2;
    const bool accept = active && !open;
    acc.add(pair_term(r2, e.w, accept, P, nnear));
    myskip = accept ? a.y : myskip;
    const uint64_t bo = ballot(active && open);
    steps++;
    visits += __popcll(ballot(active));
    accepts += __popcll(ballot(accept));
    // entries of a subtree follow it, so skip > i; max() keeps the walk finite regardless
    i = uniform(bo ? i + 1 : max(a.y, i + 1));
  }
  st[STAT_TREE_STEPS] += steps;
  st[STAT_TREE_VISITS] += visits;
  st[STAT_TREE_ACCEPT] += accepts;
  st[STAT_NEAR] += nnear;
}

HBTX_DEV void direct_sum(const WalkArgs &A, int eb, int ee, bool valid, float px, float py, float pz, Kahan &acc,
                         unsigned long long st[NSTATS])
{
  const DevParams &P = A.P;
  unsigned long long nnear = 0;
  int i = eb;
  for (; i + 4 <= ee; i += 4)
  {
    float part = 0.f;
#pragma unroll
    for (int k = 0; k < 4; k++)
    {
      const float4 e = const_load(A.ent_pm, i + k);
      float dx = e.x - px, dy = e.y - py, dz = e.z - pz;
      if (P.periodic)
      {
        dx = nearest_f(dx, P.box, P.box_half);
        dy = nearest_f(dy, P.box, P.box_half);
        dz = nearest_f(dz, P.box, P.box_half);
      }
      part += pair_term(dx * dx + dy * dy + dz * dz, e.w, valid, P, nnear);
    }
    acc.add(part);
  }
  for (; i < ee; i++)
  {
    const float4 e = const_load(A.ent_pm, i);
    float dx = e.x - px, dy = e.y - py, dz = e.z - pz;
    if (P.periodic)
    {
      dx = nearest_f(dx, P.box, P.box_half);
      dy = nearest_f(dy, P.box, P.box_half);
      dz = nearest_f(dz, P.box, P.box_half);
    }
    acc.add(pair_term(dx * dx + dy * dy + dz * dz, e.w, valid, P, nnear));
  }
  st[STAT_DIRECT_PAIRS] += (unsigned long long)__popcll(ballot(valid)) * (unsigned long long)(ee - eb);
  st[STAT_NEAR] += nnear;
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

__global__ __launch_bounds__(64) void k_walk_energy(WalkArgs A)
{
  const DevParams &P = A.P;
  const int lane = lane_id();
  unsigned long long st[NSTATS];
#pragma unroll
  for (int k = 0; k < NSTATS; k++)
    st[k] = 0;
  for (;;)
  {
    int it = 0;
    if (lane == 0)
      it = atomicAdd(A.counter, 1);
    it = uniform(__shfl(it, 0, HBTX_WAVE));
    if (it >= *A.nitems)
      break;
    const int2 item = A.items[it];
    const int s = uniform(item.x);
    const int t = item.y + lane;
    const int nt = A.ntgt[s];
    const bool valid = t < nt;
    const int p = valid ? A.cur[A.off[s] + t] : 0;
    float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid)
      x = A.posm[p];
    const int eb = uniform(A.ent_begin[s]), ee = uniform(A.ent_end[s]);
    Kahan acc;
    acc.init();
    if (A.treemode[s])
      tree_walk(A, eb, ee, valid, x.x, x.y, x.z, acc, st);
    else
      direct_sum(A, eb, ee, valid, x.x, x.y, x.z, acc, st);
    st[STAT_TARGETS] += __popcll(ballot(valid));
    st[STAT_ITEMS] += 1;
    if (!valid)
      continue;
    const int md = A.mode[s];
    const float self = (md == MODE_FULL) ? (A.selfm[p] / P.eps) : 0.f;
    const double phi = ((double)self + (double)acc.s - (double)acc.c) * P.G_over_a;
    if (A.phi_out)
      A.phi_out[p] = (float)phi;
    const float4 v = A.velu[p];
    const float *rf = A.ref + 6 * s;
    if (md == MODE_FULL)
    {
      float dv[3];
      rel_vel(P, x.x, x.y, x.z, v.x, v.y, v.z, rf, rf + 3, dv);
      const float k2 = dv[0] * dv[0] + dv[1] * dv[1] + dv[2] * dv[2];
      float E = (float)((double)k2 * 0.5 + phi);
      if (P.thermal)
        E += v.w;
      A.energy[p] = E;
    }
    else
    {
      const float *of = A.oldref + 6 * s;
      float dV[3], ov[3];
      rel_vel(P, of[0], of[1], of[2], of[3], of[4], of[5], rf, rf + 3, dV);
      const float dK = (float)(0.5 * (double)(dV[0] * dV[0] + dV[1] * dV[1] + dV[2] * dV[2]));
      rel_vel(P, x.x, x.y, x.z, v.x, v.y, v.z, of, of + 3, ov);
      const float tkin = (ov[0] * dV[0] + ov[1] * dV[1] + ov[2] * dV[2]) + dK;
      A.energy[p] = (float)((double)A.energy[p] + ((double)tkin - phi));
    }
  }
  if (A.stats && lane == 0)
  {
#pragma unroll
    for (int k = 0; k < NSTATS; k++)
      if (st[k])
        atomicAdd(A.stats + k, st[k]);
  }
}

} // namespace hbtx
