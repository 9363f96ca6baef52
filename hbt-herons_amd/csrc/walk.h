// Potential + binding-energy kernel (one wave = 64 targets of one subhalo).
//
// Tree mode: each lane applies the reference's own opening test to its own
// target (open if len^2 > r^2 theta^2 with r to the cell's centre of mass,
// src/gravity_tree.cpp:139, stored as r^2 < crit2 = len^2/theta^2) over the
// depth-first walk records (16-byte particles, 24-byte cells), so that every lane
// accepts exactly the records its own walk accepts (src/gravity_tree.cpp:82-166),
// while the 64 lanes on spatially adjacent targets share one traversal: the
// batched stack walk below (tree_walk_stack).
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
  int32_t *overflow;        // the round's overflow word (WALK_STACK_OVERFLOW)
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

// Batched stack walk: lane-exact opening decisions, with the records a wave
// visits gathered 64 at a time instead of one dependent load per record.
//
// A stack entry (o, e, mask) stands for the sibling chain of records from byte
// offset o up to e, and mask holds the lanes that reached it (that opened every
// ancestor).  Processing record o: each reached lane applies its own opening
// test; a lane accepts the record (adds its term) unless it opens it; the chain's
// rest (o's skip, up to e) is pushed with the same mask, and an opened cell pushes
// its children's chain (o + 24, skip) with the mask of the lanes that opened it.
// A lane therefore accepts exactly the records its own walk accepts, and the wave
// visits exactly the records some lane reaches (the union of its lanes' walks);
// a batch pops up to 64 entries, loads their records together (one vector
// gather), stages them in LDS and then evaluates them back to back with no memory
// access on the chain between one record's test and the next.
//
// Capacity: a batch of B entries pushes at most 2 B, so B is cut to keep the stack
// WSTACK_HEADROOM below its capacity; once the room is gone single entries are
// popped, a depth-first descent that adds at most one entry per level (the
// chain's rest) plus the entry it stands on.  A tree has at most 22 cell levels
// (21-bit keys per axis under the root), so the stack never holds more than
// WSTACK_CAP - WSTACK_HEADROOM + 24 entries; a push beyond the capacity (never
// taken) raises the overflow word instead of writing.
constexpr int WSTACK_CAP = 160;
constexpr int WSTACK_HEADROOM = 32;

// Waves per walk workgroup.  Every wave walks on its own (its own items and its
// own LDS stack, no workgroup barrier); one, two and four waves a workgroup
// time the same, as a CU holds 32 one-wave workgroups with this LDS size.
#ifndef HBTX_WALK_WG_WAVES
#define HBTX_WALK_WG_WAVES 2
#endif
constexpr int WALK_WG_THREADS = HBTX_WAVE * HBTX_WALK_WG_WAVES;

// m's bit for this lane ? a : b, with m a wave-uniform lane mask: one
// v_cndmask on the mask's SGPR pair, where a select on a per-lane bool would first
// turn the mask back into that bool with vector instructions.
HBTX_DEV float lane_select(uint64_t m, float a, float b)
{
  float r;
  asm("v_cndmask_b32_e64 %0, %1, %2, %3" : "=v"(r) : "v"(b), "v"(a), "s"(m));
  return r;
}

// acc = -a * b + acc on the lanes of the wave-uniform mask m, the others keeping
// acc: one FMA under a narrowed exec mask (restored in the same block), where a
// select would take a second vector instruction.
HBTX_DEV void fma_neg_on(uint64_t m, float &acc, float a, float b)
{
  uint64_t save;
  asm volatile("s_mov_b64 %1, exec\n\t"
               "s_and_b64 exec, exec, %4\n\t"
               "v_fma_f32 %0, -%2, %3, %0\n\t"
               "s_mov_b64 exec, %1"
               : "+v"(acc), "=&s"(save)
               : "v"(a), "v"(b), "s"(m)
               : "scc");
}

// Orders one wave's LDS writes before its later LDS reads by other lanes: the
// LDS serves a wave's instructions in order, so only the compiler's reordering
// has to be held back (no workgroup barrier: the waves of a walk workgroup are
// independent).
HBTX_DEV void wave_lds_sync()
{
  __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
  __builtin_amdgcn_wave_barrier();
  __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

struct WalkLds
{
  int4 stk[WSTACK_CAP];  // (record offset, chain end, reached mask low, high word)
  float4 rec[HBTX_WAVE + 1]; // a batch's records: x, y, z, m (then the lanes that opened each)
  float4 aux[HBTX_WAVE + 1]; // crit2 (0 for a particle), mask low, crit2, mask high
  // (one spare slot each: the record after a batch's last is read ahead and not used)
};

#define WALK_STACK_OVERFLOW 4 // overflow-word bit: the walk stack was full (never expected)

template <bool PERIODIC, bool STATS>
HBTX_DEV void tree_walk_stack(const uint2 *wt, int ub, int ue, bool valid, float px, float py, float pz,
                              const DevParams &P, Kahan &acc, WalkCounts &wc, WalkLds &L, int32_t *overflow)
{
  const uint2 *tree = wt + ub;
  const int32_t span = 8 * (ue - ub);
  const int lane = lane_id();
  const uint64_t vmask = wave_mask(valid);
  if (vmask == 0ull)
    return;
  // lanes 0-31 read (crit2, mask low), lanes 32-63 (crit2, mask high) of a staged record
  const float2 *aw = reinterpret_cast<const float2 *>(L.aux) + (lane >> 5);
  // the staged records and open masks through VGPR base addresses (a zero the
  // compiler cannot see through), so that record j of an unrolled group is an
  // immediate offset instead of an address moved from an SGPR per record
  int vz;
  asm volatile("v_mov_b32 %0, 0" : "=v"(vz));
  const float4 *rec = L.rec + vz;
  uint2 *omw = reinterpret_cast<uint2 *>(L.rec) + vz;
  const uint32_t bit = 1u << (lane & 31);
  if (lane == 0)
    L.stk[0] = make_int4(0, span, (int)(uint32_t)vmask, (int)(uint32_t)(vmask >> 32));
  wave_lds_sync();
  int top = 1;
  float blk = 0.f;
  int nblk = 0;
  while (top > 0)
  {
    const int room = WSTACK_CAP - WSTACK_HEADROOM - top;
    const int nb = uniform(min(min(HBTX_WAVE, top), max(1, room)));
    top -= nb;
    const bool mine = lane < nb;
    int4 it = make_int4(0, 0, 0, 0);
    float4 a = make_float4(0.f, 0.f, 0.f, 0.f);
    float c2 = 0.f;
    int32_t nx = 0;
    bool cell = false;
    if (mine)
    {
      it = L.stk[top + lane];
      // 24 bytes whatever the record: a particle's last 8 are the next record's
      // (or the spare units past the records' end) and are not used, so that
      // both loads go out together
      const uint2 *rp = tree + (it.x >> 3);
      const uint2 u0 = rp[0], u1 = rp[1], u2 = rp[2];
      a = make_float4(__uint_as_float(u0.x), __uint_as_float(u0.y), __uint_as_float(u1.x),
                      __uint_as_float(u1.y));
      cell = (int32_t)u1.y < 0;
      c2 = cell ? __uint_as_float(u2.x) : 0.f;
      nx = cell ? (int32_t)u2.y : it.x + 16;
    }
    // staged cells first, then particles (the order within a batch is free): a
    // particle is never opened, so its visit skips the opening test and mask
    const uint64_t lt = lanemask_lt();
    const uint64_t cmask = wave_mask(mine && cell), pmask = wave_mask(mine && !cell);
    const int ncell = __popcll(cmask);
    const int slot = cell ? __popcll(cmask & lt) : ncell + __popcll(pmask & lt);
    if (mine)
    {
      L.rec[slot] = make_float4(a.x, a.y, a.z, fabsf(a.w));
      L.aux[slot] = make_float4(c2, __int_as_float(it.z), c2, __int_as_float(it.w));
    }
    wave_lds_sync();
    // Record j of the batch against every lane's target.  The lane predicates are
    // single compares whose masks the scalar unit combines: reached R, opened Lt
    // (r2 < crit2), near N (r2 < h2).  An accepted lane (R, not Lt) adds -m/r,
    // or the spline where N; the sum runs under the accepting lanes' exec mask,
    // so that a lane that does not accept never uses its 1/r (infinite where the
    // record is the target itself).  A cell's visit returns the lanes that opened it.
    float4 rn = rec[0];
    float2 wn = aw[0];
    auto visit = [&](int j, bool is_cell) -> uint64_t {
      // record j was read ahead; read record j + 1 while it is evaluated
      const float4 r = rn;
      const float2 w = wn;
      rn = rec[j + 1];
      wn = aw[2 * (j + 1)];
      const float r2 = sep2<PERIODIC>(r.x, r.y, r.z, px, py, pz, P);
      const uint64_t rm = __builtin_amdgcn_uicmp(__float_as_uint(w.y) & bit, 0u, 33 /* ne */);
      const uint64_t lm = is_cell ? __builtin_amdgcn_fcmpf(r2, w.x, 4 /* ordered < */) : 0ull;
      const uint64_t nm = __builtin_amdgcn_fcmpf(r2, P.h2, 4 /* ordered < */);
      const float rinv = __builtin_amdgcn_rsqf(r2);
      fma_neg_on(rm & ~(lm | nm), blk, r.w, rinv);
      const uint64_t na = rm & nm & ~lm;
      if (na != 0ull)
      {
        if (STATS)
          wc.nnear++;
        // at zero distance (the target's own record, or a coincident source) the
        // spline is its value at u = 0, m hinv (-2.8), as spline_near forms it
        const uint64_t zm = __builtin_amdgcn_fcmpf(r2, 0.f, 1 /* ordered == */);
        const float tn = ((na & ~zm) == 0ull) ? (r.w * P.hinv) * -2.8f : spline_near(r2, rinv, r.w, P.hinv);
        blk += lane_select(na, tn, 0.f);
      }
      if (STATS)
      {
        wc.steps++;
        wc.visits += __popcll(rm);
        wc.accepts += __popcll(rm & ~lm);
      }
      return rm & lm;
    };
    // plain FP32 sums over blocks of at least 16 records, Kahan-compensated
    // across blocks
    auto block_end = [&](int n) {
      nblk += n;
      if (nblk >= 16)
      {
        acc.add(blk);
        blk = 0.f;
        nblk = 0;
      }
    };
    // cells: the open masks go to the slots of records already read (record j's
    // first 8 bytes)
    int j = 0;
    for (; j + 4 <= ncell; j += 4)
    {
      const uint64_t o0 = visit(j, true), o1 = visit(j + 1, true), o2 = visit(j + 2, true),
                     o3 = visit(j + 3, true);
      if (lane == 0)
      {
        omw[2 * j] = make_uint2((uint32_t)o0, (uint32_t)(o0 >> 32));
        omw[2 * j + 2] = make_uint2((uint32_t)o1, (uint32_t)(o1 >> 32));
        omw[2 * j + 4] = make_uint2((uint32_t)o2, (uint32_t)(o2 >> 32));
        omw[2 * j + 6] = make_uint2((uint32_t)o3, (uint32_t)(o3 >> 32));
      }
      block_end(4);
    }
    for (; j < ncell; j++)
    {
      const uint64_t o = visit(j, true);
      if (lane == 0)
        omw[2 * j] = make_uint2((uint32_t)o, (uint32_t)(o >> 32));
      block_end(1);
    }
    // particles
    for (; j + 4 <= nb; j += 4)
    {
      visit(j, false);
      visit(j + 1, false);
      visit(j + 2, false);
      visit(j + 3, false);
      block_end(4);
    }
    for (; j < nb; j++)
    {
      visit(j, false);
      block_end(1);
    }
    // push the chains' rests and the opened cells' children
    wave_lds_sync();
    // (the popped entries are still in place: nothing is pushed before this point)
    if (mine)
      it = L.stk[top + lane];
    const bool is_cell = slot < ncell;
    const uint2 om = reinterpret_cast<const uint2 *>(L.rec)[2 * slot];
    const bool sib = mine && nx < it.y;
    const bool kid = mine && is_cell && (om.x | om.y) != 0u;
    const uint64_t msib = wave_mask(sib), mkid = wave_mask(kid);
    int pos = top + __popcll(msib & lt) + __popcll(mkid & lt);
    const int pushed = __popcll(msib) + __popcll(mkid);
    if (top + pushed > WSTACK_CAP)
    {
      if (lane == 0 && overflow)
        atomicOr(overflow, WALK_STACK_OVERFLOW);
      break;
    }
    if (sib)
      L.stk[pos++] = make_int4(nx, it.y, it.z, it.w);
    if (kid)
      L.stk[pos] = make_int4(it.x + 24, nx, (int)om.x, (int)om.y);
    top = uniform(top + pushed);
    wave_lds_sync();
  }
  acc.add(blk);
}

template <bool PERIODIC>
HBTX_DEV float direct_term(const float4 e, float px, float py, float pz, const DevParams &P, bool &near)
{
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

// Four particle records (64 bytes at an 8-byte aligned unit) in one scalar load.
HBTX_DEV void const_load_4rec(const uint2 *p, float4 r[4])
{
  typedef int v16i __attribute__((ext_vector_type(16)));
  typedef const __attribute__((address_space(4))) v16i c16;
  const v16i a = *(const c16 *)(uintptr_t)p;
#pragma unroll
  for (int k = 0; k < 4; k++)
    r[k] = make_float4(__int_as_float(a[4 * k]), __int_as_float(a[4 * k + 1]), __int_as_float(a[4 * k + 2]),
                       __int_as_float(a[4 * k + 3]));
}

// Direct mode: the records are particles only (two units each), summed in
// groups of four (one scalar load each), Kahan-compensated across groups.
template <bool PERIODIC>
HBTX_DEV void direct_sum(const uint2 *wt, int ub, int ue, float px, float py, float pz, const DevParams &P,
                         Kahan &acc, WalkCounts &wc)
{
  const uint2 *ent = wt + ub;
  const int ee = (ue - ub) / 2;
  int i = 0;
  for (; i + 4 <= ee; i += 4)
  {
    float4 r[4];
    const_load_4rec(ent + 2 * i, r);
    bool nr = false;
    float part = direct_term<PERIODIC>(r[0], px, py, pz, P, nr);
    part += direct_term<PERIODIC>(r[1], px, py, pz, P, nr);
    part += direct_term<PERIODIC>(r[2], px, py, pz, P, nr);
    part += direct_term<PERIODIC>(r[3], px, py, pz, P, nr);
    wc.nnear += nr;
    acc.add(part);
  }
  for (; i < ee; i++)
  {
    bool nr = false;
    acc.add(direct_term<PERIODIC>(const_load_f4u(ent + 2 * i), px, py, pz, P, nr));
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
__global__ __launch_bounds__(WALK_WG_THREADS, 8) void k_walk_energy(WalkArgs A)
{
  const DevParams &P = A.P;
  const int lane = lane_id();
  __shared__ WalkLds wls[HBTX_WALK_WG_WAVES];
  WalkLds &wl = wls[threadIdx.x / HBTX_WAVE];
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
      {
        tree_walk_stack<PERIODIC, STATS>(A.wt, eb, ee, valid, x.x, x.y, x.z, P, acc, wc, wl, A.overflow);
      }
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
