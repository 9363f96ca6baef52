// Per-round kernels of the batched unbinding (Subhalo_t::Unbind,
// src/subhalo_unbind.cpp:463-808, advanced for every subhalo at once).
//
// Layout: particles of subhalo s occupy global indices [off[s], off[s+1]) in
// input order ("slots").  state[p] = 0 while p is a candidate (in the bound set
// being iterated), r > 0 once p was removed in round r.  cur[] holds, in slot
// layout, the current candidates of each subhalo in walk order (the key order
// of the last full tree), compacted after every round.
#pragma once
#include "common.h"
#include "scan.h"
#include "hilbert.h"

namespace hbtx
{

struct SubArrays
{
  int32_t nsub;
  const int32_t *off;    // [nsub+1]
  const int32_t *sub_of; // [N]
  int32_t *status, *corr, *iters, *mode, *treemode;
  int32_t *nlast, *nbound, *nrem, *nuns, *ntgt, *nsrc;
  float *mlast, *mbound;
  float *ref, *oldref, *mb; // [nsub*6]
  float *upscale;           // [nsub*6]
  int32_t *src_off, *ent_begin, *ent_end;
  int32_t *nitems;          // [nsub]
  double *cube;             // [nsub*4] centre xyz, lenhalf
  double *len0;             // [nsub]
  uint32_t *bbox;           // [nsub*6] ordered min xyz, max xyz
  int32_t *bnd_cnt, *bnd_tr;
  int32_t *first_idx;       // smallest candidate index (periodic reference particle)
  int32_t *active_count;    // [1]
  int32_t *max_active_n;    // [1] particles of the largest subhalo still active after the round
  // learned warm start: warm = WARM_* state; npro = proposed particles; nadd = bound
  // particles outside the proposal (re-admitted by the warm round); reset = proposal
  // rejected, restart from all particles
  int32_t *warm, *npro, *nadd, *reset;
  float *minit;             // initial bound mass (all particles)
};

struct PartArrays
{
  const float4 *posm;
  const float4 *velu;
  const int8_t *ptype; // may be null (all type 1)
  const uint8_t *proposal; // warm-start proposal (1 = proposed bound), may be null
  int32_t *state;
  float *energy;
  float *esl; // the energies of the round's targets again, by slot (the walk writes both)
  float *selfm;
  int32_t *cur, *cur2;
};

HBTX_DEV int ptype_of(const PartArrays &Q, int p)
{
  return Q.ptype ? (int)Q.ptype[p] : 1;
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
  GRID_STRIDE(i, n) p[i] = v;
}
__global__ void k_copy_i32(const int32_t *src, int32_t *dst)
{
  if (blockIdx.x == 0 && threadIdx.x == 0)
    *dst = *src;
}
__global__ void k_fill_u64(unsigned long long *p, int64_t n, unsigned long long v)
{
  GRID_STRIDE(i, n) p[i] = v;
}

// Batch start: every particle is a candidate, frame from the caller
// (src/subhalo_unbind.cpp:503-541).
__global__ void k_init_particles(SubArrays S, PartArrays Q, int64_t n)
{
  GRID_STRIDE(p, n)
  {
    Q.state[p] = 0;
    Q.cur[p] = (int32_t)p;
    Q.energy[p] = 0.f;
    Q.selfm[p] = 0.f;
  }
}

__global__ void k_init_subs(SubArrays S, const float *ref_in, const float *mb_in, const uint8_t *has_proposal)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  const int n = S.off[s + 1] - S.off[s];
  S.status[s] = (n == 0) ? ST_CONVERGED : ST_ACTIVE;
  S.corr[s] = 0;
  S.warm[s] = WARM_OFF;
  S.npro[s] = 0;
  S.nadd[s] = 0;
  S.reset[s] = 0;
  S.iters[s] = 0;
  S.nbound[s] = n;
  S.nlast[s] = n;
  S.nrem[s] = 0;
  S.nuns[s] = 0;
  for (int k = 0; k < 6; k++)
  {
    S.ref[6 * s + k] = ref_in[6 * s + k];
    S.oldref[6 * s + k] = 0.f;
    S.mb[6 * s + k] = mb_in[6 * s + k];
    S.upscale[6 * s + k] = 1.f;
  }
}

// Start of a round: clear the active count, advance the device round counter.
__global__ void k_round_begin(int32_t *active, int32_t *round_p, int32_t *max_active_n)
{
  if (blockIdx.x == 0 && threadIdx.x == 0)
  {
    *active = 0;
    *round_p += 1;
    *max_active_n = 0;
  }
}

// Mode of the round (:563-600); subsampling keeps MaxSampleSize+Nunsample sources.
__global__ void k_plan(SubArrays S, DevParams P, const int32_t *round_p)
{
  const int round = *round_p;
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  if (S.status[s] != ST_ACTIVE)
  {
    S.mode[s] = MODE_IDLE;
    S.ntgt[s] = 0;
    S.nsrc[s] = 0;
    S.nitems[s] = 0;
    S.treemode[s] = 0;
    return;
  }
  const int nb = S.nbound[s];
  int ns;
  if (S.warm[s] == WARM_ADD)
  {
    // after an accepted warm round: the bound particles from outside the proposal
    // are the sources; frames, nlast and the correction state stay as they are
    S.mode[s] = MODE_ADD;
    ns = S.nadd[s];
  }
  else if (S.corr[s])
  {
    S.mode[s] = MODE_CORR;
    ns = S.nrem[s];
  }
  else
  {
    S.mode[s] = MODE_FULL;
    S.mlast[s] = S.mbound[s];
    // SetMassUpscaleFactor (:374-391)
    if (P.mss == 0 || (int64_t)nb < P.mss + S.nuns[s])
      ns = nb;
    else
      ns = (int)(P.mss + S.nuns[s]);
  }
  if (S.warm[s] != WARM_ADD)
    S.nlast[s] = nb;
  if (S.warm[s] == WARM_PENDING)
  {
    // warm round: every particle is a target, the proposed set is the source set,
    // and convergence is measured against the proposed set
    ns = S.npro[s];
    S.nlast[s] = ns;
  }
  S.ntgt[s] = nb;
  S.nsrc[s] = ns;
  S.treemode[s] = ((int64_t)ns > P.direct_cutoff) ? 1 : 0;
  S.nitems[s] = (nb + HBTX_WAVE - 1) / HBTX_WAVE;
  S.iters[s] = round;
}

// Candidate rank in input order (the E list order between stable partitions):
// cand[p] = 1 for candidates of full-mode subhaloes (exclusive scan -> rank).
__global__ void k_mark_candidates(SubArrays S, PartArrays Q, int64_t n, int32_t *flag)
{
  GRID_STRIDE(p, n)
  {
    const int s = S.sub_of[p];
    flag[p] = (S.mode[s] == MODE_FULL && Q.state[p] == 0) ? 1 : 0;
  }
}

// Source flags: full mode -> candidates with rank < nsrc; correction -> removed
// in the previous round.  Also the self mass used to cancel the self term.
// (also each tile's flag total in tsum[t], and 0 for the tiles [ntiles, maxt), for
// the scan over the tiles; launched with SCAN_THREADS threads a block)
__global__ __launch_bounds__(SCAN_THREADS) void k_mark_sources(SubArrays S, PartArrays Q, DevParams P,
                                                               const int4 *tiles, const int32_t *ntiles_p,
                                                               const int32_t *round_p, const int32_t *crank,
                                                               int32_t *flag, int32_t *tsum, int32_t maxt)
{
  __shared__ int32_t lds[SCAN_THREADS / HBTX_WAVE];
  const int round = *round_p;
  TILE_LOOP(tiles, ntiles_p)
  {
    int32_t cnt = 0;
    const int4 tl = tiles[t_];
    const int s = tl.x;
    const int md = S.mode[s];
    const int warm = S.warm[s];
    const int nsrc = S.nsrc[s];
    const int rank0 = crank ? crank[S.off[s]] : 0;
    for (int p = tl.y + threadIdx.x; p < tl.z; p += blockDim.x)
    {
      int f = 0;
      float sm = 0.f;
      const int st = Q.state[p];
      if (md == MODE_FULL && st == 0 && warm == WARM_PENDING)
      {
        if (Q.proposal[p])
        {
          f = 1;
          sm = Q.posm[p].w;
        }
      }
      else if (md == MODE_ADD)
      {
        // the bound particles the warm round found outside the proposal (each
        // also a target: its self term is cancelled)
        if (st == 0 && !Q.proposal[p])
        {
          f = 1;
          sm = Q.posm[p].w;
        }
      }
      else if (md == MODE_FULL && st == 0)
      {
        // candidate rank in E-list order (only with subsampling; else every candidate is a source)
        const int rank = crank ? crank[p] - rank0 : 0;
        if (rank < nsrc)
        {
          f = 1;
          const float m = Q.posm[p].w;
          const float *up = S.upscale + 6 * s;
          const int t = ptype_of(Q, p);
          sm = P.dm_only ? m * up[1] : (donot_subsample(P, t) ? m : m * up[t]);
        }
      }
      else if (md == MODE_CORR && warm == WARM_CORR)
        // first correction after a warm round: the proposed particles the warm round
        // removed (the proposal was its source set), and any the MODE_ADD round
        // (round 2, if it ran) left unbound
        f = ((st == 1 && Q.proposal[p]) || (st == 2 && round == 3)) ? 1 : 0;
      else if (md == MODE_CORR && st == round - 1)
        f = 1;
      flag[p] = f;
      cnt += f;
      if (md != MODE_CORR) // (a correction's walk reads no self mass)
        Q.selfm[p] = sm;
    }
    int32_t tot;
    block_excl_scan(cnt, lds, &tot);
    if (threadIdx.x == 0)
      tsum[t_] = tot;
  }
  for (int t = *ntiles_p + blockIdx.x * blockDim.x + threadIdx.x; t < maxt; t += gridDim.x * blockDim.x)
    tsum[t] = 0;
}

__global__ void k_sub_src(SubArrays S, const int32_t *sub_first)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  S.src_off[s] = (S.mode[s] != MODE_IDLE) ? sub_first[s] : 0;
  const uint32_t big = float_to_ordered(3.0e38f), small = float_to_ordered(-3.0e38f);
  for (int k = 0; k < 3; k++)
  {
    S.bbox[6 * s + k] = big;
    S.bbox[6 * s + 3 + k] = small;
  }
}

// Bounding box of each tree-mode subhalo's sources (src/oct_tree.tpp:344-355).
// Waves whose lanes share one subhalo reduce first, then one atomic per wave.
__global__ void k_bbox(SubArrays S, const float4 *posm, const int32_t *src, const int32_t *s_total)
{
  const int ST = *s_total;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < ST; base += (int64_t)gridDim.x * blockDim.x)
  {
    const int64_t j = base + threadIdx.x;
    const bool valid = j < ST;
    int s = -1;
    float4 x = make_float4(0.f, 0.f, 0.f, 0.f);
    if (valid)
    {
      const int p = src[j];
      s = S.sub_of[p];
      if (!S.treemode[s])
        s = -1;
      else
        x = posm[p];
    }
    const int s0 = __shfl(s, 0, HBTX_WAVE);
    const bool same = ballot(s != s0) == 0ull;
    if (same)
    {
      if (s0 < 0)
        continue;
      float mn[3] = {x.x, x.y, x.z}, mx[3] = {x.x, x.y, x.z};
#pragma unroll
      for (int d = 1; d < HBTX_WAVE; d <<= 1)
#pragma unroll
        for (int k = 0; k < 3; k++)
        {
          mn[k] = fminf(mn[k], __shfl_xor(mn[k], d, HBTX_WAVE));
          mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], d, HBTX_WAVE));
        }
      if (lane_id() == 0)
        for (int k = 0; k < 3; k++)
        {
          atomicMin(S.bbox + 6 * s0 + k, float_to_ordered(mn[k]));
          atomicMax(S.bbox + 6 * s0 + 3 + k, float_to_ordered(mx[k]));
        }
    }
    else if (s >= 0)
    {
      const float xs[3] = {x.x, x.y, x.z};
      for (int k = 0; k < 3; k++)
      {
        atomicMin(S.bbox + 6 * s + k, float_to_ordered(xs[k]));
        atomicMax(S.bbox + 6 * s + 3 + k, float_to_ordered(xs[k]));
      }
    }
  }
}

// Bounding box per tile of a dense source range (tiles never span subhaloes):
// block min/max reduction, then six atomics per tile (min/max are order-free).
__global__ __launch_bounds__(256) void k_bbox_tiles(SubArrays S, const float4 *posm, const int32_t *src,
                                                    const int4 *tiles, const int32_t *ntiles_p)
{
  __shared__ float lds[6][256 / HBTX_WAVE];
  const int nt = *ntiles_p;
  for (int t = blockIdx.x; t < nt; t += gridDim.x)
  {
    const int4 tl = tiles[t];
    const int s = tl.x;
    float mn[3] = {3.0e38f, 3.0e38f, 3.0e38f}, mx[3] = {-3.0e38f, -3.0e38f, -3.0e38f};
    for (int j = tl.y + threadIdx.x; j < tl.z; j += blockDim.x)
    {
      const float4 x = posm[src[j]];
      mn[0] = fminf(mn[0], x.x);
      mn[1] = fminf(mn[1], x.y);
      mn[2] = fminf(mn[2], x.z);
      mx[0] = fmaxf(mx[0], x.x);
      mx[1] = fmaxf(mx[1], x.y);
      mx[2] = fmaxf(mx[2], x.z);
    }
#pragma unroll
    for (int d = 1; d < HBTX_WAVE; d <<= 1)
#pragma unroll
      for (int k = 0; k < 3; k++)
      {
        mn[k] = fminf(mn[k], __shfl_xor(mn[k], d, HBTX_WAVE));
        mx[k] = fmaxf(mx[k], __shfl_xor(mx[k], d, HBTX_WAVE));
      }
    const int w = threadIdx.x / HBTX_WAVE;
    if (lane_id() == 0)
      for (int k = 0; k < 3; k++)
      {
        lds[k][w] = mn[k];
        lds[3 + k][w] = mx[k];
      }
    __syncthreads();
    if (threadIdx.x == 0)
    {
      for (int ww = 1; ww < (int)blockDim.x / HBTX_WAVE; ww++)
        for (int k = 0; k < 3; k++)
        {
          lds[k][0] = fminf(lds[k][0], lds[k][ww]);
          lds[3 + k][0] = fmaxf(lds[3 + k][0], lds[3 + k][ww]);
        }
      for (int k = 0; k < 3; k++)
      {
        atomicMin(S.bbox + 6 * s + k, float_to_ordered(lds[k][0]));
        atomicMax(S.bbox + 6 * s + 3 + k, float_to_ordered(lds[3 + k][0]));
      }
    }
    __syncthreads();
  }
}

// Dense source ranges of tree-mode subhaloes (for the bounding-box tiles).
__global__ void k_tree_src_ranges(SubArrays S, int32_t *begin, int32_t *len)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  begin[s] = S.src_off[s];
  len[s] = (S.mode[s] != MODE_IDLE && S.treemode[s]) ? S.nsrc[s] : 0;
}

// Root cube: Len = max extent, Center = bbox midpoint (src/oct_tree.tpp:357-365).
__global__ void k_cube(SubArrays S)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub || !S.treemode[s] || S.mode[s] == MODE_IDLE)
    return;
  double xmin[3], xmax[3];
  for (int k = 0; k < 3; k++)
  {
    xmin[k] = ordered_to_float(S.bbox[6 * s + k]);
    xmax[k] = ordered_to_float(S.bbox[6 * s + 3 + k]);
  }
  double Len = xmax[0] - xmin[0];
  for (int j = 1; j < 3; j++)
    if ((xmax[j] - xmin[j]) > Len)
      Len = xmax[j] - xmin[j];
  for (int j = 0; j < 3; j++)
    S.cube[4 * s + j] = 0.5 * (xmax[j] + xmin[j]);
  S.cube[4 * s + 3] = 0.5 * Len;
  S.len0[s] = Len;
}

__global__ void k_keys(SubArrays S, DevParams P, const float4 *posm, const int32_t *src, const int32_t *src_sub,
                       const int32_t *s_total, uint64_t *keys)
{
  const int ST = *s_total;
  GRID_STRIDE(j, ST)
  {
    const int s = src_sub[j];
    if (!S.treemode[s])
    {
      keys[j] = 0ull;
      continue;
    }
    const int p = src[j];
    const float4 x = posm[p];
    const double *c = S.cube + 4 * s;
    const uint64_t k = octree_key(x.x, x.y, x.z, c[0], c[1], c[2], c[3]);
    keys[j] = hash_digits(k, resolution_depth(c[3], 0.5 * P.res), (uint32_t)(p - S.off[s]));
  }
}

// Sort segments: tree-mode sources only.
__global__ void k_sort_segments(SubArrays S, int32_t *seg_begin, int32_t *seg_len)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  seg_begin[s] = S.src_off[s];
  seg_len[s] = (S.mode[s] != MODE_IDLE && S.treemode[s]) ? S.nsrc[s] : 0;
}

// Walk order of a full round: when every target is a source, the targets take
// the key order of the freshly sorted sources (spatially coherent waves); the
// compactions of later rounds keep that order.
__global__ void k_targets_from_sources(SubArrays S, PartArrays Q, const int32_t *src, const int32_t *src_sub,
                                       const int32_t *s_total)
{
  const int ST = *s_total;
  GRID_STRIDE(j, ST)
  {
    const int s = src_sub[j];
    if ((S.mode[s] == MODE_FULL || S.mode[s] == MODE_REFINE) && S.treemode[s] && S.nsrc[s] == S.ntgt[s])
      Q.cur[S.off[s] + (int)(j - S.src_off[s])] = src[j];
  }
}

// Per subhalo, the walk-record units of its tree, clamped to the record buffer (3
// cap units) so that an overflowing round, reported to the host, never reads past
// the buffer.  The walk addresses a tree's records by int32 byte offsets: a tree
// over WALK_MAX_SPAN_UNITS units raises overflow bit 2 (reported as a capacity
// error) and walks nothing.
__global__ void k_sub_entries(SubArrays S, const int32_t *cscan, int32_t cap, int32_t *overflow)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  const int b = S.src_off[s], e = b + S.nsrc[s];
  const int ub = min(wrec_unit(b, cscan[b]), 3 * cap);
  int ue = min(wrec_unit(e, cscan[e]), 3 * cap);
  if (ue - ub > WALK_MAX_SPAN_UNITS)
  {
    atomicOr(overflow, 2);
    ue = ub;
  }
  S.ent_begin[s] = ub;
  S.ent_end[s] = ue;
}

// Entries of the round's trees: e_total = min(E, cap), overflow raised when E
// exceeds the capacity.
__global__ void k_entry_total(const int32_t *s_total, const int32_t *cscan, int32_t *e_total, int32_t cap,
                              int32_t *overflow)
{
  if (blockIdx.x == 0 && threadIdx.x == 0)
  {
    const int S = *s_total;
    const int E = S + cscan[S];
    *e_total = min(E, cap);
    if (E > cap)
      atomicOr(overflow, 1);
  }
}

// Work items in subhalo size order (largest first).
__global__ void k_items_count(SubArrays S, const int32_t *order, int32_t *cnt)
{
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i < S.nsub)
    cnt[i] = S.nitems[order[i]];
}

// one wave per subhalo writes its items
__global__ void k_items_emit(SubArrays S, const int32_t *order, const int32_t *iscan, int2 *items)
{
  const int i = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / HBTX_WAVE);
  if (i >= S.nsub)
    return;
  const int s = order[i];
  const int n = S.nitems[s];
  const int b = iscan[i];
  for (int k = lane_id(); k < n; k += HBTX_WAVE)
    items[b + k] = make_int2(s, k * HBTX_WAVE);
}

// Partition flags over slots: low 32 bits = bound, high 32 bits = bound tracer.
// In a warm round every particle with E < 0 stays, proposed or not; the bound
// particles outside the proposal are counted in warm_np[s] (their potential is
// added to the others' energies by the MODE_ADD round that follows).
// (also each tile's flag total in tsum[t], and 0 for the tiles [ntiles, maxt);
// launched with SCAN_THREADS threads a block)
__global__ __launch_bounds__(SCAN_THREADS) void k_partition_flags(SubArrays S, PartArrays Q, DevParams P,
                                                                  const int4 *tiles, const int32_t *ntiles_p,
                                                                  int64_t *flag, int32_t *warm_np, int64_t *tsum,
                                                                  int32_t maxt)
{
  __shared__ int64_t lds[SCAN_THREADS / HBTX_WAVE];
  TILE_LOOP(tiles, ntiles_p)
  {
    int64_t cnt = 0;
    const int4 tl = tiles[t_];
    const int s = tl.x;
    const int end = min(tl.z, S.off[s] + S.ntgt[s]);
    const bool warm = warm_np && S.warm[s] == WARM_PENDING;
    for (int q = tl.y + threadIdx.x; q < tl.z; q += blockDim.x)
    {
      int64_t f = 0;
      bool outside = false; // bound, but not in the proposal
      // the energy by slot: a contiguous read (the particle's own is a gather)
      if (q < end && Q.esl[q] < 0.f)
      {
        const int p = (warm || Q.ptype) ? Q.cur[q] : 0;
        outside = warm && !Q.proposal[p];
        f = 1 | ((int64_t)(is_tracer(P, ptype_of(Q, p)) ? 1 : 0) << 32);
      }
      flag[q] = f;
      cnt += f;
      // the tile is one subhalo's: one counter update per wave
      const uint64_t m = ballot(outside);
      if (outside && (m & lanemask_lt()) == 0ull)
        atomicAdd(warm_np + s, (int32_t)__popcll(m));
    }
    int64_t tot;
    block_excl_scan(cnt, lds, &tot);
    if (threadIdx.x == 0)
      tsum[t_] = tot;
  }
  for (int t = *ntiles_p + blockIdx.x * blockDim.x + threadIdx.x; t < maxt; t += gridDim.x * blockDim.x)
    tsum[t] = 0;
}

__global__ void k_copy_cur(SubArrays S, PartArrays Q, const int4 *tiles, const int32_t *ntiles_p,
                           const int64_t *sub_total)
{
  TILE_LOOP(tiles, ntiles_p)
  {
    const int4 tl = tiles[t_];
    const int s = tl.x;
    const int nb = (int)(sub_total[s] & 0xffffffffll);
    const int end = min(tl.z, S.off[s] + nb);
    for (int q = tl.y + threadIdx.x; q < end; q += blockDim.x)
      Q.cur[q] = Q.cur2[q];
  }
}

// Disruption test and correction decision (:625-683).
__global__ void k_round_decide(SubArrays S, DevParams P, const int64_t *sub_total, const int32_t *warm_np)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub || S.mode[s] == MODE_IDLE)
    return;
  const int64_t d = sub_total[s];
  const int nb = (int)(d & 0xffffffffll);
  const int ntr = P.dm_only ? nb : (int)(d >> 32);
  if (S.mode[s] == MODE_ADD)
  {
    // The added potential only deepens the energies, so every candidate normally
    // stays; any that did not are sources of the correction that follows, with the
    // proposed particles the warm round removed.  Frames and the correction state
    // are left as the warm round set them.
    S.nrem[s] += S.nbound[s] - nb;
    S.nbound[s] = nb;
    S.bnd_cnt[s] = nb;
    S.bnd_tr[s] = ntr;
    S.warm[s] = WARM_CORR; // (the disruption test follows the correction)
    return;
  }
  const int nlast = S.nlast[s];
  S.nbound[s] = nb;
  S.nrem[s] = nlast - nb;
  S.bnd_cnt[s] = nb;
  S.bnd_tr[s] = ntr;
  if (S.warm[s] == WARM_CORR)
    S.warm[s] = WARM_DONE; // this was the first correction after the warm round
  if (S.warm[s] == WARM_PENDING)
  {
    // Proposal verification: the proposed set's exact potential is evaluated at
    // every particle, and every particle with E < 0 stays, proposed or not.
    // Accepted when the bound particles from outside the proposal number at most
    // 0.1% of the bound set, at least half of the proposal is bound and the
    // subhalo is well above the disruption limit (twice MinNumPartOfSub and
    // MinNumTracerPartOfSub).  The outside ones' potential is then added to every
    // candidate's energy (a MODE_ADD round) and the correction sequence continues
    // from the bound set.  Otherwise the plain iteration restarts from every
    // particle.
    const int np = warm_np ? warm_np[s] : 0; // bound, outside the proposal
    const int nin = nb - np;                 // bound and proposed
    const int n = S.off[s + 1] - S.off[s];
    const int npro = S.npro[s];
    const bool accept = (int64_t)nin >= 2 * P.min_part && (int64_t)ntr >= 2 * P.min_tracer &&
                        2 * (int64_t)nin >= (int64_t)npro && (int64_t)np * 1000 <= (int64_t)nb;
    S.warm[s] = WARM_DONE;
    if (accept)
    {
      // the warm round's frame is the proposal's: at least one more round follows so
      // that the final energies are taken in the bound set's frame
      const int nrem = npro - nin; // proposed but unbound: the first correction's sources
      S.nrem[s] = nrem;
      S.nlast[s] = n;
      S.nadd[s] = np;
      if (nrem < nb)
      {
        S.corr[s] = 1;
        S.warm[s] = (np > 0) ? WARM_ADD : WARM_CORR;
        for (int k = 0; k < 6; k++)
          S.oldref[6 * s + k] = S.ref[6 * s + k];
      }
    }
    else
    {
      S.reset[s] = 1;
      S.nlast[s] = n;
    }
    return;
  }
  if ((int64_t)nb < P.min_part || (int64_t)ntr < P.min_tracer)
  {
    S.status[s] = ST_DISRUPTED;
    for (int k = 0; k < 6; k++)
      S.ref[6 * s + k] = S.mb[6 * s + k];
    S.nbound[s] = 0;
    S.mbound[s] = 0.f;
    return;
  }
  const int ndiff = nlast - nb;
  if (ndiff < nb)
    if (P.mss <= 0 || (int64_t)ndiff < P.mss)
    {
      S.corr[s] = 1;
      for (int k = 0; k < 6; k++)
        S.oldref[6 * s + k] = S.ref[6 * s + k];
    }
}

// Rejected proposals: back to the starting state of a plain unbinding.
__global__ void k_warm_reset_particles(SubArrays S, PartArrays Q, int64_t n)
{
  GRID_STRIDE(p, n)
  {
    const int s = S.sub_of[p];
    if (S.reset[s])
    {
      Q.state[p] = 0;
      Q.cur[p] = (int32_t)p;
    }
  }
}

__global__ void k_warm_reset_subs(SubArrays S, const float *ref_in)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub || !S.reset[s])
    return;
  const int n = S.off[s + 1] - S.off[s];
  S.nbound[s] = n;
  S.nlast[s] = n;
  S.mbound[s] = S.minit[s];
  S.corr[s] = 0;
  for (int k = 0; k < 6; k++)
    S.ref[6 * s + k] = ref_in[6 * s + k];
  S.status[s] = ST_ACTIVE;
  S.reset[s] = 0;
}

// Warm-round walk order: every target sorted by its key in the proposal's tree cube
// (top 10 levels).
__global__ void k_warm_keys(SubArrays S, PartArrays Q, int64_t n, uint64_t *keys, int32_t *seg_begin, int32_t *seg_len)
{
  GRID_STRIDE(q, n)
  {
    const int s = S.sub_of[q];
    if (S.warm[s] == WARM_PENDING && S.mode[s] == MODE_FULL && S.treemode[s] && S.nsrc[s] < S.ntgt[s])
    {
      const float4 x = Q.posm[Q.cur[q]];
      const double *c = S.cube + 4 * s;
      // the top 10 levels are enough to group targets for the walk (cells below
      // 1/1024 of the cube hold particles closer than the softening), along the
      // Hilbert curve, whose runs of 64 cells are more compact than the Morton
      // curve's
      keys[q] = hilbert3_10(octree_key(x.x, x.y, x.z, c[0], c[1], c[2], c[3]) >> (3 * (HBTX_KEY_LEVELS - 10)));
    }
  }
}

__global__ void k_warm_segments(SubArrays S, int32_t *seg_begin, int32_t *seg_len)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  seg_begin[s] = S.off[s];
  seg_len[s] = (S.warm[s] == WARM_PENDING && S.mode[s] == MODE_FULL && S.treemode[s] && S.nsrc[s] < S.ntgt[s])
                 ? (S.off[s + 1] - S.off[s])
                 : 0;
}

// ---------------------------------------------------------------------------
// Deterministic segmented reductions over slot ranges [off[s], off[s]+len[s]):
// fixed tiles, block tree-reduction per tile, per-subhalo sum of tiles in order.
// ---------------------------------------------------------------------------
#define RED_THREADS 256
#define RED_TILE 4096
#define RED_NV 20

// Sum the first NV of the RED_NV values over a 256-thread block (thread 0 gets
// the totals, in wave order).  A wave whose lanes saw no element of the tile
// holds zeros: it skips the shuffles, since its sum is zero.
template <int NV>
HBTX_DEV void block_reduce_vec(double v[RED_NV], double *lds /* [RED_THREADS/64][RED_NV] */, bool wave_has_data)
{
  if (wave_has_data)
  {
#pragma unroll
    for (int k = 0; k < NV; k++)
    {
      double x = v[k];
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1)
        x += __shfl_down(x, d, HBTX_WAVE);
      v[k] = x;
    }
  }
  const int w = threadIdx.x / HBTX_WAVE;
  if (lane_id() == 0)
#pragma unroll
    for (int k = 0; k < NV; k++)
      lds[w * RED_NV + k] = v[k];
  __syncthreads();
  if (threadIdx.x == 0)
#pragma unroll
    for (int k = 0; k < NV; k++)
    {
      double x = 0.;
      for (int ww = 0; ww < RED_THREADS / HBTX_WAVE; ww++)
        x += lds[ww * RED_NV + k];
      v[k] = x;
    }
  __syncthreads();
}

// does the wave of this thread hold elements of the tile [b, e) (strided by the block)?
HBTX_DEV bool wave_in_tile(int b, int e)
{
  return b + (int)(threadIdx.x & ~(HBTX_WAVE - 1)) < e;
}

// A tile [b, e) of at most RED_THREADS slots reduced by one wave, with no barrier:
// the lanes take the block's waves in turn (slot b + 64 w + lane), reduce each with
// the same shuffle tree and add the wave sums in wave order, so lane 0 gets exactly
// the totals block_reduce_vec gives when each thread of a block holds one slot.
// acc(q, v) adds slot q's terms to v.
template <int NV, class Acc>
HBTX_DEV void wave_reduce_tile(int b, int e, double tot[RED_NV], Acc acc)
{
#pragma unroll
  for (int k = 0; k < RED_NV; k++)
    tot[k] = 0.;
  for (int w = 0; w < RED_THREADS / HBTX_WAVE && b + w * HBTX_WAVE < e; w++)
  {
    double v[RED_NV];
#pragma unroll
    for (int k = 0; k < RED_NV; k++)
      v[k] = 0.;
    const int q = b + w * HBTX_WAVE + lane_id();
    if (q < e)
      acc(q, v);
#pragma unroll
    for (int k = 0; k < NV; k++)
    {
      double x = v[k];
#pragma unroll
      for (int d = 32; d >= 1; d >>= 1)
        x += __shfl_down(x, d, HBTX_WAVE);
      tot[k] += x;
    }
  }
}

__global__ void k_red_ntiles(const int32_t *len, int32_t nsub, int32_t *ntile)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s < nsub)
    ntile[s] = (len[s] + RED_TILE - 1) / RED_TILE;
}

__global__ void k_red_tiles(const int32_t *begin, const int32_t *len, const int32_t *toff, int32_t nsub, int4 *tiles)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsub)
    return;
  const int b = begin[s], e = b + len[s];
  int t = toff[s];
  for (int x = b; x < e; x += RED_TILE, t++)
    tiles[t] = make_int4(s, x, min(e, x + RED_TILE), 0);
}

// Centre-of-mass frame of the bound candidates (CentreOfMassReferenceFrame, :151-198),
// summed in input order over particles still flagged as candidates:
// v[0]=M, v[1..3]=sum m x (float products, or double NEAREST offsets when periodic), v[4..6]=sum m v
struct ComOp
{
  static constexpr int NV = 7; // mass, mass-weighted position, mass-weighted velocity
  HBTX_DEV static void acc(const SubArrays &S, const PartArrays &Q, const DevParams &P, int s, int q, double v[RED_NV])
  {
    const int p = q;
    if (Q.state[p] != 0)
      return;
    const float4 x = Q.posm[p];
    const float4 u = Q.velu[p];
    const float m = x.w;
    v[0] += m;
    if (P.periodic)
    {
      const float4 r0 = Q.posm[S.first_idx[s]];
      v[1] += nearest_d((double)x.x - (double)r0.x, P.box, P.box_half) * m;
      v[2] += nearest_d((double)x.y - (double)r0.y, P.box, P.box_half) * m;
      v[3] += nearest_d((double)x.z - (double)r0.z, P.box, P.box_half) * m;
    }
    else
    {
      v[1] += x.x * m;
      v[2] += x.y * m;
      v[3] += x.z * m;
    }
    v[4] += u.x * m;
    v[5] += u.y * m;
    v[6] += u.z * m;
  }
};

template <class Op>
__global__ __launch_bounds__(RED_THREADS) void k_red_tile(SubArrays S, PartArrays Q, DevParams P, const int4 *tiles,
                                                         const int32_t *ntiles_p, double *partial)
{
  __shared__ double lds[(RED_THREADS / HBTX_WAVE) * RED_NV];
  const int nt = *ntiles_p;
  // tiles of at most RED_THREADS slots (most subhaloes): one wave each
  const int wpb = RED_THREADS / HBTX_WAVE;
  for (int t = uniform((int)(blockIdx.x * wpb + threadIdx.x / HBTX_WAVE)); t < nt; t += gridDim.x * wpb)
  {
    const int4 tl = tiles[t];
    if (tl.z - tl.y > RED_THREADS)
      continue;
    double v[RED_NV];
    wave_reduce_tile<Op::NV>(tl.y, tl.z, v, [&](int q, double *vv) { Op::acc(S, Q, P, tl.x, q, vv); });
    if (lane_id() == 0)
      for (int k = 0; k < RED_NV; k++)
        partial[(int64_t)t * RED_NV + k] = v[k];
  }
  // larger tiles: the whole block
  for (int t = blockIdx.x; t < nt; t += gridDim.x)
  {
    const int4 tl = tiles[t];
    if (tl.z - tl.y <= RED_THREADS)
      continue;
    double v[RED_NV];
#pragma unroll
    for (int k = 0; k < RED_NV; k++)
      v[k] = 0.;
    for (int q = tl.y + threadIdx.x; q < tl.z; q += RED_THREADS)
      Op::acc(S, Q, P, tl.x, q, v);
    block_reduce_vec<Op::NV>(v, lds, wave_in_tile(tl.y, tl.z));
    if (threadIdx.x == 0)
      for (int k = 0; k < RED_NV; k++)
        partial[(int64_t)t * RED_NV + k] = v[k];
  }
}

// One thread per (subhalo, value): the tiles' partial sums in tile order, with
// reads and writes contiguous across the RED_NV values of a subhalo.
__global__ void k_red_finalize(const int32_t *toff, const int32_t *ntile, int32_t nsub, const double *partial,
                               double *out)
{
  const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= (int64_t)nsub * RED_NV)
    return;
  const int s = (int)(i / RED_NV), k = (int)(i - (int64_t)s * RED_NV);
  const int t0 = toff[s], n = ntile[s];
  double v = 0.;
  for (int t = t0; t < t0 + n; t++)
    v += partial[(int64_t)t * RED_NV + k];
  out[i] = v;
}

inline int red_finalize_blocks(int32_t nsub)
{
  return (int)(((int64_t)nsub * RED_NV + 255) / 256);
}

// Reduction range of the CoM: bound candidates of subhaloes still active after the decision.
__global__ void k_com_ranges(SubArrays S, int32_t *begin, int32_t *len)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  begin[s] = S.off[s];
  len[s] = (S.mode[s] != MODE_IDLE && S.status[s] == ST_ACTIVE) ? (S.off[s + 1] - S.off[s]) : 0;
}

// smallest input index among bound candidates (the E list's first particle)
__global__ void k_first_index(SubArrays S, PartArrays Q, int64_t n)
{
  GRID_STRIDE(q, n)
  {
    const int s = S.sub_of[q];
    const int k = (int)(q - S.off[s]);
    if (S.mode[s] != MODE_IDLE && S.status[s] == ST_ACTIVE && k < S.nbound[s])
      atomicMin(S.first_idx + s, Q.cur[q]);
  }
}

// Frame update and convergence test (:687-690).
__global__ void k_round_finish(SubArrays S, DevParams P, const double *red)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub || S.mode[s] == MODE_IDLE || S.status[s] != ST_ACTIVE)
    return;
  const double *v = red + (int64_t)s * RED_NV;
  const double M = v[0];
  double wp[3] = {v[1] / M, v[2] / M, v[3] / M};
  if (P.periodic)
  {
    // Q.posm is not visible here; the reference position is added on the host-side
    // kernel k_round_finish_periodic.
  }
  for (int k = 0; k < 3; k++)
  {
    S.ref[6 * s + k] = (float)wp[k];
    S.ref[6 * s + 3 + k] = (float)(v[4 + k] / M);
  }
  S.mbound[s] = (float)M;
  const int64_t nb = S.nbound[s];
  const int64_t nlast = S.nlast[s];
  // (no convergence before the MODE_ADD round has added the outside particles' potential)
  if (!S.reset[s] && S.mode[s] != MODE_ADD && S.warm[s] != WARM_ADD && nb >= nlast * P.prec)
    S.status[s] = ST_CONVERGED;
  else
  {
    atomicAdd(S.active_count, 1);
    atomicMax(S.max_active_n, S.off[s + 1] - S.off[s]);
  }
}

__global__ void k_round_finish_periodic(SubArrays S, PartArrays Q, DevParams P, const double *red)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub || S.mode[s] == MODE_IDLE || S.status[s] == ST_DISRUPTED)
    return;
  const double *v = red + (int64_t)s * RED_NV;
  const double M = v[0];
  const float4 r0 = Q.posm[S.first_idx[s]];
  S.ref[6 * s + 0] = (float)(v[1] / M + (double)r0.x);
  S.ref[6 * s + 1] = (float)(v[2] / M + (double)r0.y);
  S.ref[6 * s + 2] = (float)(v[3] / M + (double)r0.z);
}

// ---------------------------------------------------------------------------
// Final ordering and per-subhalo properties
// ---------------------------------------------------------------------------

// Sort key of every particle: bound first by energy, then the unbound batches
// with the latest round first, each by energy (:673, :710, :744-750).
// group above the energy bits: 0 for the bound particles, then the unbound
// batches latest first (removed in round r of `rounds`: rounds + 1 - r)
__global__ void k_final_keys(SubArrays S, PartArrays Q, int64_t n, int32_t rounds, uint64_t *keys, int32_t *vals)
{
  GRID_STRIDE(p, n)
  {
    const int s = S.sub_of[p];
    uint64_t key;
    if (S.status[s] == ST_CONVERGED)
    {
      const int st = Q.state[p];
      const uint64_t group = (st == 0) ? 0ull : (uint64_t)(rounds + 1 - st);
      key = (group << 32) | (uint64_t)float_to_ordered(Q.energy[p]);
    }
    else
      key = (uint64_t)(p - S.off[s]);
    keys[p] = key;
    vals[p] = (int32_t)p;
  }
}

__global__ void k_final_segments(SubArrays S, int32_t *seg_begin, int32_t *seg_len)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  seg_begin[s] = S.off[s];
  seg_len[s] = (S.status[s] == ST_CONVERGED) ? (S.off[s + 1] - S.off[s]) : 0;
}

// AverageKinematics (:200-249) and CountParticleTypes sums over the bound particles
// of converged subhaloes (input order): v[0]=sum E m, v[1]=sum u m, v[2]=sum K m (2x),
// v[3..5]=L m, v[6]=M, v[7..12]=per-type mass, v[13..18]=per-type count.

#define FIN_BATCH 4
// TYPED false (no particle types: every particle is type 1): the per-type sums are
// the type-1 mass, equal to v[6] term for term, and the count, so only those
// eight sums are carried (fewer registers) and the other type sums are zero.
template <bool TYPED>
__global__ __launch_bounds__(RED_THREADS) void k_final_tile(SubArrays S, PartArrays Q, DevParams P, const int4 *tiles,
                                                           const int32_t *ntiles_p, double *partial)
{
  __shared__ double lds[(RED_THREADS / HBTX_WAVE) * RED_NV];
  const int nt = *ntiles_p;
  for (int t = blockIdx.x; t < nt; t += gridDim.x)
  {
    const int4 tl = tiles[t];
    const int s = tl.x;
    double v[RED_NV];
#pragma unroll
    for (int k = 0; k < RED_NV; k++)
      v[k] = 0.;
    const float *rf = S.ref + 6 * s;
    // FIN_BATCH of this thread's slots at a time (the same order): their states,
    // then the bound ones' data, all loads of a batch in flight together
    for (int p0 = tl.y + threadIdx.x; p0 < tl.z; p0 += FIN_BATCH * RED_THREADS)
    {
      bool in[FIN_BATCH];
      float4 xb[FIN_BATCH], ub[FIN_BATCH];
      float eb[FIN_BATCH];
      int tb[FIN_BATCH];
#pragma unroll
      for (int r = 0; r < FIN_BATCH; r++)
      {
        const int p = p0 + r * RED_THREADS;
        in[r] = p < tl.z && Q.state[p] == 0;
      }
#pragma unroll
      for (int r = 0; r < FIN_BATCH; r++)
      {
        const int p = in[r] ? p0 + r * RED_THREADS : p0;
        xb[r] = Q.posm[p];
        ub[r] = Q.velu[p];
        eb[r] = Q.energy[p];
        tb[r] = TYPED ? ptype_of(Q, p) : 1;
      }
#pragma unroll
      for (int r = 0; r < FIN_BATCH; r++)
      {
        if (!in[r])
          continue;
        const float4 x = xb[r];
        const float4 u = ub[r];
        const float m = x.w;
        const int ty = tb[r];
        v[0] += eb[r] * m;
        if (P.thermal)
          v[1] += u.w * m;
        double dx[3] = {(double)(x.x - rf[0]), (double)(x.y - rf[1]), (double)(x.z - rf[2])};
        double dv[3];
        const double xv[3] = {u.x, u.y, u.z};
        for (int j = 0; j < 3; j++)
        {
          if (P.periodic)
            dx[j] = nearest_d(dx[j], P.box, P.box_half);
          dx[j] *= P.a;
          dv[j] = (double)((float)xv[j] - rf[3 + j]) + P.hz * dx[j];
          v[2] += dv[j] * dv[j] * m;
        }
        v[3] += (dx[1] * dv[2] - dx[2] * dv[1]) * m;
        v[4] += (dx[2] * dv[0] - dx[0] * dv[2]) * m;
        v[5] += (dx[0] * dv[1] - dx[1] * dv[0]) * m;
        v[6] += m;
        if (TYPED)
        {
#pragma unroll
          for (int k = 0; k < 6; k++)
          {
            v[7 + k] += (ty == k) ? (double)m : 0.;
            v[13 + k] += (ty == k) ? 1.0 : 0.;
          }
        }
        else
          v[7] += 1.0;
      }
    }
    if (TYPED)
      block_reduce_vec<19>(v, lds, wave_in_tile(tl.y, tl.z)); // v[0..18]
    else
    {
      block_reduce_vec<8>(v, lds, wave_in_tile(tl.y, tl.z)); // v[0..7]
      const double cnt = v[7];
#pragma unroll
      for (int k = 7; k < 19; k++)
        v[k] = 0.;
      v[8] = v[6];
      v[14] = cnt;
    }
    if (threadIdx.x == 0)
      for (int k = 0; k < RED_NV; k++)
        partial[(int64_t)t * RED_NV + k] = v[k];
  }
}

// TracerIndex: first tracer in final order (CountParticleTypes, src/subhalo.cpp:311-392);
// one wave per converged subhalo.
__global__ void k_first_tracer(SubArrays S, PartArrays Q, DevParams P, const int32_t *vals, int32_t *first_tracer)
{
  const int s = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / HBTX_WAVE);
  if (s >= S.nsub)
    return;
  int ft = 0x7fffffff;
  if (S.status[s] == ST_CONVERGED)
  {
    const int b = S.off[s], nb = S.nbound[s];
    for (int base = 0; base < nb; base += HBTX_WAVE)
    {
      const int k = base + lane_id();
      const bool tr = k < nb && is_tracer(P, ptype_of(Q, vals[b + k]));
      const uint64_t m = ballot(tr);
      if (m)
      {
        ft = base + __ffsll((long long)m) - 1;
        break;
      }
    }
  }
  if (lane_id() == 0)
    first_tracer[s] = ft;
}

// Core phase space (GetCorePhaseSpaceProperties, src/subhalo_merge.cpp:94-181):
// the first CoreSize tracers in final order, then non-tracers; one wave per subhalo
// walks the final order in chunks of 64.  out[13*s..]: m, m x[3], m x^2[3], m v[3], m v^2[3].
__global__ void k_core_wave(SubArrays S, PartArrays Q, DevParams P, const int32_t *vals, double *out)
{
  const int s = (int)(((int64_t)blockIdx.x * blockDim.x + threadIdx.x) / HBTX_WAVE);
  if (s >= S.nsub)
    return;
  double v[13];
  for (int k = 0; k < 13; k++)
    v[k] = 0.;
  if (S.status[s] == ST_CONVERGED && S.nbound[s] > 0)
  {
    const int b = S.off[s], nb = S.nbound[s];
    const int64_t core = max((int64_t)(nb * P.core_factor), P.core_min);
    const float4 o = Q.posm[vals[b]];
    const double os[3] = {o.x, o.y, o.z};
    int64_t num = 0;
    for (int pass = 0; pass < 2 && num < core; pass++)
      for (int base = 0; base < nb && num < core; base += HBTX_WAVE)
      {
        const int k = base + lane_id();
        const int p = (k < nb) ? vals[b + k] : 0;
        const bool tr = (k < nb) && is_tracer(P, ptype_of(Q, p));
        const bool want = (k < nb) && (pass == 0 ? tr : !tr);
        const uint64_t m = ballot(want);
        const int64_t rank = __popcll(m & lanemask_lt());
        if (want && num + rank < core)
        {
          const float4 x = Q.posm[p];
          const float4 u = Q.velu[p];
          const double ms = x.w;
          const double xs[3] = {x.x, x.y, x.z}, vs[3] = {u.x, u.y, u.z};
          v[0] += ms;
          for (int d = 0; d < 3; d++)
          {
            double dx = xs[d];
            if (P.periodic)
              dx = nearest_d(dx - os[d], P.box, P.box_half);
            v[1 + d] += ms * dx;
            v[4 + d] += ms * dx * dx;
            v[7 + d] += ms * vs[d];
            v[10 + d] += ms * vs[d] * vs[d];
          }
        }
        num += min((int64_t)__popcll(m), core - num);
      }
  }
  for (int k = 0; k < 13; k++)
  {
    double x = v[k];
    for (int d = 32; d >= 1; d >>= 1)
      x += __shfl_down(x, d, HBTX_WAVE);
    if (lane_id() == 0)
      out[(int64_t)RED_NV * s + k] = x;
  }
}

__global__ void k_final_ranges(SubArrays S, int32_t *begin, int32_t *len)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  begin[s] = S.off[s];
  len[s] = (S.status[s] == ST_CONVERGED) ? (S.off[s + 1] - S.off[s]) : 0;
}

// ParticlePotentialEnergies (GetPotentialEnergy, :114-135) of the bound part.
__global__ void k_potential_energy(SubArrays S, PartArrays Q, DevParams P, int64_t n, const int32_t *vals, float *out)
{
  GRID_STRIDE(q, n)
  {
    const int s = S.sub_of[q];
    const int k = (int)(q - S.off[s]);
    float E = 0.f;
    if (S.status[s] == ST_CONVERGED && k < S.nbound[s])
    {
      const int p = vals[q];
      const float *rf = S.ref + 6 * s;
      const float4 x = Q.posm[p];
      const float4 u = Q.velu[p];
      E = Q.energy[p];
      if (P.thermal)
        E -= u.w;
      const float xs[3] = {x.x, x.y, x.z}, vs[3] = {u.x, u.y, u.z};
      for (int j = 0; j < 3; j++)
      {
        double dx = (double)(xs[j] - rf[j]);
        if (P.periodic)
          dx = nearest_d(dx, P.box, P.box_half);
        dx *= P.a;
        const double dv = (double)(vs[j] - rf[3 + j]) + P.hz * dx;
        E = (float)((double)E - 0.5 * dv * dv);
      }
    }
    out[q] = E;
  }
}

} // namespace hbtx
