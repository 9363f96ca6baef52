// GPU Barnes-Hut octree with the reference's structure (single-particle leaves,
// bounding cube, len halving rule), laid out in depth-first preorder.
//
// Sources of every subhalo are contiguous in a dense list and sorted by their
// 63-bit octree key (tree cube of that subhalo).  For sorted sources j of one
// segment, c_j = number of leading levels shared by keys j and j+1; the cells
// of the reference octree are exactly the key prefixes shared by >= 2 sources,
// and the cells that start at source j are the depths (c_{j-1}, c_j].  With
// cnt_j = max(0, c_j - c_{j-1}) and cscan its exclusive prefix sum, entry
// positions are closed-form:
//   cell of depth l starting at j : j + cscan[j] + (l - c_{j-1} - 1)
//   particle j                    : j + cscan[j] + cnt_j
// and every entry's "skip" (first entry after its subtree: the reference's
// sibling link, src/gravity_tree.cpp:50-80) is the position of the entry group
// that follows the last source of its key range.
//
// Entry records (struct Entry): centre of mass / position, mass, squared
// opening radius len^2/theta^2, len^2, skip; ent_depth = depth (cells) or -1
// (particles).  A particle is an entry with opening radius 0 (never opened), so
// the walk treats both uniformly.
#pragma once
#include "common.h"

namespace hbtx
{

struct TreeArgs
{
  const int32_t *src;        // dense source list (global particle index), sorted per segment
  const uint64_t *keys;      // sorted keys, aligned with src
  const int32_t *src_sub;    // subhalo of every dense source
  const int32_t *src_off;    // per subhalo: first dense source
  const int32_t *nsrc;       // per subhalo: number of sources
  const int32_t *treemode;   // per subhalo: 1 tree, 0 direct (particle entries only)
  const int32_t *mode;       // per subhalo: MODE_FULL / MODE_CORR
  const double *len0;        // per subhalo: root len (cube side, src/oct_tree.tpp:357-360)
  const float4 *posm;        // particle position + mass
  const int8_t *ptype;       // particle type or null
  const float *upscale;      // per subhalo [6] mass upscale factors
  char2 *lev;                // per dense source: (c_{j-1}, c_j) of its tree, (0, 0) in direct mode
  const int32_t *cnt;        // cells starting at each dense source
  const int32_t *cscan;      // exclusive scan of cnt (size >= S+1)
  const int32_t *s_total;    // number of dense sources
  Entry *ent;
  int8_t *ent_depth;
  uint2 *wt;                 // walk records (units of 8 bytes)
  int32_t *cell_u;           // per entry index: unit of a cell's walk record
  int32_t cap;               // entry capacity (walk records: 3 cap units)
  int32_t *overflow;         // bit 1 set if entries would exceed cap
  DevParams P;
};

// Source mass as the reference tree sees it: GetMass with the upscale factors
// during a full iteration (src/subhalo_unbind.cpp:91-101), the true mass in a
// correction tree (a fresh EnergySnapshot_t, :573-575).
HBTX_DEV float source_mass(const TreeArgs &A, int p, int s)
{
  const float m = A.posm[p].w;
  if (A.mode[s] != MODE_FULL)
    return m;
  const float *up = A.upscale + 6 * s;
  if (A.P.dm_only)
    return m * up[1];
  const int t = A.ptype ? A.ptype[p] : 1;
  if (donot_subsample(A.P, t))
    return m;
  return m * up[t];
}

HBTX_DEV void pair_levels(const TreeArgs &A, int j, int b, int e, int &cprev, int &cj)
{
  cj = (j + 1 < e) ? common_levels(A.keys[j], A.keys[j + 1]) : ((j == b) ? 0 : -1);
  cprev = (j > b) ? common_levels(A.keys[j - 1], A.keys[j]) : -1;
}

// (also records the deepest cell level in *maxdepth)
__global__ void k_tree_count(TreeArgs A, int32_t *__restrict__ cnt, int32_t *maxdepth)
{
  const int S = *A.s_total;
  int depth = 0;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < S; j += gridDim.x * blockDim.x)
  {
    int c = 0;
    const int s = A.src_sub[j];
    char2 lv = make_char2(0, 0);
    if (A.treemode[s])
    {
      const int b = A.src_off[s], e = b + A.nsrc[s];
      int cprev, cj;
      pair_levels(A, j, b, e, cprev, cj);
      c = max(0, cj - cprev);
      if (c > 0)
        depth = max(depth, cj);
      lv = make_char2((signed char)cprev, (signed char)cj);
    }
    A.lev[j] = lv;
    cnt[j] = c;
  }
  // deepest cell level: wave maximum, and an atomic only from waves that raise the
  // current value (one shared word: an atomic from every wave serialises)
#pragma unroll
  for (int d = 1; d < HBTX_WAVE; d <<= 1)
    depth = max(depth, __shfl_xor(depth, d, HBTX_WAVE));
  if (lane_id() == 0 && depth > 0 && depth > __hip_atomic_load(maxdepth, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT))
    atomicMax(maxdepth, depth);
}

HBTX_DEV void tree_emit_one(const TreeArgs &A, int j)
{
  const int p = A.src[j];
  const int s = A.src_sub[j];
  const int P = j + A.cscan[j];
  const float4 x = A.posm[p];
  const float m = source_mass(A, p, s);
  const int b = A.src_off[s], e = b + A.nsrc[s];
  const int t0 = b + A.cscan[b]; // first entry of this subhalo's tree
  const int u0 = wrec_unit(b, A.cscan[b]), uj = wrec_unit(j, A.cscan[j]);
  if (!A.treemode[s])
  {
    if (P < A.cap)
    {
      A.ent[P] = Entry{x.x, x.y, x.z, m, 0.f, (P + 1 - t0) * (int)sizeof(Entry), 0.f, P + 1};
      A.ent_depth[P] = -1;
      A.wt[uj] = make_uint2(__float_as_uint(x.x), __float_as_uint(x.y));
      A.wt[uj + 1] = make_uint2(__float_as_uint(x.z), __float_as_uint(m));
    }
    else
      atomicOr(A.overflow, 1);
    return;
  }
  const char2 lv = A.lev[j];
  const int cprev = lv.x, cj = lv.y;
  const int nn = max(0, cj - cprev);
  const int q = P + nn;
  if (q >= A.cap)
  {
    atomicOr(A.overflow, 1);
    return;
  }
  A.ent[q] = Entry{x.x, x.y, x.z, m, 0.f, (q + 1 - t0) * (int)sizeof(Entry), 0.f, q + 1};
  A.ent_depth[q] = -1;
  const int up = uj + 3 * nn;
  A.wt[up] = make_uint2(__float_as_uint(x.x), __float_as_uint(x.y));
  A.wt[up + 1] = make_uint2(__float_as_uint(x.z), __float_as_uint(m));
  if (nn == 0)
    return;
  // len of the first depth, then the reference halving rule (src/gravity_tree.cpp:38-41)
  double L = A.len0[s];
  for (int l = 0; l < cprev + 1; l++)
    L = (L >= A.P.res) ? L * 0.5 : L;
  const uint64_t kj = A.keys[j];
  int hi_l = e; // cells nest: a deeper cell ends no later than the one above it
  for (int l = cprev + 1; l <= cj; l++)
  {
    const int pos = P + (l - cprev - 1);
    int bend;
    if (l == 0)
      bend = e - 1;
    else
    {
      // the cell's last source: gallop from j (most cells are small), then bisect
      const uint64_t pj = key_prefix(kj, l);
      int lo = j + 1, hi = hi_l;
      for (int span = 1; lo < hi; span <<= 1)
      {
        const int probe = min(j + span, hi - 1);
        if (key_prefix(A.keys[probe], l) > pj)
        {
          hi = probe;
          break;
        }
        lo = probe + 1;
      }
      while (lo < hi)
      {
        const int mid = (lo + hi) >> 1;
        if (key_prefix(A.keys[mid], l) > pj)
          hi = mid;
        else
          lo = mid + 1;
      }
      bend = lo - 1;
      hi_l = bend + 1;
    }
    const int skip = (bend + 1) + A.cscan[bend + 1];
    const float lf = (float)L;
    const float len2 = lf * lf;
    // open if len^2 > r^2 theta^2 (src/gravity_tree.cpp:139)  <=>  r^2 < len^2 / theta^2
    const float crit2 = (A.P.theta2 > 0.f) ? len2 / A.P.theta2 : __int_as_float(0x7f800000);
    A.ent[pos] = Entry{0.f, 0.f, 0.f, 0.f, crit2, (skip - t0) * (int)sizeof(Entry), len2, skip};
    A.ent_depth[pos] = (int8_t)l;
    // walk record: position and mass come with the moments
    const int uc = uj + 3 * (l - cprev - 1);
    A.cell_u[pos] = uc;
    A.wt[uc + 2] = make_uint2(__float_as_uint(crit2),
                              (uint32_t)(8 * (wrec_unit(bend + 1, A.cscan[bend + 1]) - u0)));
    L = (L >= A.P.res) ? L * 0.5 : L;
  }
}

__global__ void k_tree_emit(TreeArgs A)
{
  const int S = *A.s_total;
  for (int j = blockIdx.x * blockDim.x + threadIdx.x; j < S; j += gridDim.x * blockDim.x)
    tree_emit_one(A, j);
}

// Cells bucketed by depth, so that each depth's moment pass touches only its own
// cells.  Pass 1 counts cells per depth (block histograms in LDS); after the
// offsets, pass 2 writes every cell's index into its depth's bucket: a wave
// takes one LDS place per distinct depth among its lanes (ballot match), the
// block reserves one range per depth, and each lane writes at its rank.
#define DEPTH_BINS 32
#define DEPTH_ITEMS 4
__global__ __launch_bounds__(256) void k_depth_count(const int8_t *__restrict__ ent_depth,
                                                    const int32_t *__restrict__ e_total_p, int32_t *__restrict__ cnt)
{
  __shared__ int32_t h[DEPTH_BINS];
  if (threadIdx.x < DEPTH_BINS)
    h[threadIdx.x] = 0;
  __syncthreads();
  const int E = *e_total_p;
  for (int q = blockIdx.x * blockDim.x + threadIdx.x; q < E; q += gridDim.x * blockDim.x)
  {
    const int d = ent_depth[q];
    if (d >= 0)
      atomicAdd(&h[d], 1);
  }
  __syncthreads();
  if (threadIdx.x < DEPTH_BINS && h[threadIdx.x])
    atomicAdd(cnt + threadIdx.x, h[threadIdx.x]);
}

// off[d] = first slot of depth d (exclusive prefix of cnt), cursor zeroed; one thread.
__global__ void k_depth_offsets(const int32_t *__restrict__ cnt, int32_t *__restrict__ off,
                                int32_t *__restrict__ cursor)
{
  if (threadIdx.x != 0 || blockIdx.x != 0)
    return;
  int run = 0;
  for (int d = 0; d < DEPTH_BINS; d++)
  {
    off[d] = run;
    run += cnt[d];
    cursor[d] = 0;
  }
  off[DEPTH_BINS] = run;
}

__global__ __launch_bounds__(256) void k_depth_scatter(const int8_t *__restrict__ ent_depth,
                                                      const int32_t *__restrict__ e_total_p,
                                                      const int32_t *__restrict__ off, int32_t *__restrict__ cursor,
                                                      int32_t *__restrict__ cells)
{
  __shared__ int32_t h[DEPTH_BINS], base[DEPTH_BINS];
  const int E = *e_total_p;
  const int lane = lane_id();
  const int per_block = DEPTH_ITEMS * 256;
  // every block runs the same number of iterations (the barriers are uniform)
  for (int q0 = blockIdx.x * per_block; q0 < E; q0 += gridDim.x * per_block)
  {
    if (threadIdx.x < DEPTH_BINS)
      h[threadIdx.x] = 0;
    __syncthreads();
    int d[DEPTH_ITEMS], slot[DEPTH_ITEMS];
#pragma unroll
    for (int r = 0; r < DEPTH_ITEMS; r++)
    {
      const int q = q0 + r * 256 + (int)threadIdx.x;
      d[r] = (q < E) ? ent_depth[q] : -1;
      slot[r] = 0;
      uint64_t todo = ballot(d[r] >= 0);
      while (todo)
      {
        const int leader = __ffsll((long long)todo) - 1;
        const int d0 = __shfl(d[r], leader, HBTX_WAVE);
        const uint64_t grp = ballot(d[r] == d0);
        int b = 0;
        if (lane == leader)
          b = atomicAdd(&h[d0], __popcll(grp));
        b = __shfl(b, leader, HBTX_WAVE);
        if (d[r] == d0)
          slot[r] = b + __popcll(grp & lanemask_lt());
        todo &= ~grp;
      }
    }
    __syncthreads();
    if (threadIdx.x < DEPTH_BINS)
      base[threadIdx.x] = h[threadIdx.x] ? atomicAdd(cursor + threadIdx.x, h[threadIdx.x]) : 0;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < DEPTH_ITEMS; r++)
      if (d[r] >= 0)
        cells[off[d[r]] + base[d[r]] + slot[r]] = q0 + r * 256 + (int)threadIdx.x;
    __syncthreads();
  }
}

// Mass and centre of mass of every cell at one depth from its children, in
// child (sub-index) order with double accumulation of the float child values,
// as UpdateInternalNodes/ProcessNode do (src/gravity_tree.cpp:25-80).  Also
// fills the cell's walk record.
__global__ void k_tree_moments(TreeArgs A, const int32_t *__restrict__ cells, const int32_t *__restrict__ off,
                               const int32_t *__restrict__ e_total_p, int depth)
{
  const int E = *e_total_p;
  const int b = off[depth], e = off[depth + 1];
  Entry *__restrict__ ent = A.ent;
  for (int k = b + blockIdx.x * blockDim.x + threadIdx.x; k < e; k += gridDim.x * blockDim.x)
  {
    const int q = cells[k];
    const int end = min(ent[q].skip, E);
    double m = 0., cx = 0., cy = 0., cz = 0.;
    int c = q + 1;
    while (c < end)
    {
      const Entry en = ent[c];
      const double cm = en.m;
      m += cm;
      cx += (double)en.x * cm;
      cy += (double)en.y * cm;
      cz += (double)en.z * cm;
      c = max(en.skip, c + 1);
    }
    const float fx = (float)(cx / m), fy = (float)(cy / m), fz = (float)(cz / m), fm = (float)m;
    ent[q].x = fx;
    ent[q].y = fy;
    ent[q].z = fz;
    ent[q].m = fm;
    const int u = A.cell_u[q];
    A.wt[u] = make_uint2(__float_as_uint(fx), __float_as_uint(fy));
    A.wt[u + 1] = make_uint2(__float_as_uint(fz), __float_as_uint(fm) | 0x80000000u); // sign bit: a cell
  }
}

} // namespace hbtx
