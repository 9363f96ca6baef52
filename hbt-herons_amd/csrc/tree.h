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
// Entries: ent = centre of mass / position and mass, ent_depth = depth (cells)
// or -1 (particles), eskip = skip (cells).  The walk records (wt) carry, per
// cell, the squared opening radius len^2/theta^2 and the skip as a byte offset;
// a particle record has no opening radius (never opened), so the walk treats
// both uniformly.
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
  float4 *ent;               // per entry: (x, y, z, m)
  int32_t *eskip;            // per cell entry: first entry after its subtree
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

__global__ void k_tree_count(TreeArgs A, int32_t *__restrict__ cnt)
{
  const int S = *A.s_total;
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
      lv = make_char2((signed char)cprev, (signed char)cj);
    }
    A.lev[j] = lv;
    cnt[j] = c;
  }
}

HBTX_DEV void tree_emit_one(const TreeArgs &A, int j)
{
  const int p = A.src[j];
  const int s = A.src_sub[j];
  const int P = j + A.cscan[j];
  const float4 x = A.posm[p];
  const float m = source_mass(A, p, s);
  const int b = A.src_off[s], e = b + A.nsrc[s];
  const int u0 = wrec_unit(b, A.cscan[b]), uj = wrec_unit(j, A.cscan[j]);
  if (!A.treemode[s])
  {
    if (P < A.cap)
    {
      A.ent[P] = make_float4(x.x, x.y, x.z, m);
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
  A.ent[q] = make_float4(x.x, x.y, x.z, m);
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
    A.eskip[pos] = skip; // the moments pass writes the cell's ent
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

// Cell moments (mass and centre of mass) bottom-up, in child (sub-index) order
// with double accumulation of the float child values, as UpdateInternalNodes /
// ProcessNode do (src/gravity_tree.cpp:25-80), without one launch per depth.
//
// The entries are cut into windows of MOM_WIN consecutive entries.
// k_moments_windows stages one window's entries in LDS and forms, depth by depth
// with block barriers, the moments of every cell whose subtree [q, skip) lies in
// the window.  A cell that starts in a window and ends beyond it contains the
// window's last entry, so such cells are nested, at most one per depth: the
// window records them in its MOM_SLOTS slots (-1 where none).
// k_moments_groups takes MOM_GROUP consecutive units of the level below (windows,
// then groups of windows) per block and forms the recorded cells that end in its
// range, depth by depth: their children are particles, cells of one window, or
// recorded cells one depth deeper that start in the same range.  Cells that end
// beyond the range are again nested and go up one level.  The last level covers
// every entry, so every cell is formed exactly once.
#define MOM_WIN 1024
#define MOM_GROUP 256
#define MOM_SLOTS (HBTX_KEY_LEVELS + 1)

HBTX_DEV void store_cell(const TreeArgs &A, int q, float fx, float fy, float fz, float fm)
{
  A.ent[q] = make_float4(fx, fy, fz, fm);
  const int u = A.cell_u[q];
  A.wt[u] = make_uint2(__float_as_uint(fx), __float_as_uint(fy));
  A.wt[u + 1] = make_uint2(__float_as_uint(fz), __float_as_uint(fm) | 0x80000000u); // sign bit: a cell
}

#define MOM_THREADS 256
#define MOM_PER (MOM_WIN / MOM_THREADS)
__global__ __launch_bounds__(MOM_THREADS) void k_moments_windows(TreeArgs A, const int32_t *__restrict__ e_total_p,
                                                                 int32_t *__restrict__ slots)
{
  __shared__ float4 xm[MOM_WIN];
  __shared__ int32_t sk[MOM_WIN];
  __shared__ int32_t ws[MOM_SLOTS];
  __shared__ uint32_t depths_s;
  const int E = *e_total_p;
  const int nwin = (E + MOM_WIN - 1) / MOM_WIN;
  const int t = threadIdx.x;
  // every block runs the same number of iterations (the barriers are uniform)
  for (int w = blockIdx.x; w < nwin; w += gridDim.x)
  {
    const int a = w * MOM_WIN, nw = min(MOM_WIN, E - a);
    if (t < MOM_SLOTS)
      ws[t] = -1;
    if (t == 0)
      depths_s = 0u;
    __syncthreads();
    // stage the window: this thread's entries i = r * MOM_THREADS + t; the
    // depth and subtree end of each stay in registers (-1: a particle, or a cell
    // that ends beyond the window)
    int dep[MOM_PER], end[MOM_PER];
    uint32_t depths = 0u;
#pragma unroll
    for (int r = 0; r < MOM_PER; r++)
    {
      const int i = r * MOM_THREADS + t;
      dep[r] = -1;
      end[r] = 0;
      if (i < nw)
      {
        const float4 v = A.ent[a + i];
        const int ks = A.eskip[a + i]; // (unset for a particle)
        const int d = A.ent_depth[a + i];
        const int k = (d >= 0) ? min(ks, E) - a : i + 1;
        xm[i] = v;
        sk[i] = k;
        end[r] = k;
        if (d >= 0)
        {
          if (k <= nw)
          {
            dep[r] = d;
            depths |= 1u << d;
          }
          else
            ws[d] = a + i;
        }
      }
    }
#pragma unroll
    for (int o = 1; o < HBTX_WAVE; o <<= 1)
      depths |= (uint32_t)__shfl_xor((int)depths, o, HBTX_WAVE);
    if (lane_id() == 0 && depths)
      atomicOr(&depths_s, depths);
    __syncthreads();
    // the depths present, deepest first: a cell's child cells are one depth deeper
    uint32_t todo = depths_s;
    while (todo)
    {
      const int d = 31 - __clz((int)todo);
      todo &= ~(1u << d);
#pragma unroll
      for (int r = 0; r < MOM_PER; r++)
      {
        if (dep[r] != d)
          continue;
        const int i = r * MOM_THREADS + t;
        double m = 0., cx = 0., cy = 0., cz = 0.;
        for (int c = i + 1; c < end[r]; c = max(sk[c], c + 1))
        {
          const float4 v = xm[c];
          const double cm = v.w;
          m += cm;
          cx += (double)v.x * cm;
          cy += (double)v.y * cm;
          cz += (double)v.z * cm;
        }
        xm[i] = make_float4((float)(cx / m), (float)(cy / m), (float)(cz / m), (float)m);
      }
      __syncthreads();
    }
    // the window's cells to HBM, once
#pragma unroll
    for (int r = 0; r < MOM_PER; r++)
      if (dep[r] >= 0)
      {
        const int i = r * MOM_THREADS + t;
        const float4 v = xm[i];
        store_cell(A, a + i, v.x, v.y, v.z, v.w);
      }
    if (t < MOM_SLOTS)
      slots[(int64_t)w * MOM_SLOTS + t] = ws[t];
    __syncthreads(); // the next window reuses the LDS
  }
}

// One level above the windows: units of unit_span entries below, MOM_GROUP units a block.
__global__ __launch_bounds__(MOM_GROUP) void k_moments_groups(TreeArgs A, const int32_t *__restrict__ e_total_p,
                                                              int64_t unit_span, const int32_t *__restrict__ slots_in,
                                                              int32_t *__restrict__ slots_out)
{
  __shared__ int32_t ws[MOM_SLOTS];
  const int E = *e_total_p;
  const int64_t gspan = unit_span * MOM_GROUP;
  const int64_t nunits = (E + unit_span - 1) / unit_span;
  const int64_t ngroups = (E + gspan - 1) / gspan;
  for (int64_t g = blockIdx.x; g < ngroups; g += gridDim.x)
  {
    const int64_t gend = min((int64_t)E, (g + 1) * gspan);
    if (threadIdx.x < MOM_SLOTS)
      ws[threadIdx.x] = -1;
    __syncthreads();
    const int64_t u = g * MOM_GROUP + threadIdx.x;
    for (int d = MOM_SLOTS - 1; d >= 0; d--)
    {
      const int q = (u < nunits) ? slots_in[u * MOM_SLOTS + d] : -1;
      if (q >= 0)
      {
        const int end = min(A.eskip[q], E);
        if (end > gend)
          ws[d] = q;
        else
        {
          double m = 0., cx = 0., cy = 0., cz = 0.;
          int c = q + 1;
          while (c < end)
          {
            const float4 en = A.ent[c];
            const double cm = en.w;
            m += cm;
            cx += (double)en.x * cm;
            cy += (double)en.y * cm;
            cz += (double)en.z * cm;
            const int dc = A.ent_depth[c], kc = A.eskip[c]; // (both loads in flight together)
            c = (dc >= 0) ? max(kc, c + 1) : c + 1;
          }
          store_cell(A, q, (float)(cx / m), (float)(cy / m), (float)(cz / m), (float)m);
        }
      }
      __syncthreads(); // this depth's cells are written before the next shallower depth reads them
    }
    if (threadIdx.x < MOM_SLOTS)
      slots_out[g * MOM_SLOTS + threadIdx.x] = ws[threadIdx.x];
    __syncthreads();
  }
}

} // namespace hbtx
