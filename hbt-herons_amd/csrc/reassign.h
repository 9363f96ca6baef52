// Particle reassignment: SubhaloSnapshot_t::ReassignParticles
// (src/subhalo_reassign_gas.cpp:65-197) with the neighbour search of
// GeoTree_t::NearestNeighbours (src/geometric_tree.cpp:108-141) and
// GeoTree_t::Search (:143-210).
//
// The reference searches a growing sphere, radius r_0 = rguess,
// r_{j+1} = (float)(r_j * 1.26), and collects the tracers inside EVERY sphere
// into one bounded max-heap of capacity k, stopping after the first search j = J
// where the heap holds min(n_tracers, k) entries.  The spheres are nested, so a
// tracer t first found at search j_t sits in the heap's input J - j_t + 1 times,
// and the heap ends holding the k smallest entries of that multiset.  Every
// tracer in it is among the k nearest tracers of the particle, so the device
// computes the k nearest tracers once (a union walk over a per-host-halo box
// hierarchy), then j_t for each of them (binary search over the table of radii,
// with the reference's per-axis and sphere tests), J, and the multiset's
// leading k entries.
//
// Layout per host halo g (particles [goff[g], goff[g+1])):
//  * tracers sorted by a Morton key of their offset from the halo's first
//    particle (periodic minimum image), records trec = (x, y, z, particle index bits);
//  * leaves of RS_LEAF consecutive sorted tracers, RS_FAN leaves per node one
//    level up, until one root: levels stored level by level, boxes (lo, hi) in
//    offset coordinates widened by a rounding margin;
//  * non-tracers sorted by the same key and cut into tiles of one wave.
#pragma once
#include "common.h"

namespace hbtx
{
#define RS_LEAF 32
#define RS_FAN 8
#define RS_LEVELS 9 // RS_LEAF * RS_FAN^(RS_LEVELS-1) = 2^29 tracers per host halo
#define RS_RADII 512
#define RS_THREADS 256

struct RsGroup
{
  int32_t tb, te;  // tracer range in the sorted tracer list
  int32_t qb, qe;  // non-tracer range in the sorted query list
  float p0x, p0y, p0z;
  float margin;    // box widening that covers the rounding of offsets
  int32_t wrap;    // offsets may differ from the minimum image (periodic, extent >= box/4)
  int32_t top;     // level of the root (-1: no tracers)
  int32_t pad0, pad1; // (48-byte records)
};

struct RsArgs
{
  int32_t n, ngroups;
  const float *pos;      // [3n]
  const uint8_t *tracer; // [n]
  const int32_t *sub;    // [n]
  const int32_t *goff;   // [G+1]
  float box, half;
  int32_t periodic;
  int32_t k;
  int32_t *gid;          // [n]
  int32_t *tflag;        // [n+1] tracer flags, then their exclusive scan
  RsGroup *grp;          // [G]
  uint32_t *bbox;        // [G*7]: ordered lo xyz, ordered hi xyz, max |offset| bits
  int32_t *cnt;          // [RS_LEVELS*G] nodes per level and host halo
  int32_t *loff;         // [RS_LEVELS*(G+1)] exclusive scans of cnt per level
  int32_t *qtiles_n;     // [G+1] wave tiles per host halo, then their scan
  uint64_t *tkey, *qkey; // sorted keys
  int32_t *tval, *qval;  // particle index per sorted slot
  int32_t *seg;          // [4G]: tracer begins | tracer counts | query begins | query counts
  float4 *trec;          // [nt]
  float4 *nodes;         // 2 per node: lo, hi
  const int64_t *lbase;  // [RS_LEVELS] first node of every level
  int32_t *out;          // [n]
};

HBTX_DEV float rs_offset(float x, float x0, const RsArgs &A)
{
  double d = (double)x - (double)x0;
  if (A.periodic)
    d = nearest_d(d, (double)A.box, (double)A.half);
  return (float)d;
}

HBTX_DEV int rs_group_of(const int32_t *goff, int ng, int i)
{
  int lo = 0, hi = ng; // goff[lo] <= i < goff[hi]
  while (hi - lo > 1)
  {
    const int mid = (lo + hi) >> 1;
    if (goff[mid] <= i)
      lo = mid;
    else
      hi = mid;
  }
  return lo;
}

HBTX_DEV uint64_t rs_spread(uint32_t v)
{
  uint64_t x = v & 0x1fffffu;
  x = (x | (x << 32)) & 0x1f00000000ffffull;
  x = (x | (x << 16)) & 0x1f0000ff0000ffull;
  x = (x | (x << 8)) & 0x100f00f00f00f00full;
  x = (x | (x << 4)) & 0x10c30c30c30c30c3ull;
  x = (x | (x << 2)) & 0x1249249249249249ull;
  return x;
}

// Pass 1 over particles: host halo, tracer flag, output cleared.
__global__ void k_rs_flags(RsArgs A)
{
  GRID_STRIDE(i, A.n)
  {
    A.gid[i] = rs_group_of(A.goff, A.ngroups, (int)i);
    A.tflag[i] = A.tracer[i] ? 1 : 0;
    A.out[i] = -1;
  }
}

// Per host halo: ranges, frame origin, cleared box accumulators, wave tiles.
__global__ void k_rs_groups(RsArgs A)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.ngroups)
    return;
  const int a = A.goff[g], e = A.goff[g + 1];
  RsGroup G;
  G.tb = A.tflag[a];
  G.te = A.tflag[e];
  G.qb = a - G.tb;
  G.qe = e - G.te;
  G.p0x = (e > a) ? A.pos[3 * (int64_t)a] : 0.f;
  G.p0y = (e > a) ? A.pos[3 * (int64_t)a + 1] : 0.f;
  G.p0z = (e > a) ? A.pos[3 * (int64_t)a + 2] : 0.f;
  G.margin = 0.f;
  G.wrap = 0;
  G.top = -1;
  G.pad0 = G.pad1 = 0;
  A.grp[g] = G;
  uint32_t *bb = A.bbox + 7 * (int64_t)g;
  for (int d = 0; d < 3; d++)
  {
    bb[d] = float_to_ordered(__int_as_float(0x7f800000));      // +inf
    bb[3 + d] = float_to_ordered(__int_as_float(0xff800000));  // -inf
  }
  bb[6] = 0u;
  const int nt = G.te - G.tb, nq = G.qe - G.qb;
  A.qtiles_n[g] = (nt > 0) ? (nq + HBTX_WAVE - 1) / HBTX_WAVE : 0;
  A.seg[g] = G.tb;
  A.seg[A.ngroups + g] = nt;
  A.seg[2 * A.ngroups + g] = G.qb;
  A.seg[3 * A.ngroups + g] = nq;
  int c = (nt + RS_LEAF - 1) / RS_LEAF;
  for (int l = 0; l < RS_LEVELS; l++)
  {
    A.cnt[l * A.ngroups + g] = c;
    c = (c > 1) ? (c + RS_FAN - 1) / RS_FAN : 0;
  }
}

HBTX_DEV void rs_wave_min_max(float &mn, float &mx)
{
#pragma unroll
  for (int d = 1; d < HBTX_WAVE; d <<= 1)
  {
    mn = fminf(mn, __shfl_xor(mn, d, HBTX_WAVE));
    mx = fmaxf(mx, __shfl_xor(mx, d, HBTX_WAVE));
  }
}

// Pass 2 over particles: tracer / non-tracer lists (stable), box of the tracers'
// offsets and largest |offset| per host halo.  A block takes RS_CHUNK consecutive
// particles; every thread keeps running extrema for the host halo it is in and
// flushes them with atomics when the host halo changes, and at the end the block
// combines them first when all its threads end in one host halo (the usual case),
// so a host halo receives about one update per chunk.
#define RS_CHUNK (RS_THREADS * 16)
struct RsBox
{
  float l[3], h[3], e;
  HBTX_DEV void reset()
  {
    for (int d = 0; d < 3; d++)
    {
      l[d] = INFINITY;
      h[d] = -INFINITY;
    }
    e = 0.f;
  }
  HBTX_DEV void flush(uint32_t *bb) const
  {
    if (l[0] <= h[0])
      for (int d = 0; d < 3; d++)
      {
        atomicMin(bb + d, float_to_ordered(l[d]));
        atomicMax(bb + 3 + d, float_to_ordered(h[d]));
      }
    atomicMax(bb + 6, __float_as_uint(e));
  }
};

__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(RsArgs A)
{
  __shared__ float red[7][RS_THREADS / HBTX_WAVE];
  __shared__ int first_s, same_s;
  const int64_t nchunk = ((int64_t)A.n + RS_CHUNK - 1) / RS_CHUNK;
  for (int64_t ch = blockIdx.x; ch < nchunk; ch += gridDim.x)
  {
    RsBox B;
    B.reset();
    int cur = -1;
    for (int r = 0; r < RS_CHUNK / RS_THREADS; r++)
    {
      const int64_t i = ch * RS_CHUNK + (int64_t)r * RS_THREADS + threadIdx.x;
      if (i >= A.n)
        break;
      const int g = A.gid[i];
      if (g != cur)
      {
        if (cur >= 0)
          B.flush(A.bbox + 7 * (int64_t)cur);
        B.reset();
        cur = g;
      }
      const RsGroup &G = A.grp[g];
      const bool tr = A.tracer[i] != 0;
      const int s = A.tflag[i];
      if (tr)
        A.tval[s] = (int32_t)i;
      else
        A.qval[i - s] = (int32_t)i;
      const float o[3] = {rs_offset(A.pos[3 * i], G.p0x, A), rs_offset(A.pos[3 * i + 1], G.p0y, A),
                          rs_offset(A.pos[3 * i + 2], G.p0z, A)};
      for (int d = 0; d < 3; d++)
      {
        if (tr)
        {
          B.l[d] = fminf(B.l[d], o[d]);
          B.h[d] = fmaxf(B.h[d], o[d]);
        }
        B.e = fmaxf(B.e, fabsf(o[d]));
      }
    }
    // block combine when every thread ends in the host halo of thread 0
    if (threadIdx.x == 0)
    {
      first_s = cur;
      same_s = 1;
    }
    __syncthreads();
    const int first = first_s;
    if (cur != first)
      same_s = 0; // every writer stores the same value
    __syncthreads();
    const bool same_all = same_s != 0; // uniform over the block
    __syncthreads();
    if (same_all && first >= 0)
    {
      float v[7] = {B.l[0], B.l[1], B.l[2], -B.h[0], -B.h[1], -B.h[2], -B.e};
#pragma unroll
      for (int j = 0; j < 7; j++)
        for (int d = 1; d < HBTX_WAVE; d <<= 1)
          v[j] = fminf(v[j], __shfl_xor(v[j], d, HBTX_WAVE));
      __syncthreads();
      if (lane_id() == 0)
        for (int j = 0; j < 7; j++)
          red[j][threadIdx.x / HBTX_WAVE] = v[j];
      __syncthreads();
      if (threadIdx.x == 0)
      {
        RsBox T;
        for (int d = 0; d < 3; d++)
        {
          T.l[d] = red[d][0];
          T.h[d] = -red[3 + d][0];
        }
        T.e = -red[6][0];
        for (int w = 1; w < RS_THREADS / HBTX_WAVE; w++)
        {
          for (int d = 0; d < 3; d++)
          {
            T.l[d] = fminf(T.l[d], red[d][w]);
            T.h[d] = fmaxf(T.h[d], -red[3 + d][w]);
          }
          T.e = fmaxf(T.e, -red[6][w]);
        }
        T.flush(A.bbox + 7 * (int64_t)first);
      }
    }
    else if (cur >= 0)
      B.flush(A.bbox + 7 * (int64_t)cur);
    __syncthreads();
  }
}

// Morton cube of a host halo: origin and cells per unit length.  Keys use
// RS_KEY_BITS bits per axis: they order the tracers into compact leaves and the
// queries into compact waves (any order is correct; a coarser key only sorts faster).
#ifndef RS_KEY_BITS
#define RS_KEY_BITS 10
#endif
HBTX_DEV void rs_cube(const uint32_t *bb, float &x0, float &y0, float &z0, double &scale)
{
  x0 = ordered_to_float(bb[0]);
  y0 = ordered_to_float(bb[1]);
  z0 = ordered_to_float(bb[2]);
  const double side = fmax((double)ordered_to_float(bb[3]) - x0,
                           fmax((double)ordered_to_float(bb[4]) - y0, (double)ordered_to_float(bb[5]) - z0));
  scale = (side > 0.0) ? (double)((1 << RS_KEY_BITS) - 1) / side : 0.0;
}

HBTX_DEV uint32_t rs_cell(float o, float o0, double scale)
{
  const double c = ((double)o - (double)o0) * scale;
  return (uint32_t)fmin(fmax(c, 0.0), (double)((1 << RS_KEY_BITS) - 1));
}

HBTX_DEV uint64_t rs_key(float ox, float oy, float oz, const uint32_t *bb)
{
  float x0, y0, z0;
  double scale;
  rs_cube(bb, x0, y0, z0, scale);
  return rs_spread(rs_cell(ox, x0, scale)) | (rs_spread(rs_cell(oy, y0, scale)) << 1) |
         (rs_spread(rs_cell(oz, z0, scale)) << 2);
}

// Pass 3: sort keys of tracers and non-tracers; per host halo the margin and
// wrap flag.
__global__ void k_rs_keys(RsArgs A, int32_t nt_all)
{
  GRID_STRIDE(s, A.n)
  {
    const bool trs = s < nt_all;
    const int i = trs ? A.tval[s] : A.qval[s - nt_all];
    const int g = A.gid[i];
    const RsGroup &G = A.grp[g];
    const float ox = rs_offset(A.pos[3 * (int64_t)i], G.p0x, A);
    const float oy = rs_offset(A.pos[3 * (int64_t)i + 1], G.p0y, A);
    const float oz = rs_offset(A.pos[3 * (int64_t)i + 2], G.p0z, A);
    const uint64_t key = rs_key(ox, oy, oz, A.bbox + 7 * (int64_t)g);
    if (trs)
      A.tkey[s] = key;
    else
      A.qkey[s - nt_all] = key;
  }
}

__global__ void k_rs_group_finish(RsArgs A)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.ngroups)
    return;
  RsGroup &G = A.grp[g];
  const float ext = __uint_as_float(A.bbox[7 * (int64_t)g + 6]);
  // offsets are float roundings of exact differences (relative error 2^-24);
  // the widening also absorbs the float arithmetic of the box distance
  G.margin = ext * 1.0e-6f + 1.0e-30f;
  G.wrap = (A.periodic && 4.0 * (double)ext + G.margin >= (double)A.half) ? 1 : 0;
  int top = -1;
  for (int l = 0; l < RS_LEVELS; l++)
    if (A.cnt[l * A.ngroups + g] == 1)
    {
      top = l;
      break;
    }
  G.top = top;
}

// Tracer records in sorted order: original position and particle index.
__global__ void k_rs_records(RsArgs A, int32_t nt_all)
{
  GRID_STRIDE(s, nt_all)
  {
    const int i = A.tval[s];
    A.trec[s] = make_float4(A.pos[3 * (int64_t)i], A.pos[3 * (int64_t)i + 1], A.pos[3 * (int64_t)i + 2],
                            __int_as_float(i));
  }
}

// Boxes of one level: leaves (level 0) from their tracers, upper levels from
// their children.
__global__ void k_rs_level(RsArgs A, int level, int32_t total)
{
  GRID_STRIDE(x, total)
  {
    const int32_t *lo = A.loff + (int64_t)level * (A.ngroups + 1);
    const int g = rs_group_of(lo, A.ngroups, (int)x);
    const int i = (int)x - lo[g];
    const RsGroup &G = A.grp[g];
    float l0 = INFINITY, l1 = INFINITY, l2 = INFINITY, h0 = -INFINITY, h1 = -INFINITY, h2 = -INFINITY;
    if (level == 0)
    {
      const int b = G.tb + i * RS_LEAF, e = min(G.te, b + RS_LEAF);
      for (int t = b; t < e; t++)
      {
        const float4 r = A.trec[t];
        const float ox = rs_offset(r.x, G.p0x, A), oy = rs_offset(r.y, G.p0y, A), oz = rs_offset(r.z, G.p0z, A);
        l0 = fminf(l0, ox);
        l1 = fminf(l1, oy);
        l2 = fminf(l2, oz);
        h0 = fmaxf(h0, ox);
        h1 = fmaxf(h1, oy);
        h2 = fmaxf(h2, oz);
      }
      l0 -= G.margin;
      l1 -= G.margin;
      l2 -= G.margin;
      h0 += G.margin;
      h1 += G.margin;
      h2 += G.margin;
    }
    else
    {
      const int32_t *clo = A.loff + (int64_t)(level - 1) * (A.ngroups + 1);
      const int nc = A.cnt[(level - 1) * A.ngroups + g];
      const int64_t cb = A.lbase[level - 1] + clo[g];
      for (int c = RS_FAN * i; c < min(nc, RS_FAN * i + RS_FAN); c++)
      {
        const float4 a = A.nodes[2 * (cb + c)], z = A.nodes[2 * (cb + c) + 1];
        l0 = fminf(l0, a.x);
        l1 = fminf(l1, a.y);
        l2 = fminf(l2, a.z);
        h0 = fmaxf(h0, z.x);
        h1 = fmaxf(h1, z.y);
        h2 = fmaxf(h2, z.z);
      }
    }
    const int64_t node = A.lbase[level] + x;
    A.nodes[2 * node] = make_float4(l0, l1, l2, 0.f);
    A.nodes[2 * node + 1] = make_float4(h0, h1, h2, 0.f);
  }
}

// Sizes the host needs: tracers, nodes per level, wave tiles.
__global__ void k_rs_totals(RsArgs A, int32_t *tot)
{
  const int l = threadIdx.x;
  if (l < RS_LEVELS)
    tot[l] = A.loff[(int64_t)l * (A.ngroups + 1) + A.ngroups];
  if (l == RS_LEVELS)
    tot[l] = A.tflag[A.n];
  if (l == RS_LEVELS + 1)
    tot[l] = A.qtiles_n[A.ngroups];
}

// tile t = (host halo, first query, end query, 0), one thread per tile
__global__ void k_rs_tiles(RsArgs A, int4 *tiles, int32_t ntiles)
{
  GRID_STRIDE(t, ntiles)
  {
    const int g = rs_group_of(A.qtiles_n, A.ngroups, (int)t);
    const RsGroup &G = A.grp[g];
    const int q = G.qb + ((int)t - A.qtiles_n[g]) * HBTX_WAVE;
    tiles[t] = make_int4(g, q, min(G.qe, q + HBTX_WAVE), 0);
  }
}

// Squared pair distance as the reference computes it: double differences of
// the float coordinates, NEAREST per axis, rounded to float for the heap
// (src/geometric_tree.cpp:158-181; LocatedParticle_t::d2 is HBTReal).
HBTX_DEV double rs_r2(float tx, float ty, float tz, float qx, float qy, float qz, const RsArgs &A, double &dx,
                      double &dy, double &dz)
{
#pragma clang fp contract(off) // no fused multiply-add: the sum rounds as the host's does
  dx = (double)tx - (double)qx;
  dy = (double)ty - (double)qy;
  dz = (double)tz - (double)qz;
  if (A.periodic)
  {
    dx = nearest_d(dx, (double)A.box, (double)A.half);
    dy = nearest_d(dy, (double)A.box, (double)A.half);
    dz = nearest_d(dz, (double)A.box, (double)A.half);
  }
  return dx * dx + dy * dy + dz * dz;
}

// Does the sphere of radius r about the query collect the tracer?  The per-axis
// tests, then r2 < (double)(r * r) with the product in float
// (src/geometric_tree.cpp:148,160-181).
HBTX_DEV bool rs_inside(double dx, double dy, double dz, double r2, float r)
{
  const double rd = (double)r;
  if (dx > rd || dx < -rd || dy > rd || dy < -rd || dz > rd || dz < -rd)
    return false;
  const float rr = r * r;
  return r2 < (double)rr;
}

// Lower bound of the distance from a query offset to a box, one axis.
HBTX_DEV float rs_gap(float q, float lo, float hi)
{
  return fmaxf(0.f, fmaxf(lo - q, q - hi));
}

HBTX_DEV float rd_lane(float v, int l)
{
  return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), l));
}
HBTX_DEV int rd_lane(int v, int l)
{
  return __builtin_amdgcn_readlane(v, l);
}
HBTX_DEV float wr_lane(float v, int l, float old)
{
  return (lane_id() == l) ? v : old;
}
HBTX_DEV int wr_lane(int v, int l, int old)
{
  return (lane_id() == l) ? v : old;
}

// Squared lower bound of the distance from a query offset to a box.
HBTX_DEV float rs_box_gap2(float ux, float uy, float uz, float lx, float ly, float lz, float hx, float hy,
                           float hz, bool wrap, float box)
{
  float gx = rs_gap(ux, lx, hx), gy = rs_gap(uy, ly, hy), gz = rs_gap(uz, lz, hz);
  if (wrap)
  {
    gx = fminf(gx, fminf(rs_gap(ux + box, lx, hx), rs_gap(ux - box, lx, hx)));
    gy = fminf(gy, fminf(rs_gap(uy + box, ly, hy), rs_gap(uy - box, ly, hy)));
    gz = fminf(gz, fminf(rs_gap(uz + box, lz, hz), rs_gap(uz - box, lz, hz)));
    // q +- box is rounded to the float spacing of box
    const float slack = box * 2.5e-7f;
    gx = fmaxf(0.f, gx - slack);
    gy = fmaxf(0.f, gy - slack);
    gz = fmaxf(0.f, gz - slack);
  }
  return (gx * gx + gy * gy + gz * gz) * (1.0f - 1.0e-6f);
}

// Neighbours are ordered by (squared distance, particle index), one 64-bit key
// (the float bits of a non-negative distance order as integers): the reference
// orders exactly equal float distances by its tree's traversal order and its
// heap's internals, which no other search reproduces.
HBTX_DEV uint64_t rs_nkey(float d2, int pid)
{
  return ((uint64_t)__float_as_uint(d2) << 32) | (uint32_t)pid;
}
HBTX_DEV float rs_nkey_d2(uint64_t key)
{
  return __uint_as_float((uint32_t)(key >> 32));
}

// Union walk of one wave over its host halo's boxes.  Every lane keeps the k
// nearest (keys ascending in key[K-k..K-1]; key[0..K-k) hold 0 and never move).  The wave first reads the leaves its queries' keys fall in (the
// queries and tracers share one Morton order, so these are a short run), then
// walks the boxes depth first with an explicit stack held across the lanes
// (entry e in lane e: level, index and box; at most 8 + 7 (levels - 2) entries):
// a popped box is opened when any lane's k-th distance reaches it, its children's
// boxes arrive in one vector load, and a leaf's 32 tracers arrive in one vector
// load and are broadcast lane by lane.  A float lower bound screens every pair
// before the exact double distance.  Then the move rule per lane.
template <int K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_knn(RsArgs A, const int4 *__restrict__ tiles, int32_t ntiles,
                                                       const float *__restrict__ radii)
{
  __shared__ float rad[RS_RADII];
  for (int j = threadIdx.x; j < RS_RADII; j += RS_THREADS)
    rad[j] = radii[j];
  __syncthreads();
  const int lane = lane_id();
  const int wpb = RS_THREADS / HBTX_WAVE;
  const int k = A.k;
  for (int t = uniform(blockIdx.x * wpb + (int)(threadIdx.x / HBTX_WAVE)); t < ntiles; t += gridDim.x * wpb)
  {
    const int4 tl = tiles[t];
    const int g = uniform(tl.x);
    const RsGroup G = A.grp[g];
    const int q = tl.y + lane;
    const int nact = uniform(tl.z - tl.y); // active lanes are 0 .. nact-1
    const bool active = lane < nact;
    const int pi = active ? A.qval[q] : A.qval[tl.y];
    const float qx = A.pos[3 * (int64_t)pi], qy = A.pos[3 * (int64_t)pi + 1], qz = A.pos[3 * (int64_t)pi + 2];
    const int own = A.sub[pi];
    const float ux = rs_offset(qx, G.p0x, A), uy = rs_offset(qy, G.p0y, A), uz = rs_offset(qz, G.p0z, A);
    const int nt = G.te - G.tb;
    const bool wrap = G.wrap != 0;
    // lane l < RS_LEVELS: first node and node count of level l in this host halo
    int lb_lo = 0, lb_hi = 0, lcnt = 0;
    if (lane < RS_LEVELS)
    {
      const int64_t b = A.lbase[lane] + A.loff[(int64_t)lane * (A.ngroups + 1) + g];
      lb_lo = (int)(uint32_t)b;
      lb_hi = (int)(b >> 32);
      lcnt = A.cnt[lane * A.ngroups + g];
    }
    auto level_base = [&](int l) -> int64_t {
      return ((int64_t)rd_lane(lb_hi, l) << 32) | (uint32_t)rd_lane(lb_lo, l);
    };
    uint64_t key[K];
#pragma unroll
    for (int j = 0; j < K; j++)
      key[j] = (j < K - k) ? 0ull : ~0ull;
    float kth = INFINITY; // distance part of key[K-1]
    auto insert = [&](uint64_t x) {
#pragma unroll
      for (int j = K - 1; j > 0; j--)
      {
        const bool sh = x < key[j - 1];
        const bool pl = !sh && x < key[j];
        key[j] = sh ? key[j - 1] : (pl ? x : key[j]);
      }
      if (x < key[0])
        key[0] = x;
      kth = (key[K - 1] == ~0ull) ? INFINITY : rs_nkey_d2(key[K - 1]); // +inf until k are held
    };
    const float slack = A.periodic ? A.box * 2.5e-7f : 0.f;
    auto consider = [&](float tx, float ty, float tz, int pid) {
      // float lower bound: each difference is rounded once (and wrapped with an
      // error below the float spacing of box), the sum once more
      float ex = tx - qx, ey = ty - qy, ez = tz - qz;
      if (A.periodic)
      {
        ex = fmaxf(fabsf(nearest_f(ex, A.box, A.half)) - slack, 0.f);
        ey = fmaxf(fabsf(nearest_f(ey, A.box, A.half)) - slack, 0.f);
        ez = fmaxf(fabsf(nearest_f(ez, A.box, A.half)) - slack, 0.f);
      }
      const float e2 = (ex * ex + ey * ey + ez * ez) * (1.0f - 4.0e-6f);
      if (active && e2 <= kth)
      {
        double dx, dy, dz;
        const uint64_t x = rs_nkey((float)rs_r2(tx, ty, tz, qx, qy, qz, A, dx, dy, dz), pid);
        if (x < key[K - 1])
          insert(x);
      }
    };
    auto leaf = [&](int li) {
      const int b = G.tb + li * RS_LEAF, n = min(G.te - b, RS_LEAF);
      float4 r = make_float4(0.f, 0.f, 0.f, 0.f);
      if (lane < n)
        r = A.trec[b + lane];
      for (int j = 0; j < n; j++)
        consider(rd_lane(r.x, j), rd_lane(r.y, j), rd_lane(r.z, j), __float_as_int(rd_lane(r.w, j)));
    };
    // the leaves the queries' keys fall in
    int home = 0;
    {
      const uint64_t key = A.qkey[active ? q : tl.y];
      int lo = G.tb, hi = G.te;
      while (lo < hi)
      {
        const int mid = (lo + hi) >> 1;
        if (A.tkey[mid] < key)
          lo = mid + 1;
        else
          hi = mid;
      }
      home = min((lo - G.tb) / RS_LEAF, rd_lane(lcnt, 0) - 1);
    }
    int h[4];
    {
      const int h0 = rd_lane(home, 0), h3 = rd_lane(home, nact - 1);
      if (h3 - h0 < 4)
      {
#pragma unroll
        for (int j = 0; j < 4; j++)
          h[j] = (h0 + j <= h3) ? h0 + j : -1;
      }
      else
      {
        h[0] = h0;
        h[1] = rd_lane(home, (nact - 1) / 3);
        h[2] = rd_lane(home, 2 * (nact - 1) / 3);
        h[3] = h3;
        h[1] = (h[1] == h[0]) ? -1 : h[1];
        h[2] = (h[2] == h[0] || h[2] == h[1]) ? -1 : h[2];
        h[3] = (h[3] == h[0] || h[3] == h[1] || h[3] == h[2]) ? -1 : h[3];
      }
    }
#pragma unroll
    for (int j = 0; j < 4; j++)
      if (h[j] >= 0)
        leaf(h[j]);
    // depth-first walk from the root
    int s_code = 0;
    float s_lx = 0.f, s_ly = 0.f, s_lz = 0.f, s_hx = 0.f, s_hy = 0.f, s_hz = 0.f;
    int sp = 0;
    auto push = [&](int code, float lx, float ly, float lz, float hx, float hy, float hz) {
      s_code = wr_lane(code, sp, s_code);
      s_lx = wr_lane(lx, sp, s_lx);
      s_ly = wr_lane(ly, sp, s_ly);
      s_lz = wr_lane(lz, sp, s_lz);
      s_hx = wr_lane(hx, sp, s_hx);
      s_hy = wr_lane(hy, sp, s_hy);
      s_hz = wr_lane(hz, sp, s_hz);
      sp++;
    };
    {
      const int64_t root = level_base(G.top);
      const float4 bl = A.nodes[2 * root], bh = A.nodes[2 * root + 1];
      push(G.top << 24, bl.x, bl.y, bl.z, bh.x, bh.y, bh.z);
    }
    while (sp > 0)
    {
      sp--;
      const int code = rd_lane(s_code, sp);
      const int l = code >> 24, i = code & 0xffffff;
      const float gap2 = rs_box_gap2(ux, uy, uz, rd_lane(s_lx, sp), rd_lane(s_ly, sp), rd_lane(s_lz, sp),
                                     rd_lane(s_hx, sp), rd_lane(s_hy, sp), rd_lane(s_hz, sp), wrap, A.box);
      if (!ballot(active && gap2 <= kth))
        continue;
      if (l == 0)
      {
        if (i != h[0] && i != h[1] && i != h[2] && i != h[3])
          leaf(i);
        continue;
      }
      const int nc = min(rd_lane(lcnt, l - 1) - RS_FAN * i, RS_FAN);
      const int64_t cb = level_base(l - 1) + RS_FAN * i;
      float4 cl = make_float4(0.f, 0.f, 0.f, 0.f), ch = cl;
      if (lane < nc)
      {
        cl = A.nodes[2 * (cb + lane)];
        ch = A.nodes[2 * (cb + lane) + 1];
      }
      for (int c = nc - 1; c >= 0; c--) // child 0 ends on top
      {
        const float lx = rd_lane(cl.x, c), ly = rd_lane(cl.y, c), lz = rd_lane(cl.z, c);
        const float hx = rd_lane(ch.x, c), hy = rd_lane(ch.y, c), hz = rd_lane(ch.z, c);
        const float cg = rs_box_gap2(ux, uy, uz, lx, ly, lz, hx, hy, hz, wrap, A.box);
        if (ballot(active && cg <= kth))
          push(((l - 1) << 24) | (RS_FAN * i + c), lx, ly, lz, hx, hy, hz);
      }
    }
    if (!active)
      continue;
    // j_t of each neighbour: the first search whose sphere collects it
    const int nr = min(nt, k);
    int je[K], se[K];
    int jmin = RS_RADII;
#pragma unroll
    for (int j = 0; j < K; j++)
    {
      je[j] = RS_RADII;
      se[j] = -1;
      if (j >= K - k && key[j] != ~0ull)
      {
        const int64_t p = (int32_t)(uint32_t)key[j];
        double dx, dy, dz;
        const double r2 = rs_r2(A.pos[3 * p], A.pos[3 * p + 1], A.pos[3 * p + 2], qx, qy, qz, A, dx, dy, dz);
        int a = 0, b = RS_RADII - 1; // smallest index that collects it (the last entry always does)
        while (a < b)
        {
          const int m = (a + b) >> 1;
          if (rs_inside(dx, dy, dz, r2, rad[m]))
            b = m;
          else
            a = m + 1;
        }
        je[j] = a;
        se[j] = A.sub[p];
        jmin = min(jmin, a);
      }
    }
    // J: the first search after which the heap holds nr entries
    int J = jmin;
    for (int it = 0; it <= K; it++)
    {
      int S = 0;
#pragma unroll
      for (int j = 0; j < K; j++)
        S += (je[j] <= J) ? (J - je[j] + 1) : 0;
      if (S >= nr)
        break;
      J++;
    }
    // the heap: leading k entries of the multiset, nearest first
    int taken = 0, first = -1;
    bool found = false;
#pragma unroll
    for (int j = 0; j < K; j++)
    {
      if (je[j] <= J && taken < k)
      {
        taken += J - je[j] + 1;
        if (first < 0)
          first = se[j];
        found = found || (se[j] == own);
      }
    }
    if (!found && first >= 0 && first != own)
      A.out[pi] = first;
  }
}

} // namespace hbtx
