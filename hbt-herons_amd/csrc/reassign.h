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
//  * tracers sorted by a 63-bit Morton key of their offset from the halo's first
//    particle (periodic minimum image), records trec = (x, y, z, subhalo bits);
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
  int32_t pad0, pad1;
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
// offsets and largest |offset| per host halo (wave-aggregated atomics when the
// wave lies in one host halo).
__global__ __launch_bounds__(RS_THREADS) void k_rs_scatter(RsArgs A)
{
  const int64_t stride = (int64_t)gridDim.x * blockDim.x;
  for (int64_t base = (int64_t)blockIdx.x * blockDim.x; base < A.n; base += stride)
  {
    const int64_t i = base + threadIdx.x;
    const bool valid = i < A.n;
    int g = valid ? A.gid[i] : -1;
    bool tr = false;
    float ox = 0.f, oy = 0.f, oz = 0.f;
    if (valid)
    {
      const RsGroup &G = A.grp[g];
      tr = A.tracer[i] != 0;
      const int s = A.tflag[i];
      if (tr)
        A.tval[s] = (int32_t)i;
      else
        A.qval[i - s] = (int32_t)i;
      ox = rs_offset(A.pos[3 * i], G.p0x, A);
      oy = rs_offset(A.pos[3 * i + 1], G.p0y, A);
      oz = rs_offset(A.pos[3 * i + 2], G.p0z, A);
    }
    const float ext = fmaxf(fabsf(ox), fmaxf(fabsf(oy), fabsf(oz)));
    const int g0 = __shfl(g, 0, HBTX_WAVE);
    const bool same = ballot(valid && g == g0) == ballot(valid);
    if (same)
    {
      // one host halo: reduce over the wave, lane 0 updates
      float lx = tr ? ox : INFINITY, hx = tr ? ox : -INFINITY;
      float ly = tr ? oy : INFINITY, hy = tr ? oy : -INFINITY;
      float lz = tr ? oz : INFINITY, hz = tr ? oz : -INFINITY;
      float e0 = valid ? ext : 0.f, e1 = e0;
      rs_wave_min_max(lx, hx);
      rs_wave_min_max(ly, hy);
      rs_wave_min_max(lz, hz);
      rs_wave_min_max(e1, e0);
      if (lane_id() == 0 && g0 >= 0)
      {
        uint32_t *bb = A.bbox + 7 * (int64_t)g0;
        if (lx <= hx)
        {
          atomicMin(bb + 0, float_to_ordered(lx));
          atomicMin(bb + 1, float_to_ordered(ly));
          atomicMin(bb + 2, float_to_ordered(lz));
          atomicMax(bb + 3, float_to_ordered(hx));
          atomicMax(bb + 4, float_to_ordered(hy));
          atomicMax(bb + 5, float_to_ordered(hz));
        }
        atomicMax(bb + 6, __float_as_uint(e0));
      }
    }
    else if (valid)
    {
      uint32_t *bb = A.bbox + 7 * (int64_t)g;
      if (tr)
      {
        atomicMin(bb + 0, float_to_ordered(ox));
        atomicMin(bb + 1, float_to_ordered(oy));
        atomicMin(bb + 2, float_to_ordered(oz));
        atomicMax(bb + 3, float_to_ordered(ox));
        atomicMax(bb + 4, float_to_ordered(oy));
        atomicMax(bb + 5, float_to_ordered(oz));
      }
      atomicMax(bb + 6, __float_as_uint(ext));
    }
  }
}

// Morton cube of a host halo: origin and cells per unit length.
HBTX_DEV void rs_cube(const uint32_t *bb, float &x0, float &y0, float &z0, double &scale)
{
  x0 = ordered_to_float(bb[0]);
  y0 = ordered_to_float(bb[1]);
  z0 = ordered_to_float(bb[2]);
  const double side = fmax((double)ordered_to_float(bb[3]) - x0,
                           fmax((double)ordered_to_float(bb[4]) - y0, (double)ordered_to_float(bb[5]) - z0));
  scale = (side > 0.0) ? 2097151.0 / side : 0.0;
}

HBTX_DEV uint32_t rs_cell(float o, float o0, double scale)
{
  const double c = ((double)o - (double)o0) * scale;
  return (uint32_t)fmin(fmax(c, 0.0), 2097151.0);
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

// Tracer records in sorted order: original position and subhalo.
__global__ void k_rs_records(RsArgs A, int32_t nt_all)
{
  GRID_STRIDE(s, nt_all)
  {
    const int i = A.tval[s];
    A.trec[s] = make_float4(A.pos[3 * (int64_t)i], A.pos[3 * (int64_t)i + 1], A.pos[3 * (int64_t)i + 2],
                            __int_as_float(A.sub[i]));
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

// tile = (host halo, first query, end query, 0)
__global__ void k_rs_tiles(RsArgs A, int4 *tiles)
{
  const int g = blockIdx.x * blockDim.x + threadIdx.x;
  if (g >= A.ngroups)
    return;
  const RsGroup &G = A.grp[g];
  int t = A.qtiles_n[g];
  if (G.te == G.tb)
    return;
  for (int q = G.qb; q < G.qe; q += HBTX_WAVE, t++)
    tiles[t] = make_int4(g, q, min(G.qe, q + HBTX_WAVE), 0);
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

// Union walk of one wave over its host halo's boxes: every lane keeps the k
// smallest squared distances (ascending in d[K-k..K-1]; d[0..K-k) hold -inf and
// never move), a box is opened when any lane's k-th distance reaches it, and
// leaves are read at wave-uniform addresses.  Then the move rule per lane.
template <int K>
__global__ __launch_bounds__(RS_THREADS) void k_rs_knn(RsArgs A, const int4 *__restrict__ tiles,
                                                       int32_t ntiles, const float *__restrict__ radii)
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
    const bool active = q < tl.z;
    const int pi = active ? A.qval[q] : 0;
    const float qx = A.pos[3 * (int64_t)pi], qy = A.pos[3 * (int64_t)pi + 1], qz = A.pos[3 * (int64_t)pi + 2];
    const int own = A.sub[pi];
    const float ux = rs_offset(qx, G.p0x, A), uy = rs_offset(qy, G.p0y, A), uz = rs_offset(qz, G.p0z, A);
    const int nt = G.te - G.tb;
    float d[K];
    int id[K];
#pragma unroll
    for (int j = 0; j < K; j++)
    {
      d[j] = (j < K - k) ? -INFINITY : INFINITY;
      id[j] = -1;
    }
    auto insert = [&](float x, int tid) {
#pragma unroll
      for (int j = K - 1; j > 0; j--)
      {
        const bool sh = d[j - 1] > x;
        const bool pl = !sh && d[j] > x;
        d[j] = sh ? d[j - 1] : (pl ? x : d[j]);
        id[j] = sh ? id[j - 1] : (pl ? tid : id[j]);
      }
      if (d[0] > x)
      {
        d[0] = x;
        id[0] = tid;
      }
    };
    auto consider = [&](int s) {
      const float4 r = A.trec[s];
      double dx, dy, dz;
      const float d2 = (float)rs_r2(r.x, r.y, r.z, qx, qy, qz, A, dx, dy, dz);
      if (d2 < d[K - 1])
        insert(d2, s);
    };
    // home leaf: where the query's key falls among the sorted tracer keys
    const int c0 = A.cnt[g];
    int home = -1;
    if (active)
    {
      const uint64_t key = A.qkey[q];
      int lo = G.tb, hi = G.te;
      while (lo < hi)
      {
        const int mid = (lo + hi) >> 1;
        if (A.tkey[mid] < key)
          lo = mid + 1;
        else
          hi = mid;
      }
      home = min((lo - G.tb) / RS_LEAF, c0 - 1);
      const int b = G.tb + home * RS_LEAF, e = min(G.te, b + RS_LEAF);
      for (int s = b; s < e; s++)
        consider(s);
    }
    // stackless walk: node (l, i); children of (l, i) are (l-1, 8i ..)
    const int top = G.top;
    int l = top, i = 0;
    while (true)
    {
      const int32_t *lo = A.loff + (int64_t)l * (A.ngroups + 1);
      const int64_t node = A.lbase[l] + lo[g] + i;
      const float4 bl = A.nodes[2 * node], bh = A.nodes[2 * node + 1];
      float gx = rs_gap(ux, bl.x, bh.x), gy = rs_gap(uy, bl.y, bh.y), gz = rs_gap(uz, bl.z, bh.z);
      if (G.wrap)
      {
        gx = fminf(gx, fminf(rs_gap(ux + A.box, bl.x, bh.x), rs_gap(ux - A.box, bl.x, bh.x)));
        gy = fminf(gy, fminf(rs_gap(uy + A.box, bl.y, bh.y), rs_gap(uy - A.box, bl.y, bh.y)));
        gz = fminf(gz, fminf(rs_gap(uz + A.box, bl.z, bh.z), rs_gap(uz - A.box, bl.z, bh.z)));
        // q +- box is rounded to the float spacing of box
        const float slack = A.box * 2.5e-7f;
        gx = fmaxf(0.f, gx - slack);
        gy = fmaxf(0.f, gy - slack);
        gz = fmaxf(0.f, gz - slack);
      }
      const float gap2 = (gx * gx + gy * gy + gz * gz) * (1.0f - 1.0e-6f);
      const bool need = active && gap2 <= d[K - 1];
      if (ballot(need))
      {
        if (l > 0)
        {
          l--;
          i *= RS_FAN;
          continue;
        }
        // the lanes whose home leaf this is have read it already
        const bool mine = need && i != home;
        if (ballot(mine))
        {
          const int b = G.tb + i * RS_LEAF, e = min(G.te, b + RS_LEAF);
          for (int s = b; s < e; s++)
          {
            const float4 r = A.trec[s];
            double dx, dy, dz;
            const float d2 = (float)rs_r2(r.x, r.y, r.z, qx, qy, qz, A, dx, dy, dz);
            if (mine && d2 < d[K - 1])
              insert(d2, s);
          }
        }
      }
      // next node: the following sibling, else up
      bool done = false;
      while (true)
      {
        if (l == top)
        {
          done = true;
          break;
        }
        const int cl = A.cnt[l * A.ngroups + g];
        if ((i % RS_FAN) != RS_FAN - 1 && i + 1 < cl)
        {
          i++;
          break;
        }
        i /= RS_FAN;
        l++;
      }
      if (done)
        break;
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
      if (j >= K - k && id[j] >= 0)
      {
        const float4 r = A.trec[id[j]];
        double dx, dy, dz;
        const double r2 = rs_r2(r.x, r.y, r.z, qx, qy, qz, A, dx, dy, dz);
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
        se[j] = __float_as_int(r.w);
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
