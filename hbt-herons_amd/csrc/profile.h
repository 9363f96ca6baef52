// Post-unbinding properties of every subhalo's bound set, centred on its most
// bound particle (final slot 0, src/subhalo_unbind.cpp:754):
//   Subhalo_t::CalculateProfileProperties   src/subhalo.cpp:175-243
//     radii sorted, cumulative mass (double sums stored as float), radii below
//     the softening raised to it, v^2 ~ M(<r)/r: Rmax and Vmax at the first
//     maximum, RHalf = r[N/2], REnclose = r[N-1];
//   Snapshot_t::SphericalOverdensitySize    src/snapshot.cpp:258-282
//     outermost particle with M(<r) > rho_200 r^3;
//   Subhalo_t::CalculateShape               src/subhalo.cpp:245-309
//     plain and 1/r^2-weighted second moments over the bound particles but the
//     centre, float products summed in double, over Mbound.
// Every subhalo's bound particles are sorted by radius in one segmented sort;
// tiles of RED_TILE sorted slots then carry the cumulative mass (the tile offsets
// are scanned per subhalo, so no sum runs across subhaloes) and the per-tile
// candidates for the maxima and the shape sums, combined per subhalo at the end.
#pragma once
#include "common.h"
#include "unbind.h"

namespace hbtx
{
// per-tile slots of the reduction scratch (RED_NV doubles per tile)
enum
{
  PROF_SHAPE = 0,  // 12 shape sums: I {xx, xy, xz, yy, yz, zz}, then the weighted ones
  PROF_BESTV = 12, // largest v of the tile
  PROF_BESTQ = 13, // its slot (first of equal maxima)
  PROF_SOQ = 14,   // last slot meeting the overdensity, -1 if none
  PROF_TSUM = 15,  // mass of the tile
  PROF_TPRE = 16   // mass of the subhalo's earlier tiles
};

struct ProfArgs
{
  const int32_t *off;           // per subhalo first slot
  const int32_t *ford;          // final local order per slot
  const float4 *posm;
  const hbtx_sub_result *res;   // unbinding results (nbound, mbound, most bound position)
  const int32_t *seg_len;       // bound particles profiled per subhalo (0 when Nbound <= 1)
  const uint64_t *keys;         // radius bits per sorted slot
  const int32_t *vals;          // particle per sorted slot
  float *rc, *mc;               // per sorted slot: radius (softening floor), cumulative mass
  double *part;                 // per tile [RED_NV]
  DevParams P;
  float rho;                    // 200 rho_crit a^3 (float, src/snapshot.cpp:268-269)
  float vel_unit;               // G / a (float)
};

__global__ void k_prof_segments(int32_t nsub, const int32_t *off, const hbtx_sub_result *res, int32_t *seg_begin,
                                int32_t *seg_len)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsub)
    return;
  const int64_t nb = res[s].nbound;
  seg_begin[s] = off[s];
  seg_len[s] = (nb > 1 && !res[s].disrupted) ? (int32_t)nb : 0;
}

// PeriodicDistance(centre, x) (src/config_parser.h:220-233) as the sort key.
__global__ void k_prof_keys(ProfArgs A, int64_t n, const int32_t *sub_of, uint64_t *keys, int32_t *vals)
{
#pragma clang fp contract(off) // float expressions rounded as written, as on the host
  GRID_STRIDE(q, n)
  {
    const int s = sub_of[q];
    const int a = A.off[s];
    if (q - a >= A.seg_len[s])
      continue;
    const int p = a + A.ford[q];
    const float4 x = A.posm[p];
    const float *c = A.res[s].mostbound_pos;
    float dx = c[0] - x.x, dy = c[1] - x.y, dz = c[2] - x.z;
    if (A.P.periodic)
    {
      dx = nearest_f(dx, A.P.box, A.P.box_half);
      dy = nearest_f(dy, A.P.box, A.P.box_half);
      dz = nearest_f(dz, A.P.box, A.P.box_half);
    }
    keys[q] = (uint64_t)__float_as_uint(sqrtf(dx * dx + dy * dy + dz * dz));
    vals[q] = p;
  }
}

__global__ __launch_bounds__(RED_THREADS) void k_prof_tile_sum(ProfArgs A, const int4 *tiles, const int32_t *ntiles_p)
{
  __shared__ double lds[RED_THREADS / HBTX_WAVE];
  const int nt = *ntiles_p;
  for (int t = blockIdx.x; t < nt; t += gridDim.x)
  {
    const int4 tl = tiles[t];
    double m = 0.;
    for (int q = tl.y + threadIdx.x; q < tl.z; q += RED_THREADS)
      m += (double)A.posm[A.vals[q]].w;
    double tot;
    block_excl_scan(m, lds, &tot);
    if (threadIdx.x == 0)
      A.part[(int64_t)t * RED_NV + PROF_TSUM] = tot;
  }
}

// one thread per subhalo: mass of the subhalo's earlier tiles, in tile order
__global__ void k_prof_tile_prefix(int32_t nsub, const int32_t *toff, const int32_t *ntile, double *part)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsub)
    return;
  double run = 0.;
  for (int t = toff[s]; t < toff[s] + ntile[s]; t++)
  {
    part[(int64_t)t * RED_NV + PROF_TPRE] = run;
    run += part[(int64_t)t * RED_NV + PROF_TSUM];
  }
}

#define PROF_ITEMS (RED_TILE / RED_THREADS)
__global__ __launch_bounds__(RED_THREADS) void k_prof_tile_main(ProfArgs A, const int4 *tiles, const int32_t *ntiles_p)
{
#pragma clang fp contract(off) // float expressions rounded as written, as on the host
  __shared__ double lds[RED_THREADS / HBTX_WAVE];
  __shared__ double lsum[RED_THREADS / HBTX_WAVE][12];
  __shared__ float lbv[RED_THREADS / HBTX_WAVE];
  __shared__ int lbq[RED_THREADS / HBTX_WAVE], lso[RED_THREADS / HBTX_WAVE];
  const int nt = *ntiles_p;
  const DevParams &P = A.P;
  const int w = threadIdx.x / HBTX_WAVE;
  for (int t = blockIdx.x; t < nt; t += gridDim.x)
  {
    const int4 tl = tiles[t];
    const int s = tl.x;
    const int a = A.off[s];
    const int p0 = a + A.ford[a]; // the centre, left out of the shape sums (src/subhalo.cpp:262)
    const float *cen = A.res[s].mostbound_pos;
    const int q0 = tl.y + PROF_ITEMS * threadIdx.x;
    double msum = 0.;
#pragma unroll
    for (int k = 0; k < PROF_ITEMS; k++)
      if (q0 + k < tl.z)
        msum += (double)A.posm[A.vals[q0 + k]].w;
    double tot;
    double run = block_excl_scan(msum, lds, &tot) + A.part[(int64_t)t * RED_NV + PROF_TPRE];
    double sh[12];
#pragma unroll
    for (int k = 0; k < 12; k++)
      sh[k] = 0.;
    float bv = -1.f;
    int bq = 0x7fffffff, so = -1;
    for (int k = 0; k < PROF_ITEMS; k++)
    {
      const int q = q0 + k;
      if (q >= tl.z)
        break;
      const int p = A.vals[q];
      const float4 x = A.posm[p];
      run += (double)x.w;
      const float m = (float)run;
      float r = __uint_as_float((uint32_t)A.keys[q]);
      if (r < P.eps)
        r = P.eps;
      const float v = __fdiv_rn(m, r);
      A.rc[q] = r;
      A.mc[q] = m;
      if (v > bv)
      {
        bv = v;
        bq = q;
      }
      if (m > A.rho * r * r * r)
        so = q;
      if (p != p0)
      {
        float d0 = x.x - cen[0], d1 = x.y - cen[1], d2 = x.z - cen[2];
        if (P.periodic)
        {
          d0 = nearest_f(d0, P.box, P.box_half);
          d1 = nearest_f(d1, P.box, P.box_half);
          d2 = nearest_f(d2, P.box, P.box_half);
        }
        const float pm = x.w;
        const float x2 = d0 * d0, y2 = d1 * d1, z2 = d2 * d2;
        sh[0] += (double)(x2 * pm);
        sh[1] += (double)(d0 * d1 * pm);
        sh[2] += (double)(d0 * d2 * pm);
        sh[3] += (double)(y2 * pm);
        sh[4] += (double)(d1 * d2 * pm);
        sh[5] += (double)(z2 * pm);
        const float dr2 = __fdiv_rn(x2 + y2 + z2, pm);
        sh[6] += (double)__fdiv_rn(x2, dr2);
        sh[7] += (double)__fdiv_rn(d0 * d1, dr2);
        sh[8] += (double)__fdiv_rn(d0 * d2, dr2);
        sh[9] += (double)__fdiv_rn(y2, dr2);
        sh[10] += (double)__fdiv_rn(d1 * d2, dr2);
        sh[11] += (double)__fdiv_rn(z2, dr2);
      }
    }
    // block: shape sums, the first largest v, the last overdense slot
#pragma unroll
    for (int d = 32; d >= 1; d >>= 1)
    {
#pragma unroll
      for (int k = 0; k < 12; k++)
        sh[k] += __shfl_down(sh[k], d, HBTX_WAVE);
      const float ov = __shfl_down(bv, d, HBTX_WAVE);
      const int oq = __shfl_down(bq, d, HBTX_WAVE);
      if (ov > bv || (ov == bv && oq < bq))
      {
        bv = ov;
        bq = oq;
      }
      so = max(so, __shfl_down(so, d, HBTX_WAVE));
    }
    if (lane_id() == 0)
    {
#pragma unroll
      for (int k = 0; k < 12; k++)
        lsum[w][k] = sh[k];
      lbv[w] = bv;
      lbq[w] = bq;
      lso[w] = so;
    }
    __syncthreads();
    if (threadIdx.x == 0)
    {
      double *o = A.part + (int64_t)t * RED_NV;
      for (int k = 0; k < 12; k++)
      {
        double x = 0.;
        for (int ww = 0; ww < RED_THREADS / HBTX_WAVE; ww++)
          x += lsum[ww][k];
        o[PROF_SHAPE + k] = x;
      }
      float v = lbv[0];
      int q = lbq[0], sq = lso[0];
      for (int ww = 1; ww < RED_THREADS / HBTX_WAVE; ww++)
      {
        if (lbv[ww] > v || (lbv[ww] == v && lbq[ww] < q))
        {
          v = lbv[ww];
          q = lbq[ww];
        }
        sq = max(sq, lso[ww]);
      }
      o[PROF_BESTV] = (double)v;
      o[PROF_BESTQ] = (double)q;
      o[PROF_SOQ] = (double)sq;
    }
    __syncthreads();
  }
}

// one thread per subhalo: combine the tiles in order and write the properties
__global__ void k_prof_finish(ProfArgs A, int32_t nsub, const int32_t *toff, const int32_t *ntile,
                              hbtx_sub_profile *out)
{
#pragma clang fp contract(off) // float expressions rounded as written, as on the host
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= nsub)
    return;
  hbtx_sub_profile o;
  float *f = &o.rmax_comoving;
  for (int k = 0; k < 18; k++)
    f[k] = 0.f;
  o.has_r200 = 0;
  o.reserved = 0;
  const int nb = A.seg_len[s];
  if (nb > 1)
  {
    const int a = A.off[s];
    double sh[12];
    for (int k = 0; k < 12; k++)
      sh[k] = 0.;
    float bv = -1.f;
    int bq = a, so = -1;
    for (int t = toff[s]; t < toff[s] + ntile[s]; t++)
    {
      const double *pt = A.part + (int64_t)t * RED_NV;
      for (int k = 0; k < 12; k++)
        sh[k] += pt[PROF_SHAPE + k];
      const float v = (float)pt[PROF_BESTV];
      if (v > bv)
      {
        bv = v;
        bq = (int)pt[PROF_BESTQ];
      }
      so = max(so, (int)pt[PROF_SOQ]);
    }
    o.rmax_comoving = A.rc[bq];
    o.vmax_physical = sqrtf(bv * A.vel_unit);
    o.r_half_comoving = A.rc[a + nb / 2];
    o.r_enclose_comoving = A.rc[a + nb - 1];
    if (so >= 0)
    {
      const float m = A.mc[so];
      o.bound_m200_crit = m;
      o.bound_r200_crit_comoving = (float)pow((double)__fdiv_rn(m, A.rho), 1.0 / 3);
      o.has_r200 = 1;
    }
    const float mb = A.res[s].mbound;
    for (int k = 0; k < 6; k++)
    {
      o.inertial_tensor[k] = __fdiv_rn((float)sh[k], mb);
      o.inertial_tensor_weighted[k] = __fdiv_rn((float)sh[6 + k], mb);
    }
  }
  out[s] = o;
}

} // namespace hbtx
