// Source subsampling and centre refinement of Subhalo_t::Unbind.
//
//  * PrepareParticlesForSubsampling / CountUnsampledParticles
//    (src/subhalo_unbind.cpp:430-461): the caller delivers each subhalo with its
//    do-not-subsample particles first and the rest shuffled; Nunsample is the
//    length of that leading run, and after every partition the number of such
//    particles still bound.
//  * SetMassUpscaleFactor / GetMassUpscaleFactor(PerParticleType) (:290-391):
//    mass sums over candidate-rank ranges [0,Nunsample), [Nunsample,
//    Nunsample+MaxSampleSize) and the rest, in input order (the E-list order
//    between stable partitions).
//  * RefineBindingEnergyOrder (:395-425, called at :715-738): self-binding
//    energies of the max(MaxSampleSize, f*Nbound) most bound particles with true
//    masses reorder them; stored energies are kept.
#pragma once
#include "common.h"
#include "unbind.h"

namespace hbtx
{

// Initial bound mass (AccumulateMass over all particles, :537-541) and the
// first particle that may be subsampled (end of the leading do-not-subsample run).
// PERIODIC: the box is periodic (the proposed set's frame then wraps about its first particle)
template <bool PERIODIC>
__global__ __launch_bounds__(256) void k_init_mass_tiles(SubArrays S, PartArrays Q, DevParams P, const int4 *tiles,
                                                         const int32_t *ntiles_p, double *partial,
                                                         int32_t *first_sampled)
{
  __shared__ double lds[(256 / HBTX_WAVE) * RED_NV];
  __shared__ int32_t lmin;
  const int nt = *ntiles_p;
  for (int t = blockIdx.x; t < nt; t += gridDim.x)
  {
    const int4 tl = tiles[t];
    const int s = tl.x;
    // periodic boxes: the proposed set's positions as offsets from its first
    // particle, wrapped to the nearest image (CentreOfMassReferenceFrame, :151-198)
    const bool wrap = PERIODIC && Q.proposal && S.first_idx[s] != 0x7fffffff;
    const float4 r0 = wrap ? Q.posm[S.first_idx[s]] : make_float4(0.f, 0.f, 0.f, 0.f);
    if (threadIdx.x == 0)
      lmin = 0x7fffffff;
    double v[RED_NV];
#pragma unroll
    for (int k = 0; k < RED_NV; k++)
      v[k] = 0.;
    int mymin = 0x7fffffff;
    // four of this thread's slots at a time (the same order), their loads in flight together
    for (int p0 = tl.y + threadIdx.x; p0 < tl.z; p0 += 4 * blockDim.x)
    {
      bool in[4], pro[4];
      float4 xb[4];
      int tb[4];
#pragma unroll
      for (int r = 0; r < 4; r++)
      {
        const int p = p0 + r * blockDim.x;
        in[r] = p < tl.z;
        const int pc = in[r] ? p : p0;
        xb[r] = Q.posm[pc];
        tb[r] = ptype_of(Q, pc);
        pro[r] = in[r] && Q.proposal && Q.proposal[pc];
      }
#pragma unroll
      for (int r = 0; r < 4; r++)
      {
        if (!in[r])
          continue;
        const int p = p0 + r * blockDim.x;
        const float4 x = xb[r];
        v[0] += x.w;
        if (!donot_subsample(P, tb[r]))
          mymin = min(mymin, p - S.off[s]);
        if (pro[r])
        {
          // proposed set: count, mass, and the mass-weighted frame of the warm round
          const float4 u = Q.velu[p];
          v[1] += 1.0;
          v[2] += x.w;
          if (wrap)
          {
            v[3] += nearest_d((double)x.x - (double)r0.x, P.box, P.box_half) * x.w;
            v[4] += nearest_d((double)x.y - (double)r0.y, P.box, P.box_half) * x.w;
            v[5] += nearest_d((double)x.z - (double)r0.z, P.box, P.box_half) * x.w;
          }
          else
          {
            v[3] += x.x * x.w;
            v[4] += x.y * x.w;
            v[5] += x.z * x.w;
          }
          v[6] += u.x * x.w;
          v[7] += u.y * x.w;
          v[8] += u.z * x.w;
        }
      }
    }
    __syncthreads();
    atomicMin(&lmin, mymin);
    block_reduce_vec<9>(v, lds, wave_in_tile(tl.y, tl.z)); // v[0..8]
    if (threadIdx.x == 0)
    {
      for (int k = 0; k < RED_NV; k++)
        partial[(int64_t)t * RED_NV + k] = v[k];
      if (lmin != 0x7fffffff)
        atomicMin(first_sampled + s, lmin);
    }
    __syncthreads();
  }
}

// Smallest slot of a proposed particle per subhalo (the slots hold the input order
// when the unbinding starts), the origin of the proposed set's periodic frame.
__global__ void k_first_proposed(SubArrays S, PartArrays Q, int64_t n)
{
  GRID_STRIDE(p, n)
  {
    if (Q.proposal[p])
      atomicMin(S.first_idx + S.sub_of[p], (int32_t)p);
  }
}

__global__ void k_init_mass_finish(SubArrays S, PartArrays Q, DevParams P, const double *red,
                                   const int32_t *first_sampled, int warm_start)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  const int n = S.off[s + 1] - S.off[s];
  const double *v = red + (int64_t)RED_NV * s;
  S.mbound[s] = (float)v[0];
  S.minit[s] = (float)v[0];
  if (warm_start && n > 0 && (int64_t)v[1] >= 2 * P.min_part)
  {
    // warm round in the proposed set's own centre-of-mass frame; subhaloes whose
    // proposal is near the disruption limit (MinNumPartOfSub) iterate plainly, since
    // whether they survive depends on the iteration's path
    S.warm[s] = WARM_PENDING;
    S.npro[s] = (int)v[1];
    // (periodic: the wrapped mean offset from the first proposed particle, added back)
    const float4 r0 = P.periodic ? Q.posm[S.first_idx[s]] : make_float4(0.f, 0.f, 0.f, 0.f);
    const double o[3] = {r0.x, r0.y, r0.z};
    for (int k = 0; k < 3; k++)
    {
      S.ref[6 * s + k] = (float)(v[3 + k] / v[2] + (P.periodic ? o[k] : 0.0));
      S.ref[6 * s + 3 + k] = (float)(v[6 + k] / v[2]);
    }
  }
  // Nunsample (:522-524): only when subsampling can apply to this subhalo
  S.nuns[s] = (P.mss > 0 && n > P.mss) ? min(first_sampled[s], n) : 0;
}

// Mass sums for the upscale factors over candidate ranks (index order).
// v[0] = sum over ranks [0,Nunsample); v[1+t] = per type over [Nunsample, Nunsample+MaxSampleSize);
// v[7+t] = per type over the remaining candidates.
__global__ __launch_bounds__(256) void k_upscale_tiles(SubArrays S, PartArrays Q, DevParams P, const int4 *tiles,
                                                       const int32_t *ntiles_p, const int32_t *crank, double *partial)
{
  __shared__ double lds[(256 / HBTX_WAVE) * RED_NV];
  const int nt = *ntiles_p;
  for (int t = blockIdx.x; t < nt; t += gridDim.x)
  {
    const int4 tl = tiles[t];
    const int s = tl.x;
    const int base = crank[S.off[s]];
    const int nu = S.nuns[s];
    double v[RED_NV];
#pragma unroll
    for (int k = 0; k < RED_NV; k++)
      v[k] = 0.;
    for (int p = tl.y + threadIdx.x; p < tl.z; p += blockDim.x)
    {
      if (Q.state[p] != 0)
        continue;
      const int rank = crank[p] - base;
      const double m = Q.posm[p].w;
      const int ty = P.dm_only ? 1 : ptype_of(Q, p);
      if (rank < nu)
        v[0] += m;
      else if ((int64_t)rank < (int64_t)nu + P.mss)
      {
#pragma unroll
        for (int k = 0; k < 6; k++)
          v[1 + k] += (ty == k) ? m : 0.;
      }
      else
      {
#pragma unroll
        for (int k = 0; k < 6; k++)
          v[7 + k] += (ty == k) ? m : 0.;
      }
    }
    block_reduce_vec<13>(v, lds, wave_in_tile(tl.y, tl.z)); // v[0..12]
    if (threadIdx.x == 0)
      for (int k = 0; k < RED_NV; k++)
        partial[(int64_t)t * RED_NV + k] = v[k];
  }
}

// Ranges for the upscale sums: full-mode subhaloes that subsample this round.
__global__ void k_upscale_ranges(SubArrays S, int32_t *begin, int32_t *len)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  begin[s] = S.off[s];
  len[s] = (S.mode[s] == MODE_FULL && S.nsrc[s] < S.ntgt[s]) ? (S.off[s + 1] - S.off[s]) : 0;
}

// GetMassUpscaleFactor (:290-317) or ...PerParticleType (:323-361), in float.
__global__ void k_upscale_finish(SubArrays S, DevParams P, const double *red)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  float *up = S.upscale + 6 * s;
  if (!(S.mode[s] == MODE_FULL && S.nsrc[s] < S.ntgt[s]))
  {
    for (int k = 0; k < 6; k++)
      up[k] = 1.f;
    return;
  }
  const double *v = red + (int64_t)RED_NV * s;
  if (P.dm_only || !P.per_type)
  {
    const float munsampled = (float)v[0];
    const float mtrue = S.mlast[s] - munsampled;
    double ms = 0.;
    for (int k = 0; k < 6; k++)
      ms += v[1 + k];
    const float f = mtrue / (float)ms;
    for (int k = 0; k < 6; k++)
      up[k] = f;
    return;
  }
  float msub[6], mtot[6], f[6];
  for (int k = 0; k < 6; k++)
  {
    msub[k] = (float)v[1 + k];
    mtot[k] = (float)v[7 + k] + msub[k];
    f[k] = (msub[k] != 0.f) ? mtot[k] / msub[k] : 0.f;
  }
  float ttrue = 0.f, tscaled = 0.f;
  for (int k = 0; k < 6; k++)
  {
    ttrue += mtot[k];
    tscaled += msub[k] * f[k];
  }
  for (int k = 0; k < 6; k++)
    up[k] = f[k] * (ttrue / tscaled);
}

// CountUnsampledParticles (:446-461): do-not-subsample particles that are still bound.
__global__ void k_count_unsampled(SubArrays S, PartArrays Q, DevParams P, int64_t n, int32_t *cnt)
{
  GRID_STRIDE(p, n)
  {
    const int s = S.sub_of[p];
    if (S.mode[s] != MODE_IDLE && S.nuns[s] > 0 && Q.state[p] == 0 && donot_subsample(P, ptype_of(Q, p)))
      atomicAdd(cnt + s, 1);
  }
}

__global__ void k_set_unsampled(SubArrays S, const int32_t *cnt)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  if (S.mode[s] != MODE_IDLE && S.nuns[s] > 0)
    S.nuns[s] = min(S.nuns[s], cnt[s]);
}

// Centre refinement plan: size = max(MaxSampleSize, f * Nbound) most bound particles.
__global__ void k_refine_plan(SubArrays S, DevParams P)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  const int64_t nb = S.nbound[s];
  if (S.status[s] == ST_CONVERGED && P.refine && nb > P.mss)
  {
    int64_t size = max(P.mss, (int64_t)(P.refine_frac * nb));
    size = min(size, nb);
    S.mode[s] = MODE_REFINE;
    S.nsrc[s] = (int)size;
    S.ntgt[s] = (int)size;
    S.treemode[s] = 1;
    S.nitems[s] = (int)((size + HBTX_WAVE - 1) / HBTX_WAVE);
  }
  else
  {
    S.mode[s] = MODE_IDLE;
    S.nsrc[s] = 0;
    S.ntgt[s] = 0;
    S.treemode[s] = 0;
    S.nitems[s] = 0;
  }
}

__global__ void k_slot_of(SubArrays S, int64_t n, const int32_t *vals, int32_t *slot)
{
  GRID_STRIDE(q, n)
  {
    const int s = S.sub_of[q];
    slot[vals[q]] = (int32_t)(q - S.off[s]);
  }
}

// Refinement sources: the `size` most bound particles, true masses as self mass.
__global__ void k_mark_refine(SubArrays S, PartArrays Q, int64_t n, const int32_t *slot, int32_t *flag)
{
  GRID_STRIDE(p, n)
  {
    const int s = S.sub_of[p];
    const int f = (S.mode[s] == MODE_REFINE && slot[p] < S.nsrc[s]) ? 1 : 0;
    flag[p] = f;
    Q.selfm[p] = f ? Q.posm[p].w : 0.f;
  }
}

__global__ void k_refine_keys(SubArrays S, int64_t n, const int32_t *vals, const float *eref, uint64_t *keys,
                              int32_t *seg_begin, int32_t *seg_len)
{
  GRID_STRIDE(q, n)
  {
    const int s = S.sub_of[q];
    if (S.mode[s] == MODE_REFINE && (int)(q - S.off[s]) < S.nsrc[s])
      keys[q] = (uint64_t)float_to_ordered(eref[vals[q]]);
  }
}

__global__ void k_refine_segments(SubArrays S, int32_t *seg_begin, int32_t *seg_len)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  seg_begin[s] = S.off[s];
  seg_len[s] = (S.mode[s] == MODE_REFINE) ? S.nsrc[s] : 0;
}

} // namespace hbtx
