// Per-particle features of the learned warm start, from the current snapshot and
// the previous snapshot's unbinding outputs (ID-preserving: same slot).
//   f0 log10(r / r_s + 1e-3)      r from the previous centre of mass, r_s = previous
//                                 bound set's rms radius
//   f1 |v - v_bulk| / v_circ      v_circ = sqrt(G M_prev / (r_s a)), clipped at 20
//   f2 previous binding-energy rank / N   (1 for particles absent before)
//   f3 previous bound flag
//   f4 particle type / 5
//   f5 potential depth (E_prev - |v - v_bulk|^2 / 2) / v_circ^2: local density proxy
//   f6 log10(N_p) / 6
//   f7 E_prev / v_circ^2, clipped to [-20, 20]
#pragma once
#include "common.h"
#include "mlp.h"
#include "unbind.h"

namespace hbtx
{

struct PrevArrays
{
  const int32_t *rank;  // per particle: slot in the previous final order, -1 if absent
  const uint8_t *bound; // per particle: bound in the previous snapshot
  const float *energy;  // per particle: previous energy
  const float *com;     // per subhalo [6]: previous frame
  const float *mbound;  // per subhalo: previous bound mass
  const float *rscale;  // per subhalo: previous bound rms radius (comoving)
};

// groups (may be null): only the particles of these 32-particle groups.
__global__ void k_features(SubArrays S, PartArrays Q, DevParams P, PrevArrays V, int64_t n, float *X,
                           const int32_t *groups, int64_t ngroups)
{
  const int64_t cnt = groups ? 32 * ngroups : n;
  GRID_STRIDE(i, cnt)
  {
    const int64_t p = groups ? 32 * (int64_t)groups[i >> 5] + (i & 31) : i;
    if (p >= n)
      continue;
    const int s = S.sub_of[p];
    const int ns = S.off[s + 1] - S.off[s];
    const float *c = V.com + 6 * s;
    const float4 x = Q.posm[p];
    const float4 u = Q.velu[p];
    float dx = x.x - c[0], dy = x.y - c[1], dz = x.z - c[2];
    if (P.periodic)
    {
      dx = nearest_f(dx, P.box, P.box_half);
      dy = nearest_f(dy, P.box, P.box_half);
      dz = nearest_f(dz, P.box, P.box_half);
    }
    const float R = sqrtf(dx * dx + dy * dy + dz * dz);
    const float rs = fmaxf(V.rscale[s], 1e-6f);
    const float M = fmaxf(V.mbound[s], 1e-12f);
    const float vc2 = fmaxf((float)(P.G_over_a * (double)M / ((double)rs * P.a)), 1e-12f);
    const float vx = u.x - c[3], vy = u.y - c[4], vz = u.z - c[5];
    const float sp2 = vx * vx + vy * vy + vz * vz;
    const float e = V.energy[p];
    const int rk = V.rank[p];
    float *f = X + p * MLP_IN;
    f[0] = log10f(R / rs + 1e-3f);
    f[1] = fminf(sqrtf(sp2 / vc2), 20.f);
    f[2] = (rk >= 0 && ns > 0) ? (float)rk / (float)ns : 1.f;
    f[3] = (float)V.bound[p];
    f[4] = (float)ptype_of(Q, p) / 5.f;
    f[5] = fminf(fmaxf((e - 0.5f * sp2) / vc2, -20.f), 5.f);
    f[6] = log10f((float)max(ns, 1)) / 6.f;
    f[7] = fminf(fmaxf(e / vc2, -20.f), 20.f);
  }
}

// Proposed bound flag; subhaloes of fewer than min_particles particles get an
// empty proposal (they run the plain iteration).
__global__ void k_proposal_from_logit(const float *logit, int64_t n, float cut, const int32_t *sub_of,
                                      const int32_t *off, int64_t min_particles, uint8_t *proposal)
{
  GRID_STRIDE(p, n)
  {
    const int s = sub_of[p];
    const int64_t np = off[s + 1] - off[s];
    proposal[p] = (np >= min_particles && logit[p] > cut) ? 1 : 0;
  }
}

} // namespace hbtx
