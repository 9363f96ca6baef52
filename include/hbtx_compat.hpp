// hbtx_compat.hpp — C++ host layer over the hbtx C ABI.  It mirrors the
// reference's interfaces for the unbinding path, so that a caller holding
// HBT-HERONS objects can swap the CPU implementation for the GPU one:
//
//   hbtx::GravityTree          GravityTree_t            src/gravity_tree.h:8-20
//     Reserve / Build / EvaluatePotential / BindingEnergy / Clear
//   hbtx::Unbind(ctx, prm, sub, snapshot_id)
//                              Subhalo_t::Unbind        src/subhalo_unbind.cpp:463-808
//   hbtx::RefineParticles(ctx, prm, subs, snapshot_id)
//                              the unbinding loop of SubhaloSnapshot_t::RefineParticles,
//                              src/subhalo_unbind.cpp:855-919 (all subhaloes in one batch)
//   hbtx::RecursiveUnbind(ctx, prm, Subhalos, roots, snapshot_id, merge_trapped)
//                              Subhalo_t::RecursiveUnbind   src/subhalo_unbind.cpp:810-840
//                              with MergeRecursively / MergeTo src/subhalo_merge.cpp:15-91,183-239,
//                              one batch per level of the hierarchies
//   hbtx::TruncateSource(sub, relax_factor)
//                              Subhalo_t::TruncateSource    src/subhalo_unbind.cpp:842-852
//   hbtx::SubhaloStatistics(subs, snapshot_id)
//                              SubhaloSnapshot_t::PrintSubhaloStatistics (the counts)
//                              src/subhalo_unbind.cpp:940-980
//   hbtx::CalculateProfileProperties(ctx, prm, subs, snapshot_id, shape)
//                              Subhalo_t::CalculateProfileProperties / CalculateShape
//                              src/subhalo.cpp:175-309
//   hbtx::RefineSnapshot(ctx, prm, Subhalos, SubGroups, SubGroupsOfHeads, field, NumSubOld, ...)
//                              SubhaloSnapshot_t::RefineParticles src/subhalo_unbind.cpp:855-919
//                              (exclusive mass: hierarchies, field and new-born subhaloes, truncation)
//   hbtx::ReassignParticles(ctx, prm, Subhalos, SubGroups, num_neighbours)
//                              SubhaloSnapshot_t::ReassignParticles src/subhalo_reassign_gas.cpp:65-197
//   hbtx::KickNullParticles(sub, prm), hbtx::CountParticleTypes(sub, prm)
//                              Subhalo_t::KickNullParticles / CountParticleTypes
//                              src/subhalo.cpp:394-426, :311-392
//
// The templates accept any subhalo type with the reference's member names
// (Particles with Id / ComovingPosition / PhysicalVelocity / Mass and, when
// present, InternalEnergy / Type; Nbound, Mbound, ComovingAveragePosition, ...),
// including Subhalo_t itself.  Optional members (NboundType, MboundType,
// NumberUnbindingIterations, ParticleBindingEnergies, ParticlePotentialEnergies,
// SnapshotOfDeath, DescendantTrackId, SinkTrackId, SnapshotOfSink) are written
// when the type has them.  Errors throw hbtx::Error; the library never aborts.
#pragma once
#include "hbtx.h"

#include <algorithm>
#include <array>
#include <chrono>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <stdexcept>
#include <string>
#include <type_traits>
#include <utility>
#include <vector>

namespace hbtx
{

struct Error : std::runtime_error
{
  int code;
  Error(int c, const std::string &msg) : std::runtime_error(msg), code(c)
  {
  }
};

class Context
{
public:
  explicit Context(int device = 0)
  {
    int rc = hbtx_create(&ctx_, device);
    if (rc != HBTX_OK)
      throw Error(rc, "hbtx_create failed (no HIP device?)");
  }
  ~Context()
  {
    hbtx_destroy(ctx_);
  }
  Context(const Context &) = delete;
  Context &operator=(const Context &) = delete;
  hbtx_ctx *get() const
  {
    return ctx_;
  }
  void check(int rc) const
  {
    if (rc != HBTX_OK)
      throw Error(rc, hbtx_last_error(ctx_));
  }

private:
  hbtx_ctx *ctx_ = nullptr;
};

inline hbtx_params DefaultParams()
{
  hbtx_params p;
  hbtx_default_params(&p);
  return p;
}

namespace detail
{
// C++11 member detection (the reference builds with CMAKE_CXX_STANDARD 11)
template <class... Ts>
struct make_void
{
  typedef void type;
};
template <class... Ts>
using void_t = typename make_void<Ts...>::type;
template <bool B>
using bool_c = std::integral_constant<bool, B>;

#define HBTX_MEMBER_TRAIT(name)                                                                                  \
  template <class T, class = void>                                                                             \
  struct has_##name : std::false_type                                                                          \
  {                                                                                                            \
  };                                                                                                           \
  template <class T>                                                                                           \
  struct has_##name<T, void_t<decltype(std::declval<T>().name)>> : std::true_type                              \
  {                                                                                                            \
  };
HBTX_MEMBER_TRAIT(InternalEnergy)
HBTX_MEMBER_TRAIT(NboundType)
HBTX_MEMBER_TRAIT(MboundType)
HBTX_MEMBER_TRAIT(NumberUnbindingIterations)
HBTX_MEMBER_TRAIT(ParticleBindingEnergies)
HBTX_MEMBER_TRAIT(ParticlePotentialEnergies)
HBTX_MEMBER_TRAIT(SnapshotOfDeath)
HBTX_MEMBER_TRAIT(DescendantTrackId)
HBTX_MEMBER_TRAIT(SinkTrackId)
HBTX_MEMBER_TRAIT(SnapshotOfSink)
HBTX_MEMBER_TRAIT(Type)
HBTX_MEMBER_TRAIT(InertialTensor)
#undef HBTX_MEMBER_TRAIT
template <class T>
struct has_internal_energy : has_InternalEnergy<T>
{
};
template <class T>
struct has_type : has_Type<T>
{
};

// Optional members, by tag dispatch: the first overload of each pair runs when
// the type has the member(s), the second is a no-op.
template <class P>
float internal_energy(const P &p, std::true_type)
{
  return (float)p.InternalEnergy;
}
template <class P>
float internal_energy(const P &, std::false_type)
{
  return 0.f;
}
template <class P>
int32_t particle_type(const P &p, std::true_type)
{
  return (int32_t)p.Type;
}
template <class P>
int32_t particle_type(const P &, std::false_type)
{
  return 1;
}
template <class S>
void set_nbound_type(S &sub, const int64_t *v, std::true_type)
{
  for (int t = 0; t < 6; t++)
    sub.NboundType[t] = v ? v[t] : 0;
}
template <class S>
void set_nbound_type(S &, const int64_t *, std::false_type)
{
}
template <class S>
void set_mbound_type(S &sub, const float *v, std::true_type)
{
  for (int t = 0; t < 6; t++)
    sub.MboundType[t] = v ? v[t] : 0.f;
}
template <class S>
void set_mbound_type(S &, const float *, std::false_type)
{
}
template <class S>
void set_iterations(S &sub, int32_t it, std::true_type)
{
  sub.NumberUnbindingIterations = it;
}
template <class S>
void set_iterations(S &, int32_t, std::false_type)
{
}
// :646-668, SnapshotOfDeath when a subhalo disrupts
template <class S>
void mark_death(S &sub, int snapshot_id, std::true_type)
{
  if (sub.SnapshotOfDeath == -1)
    sub.SnapshotOfDeath = snapshot_id;
}
template <class S>
void mark_death(S &, int, std::false_type)
{
}
// :690-697, a resolved subhalo is alive again
template <class S>
void revive(S &sub, std::true_type)
{
  if (sub.SnapshotOfDeath != -1)
  {
    sub.SnapshotOfDeath = -1;
    sub.DescendantTrackId = -1;
  }
}
template <class S>
void revive(S &, std::false_type)
{
}
// :698-702, and no longer trapped
template <class S>
void clear_sink(S &sub, std::true_type)
{
  if (sub.SinkTrackId != -1)
  {
    sub.SnapshotOfSink = -1;
    sub.SinkTrackId = -1;
  }
}
template <class S>
void clear_sink(S &, std::false_type)
{
}
template <class S>
void store_binding_energies(S &sub, const float *e, int64_t nb, std::true_type)
{
  sub.ParticleBindingEnergies.assign(e, e + nb);
}
template <class S>
void store_binding_energies(S &, const float *, int64_t, std::false_type)
{
}
template <class S>
void store_potential_energies(S &sub, const float *e, int64_t nb, std::true_type)
{
  sub.ParticlePotentialEnergies.assign(e, e + nb);
}
template <class S>
void store_potential_energies(S &, const float *, int64_t, std::false_type)
{
}
template <class S>
void store_shape(S &sub, const hbtx_sub_profile &P, std::true_type)
{
  for (int k = 0; k < 6; k++)
  {
    sub.InertialTensor[k] = P.inertial_tensor[k];
    sub.InertialTensorWeighted[k] = P.inertial_tensor_weighted[k];
  }
}
template <class S>
void store_shape(S &, const hbtx_sub_profile &, std::false_type)
{
}
// per-type bound count and mass of one particle (CountParticleTypes)
template <class S, class P>
void add_type(S &sub, const P &p, std::true_type, std::true_type)
{
  sub.NboundType[(int)p.Type]++;
  sub.MboundType[(int)p.Type] += p.Mass;
}
template <class S, class P>
void add_type(S &sub, const P &p, std::true_type, std::false_type)
{
  sub.NboundType[(int)p.Type]++;
}
template <class S, class P>
void add_type(S &sub, const P &p, std::false_type, std::true_type)
{
  sub.MboundType[(int)p.Type] += p.Mass;
}
template <class S, class P>
void add_type(S &, const P &, std::false_type, std::false_type)
{
}

template <class V>
inline void put3(float *dst, const V &v)
{
  dst[0] = (float)v[0];
  dst[1] = (float)v[1];
  dst[2] = (float)v[2];
}
template <class V>
inline void get3(V &dst, const float *src)
{
  dst[0] = src[0];
  dst[1] = src[1];
  dst[2] = src[2];
}
} // namespace detail

// GravityTree_t counterpart: a GPU octree over a snapshot-like source set.
class GravityTree
{
public:
  GravityTree(Context &ctx, const hbtx_params &prm) : ctx_(ctx), prm_(prm)
  {
  }
  ~GravityTree()
  {
    Clear();
  }
  GravityTree(const GravityTree &) = delete;
  GravityTree &operator=(const GravityTree &) = delete;

  // Capacity hint (the device buffers are sized by Build).
  void Reserve(size_t n)
  {
    pos_.reserve(3 * n);
    mass_.reserve(n);
  }

  // Build(snap, num_part): snap.size(), snap.GetComovingPosition(i), snap.GetMass(i)
  // as in OctTree_t::Build (src/oct_tree.tpp:11-142); returns the number of cells.
  template <class Snapshot>
  int64_t Build(const Snapshot &snap, int64_t num_part = 0)
  {
    const int64_t n = num_part > 0 ? num_part : (int64_t)snap.size();
    pos_.resize(3 * n);
    mass_.resize(n);
    for (int64_t i = 0; i < n; i++)
    {
      detail::put3(&pos_[3 * i], snap.GetComovingPosition(i));
      mass_[i] = (float)snap.GetMass(i);
    }
    return Build(n, pos_.data(), mass_.data());
  }

  int64_t Build(int64_t n, const float *pos, const float *mass)
  {
    Clear();
    int64_t ncells = 0;
    ctx_.check(hbtx_tree_build(ctx_.get(), &prm_, n, pos, mass, &tree_, &ncells));
    return ncells;
  }

  // EvaluatePotential(x, targetMass): one point.  Batched overload below: use it
  // for many targets (one launch instead of one per point).
  template <class Xyz>
  double EvaluatePotential(const Xyz &x, float target_mass = 0.f)
  {
    float p[3];
    detail::put3(p, x);
    double out = 0.0;
    ctx_.check(hbtx_tree_potential(ctx_.get(), tree_, 1, p, &target_mass, &out));
    return out;
  }
  void EvaluatePotential(int64_t n, const float *pos, const float *target_mass, double *out)
  {
    ctx_.check(hbtx_tree_potential(ctx_.get(), tree_, n, pos, target_mass, out));
  }

  // BindingEnergy(x, v, ref_x, ref_v, targetMass) (src/gravity_tree.cpp:168-178).
  template <class Xyz, class Vel>
  double BindingEnergy(const Xyz &x, const Vel &v, const Xyz &ref_x, const Vel &ref_v, float target_mass = 0.f)
  {
    float p[3], q[3], rp[3], rv[3];
    detail::put3(p, x);
    detail::put3(q, v);
    detail::put3(rp, ref_x);
    detail::put3(rv, ref_v);
    double out = 0.0;
    ctx_.check(hbtx_tree_binding_energy(ctx_.get(), tree_, 1, p, q, rp, rv, &target_mass, &out));
    return out;
  }

  void Clear()
  {
    if (tree_)
      hbtx_tree_free(tree_);
    tree_ = nullptr;
  }

private:
  Context &ctx_;
  hbtx_params prm_;
  hbtx_tree *tree_ = nullptr;
  std::vector<float> pos_, mass_;
};

// Unbinds every subhalo in `subs` (pointers) in one batch and writes the
// outputs back with the reference's member names and bookkeeping
// (src/subhalo_unbind.cpp:646-707, :759-806).  snapshot_id is
// epoch.GetSnapshotId(), stored in SnapshotOfDeath when a subhalo disrupts.
// Subhaloes without particles keep the reference's orphan path (:471-489).
template <class SubhaloT>
void RefineParticles(Context &ctx, const hbtx_params &prm, std::vector<SubhaloT *> &subs, int snapshot_id,
                     uint32_t flags = HBTX_BATCHED, bool save_binding_energies = false,
                     bool save_potential_energies = false)
{
  using ParticleT = typename std::decay<decltype(subs[0]->Particles[0])>::type;
  const int64_t nsub = (int64_t)subs.size();
  std::vector<int64_t> off(nsub + 1, 0);
  for (int64_t s = 0; s < nsub; s++)
    off[s + 1] = off[s] + (int64_t)subs[s]->Particles.size();
  const int64_t n = off[nsub];
  std::vector<float> pos(3 * n), vel(3 * n), mass(n), u, refp(3 * nsub), refv(3 * nsub), mbp(3 * nsub),
    mbv(3 * nsub);
  std::vector<int32_t> type;
  std::vector<int64_t> id(n);
  constexpr bool thermal = detail::has_internal_energy<ParticleT>::value;
  constexpr bool typed = detail::has_type<ParticleT>::value;
  if (thermal)
    u.resize(n);
  if (typed)
    type.resize(n);
  for (int64_t s = 0; s < nsub; s++)
  {
    SubhaloT &S = *subs[s];
    detail::put3(&refp[3 * s], S.ComovingAveragePosition);
    detail::put3(&refv[3 * s], S.PhysicalAverageVelocity);
    detail::put3(&mbp[3 * s], S.ComovingMostBoundPosition);
    detail::put3(&mbv[3 * s], S.PhysicalMostBoundVelocity);
    for (int64_t i = 0; i < (int64_t)S.Particles.size(); i++)
    {
      const ParticleT &p = S.Particles[i];
      const int64_t k = off[s] + i;
      detail::put3(&pos[3 * k], p.ComovingPosition);
      detail::put3(&vel[3 * k], p.PhysicalVelocity);
      mass[k] = (float)p.Mass;
      id[k] = (int64_t)p.Id;
      if (thermal)
        u[k] = detail::internal_energy(p, detail::bool_c<thermal>());
      if (typed)
        type[k] = detail::particle_type(p, detail::bool_c<typed>());
    }
  }
  std::vector<int64_t> order(n);
  std::vector<float> energy(n), pot(save_potential_energies ? n : 0);
  std::vector<hbtx_sub_result> res(nsub);
  if (n > 0)
    ctx.check(hbtx_unbind_batch(ctx.get(), &prm, n, pos.data(), vel.data(), mass.data(), thermal ? u.data() : nullptr,
                                typed ? type.data() : nullptr, id.data(), nsub, off.data(), refp.data(),
                                refv.data(), mbp.data(), mbv.data(),
                                flags | (save_potential_energies ? HBTX_SAVE_POTENTIAL : 0u), order.data(),
                                energy.data(), save_potential_energies ? pot.data() : nullptr, res.data()));
  for (int64_t s = 0; s < nsub; s++)
  {
    SubhaloT &S = *subs[s];
    const hbtx_sub_result &R = res[s];
    const int64_t a = off[s], N = off[s + 1] - a;
    if (N > 0)
    {
      std::vector<ParticleT> permuted(N);
      for (int64_t i = 0; i < N; i++)
        permuted[i] = S.Particles[order[a + i]];
      S.Particles.swap(permuted);
    }
    S.Nbound = R.nbound;
    S.Mbound = R.mbound;
    detail::set_nbound_type(S, R.nbound_type, detail::has_NboundType<SubhaloT>());
    detail::set_mbound_type(S, R.mbound_type, detail::has_MboundType<SubhaloT>());
    detail::get3(S.ComovingAveragePosition, R.com_pos);
    detail::get3(S.PhysicalAverageVelocity, R.com_vel);
    detail::get3(S.ComovingMostBoundPosition, R.mostbound_pos);
    detail::get3(S.PhysicalMostBoundVelocity, R.mostbound_vel);
    S.SpecificSelfPotentialEnergy = R.spec_pot;
    S.SpecificSelfKineticEnergy = R.spec_kin;
    for (int k = 0; k < 3; k++)
      S.SpecificAngularMomentum[k] = R.spec_am[k];
    S.TracerIndex = R.tracer_index;
    detail::get3(S.CoreComovingPosition, R.core_pos);
    detail::get3(S.CorePhysicalVelocity, R.core_vel);
    S.CoreComovingSigmaR = R.core_sigma_r;
    S.CorePhysicalSigmaV = R.core_sigma_v;
    detail::set_iterations(S, R.iterations, detail::has_NumberUnbindingIterations<SubhaloT>());
    if (N == 0)
      continue; // orphan: nothing else changes (:471-489)
    if (R.disrupted)
    {
      // :646-668 (MostBoundParticleId keeps its previous value)
      detail::mark_death(S, snapshot_id, detail::has_SnapshotOfDeath<SubhaloT>());
    }
    else
    {
      // :690-702, :791-792
      detail::revive(S, detail::bool_c<detail::has_SnapshotOfDeath<SubhaloT>::value &&
                                       detail::has_DescendantTrackId<SubhaloT>::value>());
      detail::clear_sink(S, detail::bool_c<detail::has_SinkTrackId<SubhaloT>::value &&
                                           detail::has_SnapshotOfSink<SubhaloT>::value>());
      S.MostBoundParticleId = R.mostbound_id;
    }
    const int64_t nb = R.nbound;
    if (save_binding_energies)
      detail::store_binding_energies(S, energy.data() + a, nb, detail::has_ParticleBindingEnergies<SubhaloT>());
    if (save_potential_energies)
      detail::store_potential_energies(S, pot.data() + a, nb, detail::has_ParticlePotentialEnergies<SubhaloT>());
  }
}

// Subhalo_t::Unbind(epoch) for one subhalo: the plain, host-driven iteration.
template <class SubhaloT>
void Unbind(Context &ctx, const hbtx_params &prm, SubhaloT &sub, int snapshot_id, bool save_binding_energies = false,
            bool save_potential_energies = false)
{
  std::vector<SubhaloT *> one{&sub};
  RefineParticles(ctx, prm, one, snapshot_id, HBTX_PLAIN, save_binding_energies, save_potential_energies);
}

// SubhaloSnapshot_t::PrintSubhaloStatistics (src/subhalo_unbind.cpp:940-980):
// the counts it prints for one rank's subhaloes after RefineParticles.
struct SubhaloCounts
{
  int64_t sunk = 0;      // SnapshotOfDeath == SnapshotOfSink == now: merged this snapshot
  int64_t disrupted = 0; // born earlier, died now, not sunk
  int64_t fake = 0;      // born and died now: never self-bound
  int64_t newborn = 0;   // born now and alive
};
template <class SubhaloT>
SubhaloCounts SubhaloStatistics(const std::vector<SubhaloT> &subs, int snapshot_id)
{
  SubhaloCounts c;
  for (const SubhaloT &s : subs)
  {
    c.sunk += (s.SnapshotOfDeath == snapshot_id) & (s.SnapshotOfSink == snapshot_id);
    c.disrupted += (s.SnapshotOfBirth < snapshot_id) & (s.SnapshotOfDeath == snapshot_id) & (s.SnapshotOfSink == -1);
    c.fake += (s.SnapshotOfBirth == snapshot_id) & (s.SnapshotOfDeath == snapshot_id);
    c.newborn += (s.SnapshotOfBirth == snapshot_id) & (s.SnapshotOfDeath == -1);
  }
  return c;
}

// Subhalo_t::CalculateProfileProperties (src/subhalo.cpp:175-243) and, with
// `shape`, Subhalo_t::CalculateShape (:245-309) for every subhalo of `subs`,
// from its bound particles Particles[0, Nbound) (after Unbind: final order, the
// most bound first), in one launch sequence.  prm gives G, the scale factor, H(z)
// and the softening of the epoch.
template <class SubhaloT>
void CalculateProfileProperties(Context &ctx, const hbtx_params &prm, std::vector<SubhaloT *> &subs, int snapshot_id,
                                bool shape = true)
{
  const int64_t nsub = (int64_t)subs.size();
  if (nsub == 0)
    return;
  std::vector<int64_t> off(nsub + 1, 0), nb(nsub);
  std::vector<float> mb(nsub);
  for (int64_t s = 0; s < nsub; s++)
  {
    nb[s] = std::min<int64_t>((int64_t)subs[s]->Nbound, (int64_t)subs[s]->Particles.size());
    mb[s] = (float)subs[s]->Mbound;
    off[s + 1] = off[s] + nb[s];
  }
  std::vector<float> pos(3 * off[nsub]), mass(off[nsub]);
  for (int64_t s = 0; s < nsub; s++)
    for (int64_t i = 0; i < nb[s]; i++)
    {
      const auto &p = subs[s]->Particles[i];
      detail::put3(&pos[3 * (off[s] + i)], p.ComovingPosition);
      mass[off[s] + i] = (float)p.Mass;
    }
  std::vector<hbtx_sub_profile> pr(nsub);
  ctx.check(hbtx_profiles(ctx.get(), &prm, off[nsub], pos.data(), mass.data(), nsub, off.data(), nb.data(), mb.data(),
                          pr.data()));
  for (size_t s = 0; s < subs.size(); s++)
  {
    SubhaloT &S = *subs[s];
    const hbtx_sub_profile &P = pr[s];
    S.RmaxComoving = P.rmax_comoving;
    S.VmaxPhysical = P.vmax_physical;
    S.RHalfComoving = P.r_half_comoving;
    S.REncloseComoving = P.r_enclose_comoving;
    if (S.Nbound <= 1)
    {
      S.BoundR200CritComoving = 0.f;
      S.BoundM200Crit = 0.f;
    }
    else
    {
      if (P.has_r200)
      {
        S.BoundR200CritComoving = P.bound_r200_crit_comoving;
        S.BoundM200Crit = P.bound_m200_crit;
      }
      if (S.VmaxPhysical >= S.LastMaxVmaxPhysical)
      {
        S.SnapshotOfLastMaxVmax = snapshot_id;
        S.LastMaxVmaxPhysical = S.VmaxPhysical;
        S.RmaxComovingOfLastMaxVmax = S.RmaxComoving;
      }
    }
    if (shape)
      detail::store_shape(S, P, detail::has_InertialTensor<SubhaloT>());
  }
}

// ---------------------------------------------------------------------------
// Hierarchical unbinding
// ---------------------------------------------------------------------------

// PeriodicDistance (src/config_parser.h:220-233) in float.
template <class V>
float PeriodicDistance(const V &x, const V &y, const hbtx_params &prm)
{
  float d[3];
  for (int k = 0; k < 3; k++)
  {
    d[k] = (float)x[k] - (float)y[k];
    if (prm.periodic)
    {
      const float half = prm.box_size / 2.f;
      d[k] = (d[k] > half) ? d[k] - prm.box_size : ((d[k] < -half) ? d[k] + prm.box_size : d[k]);
    }
  }
  return std::sqrt(d[0] * d[0] + d[1] * d[1] + d[2] * d[2]);
}

// Subhalo_t::PhaseSpaceDistance (src/subhalo_merge.cpp:15-20): core position
// offset over the reference's core position dispersion plus core velocity offset
// over its velocity dispersion.
template <class SubhaloT>
float PhaseSpaceDistance(const SubhaloT &sub, const SubhaloT &ref, const hbtx_params &prm)
{
  const float dx = PeriodicDistance(ref.CoreComovingPosition, sub.CoreComovingPosition, prm);
  float dv2 = 0.f;
  for (int k = 0; k < 3; k++)
  {
    const float d = (float)ref.CorePhysicalVelocity[k] - (float)sub.CorePhysicalVelocity[k];
    dv2 += d * d;
  }
  return dx / ref.CoreComovingSigmaR + std::sqrt(dv2) / ref.CorePhysicalSigmaV;
}

// Subhalo_t::MergeTo (src/subhalo_merge.cpp:183-239): the bound particles go to
// the host; the subhalo keeps its most bound tracer as an orphan.
template <class SubhaloT>
void MergeTo(SubhaloT &sub, SubhaloT &host)
{
  if (sub.Nbound <= 1)
    return;
  host.Particles.insert(host.Particles.end(), sub.Particles.begin(), sub.Particles.begin() + sub.Nbound);
  const auto &tracer = sub.Particles[sub.TracerIndex];
  sub.MostBoundParticleId = tracer.Id;
  for (int k = 0; k < 3; k++)
  {
    sub.ComovingMostBoundPosition[k] = tracer.ComovingPosition[k];
    sub.PhysicalMostBoundVelocity[k] = tracer.PhysicalVelocity[k];
    sub.ComovingAveragePosition[k] = sub.ComovingMostBoundPosition[k];
    sub.PhysicalAverageVelocity[k] = sub.PhysicalMostBoundVelocity[k];
  }
  sub.Nbound = 0;
  sub.Particles.resize(0);
  // CountParticles (src/subhalo.cpp:435-455) over no particles
  sub.Mbound = 0.f;
  detail::set_nbound_type(sub, nullptr, detail::has_NboundType<SubhaloT>());
  detail::set_mbound_type(sub, nullptr, detail::has_MboundType<SubhaloT>());
}

// Subhalo_t::MergeRecursively (src/subhalo_merge.cpp:45-91): depth first over the
// hierarchy below Subhalos[self]; every subhalo that is not already trapped and
// overlaps the reference in phase space (distance < 2) sinks into it.  Returns
// whether a resolved subhalo (Nbound > 1) merged.
template <class SubhaloT>
bool MergeRecursively(std::vector<SubhaloT> &Subhalos, int64_t self, SubhaloT &ref, int snapshot_id,
                      bool merge_trapped, const hbtx_params &prm)
{
  bool merged = false;
  if (ref.Nbound <= 1)
    return merged;
  SubhaloT &S = Subhalos[self];
  for (size_t i = 0; i < S.NestedSubhalos.size(); i++)
    merged = MergeRecursively(Subhalos, (int64_t)S.NestedSubhalos[i], ref, snapshot_id, merge_trapped, prm) | merged;
  if (S.SinkTrackId == -1 && &S != &ref)
  {
    if (PhaseSpaceDistance(S, ref, prm) < 2.f)
    {
      // SetMergerInformation (:28-39)
      S.SnapshotOfSink = snapshot_id;
      S.SinkTrackId = ref.TrackId;
      if (S.SnapshotOfDeath == -1)
        S.SnapshotOfDeath = snapshot_id;
      merged = (S.Nbound > 1) | merged;
      if (merge_trapped)
        MergeTo(S, ref);
    }
  }
  return merged;
}

// Subhalo_t::TruncateSource (src/subhalo_unbind.cpp:842-852).
template <class SubhaloT>
void TruncateSource(SubhaloT &sub, double source_sub_relax_factor)
{
  int64_t ns = (sub.Nbound <= 1) ? (int64_t)sub.Nbound : (int64_t)(sub.Nbound * source_sub_relax_factor);
  if (ns > (int64_t)sub.Particles.size())
    ns = (int64_t)sub.Particles.size();
  sub.Particles.resize(ns);
}

// Subhalo_t::RecursiveUnbind (src/subhalo_unbind.cpp:810-840) for the
// hierarchies rooted at `roots` (each subhalo's children: its NestedSubhalos),
// with the same results as calling Subhalos[r].RecursiveUnbind(Subhalos, snap)
// for every root in turn.  A subhalo depends only on its own hierarchy below it,
// so the subhaloes are unbound one level at a time, every level in one batch:
// height 0 (no children) first, then the subhaloes whose children are done.
// Before a subhalo is unbound its children's unbound particles are appended in
// NestedSubhalos order (:823); after it, its hierarchy is checked for phase-space
// mergers, and a subhalo that gained particles by one is unbound again (:833-837).
// Trees must be disjoint; a subhalo reached twice is an error.
template <class SubhaloT, class RootList>
void RecursiveUnbind(Context &ctx, const hbtx_params &prm, std::vector<SubhaloT> &Subhalos,
                     const RootList &roots, int snapshot_id, bool merge_trapped,
                     uint32_t flags = HBTX_BATCHED, bool save_binding_energies = false,
                     bool save_potential_energies = false)
{
  const int64_t ns = (int64_t)Subhalos.size();
  std::vector<int32_t> height(ns, -1);
  std::vector<char> seen(ns, 0);
  std::vector<std::vector<int64_t>> levels;
  // iterative post-order walk of every tree
  std::vector<std::pair<int64_t, size_t>> stack;
  for (size_t ri = 0; ri < (size_t)roots.size(); ri++)
  {
    const int64_t r = (int64_t)roots[ri];
    if (r < 0 || r >= ns || seen[r])
      throw Error(HBTX_E_INVALID, "RecursiveUnbind: bad or repeated root " + std::to_string(r));
    seen[r] = 1;
    stack.push_back({r, 0});
    while (!stack.empty())
    {
      auto &top = stack.back();
      const auto &kids = Subhalos[top.first].NestedSubhalos;
      if (top.second < kids.size())
      {
        const int64_t c = (int64_t)kids[top.second++];
        if (c < 0 || c >= ns || seen[c])
          throw Error(HBTX_E_INVALID, "RecursiveUnbind: subhalo " + std::to_string(c) + " reached twice");
        seen[c] = 1;
        stack.push_back({c, 0});
        continue;
      }
      int32_t h = 0;
      for (auto c : kids)
        h = std::max(h, height[c] + 1);
      height[top.first] = h;
      if ((int32_t)levels.size() <= h)
        levels.resize(h + 1);
      levels[h].push_back(top.first);
      stack.pop_back();
    }
  }
  for (auto &level : levels)
  {
    std::vector<SubhaloT *> batch;
    batch.reserve(level.size());
    for (int64_t s : level)
    {
      SubhaloT &S = Subhalos[s];
      for (auto c : S.NestedSubhalos)
      {
        const SubhaloT &C = Subhalos[c];
        S.Particles.insert(S.Particles.end(), C.Particles.begin() + C.Nbound, C.Particles.end());
      }
      batch.push_back(&S);
    }
    RefineParticles(ctx, prm, batch, snapshot_id, flags, save_binding_energies, save_potential_energies);
    std::vector<SubhaloT *> again;
    for (int64_t s : level)
      if (MergeRecursively(Subhalos, s, Subhalos[s], snapshot_id, merge_trapped, prm) && merge_trapped)
        again.push_back(&Subhalos[s]);
    if (!again.empty())
      RefineParticles(ctx, prm, again, snapshot_id, flags, save_binding_energies, save_potential_energies);
  }
}

// SubhaloSnapshot_t::RefineParticles (src/subhalo_unbind.cpp:855-919, the
// exclusive-mass build).  For every host halo h with subhaloes, the central
// SubGroups[h][0] gets the host halo's other heads (SubGroupsOfHeads[h][1..])
// appended to its NestedSubhalos for the duration of its RecursiveUnbind;
// then the field subhaloes (MemberTable.SubGroups[-1]) and the subhaloes born
// this snapshot (index >= NumSubOld) are unbound on their own, and every
// subhalo's source is truncated (TruncateSource).  The hierarchies of all host
// haloes go level by level in shared batches (RecursiveUnbind), the field and
// new-born subhaloes in one more batch.  Returns PrintSubhaloStatistics' counts;
// *elapsed_s receives the wall time (PrintTimeImbalanceStatistics' figure on
// one rank).
template <class SubhaloT, class GroupList, class HeadList, class FieldList>
SubhaloCounts RefineSnapshot(Context &ctx, const hbtx_params &prm, std::vector<SubhaloT> &Subhalos,
                             const GroupList &SubGroups, const HeadList &SubGroupsOfHeads,
                             const FieldList &FieldSubhalos, int64_t NumSubOld, int snapshot_id,
                             bool merge_trapped, double source_sub_relax_factor, double *elapsed_s = nullptr,
                             uint32_t flags = HBTX_BATCHED, bool save_binding_energies = false,
                             bool save_potential_energies = false)
{
  const auto t0 = std::chrono::steady_clock::now();
  if ((size_t)SubGroupsOfHeads.size() != (size_t)SubGroups.size())
    throw Error(HBTX_E_INVALID, "RefineSnapshot: SubGroupsOfHeads must have one entry per host halo");
  std::vector<int64_t> roots;
  std::vector<std::pair<int64_t, size_t>> restore;
  for (size_t h = 0; h < (size_t)SubGroups.size(); h++)
  {
    if (SubGroups[h].size() == 0)
      continue;
    const int64_t c = (int64_t)SubGroups[h][0];
    auto &nests = Subhalos[c].NestedSubhalos;
    restore.push_back({c, nests.size()});
    const auto &heads = SubGroupsOfHeads[h];
    for (size_t j = 1; j < (size_t)heads.size(); j++)
      nests.push_back(heads[j]);
    roots.push_back(c);
  }
  try
  {
    RecursiveUnbind(ctx, prm, Subhalos, roots, snapshot_id, merge_trapped, flags, save_binding_energies,
                    save_potential_energies);
  }
  catch (...)
  {
    for (auto &r : restore)
      Subhalos[r.first].NestedSubhalos.resize(r.second);
    throw;
  }
  for (auto &r : restore)
    Subhalos[r.first].NestedSubhalos.resize(r.second);
  std::vector<SubhaloT *> rest;
  for (size_t i = 0; i < (size_t)FieldSubhalos.size(); i++)
    rest.push_back(&Subhalos[FieldSubhalos[i]]);
  for (int64_t s = NumSubOld; s < (int64_t)Subhalos.size(); s++)
    rest.push_back(&Subhalos[s]);
  if (!rest.empty())
    RefineParticles(ctx, prm, rest, snapshot_id, flags, save_binding_energies, save_potential_energies);
  for (auto &S : Subhalos)
    TruncateSource(S, source_sub_relax_factor);
  if (elapsed_s)
    *elapsed_s = std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
  return SubhaloStatistics(Subhalos, snapshot_id);
}

namespace detail
{
// Particle_t::IsTracer (src/snapshot.h:90-97)
template <class ParticleT>
bool IsTracer(const ParticleT &p, const hbtx_params &prm)
{
  if (prm.dm_only || !has_type<ParticleT>::value)
    return true;
  return ((1 << (int)particle_type(p, has_type<ParticleT>())) & prm.tracer_mask) != 0;
}
} // namespace detail

// Subhalo_t::CountParticleTypes (src/subhalo.cpp:311-392): bound particle counts
// and masses per type (float sums in particle order), and TracerIndex = the
// first bound tracer, else 0.  Without particle types (DM_ONLY) TracerIndex = 0.
template <class SubhaloT>
void CountParticleTypes(SubhaloT &S, const hbtx_params &prm)
{
  using ParticleT = typename std::decay<decltype(S.Particles[0])>::type;
  if (prm.dm_only || !detail::has_type<ParticleT>::value)
  {
    S.TracerIndex = 0;
    return;
  }
  detail::set_nbound_type(S, nullptr, detail::has_NboundType<SubhaloT>());
  detail::set_mbound_type(S, nullptr, detail::has_MboundType<SubhaloT>());
  int64_t tracer = -1;
  const int64_t nb = std::min<int64_t>((int64_t)S.Nbound, (int64_t)S.Particles.size());
  for (int64_t i = 0; i < nb; i++)
  {
    const auto &p = S.Particles[i];
    detail::add_type(S, p,
                     detail::bool_c<detail::has_type<ParticleT>::value && detail::has_NboundType<SubhaloT>::value>(),
                     detail::bool_c<detail::has_type<ParticleT>::value && detail::has_MboundType<SubhaloT>::value>());
    if (tracer < 0 && detail::IsTracer(p, prm))
      tracer = i;
  }
  S.TracerIndex = (tracer < 0) ? 0 : tracer;
}

// Subhalo_t::KickNullParticles (src/subhalo.cpp:394-426): drop particles whose Id
// is NullParticleId (-1), keeping the order and the bound/unbound split; returns
// the number dropped.  A no-op in DM_ONLY builds.
template <class SubhaloT>
int64_t KickNullParticles(SubhaloT &S, const hbtx_params &prm)
{
  if (prm.dm_only)
    return 0;
  const int64_t np = (int64_t)S.Particles.size();
  const int64_t nb = std::min<int64_t>((int64_t)S.Nbound, np);
  int64_t keep = 0, keep_bound = 0;
  for (int64_t i = 0; i < np; i++)
  {
    if (S.Particles[i].Id == -1)
      continue;
    if (i != keep)
      S.Particles[keep] = std::move(S.Particles[i]);
    keep++;
    if (i < nb)
      keep_bound = keep;
  }
  S.Nbound = keep_bound;
  S.Particles.resize(keep);
  CountParticleTypes(S, prm);
  return np - keep;
}

// SubhaloSnapshot_t::ReassignParticles (src/subhalo_reassign_gas.cpp:65-197) on
// one rank's host haloes: SubGroups[h] lists the subhaloes of host halo h
// (MemberTable.SubGroups).  The GPU chooses each non-tracer particle's target
// (hbtx_reassign_targets: the k = num_neighbours nearest tracers of the host
// halo as GeoTree_t::NearestNeighbours finds them from 0.1 SofteningHalo); then,
// as in the reference, a particle none of whose neighbours belongs to its own
// subhalo is appended to the nearest tracer's subhalo and nulled, host halo by
// host halo in SubGroups order, and every subhalo of a host halo that has
// non-tracers gets Nbound clamped to its size and KickNullParticles.  Returns
// the number of particles moved.
template <class SubhaloT, class GroupList>
int64_t ReassignParticles(Context &ctx, const hbtx_params &prm, std::vector<SubhaloT> &Subhalos,
                          const GroupList &SubGroups, int num_neighbours = 10, double *device_ms = nullptr)
{
  const int64_t ng = (int64_t)SubGroups.size();
  if (ng == 0)
    return 0;
  std::vector<int64_t> goff(ng + 1, 0);
  for (int64_t h = 0; h < ng; h++)
  {
    int64_t c = 0;
    for (auto s : SubGroups[h])
      c += (int64_t)Subhalos[s].Particles.size();
    goff[h + 1] = goff[h] + c;
  }
  const int64_t n = goff[ng];
  std::vector<float> pos(3 * std::max<int64_t>(n, 1));
  std::vector<uint8_t> tracer(std::max<int64_t>(n, 1));
  std::vector<int32_t> sub(std::max<int64_t>(n, 1)), target(std::max<int64_t>(n, 1));
  std::vector<char> has_gas(ng, 0);
  for (int64_t h = 0, i = 0; h < ng; h++)
    for (auto s : SubGroups[h])
      for (const auto &p : Subhalos[s].Particles)
      {
        detail::put3(&pos[3 * i], p.ComovingPosition);
        tracer[i] = detail::IsTracer(p, prm) ? 1 : 0;
        has_gas[h] |= !tracer[i];
        sub[i] = (int32_t)s;
        i++;
      }
  ctx.check(hbtx_reassign_targets(ctx.get(), &prm, n, pos.data(), tracer.data(), sub.data(), ng, goff.data(),
                                  num_neighbours, 0.f, target.data(), device_ms));
  int64_t moved = 0;
  for (int64_t h = 0; h < ng; h++)
  {
    if (!has_gas[h])
      continue; // no non-tracers: the reference skips the host halo
    std::vector<int64_t> sublen;
    for (auto s : SubGroups[h])
      sublen.push_back((int64_t)Subhalos[s].Particles.size());
    int64_t i = goff[h], k = 0;
    for (auto s : SubGroups[h])
    {
      for (int64_t j = 0; j < sublen[k]; j++, i++)
      {
        const int32_t t = target[i];
        if (t < 0)
          continue;
        auto &part = Subhalos[s].Particles[j];
        Subhalos[t].Particles.push_back(part);
        part.Id = -1;
        moved++;
      }
      k++;
    }
    for (auto s : SubGroups[h])
    {
      auto &S = Subhalos[s];
      if ((int64_t)S.Nbound > (int64_t)S.Particles.size())
        S.Nbound = S.Particles.size();
      KickNullParticles(S, prm);
    }
  }
  return moved;
}

} // namespace hbtx
