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

#include <array>
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
template <class T, class = void>
struct has_internal_energy : std::false_type
{
};
template <class T>
struct has_internal_energy<T, std::void_t<decltype(std::declval<T>().InternalEnergy)>> : std::true_type
{
};
template <class T, class = void>
struct has_type : std::false_type
{
};
template <class T>
struct has_type<T, std::void_t<decltype(std::declval<T>().Type)>> : std::true_type
{
};
#define HBTX_MEMBER_TRAIT(name)                                                                                  \
  template <class T, class = void>                                                                             \
  struct has_##name : std::false_type                                                                          \
  {                                                                                                            \
  };                                                                                                           \
  template <class T>                                                                                           \
  struct has_##name<T, std::void_t<decltype(std::declval<T>().name)>> : std::true_type                         \
  {                                                                                                            \
  };
HBTX_MEMBER_TRAIT(NboundType)
HBTX_MEMBER_TRAIT(MboundType)
HBTX_MEMBER_TRAIT(NumberUnbindingIterations)
HBTX_MEMBER_TRAIT(ParticleBindingEnergies)
HBTX_MEMBER_TRAIT(ParticlePotentialEnergies)
HBTX_MEMBER_TRAIT(SnapshotOfDeath)
HBTX_MEMBER_TRAIT(DescendantTrackId)
HBTX_MEMBER_TRAIT(SinkTrackId)
HBTX_MEMBER_TRAIT(SnapshotOfSink)
#undef HBTX_MEMBER_TRAIT

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
      if constexpr (thermal)
        u[k] = (float)p.InternalEnergy;
      if constexpr (typed)
        type[k] = (int32_t)p.Type;
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
    if constexpr (detail::has_NboundType<SubhaloT>::value)
      for (int t = 0; t < 6; t++)
        S.NboundType[t] = R.nbound_type[t];
    if constexpr (detail::has_MboundType<SubhaloT>::value)
      for (int t = 0; t < 6; t++)
        S.MboundType[t] = R.mbound_type[t];
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
    if constexpr (detail::has_NumberUnbindingIterations<SubhaloT>::value)
      S.NumberUnbindingIterations = R.iterations;
    if (N == 0)
      continue; // orphan: nothing else changes (:471-489)
    if (R.disrupted)
    {
      // :646-668 (MostBoundParticleId keeps its previous value)
      if constexpr (detail::has_SnapshotOfDeath<SubhaloT>::value)
        if (S.SnapshotOfDeath == -1)
          S.SnapshotOfDeath = snapshot_id;
    }
    else
    {
      // :690-702, :791-792
      if constexpr (detail::has_SnapshotOfDeath<SubhaloT>::value && detail::has_DescendantTrackId<SubhaloT>::value)
        if (S.SnapshotOfDeath != -1)
        {
          S.SnapshotOfDeath = -1;
          S.DescendantTrackId = -1;
        }
      if constexpr (detail::has_SinkTrackId<SubhaloT>::value && detail::has_SnapshotOfSink<SubhaloT>::value)
        if (S.SinkTrackId != -1)
        {
          S.SnapshotOfSink = -1;
          S.SinkTrackId = -1;
        }
      S.MostBoundParticleId = R.mostbound_id;
    }
    const int64_t nb = R.nbound;
    if constexpr (detail::has_ParticleBindingEnergies<SubhaloT>::value)
      if (save_binding_energies)
        S.ParticleBindingEnergies.assign(energy.begin() + a, energy.begin() + a + nb);
    if constexpr (detail::has_ParticlePotentialEnergies<SubhaloT>::value)
      if (save_potential_energies)
        S.ParticlePotentialEnergies.assign(pot.begin() + a, pot.begin() + a + nb);
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

} // namespace hbtx
