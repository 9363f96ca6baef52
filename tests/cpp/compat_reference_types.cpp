// Compile-only check of the C++ host layer against the reference's own types.
// Built with `g++ -std=c++11 -fsyntax-only` and the reference's src/ on the
// include path (the reference requires C++11, CMakeLists.txt:38-39); nothing is
// linked or run.  Each function is what a maintainer would write inside the
// reference to hand its path to the GPU:
//   refine_particles    SubhaloSnapshot_t::RefineParticles   src/subhalo_unbind.cpp:855-919
//   unbind_one          Subhalo_t::Unbind                     src/subhalo_unbind.cpp:463-808
//   tree_potential      GravityTree_t Build/EvaluatePotential src/gravity_tree.cpp:50-166
//   reassign            SubhaloSnapshot_t::ReassignParticles  src/subhalo_reassign_gas.cpp:65-197
//   properties          Subhalo_t::CalculateProfileProperties / CalculateShape, src/subhalo.cpp:175-309
#include "subhalo.h"

#include "hbtx_compat.hpp"

hbtx::SubhaloCounts refine_particles(hbtx::Context &ctx, const hbtx_params &prm, SubhaloSnapshot_t &snap,
                                     double relax_factor)
{
  MemberShipTable_t &T = snap.MemberTable;
  std::vector<HBTInt> field(T.SubGroups[-1].data(), T.SubGroups[-1].data() + T.SubGroups[-1].size());
  return hbtx::RefineSnapshot(ctx, prm, snap.Subhalos, T.SubGroups, T.SubGroupsOfHeads, field,
                              (int64_t)T.AllMembers.size(), snap.GetSnapshotId(), true, relax_factor);
}

void unbind_one(hbtx::Context &ctx, const hbtx_params &prm, Subhalo_t &sub, const Snapshot_t &epoch)
{
  hbtx::Unbind(ctx, prm, sub, epoch.GetSnapshotId(), true, true);
  std::vector<Subhalo_t *> subs(1, &sub);
  hbtx::RefineParticles(ctx, prm, subs, epoch.GetSnapshotId());
  hbtx::CountParticleTypes(sub, prm);
  hbtx::KickNullParticles(sub, prm);
  hbtx::TruncateSource(sub, 3.0);
}

double tree_potential(hbtx::Context &ctx, const hbtx_params &prm, const Snapshot_t &snap, const HBTxyz &x,
                      const HBTxyz &v, const HBTxyz &ref_x, const HBTxyz &ref_v)
{
  hbtx::GravityTree tree(ctx, prm);
  tree.Reserve(snap.size());
  tree.Build(snap);
  return tree.EvaluatePotential(x, 0.f) + tree.BindingEnergy(x, v, ref_x, ref_v, 0.f);
}

int64_t reassign(hbtx::Context &ctx, const hbtx_params &prm, SubhaloSnapshot_t &snap)
{
  return hbtx::ReassignParticles(ctx, prm, snap.Subhalos, snap.MemberTable.SubGroups, 10);
}

void properties(hbtx::Context &ctx, const hbtx_params &prm, SubhaloSnapshot_t &snap)
{
  std::vector<Subhalo_t *> subs;
  for (size_t i = 0; i < snap.Subhalos.size(); i++)
    subs.push_back(&snap.Subhalos[i]);
  hbtx::CalculateProfileProperties(ctx, prm, subs, snap.GetSnapshotId(), true);
  std::vector<HBTInt> roots(1, 0);
  hbtx::RecursiveUnbind(ctx, prm, snap.Subhalos, roots, snap.GetSnapshotId(), true);
}
