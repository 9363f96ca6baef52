// Drives the C++ host layer (include/hbtx_compat.hpp) with types that carry the
// reference's member names (Particle_t, Subhalo_t subsets) and writes the
// results for tests/test_gpu_compat.py to compare with the C ABI's.
//
// usage: compat_test <input.bin> <output.bin> [hydro]
//   input : int64 n, nsub; int64 offsets[nsub+1]; float pos[n][3], vel[n][3],
//           mass[n], ref_pos[nsub][3], ref_vel[nsub][3]
//   output: per subhalo int64 nbound, int64 mostbound_id, float mbound,
//           float com[6], int32 snapshot_of_death; then int64 final ids[n];
//           then Unbind() of subhalo 0: int64 nbound, int64 ids[n0];
//           then GravityTree over subhalo 0: double pot[k], double be[k] for the
//           first k = min(16, n0) particles; then per subhalo float
//           {Rmax, Vmax, RHalf, REnclose, Ixx, Izz weighted}, int32
//           SnapshotOfLastMaxVmax.
//   hydro: the input also holds int32 type[n], float u[n]; particles carry Type
//           and InternalEnergy, subhaloes NboundType / MboundType; unbinding with
//           thermal energy; output per subhalo int64 nbound, int64 NboundType[6],
//           float MboundType[6], float mbound, int64 TracerIndex; then int64 ids[n].
#include "hbtx_compat.hpp"

#include <array>
#include <cstdio>
#include <string>
#include <vector>

typedef std::array<float, 3> HBTxyz;

struct Particle_t
{
  int64_t Id;
  HBTxyz ComovingPosition;
  HBTxyz PhysicalVelocity;
  float Mass;
  int64_t HostId;
};

struct Subhalo_t
{
  int64_t Nbound = 0;
  float Mbound = 0;
  int64_t TracerIndex = 0;
  float SpecificSelfPotentialEnergy = 0, SpecificSelfKineticEnergy = 0, SpecificAngularMomentum[3] = {0, 0, 0};
  HBTxyz ComovingAveragePosition, PhysicalAverageVelocity, ComovingMostBoundPosition, PhysicalMostBoundVelocity;
  int64_t MostBoundParticleId = -1;
  int SnapshotOfDeath = -1, SnapshotOfSink = -1;
  int64_t SinkTrackId = -1, DescendantTrackId = -1;
  HBTxyz CoreComovingPosition, CorePhysicalVelocity;
  float CoreComovingSigmaR = 0, CorePhysicalSigmaV = 0;
  float RmaxComoving = 0, VmaxPhysical = 0, RHalfComoving = 0, REncloseComoving = 0;
  float BoundR200CritComoving = 0, BoundM200Crit = 0;
  float LastMaxVmaxPhysical = 0, RmaxComovingOfLastMaxVmax = 0;
  int SnapshotOfLastMaxVmax = -1;
  float InertialTensor[6] = {0, 0, 0, 0, 0, 0}, InertialTensorWeighted[6] = {0, 0, 0, 0, 0, 0};
  std::vector<Particle_t> Particles;
  std::vector<float> ParticleBindingEnergies;
};

struct SnapView
{
  const std::vector<Particle_t> &P;
  size_t size() const
  {
    return P.size();
  }
  const HBTxyz &GetComovingPosition(size_t i) const
  {
    return P[i].ComovingPosition;
  }
  float GetMass(size_t i) const
  {
    return P[i].Mass;
  }
};

struct HydroParticle_t
{
  int64_t Id;
  HBTxyz ComovingPosition;
  HBTxyz PhysicalVelocity;
  float Mass;
  float InternalEnergy;
  int Type;
};

struct HydroSubhalo_t
{
  int64_t Nbound = 0;
  float Mbound = 0;
  int64_t NboundType[6] = {0, 0, 0, 0, 0, 0};
  float MboundType[6] = {0, 0, 0, 0, 0, 0};
  int64_t TracerIndex = 0;
  float SpecificSelfPotentialEnergy = 0, SpecificSelfKineticEnergy = 0, SpecificAngularMomentum[3] = {0, 0, 0};
  HBTxyz ComovingAveragePosition, PhysicalAverageVelocity, ComovingMostBoundPosition, PhysicalMostBoundVelocity;
  int64_t MostBoundParticleId = -1;
  HBTxyz CoreComovingPosition, CorePhysicalVelocity;
  float CoreComovingSigmaR = 0, CorePhysicalSigmaV = 0;
  std::vector<HydroParticle_t> Particles;
};

template <class T>
static bool rd(FILE *f, T *p, size_t n)
{
  return fread(p, sizeof(T), n, f) == n;
}
template <class T>
static void wr(FILE *f, const T *p, size_t n)
{
  fwrite(p, sizeof(T), n, f);
}

static int hydro(const std::vector<int64_t> &off, const std::vector<float> &pos, const std::vector<float> &vel,
                 const std::vector<float> &mass, const std::vector<float> &rp, const std::vector<float> &rv,
                 const std::vector<int32_t> &type, const std::vector<float> &u, const char *outp)
{
  const int64_t nsub = (int64_t)off.size() - 1;
  std::vector<HydroSubhalo_t> subs(nsub);
  for (int64_t s = 0; s < nsub; s++)
  {
    HydroSubhalo_t &S = subs[s];
    for (int k = 0; k < 3; k++)
    {
      S.ComovingAveragePosition[k] = S.ComovingMostBoundPosition[k] = rp[3 * s + k];
      S.PhysicalAverageVelocity[k] = S.PhysicalMostBoundVelocity[k] = rv[3 * s + k];
    }
    for (int64_t i = off[s]; i < off[s + 1]; i++)
    {
      HydroParticle_t p;
      p.Id = i - off[s];
      for (int k = 0; k < 3; k++)
      {
        p.ComovingPosition[k] = pos[3 * i + k];
        p.PhysicalVelocity[k] = vel[3 * i + k];
      }
      p.Mass = mass[i];
      p.InternalEnergy = u[i];
      p.Type = type[i];
      S.Particles.push_back(p);
    }
  }
  try
  {
    hbtx::Context ctx(0);
    hbtx_params prm = hbtx::DefaultParams();
    prm.theta = 0.7f;
    prm.max_sample_size = 0;
    prm.periodic = 0;
    prm.direct_cutoff = 0;
    prm.dm_only = 0;
    prm.thermal = 1;
    std::vector<HydroSubhalo_t *> ptrs;
    for (auto &s : subs)
      ptrs.push_back(&s);
    hbtx::RefineParticles(ctx, prm, ptrs, 7);
  }
  catch (const hbtx::Error &e)
  {
    fprintf(stderr, "hbtx error %d: %s\n", e.code, e.what());
    return 1;
  }
  FILE *o = fopen(outp, "wb");
  for (auto &S : subs)
  {
    wr(o, &S.Nbound, 1);
    wr(o, S.NboundType, 6);
    wr(o, S.MboundType, 6);
    wr(o, &S.Mbound, 1);
    wr(o, &S.TracerIndex, 1);
  }
  for (auto &S : subs)
    for (auto &p : S.Particles)
      wr(o, &p.Id, 1);
  fclose(o);
  return 0;
}

int main(int argc, char **argv)
{
  const bool with_types = argc == 4 && std::string(argv[3]) == "hydro";
  if (argc != 3 && !with_types)
    return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f)
    return 2;
  int64_t hdr[2];
  if (!rd(f, hdr, 2))
    return 2;
  const int64_t n = hdr[0], nsub = hdr[1];
  std::vector<int64_t> off(nsub + 1);
  std::vector<float> pos(3 * n), vel(3 * n), mass(n), rp(3 * nsub), rv(3 * nsub);
  if (!rd(f, off.data(), nsub + 1) || !rd(f, pos.data(), 3 * n) || !rd(f, vel.data(), 3 * n) ||
      !rd(f, mass.data(), n) || !rd(f, rp.data(), 3 * nsub) || !rd(f, rv.data(), 3 * nsub))
    return 2;
  if (with_types)
  {
    std::vector<int32_t> type(n);
    std::vector<float> u(n);
    if (!rd(f, type.data(), n) || !rd(f, u.data(), n))
      return 2;
    fclose(f);
    return hydro(off, pos, vel, mass, rp, rv, type, u, argv[2]);
  }
  fclose(f);
  auto make = [&](std::vector<Subhalo_t> &subs) {
    subs.assign(nsub, Subhalo_t());
    for (int64_t s = 0; s < nsub; s++)
    {
      Subhalo_t &S = subs[s];
      for (int k = 0; k < 3; k++)
      {
        S.ComovingAveragePosition[k] = S.ComovingMostBoundPosition[k] = rp[3 * s + k];
        S.PhysicalAverageVelocity[k] = S.PhysicalMostBoundVelocity[k] = rv[3 * s + k];
      }
      for (int64_t i = off[s]; i < off[s + 1]; i++)
      {
        Particle_t p;
        p.Id = i - off[s];
        p.HostId = -1;
        for (int k = 0; k < 3; k++)
        {
          p.ComovingPosition[k] = pos[3 * i + k];
          p.PhysicalVelocity[k] = vel[3 * i + k];
        }
        p.Mass = mass[i];
        S.Particles.push_back(p);
      }
    }
  };
  try
  {
    hbtx::Context ctx(0);
    hbtx_params prm = hbtx::DefaultParams();
    prm.theta = 0.7f;
    prm.max_sample_size = 0;
    prm.periodic = 0;
    prm.direct_cutoff = 0;
    std::vector<Subhalo_t> subs;
    make(subs);
    std::vector<Subhalo_t *> ptrs;
    for (auto &s : subs)
      ptrs.push_back(&s);
    hbtx::RefineParticles(ctx, prm, ptrs, 7, HBTX_BATCHED, true, false);
    hbtx::CalculateProfileProperties(ctx, prm, ptrs, 7);
    FILE *o = fopen(argv[2], "wb");
    for (auto &S : subs)
    {
      wr(o, &S.Nbound, 1);
      wr(o, &S.MostBoundParticleId, 1);
      wr(o, &S.Mbound, 1);
      wr(o, S.ComovingAveragePosition.data(), 3);
      wr(o, S.PhysicalAverageVelocity.data(), 3);
      wr(o, &S.SnapshotOfDeath, 1);
      if ((int64_t)S.ParticleBindingEnergies.size() != S.Nbound)
        return 3;
    }
    for (auto &S : subs)
      for (auto &p : S.Particles)
        wr(o, &p.Id, 1);
    // the plain host-driven path for one subhalo
    std::vector<Subhalo_t> again;
    make(again);
    hbtx::Unbind(ctx, prm, again[0], 7);
    wr(o, &again[0].Nbound, 1);
    for (auto &p : again[0].Particles)
      wr(o, &p.Id, 1);
    // GravityTree_t over subhalo 0's input particles
    std::vector<Subhalo_t> fresh;
    make(fresh);
    hbtx::GravityTree tree(ctx, prm);
    tree.Reserve(fresh[0].Particles.size());
    SnapView view{fresh[0].Particles};
    tree.Build(view);
    const size_t k = std::min<size_t>(16, fresh[0].Particles.size());
    for (size_t i = 0; i < k; i++)
    {
      const Particle_t &p = fresh[0].Particles[i];
      double pot = tree.EvaluatePotential(p.ComovingPosition, p.Mass);
      wr(o, &pot, 1);
    }
    for (size_t i = 0; i < k; i++)
    {
      const Particle_t &p = fresh[0].Particles[i];
      double be = tree.BindingEnergy(p.ComovingPosition, p.PhysicalVelocity, fresh[0].ComovingAveragePosition,
                                     fresh[0].PhysicalAverageVelocity, p.Mass);
      wr(o, &be, 1);
    }
    // profile and shape of every subhalo of the batch
    for (auto &S : subs)
    {
      const float pr[6] = {S.RmaxComoving, S.VmaxPhysical, S.RHalfComoving, S.REncloseComoving,
                           S.InertialTensor[0], S.InertialTensorWeighted[5]};
      wr(o, pr, 6);
      wr(o, &S.SnapshotOfLastMaxVmax, 1);
    }
    fclose(o);
  }
  catch (const hbtx::Error &e)
  {
    fprintf(stderr, "hbtx error %d: %s\n", e.code, e.what());
    return 1;
  }
  return 0;
}
