// Drives hbtx::ReassignParticles (include/hbtx_compat.hpp) on hydro particles
// with the reference's member names and writes the resulting subhaloes for
// tests/test_gpu_reassign.py to compare with the oracle's targets applied by
// the reference's rules.
//
// usage: reassign_test <input.bin> <output.bin>
//   input : int64 nsub, ngroups, k; int32 periodic; float box, softening;
//           int64 nbound[nsub]; int64 npart[nsub]; int64 group_len[ngroups];
//           int64 members[nsub] (subhaloes host halo by host halo);
//           then per subhalo npart x {int64 id, float pos[3], float mass, int32 type}
//   output: int64 moved; per subhalo int64 nbound, int64 tracer_index,
//           int64 nbound_type[6], float mbound_type[6], int64 npart, int64 ids[npart]
#include "hbtx_compat.hpp"

#include <array>
#include <cstdio>
#include <vector>

typedef std::array<float, 3> HBTxyz;

struct Particle_t
{
  int64_t Id;
  HBTxyz ComovingPosition;
  HBTxyz PhysicalVelocity;
  float Mass;
  int Type;
};

struct Subhalo_t
{
  int64_t Nbound = 0;
  int64_t TracerIndex = 0;
  int64_t NboundType[6] = {0, 0, 0, 0, 0, 0};
  float MboundType[6] = {0, 0, 0, 0, 0, 0};
  std::vector<Particle_t> Particles;
};

template <class T>
static bool rd(FILE *f, T *p, size_t n)
{
  return fread(p, sizeof(T), n, f) == n;
}

int main(int argc, char **argv)
{
  if (argc != 3)
  {
    fprintf(stderr, "usage: %s in out\n", argv[0]);
    return 2;
  }
  FILE *f = fopen(argv[1], "rb");
  if (!f)
    return 2;
  int64_t hdr[3];
  int32_t periodic;
  float box, soft;
  if (!rd(f, hdr, 3) || !rd(f, &periodic, 1) || !rd(f, &box, 1) || !rd(f, &soft, 1))
    return 3;
  const int64_t nsub = hdr[0], ng = hdr[1], k = hdr[2];
  std::vector<int64_t> nbound(nsub), npart(nsub), glen(ng), members(nsub);
  if (!rd(f, nbound.data(), nsub) || !rd(f, npart.data(), nsub) || !rd(f, glen.data(), ng) ||
      !rd(f, members.data(), nsub))
    return 3;
  std::vector<Subhalo_t> subs(nsub);
  for (int64_t s = 0; s < nsub; s++)
  {
    subs[s].Nbound = nbound[s];
    subs[s].Particles.resize(npart[s]);
    for (auto &p : subs[s].Particles)
    {
      float buf[4];
      int32_t t;
      if (!rd(f, &p.Id, 1) || !rd(f, buf, 4) || !rd(f, &t, 1))
        return 3;
      p.ComovingPosition = {buf[0], buf[1], buf[2]};
      p.PhysicalVelocity = {0.f, 0.f, 0.f};
      p.Mass = buf[3];
      p.Type = t;
    }
  }
  fclose(f);
  std::vector<std::vector<int64_t>> groups(ng);
  for (int64_t h = 0, m = 0; h < ng; h++)
    for (int64_t j = 0; j < glen[h]; j++)
      groups[h].push_back(members[m++]);
  try
  {
    hbtx::Context ctx(0);
    hbtx_params prm = hbtx::DefaultParams();
    prm.periodic = periodic;
    prm.box_size = box;
    prm.softening = soft;
    prm.tracer_mask = 2 | 16; // dark matter and stars
    prm.dm_only = 0;
    const int64_t moved = hbtx::ReassignParticles(ctx, prm, subs, groups, (int)k);
    FILE *o = fopen(argv[2], "wb");
    if (!o)
      return 2;
    fwrite(&moved, sizeof(moved), 1, o);
    for (auto &S : subs)
    {
      const int64_t np = (int64_t)S.Particles.size();
      fwrite(&S.Nbound, sizeof(int64_t), 1, o);
      fwrite(&S.TracerIndex, sizeof(int64_t), 1, o);
      fwrite(S.NboundType, sizeof(int64_t), 6, o);
      fwrite(S.MboundType, sizeof(float), 6, o);
      fwrite(&np, sizeof(np), 1, o);
      for (auto &p : S.Particles)
        fwrite(&p.Id, sizeof(int64_t), 1, o);
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
