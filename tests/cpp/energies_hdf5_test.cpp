// The energy datasets of a subhalo snapshot file (include/hbtx_hdf5.hpp), read
// back through the HDF5 C API.
//
// usage: energies_hdf5_test selftest <file.h5>
//          synthetic subhaloes (empty ones included) written and read back: one
//          dataset over the subhaloes, VLEN of native float, every value bit for
//          bit, the vectors cleared, and the error for a short vector.
//        energies_hdf5_test gpu <input.bin> <file.h5> <output.bin>
//          hbtx::RefineParticles with both energy vectors saved, then the two
//          datasets; output: per subhalo int64 len, float binding[len], then per
//          subhalo int64 len, float potential[len], as read back from the file.
//          input as tests/cpp/compat_test.cpp's.
#include "hbtx_compat.hpp"
#include "hbtx_hdf5.hpp"

#include <array>
#include <cstdio>
#include <cstring>
#include <vector>

typedef std::array<float, 3> HBTxyz;

struct Particle_t
{
  int64_t Id;
  HBTxyz ComovingPosition;
  HBTxyz PhysicalVelocity;
  float Mass;
};

struct Subhalo_t
{
  int64_t Nbound = 0;
  float Mbound = 0;
  int64_t TracerIndex = 0;
  float SpecificSelfPotentialEnergy = 0, SpecificSelfKineticEnergy = 0, SpecificAngularMomentum[3] = {0, 0, 0};
  HBTxyz ComovingAveragePosition, PhysicalAverageVelocity, ComovingMostBoundPosition, PhysicalMostBoundVelocity;
  int64_t MostBoundParticleId = -1;
  HBTxyz CoreComovingPosition, CorePhysicalVelocity;
  float CoreComovingSigmaR = 0, CorePhysicalSigmaV = 0;
  std::vector<Particle_t> Particles;
  std::vector<float> ParticleBindingEnergies, ParticlePotentialEnergies;
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

// one dataset read back: per subhalo the values; checks rank, extent and type
static bool read_back(hid_t file, const char *name, size_t nsub, std::vector<std::vector<float>> &out)
{
  const hid_t dset = H5Dopen2(file, name, H5P_DEFAULT);
  if (dset < 0)
    return false;
  const hid_t space = H5Dget_space(dset);
  hsize_t dims[2] = {0, 0};
  const int rank = H5Sget_simple_extent_dims(space, dims, NULL);
  const hid_t ftype = H5Dget_type(dset);
  const hid_t base = H5Tget_super(ftype);
  const bool type_ok = H5Tget_class(ftype) == H5T_VLEN && H5Tequal(base, H5T_NATIVE_FLOAT) > 0;
  const hid_t mtype = H5Tvlen_create(H5T_NATIVE_FLOAT);
  std::vector<hvl_t> vl(nsub);
  bool ok = rank == 1 && dims[0] == nsub && type_ok;
  if (ok && nsub > 0)
    ok = H5Dread(dset, mtype, H5S_ALL, H5S_ALL, H5P_DEFAULT, vl.data()) >= 0;
  out.assign(nsub, std::vector<float>());
  if (ok)
  {
    for (size_t i = 0; i < nsub; i++)
    {
      const float *p = static_cast<const float *>(vl[i].p);
      out[i].assign(p, p + vl[i].len);
    }
    if (nsub > 0)
      H5Dvlen_reclaim(mtype, space, H5P_DEFAULT, vl.data());
  }
  H5Tclose(mtype);
  H5Tclose(base);
  H5Tclose(ftype);
  H5Sclose(space);
  H5Dclose(dset);
  return ok;
}

static int selftest(const char *path)
{
  std::vector<Subhalo_t> subs(6);
  const int64_t nb[6] = {5, 0, 1, 300, 0, 17};
  std::vector<std::vector<float>> be(6), pe(6);
  for (int s = 0; s < 6; s++)
  {
    subs[s].Nbound = nb[s];
    for (int64_t i = 0; i < nb[s] + 3; i++) // (vectors may hold more than Nbound values)
    {
      subs[s].ParticleBindingEnergies.push_back(-1.f / (float)(s + 1) + 1e-3f * (float)i);
      subs[s].ParticlePotentialEnergies.push_back(-2.f * (float)(s + 1) - 1e-4f * (float)(i * i));
    }
    be[s].assign(subs[s].ParticleBindingEnergies.begin(), subs[s].ParticleBindingEnergies.begin() + nb[s]);
    pe[s].assign(subs[s].ParticlePotentialEnergies.begin(), subs[s].ParticlePotentialEnergies.begin() + nb[s]);
  }
  hid_t file = H5Fcreate(path, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0)
    return 10;
  hbtx::SaveParticleEnergies(file, subs, true, true);
  for (auto &s : subs)
    if (!s.ParticleBindingEnergies.empty() || !s.ParticlePotentialEnergies.empty())
      return 11;
  // binding only: a group of its own holds no potential dataset
  std::vector<Subhalo_t> one(1);
  one[0].Nbound = 2;
  one[0].ParticleBindingEnergies.assign({-3.f, -1.f});
  const hid_t grp = H5Gcreate2(file, "binding_only", H5P_DEFAULT, H5P_DEFAULT, H5P_DEFAULT);
  hbtx::SaveParticleEnergies(grp, one, true, false);
  const bool no_pot = H5Lexists(grp, "PotentialEnergies", H5P_DEFAULT) == 0;
  H5Gclose(grp);
  // a vector shorter than Nbound is refused
  std::vector<Subhalo_t> bad(1);
  bad[0].Nbound = 4;
  bad[0].ParticleBindingEnergies.assign(3, -1.f);
  bool threw = false;
  try
  {
    hbtx::SaveParticleEnergies(file, bad, true, false);
  }
  catch (const std::runtime_error &)
  {
    threw = true;
  }
  H5Fclose(file);
  if (!no_pot)
    return 12;
  if (!threw)
    return 13;
  file = H5Fopen(path, H5F_ACC_RDONLY, H5P_DEFAULT);
  std::vector<std::vector<float>> rb, rp, r1;
  if (!read_back(file, "BindingEnergies", 6, rb) || !read_back(file, "PotentialEnergies", 6, rp) ||
      !read_back(file, "binding_only/BindingEnergies", 1, r1))
    return 14;
  H5Fclose(file);
  for (int s = 0; s < 6; s++)
  {
    if (rb[s].size() != be[s].size() || rp[s].size() != pe[s].size())
      return 15;
    if (!be[s].empty() && (memcmp(rb[s].data(), be[s].data(), 4 * be[s].size()) != 0 ||
                           memcmp(rp[s].data(), pe[s].data(), 4 * pe[s].size()) != 0))
      return 16;
  }
  if (r1.size() != 1 || r1[0] != std::vector<float>({-3.f, -1.f}))
    return 17;
  printf("ok\n");
  return 0;
}

static int gpu(const char *in, const char *h5, const char *outp)
{
  FILE *f = fopen(in, "rb");
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
  fclose(f);
  std::vector<Subhalo_t> subs(nsub);
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
      for (int k = 0; k < 3; k++)
      {
        p.ComovingPosition[k] = pos[3 * i + k];
        p.PhysicalVelocity[k] = vel[3 * i + k];
      }
      p.Mass = mass[i];
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
    std::vector<Subhalo_t *> ptrs;
    for (auto &s : subs)
      ptrs.push_back(&s);
    hbtx::RefineParticles(ctx, prm, ptrs, 7, HBTX_BATCHED, true, true);
  }
  catch (const hbtx::Error &e)
  {
    fprintf(stderr, "hbtx error %d: %s\n", e.code, e.what());
    return 1;
  }
  hid_t file = H5Fcreate(h5, H5F_ACC_TRUNC, H5P_DEFAULT, H5P_DEFAULT);
  if (file < 0)
    return 10;
  hbtx::SaveParticleEnergies(file, subs, true, true);
  H5Fclose(file);
  file = H5Fopen(h5, H5F_ACC_RDONLY, H5P_DEFAULT);
  std::vector<std::vector<float>> rb, rpot;
  if (!read_back(file, "BindingEnergies", (size_t)nsub, rb) ||
      !read_back(file, "PotentialEnergies", (size_t)nsub, rpot))
    return 14;
  H5Fclose(file);
  FILE *o = fopen(outp, "wb");
  for (const auto *set : {&rb, &rpot})
    for (const auto &v : *set)
    {
      const int64_t len = (int64_t)v.size();
      wr(o, &len, 1);
      wr(o, v.data(), v.size());
    }
  fclose(o);
  return 0;
}

int main(int argc, char **argv)
{
  if (argc == 3 && strcmp(argv[1], "selftest") == 0)
    return selftest(argv[2]);
  if (argc == 5 && strcmp(argv[1], "gpu") == 0)
    return gpu(argv[2], argv[3], argv[4]);
  return 2;
}
