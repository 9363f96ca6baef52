// Drives the hierarchical unbinding of the C++ host layer
// (hbtx::RecursiveUnbind in include/hbtx_compat.hpp, one batch per level) and
// checks it against the reference's own control flow restated here: the
// depth-first recursion of Subhalo_t::RecursiveUnbind
// (src/subhalo_unbind.cpp:810-840) with MergeRecursively / MergeTo
// (src/subhalo_merge.cpp:45-91, :183-239), one hbtx::Unbind per subhalo.  Both
// must give identical subhaloes; the level-batched result is written for
// tests/test_gpu_hierarchy.py to compare with the CPU oracle.
//
// usage: hierarchy_test <input.bin> <output.bin> [<snapshot_output.bin>]
//   input : int64 nsub, n, nroots, nedges, merge_trapped, direct_cutoff;
//           int64 offsets[nsub+1]; float pos[n][3], vel[n][3], mass[n];
//           int64 id[n]; float ref_pos[nsub][3], ref_vel[nsub][3];
//           int64 track[nsub], sink[nsub]; int64 child_off[nsub+1],
//           child[nedges]; int64 roots[nroots]; optionally int64 num_sub_old
//           (subhaloes from there on were born this snapshot; default nsub)
//   output: per subhalo int64 nbound, sink, nparticles, mostbound_id;
//           int32 snapshot_of_sink, snapshot_of_death; float core[8];
//           then every subhalo's particle ids in order; snapshot_output: the
//           same for the subhaloes of hbtx::RefineSnapshot
//   stdout: "levels L merged M identical 1 truncate_ok 1 ... snapshot_identical 1"
//   (the last: hbtx::RefineSnapshot against SubhaloSnapshot_t::RefineParticles'
//   control flow restated with the same per-subhalo recursion)
#include "hbtx_compat.hpp"

#include <array>
#include <cmath>
#include <cstdio>
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
  int64_t TrackId = -1;
  int64_t Nbound = 0;
  float Mbound = 0;
  int64_t TracerIndex = 0;
  float SpecificSelfPotentialEnergy = 0, SpecificSelfKineticEnergy = 0, SpecificAngularMomentum[3] = {0, 0, 0};
  HBTxyz ComovingAveragePosition, PhysicalAverageVelocity, ComovingMostBoundPosition, PhysicalMostBoundVelocity;
  int64_t MostBoundParticleId = -1;
  int SnapshotOfDeath = -1, SnapshotOfSink = -1, SnapshotOfBirth = 0;
  int64_t SinkTrackId = -1, DescendantTrackId = -1;
  HBTxyz CoreComovingPosition, CorePhysicalVelocity;
  float CoreComovingSigmaR = 0, CorePhysicalSigmaV = 0;
  std::vector<int64_t> NestedSubhalos;
  std::vector<Particle_t> Particles;
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

// --- the reference's sequential control flow, restated independently -------
static hbtx_params g_prm;
static int g_snap = 7;
static bool g_merge = false;
static int g_merged = 0;

static float psd(const Subhalo_t &s, const Subhalo_t &ref)
{
  float dx = 0.f, dv = 0.f;
  for (int k = 0; k < 3; k++)
  {
    const float a = ref.CoreComovingPosition[k] - s.CoreComovingPosition[k];
    const float b = ref.CorePhysicalVelocity[k] - s.CorePhysicalVelocity[k];
    dx += a * a;
    dv += b * b;
  }
  return std::sqrt(dx) / ref.CoreComovingSigmaR + std::sqrt(dv) / ref.CorePhysicalSigmaV;
}

static void merge_to(Subhalo_t &s, Subhalo_t &host)
{
  if (s.Nbound <= 1)
    return;
  host.Particles.insert(host.Particles.end(), s.Particles.begin(), s.Particles.begin() + s.Nbound);
  const Particle_t &t = s.Particles[s.TracerIndex];
  s.MostBoundParticleId = t.Id;
  s.ComovingMostBoundPosition = s.ComovingAveragePosition = t.ComovingPosition;
  s.PhysicalMostBoundVelocity = s.PhysicalAverageVelocity = t.PhysicalVelocity;
  s.Nbound = 0;
  s.Particles.clear();
  s.Mbound = 0.f;
}

static bool merge_recursively(std::vector<Subhalo_t> &subs, int64_t self, Subhalo_t &ref)
{
  bool merged = false;
  if (ref.Nbound <= 1)
    return merged;
  Subhalo_t &S = subs[self];
  for (size_t i = 0; i < S.NestedSubhalos.size(); i++)
    merged = merge_recursively(subs, S.NestedSubhalos[i], ref) | merged;
  if (S.SinkTrackId == -1 && &S != &ref && psd(S, ref) < 2.f)
  {
    S.SnapshotOfSink = g_snap;
    S.SinkTrackId = ref.TrackId;
    if (S.SnapshotOfDeath == -1)
      S.SnapshotOfDeath = g_snap;
    merged = (S.Nbound > 1) | merged;
    g_merged++;
    if (g_merge)
      merge_to(S, ref);
  }
  return merged;
}

static void recursive_unbind(hbtx::Context &ctx, std::vector<Subhalo_t> &subs, int64_t self)
{
  for (size_t i = 0; i < subs[self].NestedSubhalos.size(); i++)
  {
    const int64_t c = subs[self].NestedSubhalos[i];
    recursive_unbind(ctx, subs, c);
    Subhalo_t &S = subs[self];
    const Subhalo_t &C = subs[c];
    S.Particles.insert(S.Particles.end(), C.Particles.begin() + C.Nbound, C.Particles.end());
  }
  hbtx::Unbind(ctx, g_prm, subs[self], g_snap);
  const bool merged = merge_recursively(subs, self, subs[self]);
  if (g_merge && merged)
    hbtx::Unbind(ctx, g_prm, subs[self], g_snap);
}

static bool same(const Subhalo_t &a, const Subhalo_t &b, int64_t s)
{
  bool ok = a.Nbound == b.Nbound && a.Mbound == b.Mbound && a.SinkTrackId == b.SinkTrackId &&
            a.SnapshotOfSink == b.SnapshotOfSink && a.SnapshotOfDeath == b.SnapshotOfDeath &&
            a.MostBoundParticleId == b.MostBoundParticleId && a.Particles.size() == b.Particles.size() &&
            a.CoreComovingPosition == b.CoreComovingPosition && a.CorePhysicalVelocity == b.CorePhysicalVelocity &&
            a.CoreComovingSigmaR == b.CoreComovingSigmaR && a.ComovingAveragePosition == b.ComovingAveragePosition;
  for (size_t i = 0; ok && i < a.Particles.size(); i++)
    ok = a.Particles[i].Id == b.Particles[i].Id;
  if (!ok)
    fprintf(stderr, "subhalo %ld differs: nbound %ld/%ld sink %ld/%ld particles %zu/%zu\n", (long)s, (long)a.Nbound,
            (long)b.Nbound, (long)a.SinkTrackId, (long)b.SinkTrackId, a.Particles.size(), b.Particles.size());
  return ok;
}

// per subhalo fields, then every subhalo's particle ids (the output format above)
static void write_subs(const char *path, const std::vector<Subhalo_t> &subs)
{
  FILE *o = fopen(path, "wb");
  if (!o)
    return;
  for (auto &S : subs)
  {
    const int64_t np = (int64_t)S.Particles.size();
    wr(o, &S.Nbound, 1);
    wr(o, &S.SinkTrackId, 1);
    wr(o, &np, 1);
    wr(o, &S.MostBoundParticleId, 1);
    wr(o, &S.SnapshotOfSink, 1);
    wr(o, &S.SnapshotOfDeath, 1);
    wr(o, S.CoreComovingPosition.data(), 3);
    wr(o, S.CorePhysicalVelocity.data(), 3);
    wr(o, &S.CoreComovingSigmaR, 1);
    wr(o, &S.CorePhysicalSigmaV, 1);
  }
  for (auto &S : subs)
    for (auto &p : S.Particles)
      wr(o, &p.Id, 1);
  fclose(o);
}

int main(int argc, char **argv)
{
  if (argc != 3 && argc != 4)
    return 2;
  FILE *f = fopen(argv[1], "rb");
  if (!f)
    return 2;
  int64_t hdr[6];
  if (!rd(f, hdr, 6))
    return 2;
  const int64_t nsub = hdr[0], n = hdr[1], nroots = hdr[2], nedges = hdr[3];
  g_merge = hdr[4] != 0;
  std::vector<int64_t> off(nsub + 1), id(n), track(nsub), sink(nsub), coff(nsub + 1), child(nedges), roots(nroots);
  std::vector<float> pos(3 * n), vel(3 * n), mass(n), rp(3 * nsub), rv(3 * nsub);
  if (!rd(f, off.data(), nsub + 1) || !rd(f, pos.data(), 3 * n) || !rd(f, vel.data(), 3 * n) ||
      !rd(f, mass.data(), n) || !rd(f, id.data(), n) || !rd(f, rp.data(), 3 * nsub) || !rd(f, rv.data(), 3 * nsub) ||
      !rd(f, track.data(), nsub) || !rd(f, sink.data(), nsub) || !rd(f, coff.data(), nsub + 1) ||
      !rd(f, child.data(), nedges) || !rd(f, roots.data(), nroots))
    return 2;
  int64_t num_sub_old = nsub;
  if (!rd(f, &num_sub_old, 1))
    num_sub_old = nsub;
  fclose(f);
  auto make = [&](std::vector<Subhalo_t> &subs) {
    subs.assign(nsub, Subhalo_t());
    for (int64_t s = 0; s < nsub; s++)
    {
      Subhalo_t &S = subs[s];
      S.TrackId = track[s];
      S.SinkTrackId = sink[s];
      for (int k = 0; k < 3; k++)
      {
        S.ComovingAveragePosition[k] = S.ComovingMostBoundPosition[k] = rp[3 * s + k];
        S.PhysicalAverageVelocity[k] = S.PhysicalMostBoundVelocity[k] = rv[3 * s + k];
      }
      S.NestedSubhalos.assign(child.begin() + coff[s], child.begin() + coff[s + 1]);
      for (int64_t i = off[s]; i < off[s + 1]; i++)
      {
        Particle_t p;
        p.Id = id[i];
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
    g_prm = hbtx::DefaultParams();
    g_prm.theta = 0.7f;
    g_prm.max_sample_size = 0;
    g_prm.periodic = 0;
    g_prm.direct_cutoff = hdr[5];
    std::vector<Subhalo_t> A, B;
    make(A);
    make(B);
    hbtx::RecursiveUnbind(ctx, g_prm, A, roots, g_snap, g_merge);
    for (int64_t r : roots)
      recursive_unbind(ctx, B, r);
    bool ok = true;
    for (int64_t s = 0; s < nsub; s++)
      ok = same(A[s], B[s], s) && ok;
    // levels of the hierarchies (longest root-to-leaf path)
    std::vector<int> depth(nsub, 0);
    int levels = 0;
    std::vector<int64_t> stack(roots.begin(), roots.end());
    while (!stack.empty())
    {
      const int64_t s = stack.back();
      stack.pop_back();
      levels = std::max(levels, depth[s] + 1);
      for (auto c : A[s].NestedSubhalos)
      {
        depth[c] = depth[s] + 1;
        stack.push_back(c);
      }
    }
    // TruncateSource (src/subhalo_unbind.cpp:842-852) with SourceSubRelaxFactor 3
    bool trunc_ok = true;
    for (int64_t s = 0; s < nsub; s++)
    {
      Subhalo_t T = A[s];
      hbtx::TruncateSource(T, 3.0);
      const size_t want = (T.Nbound <= 1) ? (size_t)T.Nbound : std::min((size_t)(T.Nbound * 3), A[s].Particles.size());
      trunc_ok = trunc_ok && T.Particles.size() == want;
      for (size_t i = 0; trunc_ok && i < T.Particles.size(); i++)
        trunc_ok = T.Particles[i].Id == A[s].Particles[i].Id;
    }
    write_subs(argv[2], A);
    // PrintSubhaloStatistics' counts; subhalo 0 born this snapshot
    A[0].SnapshotOfBirth = g_snap;
    const hbtx::SubhaloCounts sc = hbtx::SubhaloStatistics(A, g_snap);
    int64_t sunk = 0, dis = 0;
    for (int64_t s = 0; s < nsub; s++)
    {
      sunk += A[s].SnapshotOfSink == g_snap && A[s].SnapshotOfDeath == g_snap;
      dis += s != 0 && A[s].SnapshotOfDeath == g_snap && A[s].SnapshotOfSink == -1;
    }
    const bool stats_ok = sc.sunk == sunk && sc.disrupted == dis &&
                          sc.newborn + sc.fake == 1 && (sc.fake == 1) == (A[0].SnapshotOfDeath == g_snap);
    // SubhaloSnapshot_t::RefineParticles (src/subhalo_unbind.cpp:855-919): the roots
    // paired into host haloes (first of a pair the central, both heads); an odd last
    // root and its hierarchy are field subhaloes, each unbound on its own
    std::vector<std::vector<int64_t>> groups, heads;
    std::vector<int64_t> field;
    for (int64_t r = 0; r + 1 < nroots; r += 2)
    {
      groups.push_back({roots[r], roots[r + 1]});
      heads.push_back({roots[r], roots[r + 1]});
    }
    if (nroots % 2)
    {
      std::vector<int64_t> st{roots[nroots - 1]};
      while (!st.empty())
      {
        const int64_t x = st.back();
        st.pop_back();
        field.push_back(x);
        for (size_t j = 0; j < child.size() && j < (size_t)(coff[x + 1] - coff[x]); j++)
          st.push_back(child[coff[x] + j]);
      }
    }
    std::vector<Subhalo_t> C, D;
    make(C);
    make(D);
    double secs = 0.0;
    hbtx::RefineSnapshot(ctx, g_prm, C, groups, heads, field, num_sub_old, g_snap, g_merge, 3.0, &secs);
    if (argc == 4)
      write_subs(argv[3], C);
    for (size_t h = 0; h < groups.size(); h++)
    {
      auto &nests = D[groups[h][0]].NestedSubhalos;
      const size_t old = nests.size();
      nests.insert(nests.end(), heads[h].begin() + 1, heads[h].end());
      recursive_unbind(ctx, D, groups[h][0]);
      nests.resize(old);
    }
    for (int64_t x : field)
      hbtx::Unbind(ctx, g_prm, D[x], g_snap);
    for (int64_t x = num_sub_old; x < nsub; x++)
      hbtx::Unbind(ctx, g_prm, D[x], g_snap);
    for (auto &S : D)
    {
      const size_t keep = (S.Nbound <= 1) ? (size_t)S.Nbound : std::min((size_t)(S.Nbound * 3), S.Particles.size());
      S.Particles.resize(keep);
    }
    bool snap_ok = true;
    for (int64_t s = 0; s < nsub; s++)
      snap_ok = same(C[s], D[s], s) && snap_ok;
    printf("levels %d merged %d identical %d truncate_ok %d stats_ok %d sunk %ld snapshot_identical %d "
           "snapshot_groups %zu field %zu\n",
           levels, g_merged, ok ? 1 : 0, trunc_ok ? 1 : 0, stats_ok ? 1 : 0, (long)sc.sunk, snap_ok ? 1 : 0,
           groups.size(), field.size());
    return (ok && snap_ok) ? 0 : 1;
  }
  catch (const hbtx::Error &e)
  {
    fprintf(stderr, "hbtx error %d: %s\n", e.code, e.what());
    return 4;
  }
}
