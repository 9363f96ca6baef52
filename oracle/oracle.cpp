// CPU oracle for the gravitational-unbinding path (test infrastructure only; see
// oracle.h).  Every function cites the reference lines whose behaviour it
// restates.  Types follow the reference's default build: HBTReal = float,
// HBTInt = int64, double walk/accumulators.
#include "oracle.h"

#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstring>
#include <queue>
#include <vector>

namespace
{
typedef std::array<float, 3> xyz;
const int TYPE_MAX = 6;

struct Cfg
{
  ounb_params p;
  double box_half;
  float theta2;     // TreeNodeOpenAngleSquare (float), src/config_parser.cpp:212
  float res;        // TreeNodeResolution = 0.1 SofteningHalo, :210
  float res_half;   // TreeNodeResolutionHalf, :211
  float eps;        // GetCurrentSoftening, :641-649
  explicit Cfg(const ounb_params &q) : p(q)
  {
    box_half = q.box_size / 2.;
    theta2 = q.theta * q.theta;
    res = q.softening * 0.1f;
    res_half = res / 2.f;
    eps = (q.max_phys_softening == -1.f) ? q.softening : std::min(q.softening, q.max_phys_softening / q.scale_factor);
  }
  // NEAREST, src/config_parser.h:218-219
  double nearest(double x) const
  {
    return (x > box_half) ? (x - p.box_size) : ((x < -box_half) ? (x + p.box_size) : x);
  }
};

struct Particle
{
  int64_t id;
  xyz x;
  xyz v;
  float m;
  float u;
  int type;
};

// Particle_t::IsTracer / DoNotSubsample, src/snapshot.h:90-105
inline int is_tracer(const Cfg &c, const Particle &p)
{
  return c.p.dm_only ? 1 : ((1 << p.type) & c.p.tracer_mask);
}
inline int do_not_subsample(const Cfg &c, const Particle &p)
{
  return c.p.dm_only ? 0 : ((1 << p.type) & c.p.donot_subsample_mask);
}

// Snapshot_t::RelativeVelocity, src/snapshot.h:135-148 (float arithmetic)
inline void relative_velocity(const Cfg &c, const xyz &tp, const xyz &tv, const xyz &rp, const xyz &rv, xyz &dv)
{
  for (int j = 0; j < 3; j++)
  {
    float dx = tp[j] - rp[j];
    if (c.p.periodic)
      dx = c.nearest(dx);
    dv[j] = tv[j] - rv[j];
    dv[j] += c.p.hz * c.p.scale_factor * dx;
  }
}

inline float vecdot(const xyz &a, const xyz &b)
{
  return a[0] * b[0] + a[1] * b[1] + a[2] * b[2];
}

// ---------------------------------------------------------------------------
// Octree with single-particle leaves (TreeCell_t union, src/oct_tree.h:29-45)
// ---------------------------------------------------------------------------
struct Cell
{
  union
  {
    int64_t sons[8];
    struct
    {
      float s[3];
      float len;
      float mass;
      int64_t sibling;
      int64_t nextnode;
    } way;
  };
  explicit Cell(int64_t i)
  {
    for (int k = 0; k < 8; k++)
      sons[k] = i;
  }
};

// Source accessor: position and (possibly upscaled) mass of tree particle i.
struct Sources
{
  virtual ~Sources() {}
  virtual const xyz &pos(int64_t i) const = 0;
  virtual float mass(int64_t i) const = 0;
};

class Tree
{
public:
  std::vector<Cell> cells;
  std::vector<int64_t> next_from_particle;
  const Sources *src = nullptr;
  int64_t np = 0;
  const Cfg *cfg = nullptr;
  uint64_t rng_state = 0x9E3779B97F4A7C15ull;

  Cell *nodes()
  {
    return cells.data() - np;
  }

  double rand01()
  {
    // xorshift64*: deterministic stand-in for drand48 (src/oct_tree.tpp:432)
    rng_state ^= rng_state >> 12;
    rng_state ^= rng_state << 25;
    rng_state ^= rng_state >> 27;
    return (double)((rng_state * 2685821657736338717ull) >> 11) * (1.0 / 9007199254740992.0);
  }

  // OctTree_t::Build, src/oct_tree.tpp:11-142
  int64_t build(const Cfg &c, const Sources &s, int64_t num_part)
  {
    cfg = &c;
    src = &s;
    np = num_part;
    next_from_particle.assign(np, -1);
    cells.clear();
    cells.reserve(std::max<int64_t>(10, np));
    double xmin[3], xmax[3], Center[3], Len, Lenhalf;
    for (int j = 0; j < 3; j++)
      xmin[j] = xmax[j] = s.pos(0)[j];
    for (int64_t i = 1; i < np; i++)
      for (int j = 0; j < 3; j++)
      {
        if (s.pos(i)[j] > xmax[j])
          xmax[j] = s.pos(i)[j];
        else if (s.pos(i)[j] < xmin[j])
          xmin[j] = s.pos(i)[j];
      }
    Len = xmax[0] - xmin[0];
    for (int j = 1; j < 3; j++)
      if ((xmax[j] - xmin[j]) > Len)
        Len = xmax[j] - xmin[j];
    for (int j = 0; j < 3; j++)
      Center[j] = 0.5 * (xmax[j] + xmin[j]);
    Lenhalf = 0.5 * Len;
    cells.emplace_back(-1);
    for (int64_t i = 0; i < np; i++)
    {
      int64_t nodeid = np;
      double lenhalf = Lenhalf, center[3] = {Center[0], Center[1], Center[2]};
      while (true)
      {
        lenhalf *= 0.5;
        int sub = 0;
        const xyz &x = s.pos(i);
        for (int j = 0; j < 3; j++)
        {
          if (x[j] > center[j])
          {
            center[j] += lenhalf;
            sub += (1 << j);
          }
          else
            center[j] -= lenhalf;
        }
        int64_t subid = nodes()[nodeid].sons[sub];
        if (subid < 0)
        {
          nodes()[nodeid].sons[sub] = i;
          break;
        }
        else if (subid < np)
        {
          int64_t newnodeid = np + (int64_t)cells.size();
          cells.emplace_back(-1);
          nodes()[nodeid].sons[sub] = newnodeid;
          nodeid = newnodeid;
          if (lenhalf < c.res_half)
          {
            sub = (int)(8.0 * rand01());
            if (sub >= 8)
              sub = 7;
          }
          else
          {
            sub = 0;
            const xyz &y = s.pos(subid);
            for (int j = 0; j < 3; j++)
              if (y[j] > center[j])
                sub += (1 << j);
          }
          nodes()[nodeid].sons[sub] = subid;
        }
        else
          nodeid = subid;
      }
    }
    update_internal(np, -1, Len);
    return (int64_t)cells.size();
  }

  // GravityTree_t::ProcessNode, src/gravity_tree.cpp:25-48
  void process(int64_t nodeid, int64_t nextid, double &mass, double com[3], double len)
  {
    if (nodeid < np)
    {
      double m = src->mass(nodeid);
      mass += m;
      const xyz &x = src->pos(nodeid);
      for (int j = 0; j < 3; j++)
        com[j] += x[j] * m;
      next_from_particle[nodeid] = nextid;
    }
    else
    {
      if (len >= cfg->res)
        update_internal(nodeid, nextid, len / 2.);
      else
        update_internal(nodeid, nextid, len);
      double m = nodes()[nodeid].way.mass;
      mass += m;
      for (int j = 0; j < 3; j++)
        com[j] += nodes()[nodeid].way.s[j] * m;
    }
  }

  // GravityTree_t::UpdateInternalNodes, src/gravity_tree.cpp:50-80
  void update_internal(int64_t no, int64_t sib, double len)
  {
    int64_t sons[8];
    double mass = 0., com[3] = {0., 0., 0.};
    for (int j = 0; j < 8; j++)
      sons[j] = nodes()[no].sons[j];
    nodes()[no].way.len = len;
    nodes()[no].way.sibling = sib;
    int i = 0;
    while (sons[i] < 0)
      i++;
    int64_t pp = sons[i];
    nodes()[no].way.nextnode = pp;
    for (i++; i < 8; i++)
      if (sons[i] >= 0)
      {
        int64_t p = pp;
        pp = sons[i];
        process(p, pp, mass, com, len);
      }
    process(pp, sib, mass, com, len);
    nodes()[no].way.mass = mass;
    for (int j = 0; j < 3; j++)
      nodes()[no].way.s[j] = com[j] / mass;
  }

  // GravityTree_t::EvaluatePotential, src/gravity_tree.cpp:82-166
  double potential(const xyz &target, float target_mass) const
  {
    const Cfg &c = *cfg;
    const Cell *nd = cells.data() - np;
    double h = 2.8 * c.eps, h_inv = 1.0 / h;
    double pot = target_mass / c.eps;
    double px = target[0], py = target[1], pz = target[2];
    int64_t no = np;
    while (no >= 0)
    {
      double dx, dy, dz, mass, r2;
      if (no < np)
      {
        const xyz &p = src->pos(no);
        dx = p[0] - px;
        dy = p[1] - py;
        dz = p[2] - pz;
        if (c.p.periodic)
        {
          dx = c.nearest(dx);
          dy = c.nearest(dy);
          dz = c.nearest(dz);
        }
        mass = src->mass(no);
        no = next_from_particle[no];
        r2 = dx * dx + dy * dy + dz * dz;
      }
      else
      {
        const Cell &n = nd[no];
        dx = n.way.s[0] - px;
        dy = n.way.s[1] - py;
        dz = n.way.s[2] - pz;
        if (c.p.periodic)
        {
          dx = c.nearest(dx);
          dy = c.nearest(dy);
          dz = c.nearest(dz);
        }
        mass = n.way.mass;
        r2 = dx * dx + dy * dy + dz * dz;
        if ((n.way.len * n.way.len) > (r2 * c.theta2))
        {
          no = n.way.nextnode;
          continue;
        }
        no = n.way.sibling;
      }
      double r = std::sqrt(r2);
      if (r >= h)
        pot -= mass / r;
      else
      {
        double u = r * h_inv, wp;
        if (u < 0.5)
          wp = -2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6));
        else
          wp = -3.2 + 0.066666666667 / u + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u)));
        pot += mass * h_inv * wp;
      }
    }
    return pot * c.p.G / c.p.scale_factor;
  }

  // GravityTree_t::BindingEnergy, src/gravity_tree.cpp:168-178
  double binding_energy(const xyz &tp, const xyz &tv, const xyz &rp, const xyz &rv, float target_mass) const
  {
    double pot = potential(tp, target_mass);
    xyz dv;
    relative_velocity(*cfg, tp, tv, rp, rv, dv);
    return vecdot(dv, dv) * 0.5 + pot;
  }
};

// ---------------------------------------------------------------------------
// Unbinding (Subhalo_t::Unbind and the EnergySnapshot_t adapter)
// ---------------------------------------------------------------------------
struct PE
{
  float E;
  int64_t idx;
};

// EnergySnapshot_t, src/subhalo_unbind.cpp:46-393
struct ESnap : public Sources
{
  PE *elist;
  int64_t n;
  const std::vector<Particle> &P;
  const Cfg &c;
  float upscale[TYPE_MAX];
  ESnap(PE *e, int64_t n_, const std::vector<Particle> &p, const Cfg &cfg) : elist(e), n(n_), P(p), c(cfg)
  {
    reset();
  }
  const Particle &part(int64_t i) const
  {
    return P[elist[i].idx];
  }
  const xyz &pos(int64_t i) const override
  {
    return part(i).x;
  }
  // GetMass, :91-101
  float mass(int64_t i) const override
  {
    const Particle &q = part(i);
    if (c.p.dm_only)
      return q.m * upscale[1];
    if (do_not_subsample(c, q))
      return q.m;
    return q.m * upscale[q.type];
  }
  void reset()
  {
    for (int k = 0; k < TYPE_MAX; k++)
      upscale[k] = 1.f;
  }
  // CentreOfMassReferenceFrame, :151-198
  double com_frame(xyz &cp, xyz &cv, int64_t num) const
  {
    double tm = 0, wp[3] = {0, 0, 0}, wv[3] = {0, 0, 0}, refp[3] = {0, 0, 0};
    if (c.p.periodic)
      for (int j = 0; j < 3; j++)
        refp[j] = pos(0)[j];
    for (int64_t i = 0; i < num; i++)
    {
      const float m = mass(i);
      const xyz &x = pos(i);
      const xyz &v = part(i).v;
      tm += m;
      for (int j = 0; j < 3; j++)
      {
        wp[j] += c.p.periodic ? c.nearest(x[j] - refp[j]) * m : x[j] * m;
        wv[j] += v[j] * m;
      }
    }
    for (int j = 0; j < 3; j++)
    {
      wp[j] /= tm;
      wv[j] /= tm;
    }
    if (c.p.periodic)
      for (int j = 0; j < 3; j++)
        wp[j] += refp[j];
    for (int j = 0; j < 3; j++)
    {
      cp[j] = wp[j];
      cv[j] = wv[j];
    }
    return tm;
  }
  // AverageKinematics, :200-249
  void average_kinematics(float &spot, float &skin, float sam[3], int64_t num, const xyz &rp, const xyz &rv) const
  {
    if (num <= 1)
    {
      spot = skin = 0.f;
      sam[0] = sam[1] = sam[2] = 0.f;
      return;
    }
    double E = 0., K = 0., AMx = 0., AMy = 0., AMz = 0., M = 0.;
    for (int64_t i = 0; i < num; i++)
    {
      float m = mass(i);
      E += elist[i].E * m;
      if (c.p.thermal)
        E -= part(i).u * m;
      const xyz &x = pos(i);
      const xyz &v = part(i).v;
      double dx[3], dv[3];
      for (int j = 0; j < 3; j++)
      {
        dx[j] = x[j] - rp[j];
        if (c.p.periodic)
          dx[j] = c.nearest(dx[j]);
        dx[j] *= c.p.scale_factor;
        dv[j] = v[j] - rv[j] + c.p.hz * dx[j];
        K += dv[j] * dv[j] * m;
      }
      AMx += (dx[1] * dv[2] - dx[2] * dv[1]) * m;
      AMy += (dx[2] * dv[0] - dx[0] * dv[2]) * m;
      AMz += (dx[0] * dv[1] - dx[1] * dv[0]) * m;
      M += m;
    }
    E /= M;
    K *= 0.5 / M;
    spot = E - K;
    skin = K;
    sam[0] = AMx / M;
    sam[1] = AMy / M;
    sam[2] = AMz / M;
  }
  // AccumulateMass overloads, :253-285
  void acc_mass(int64_t a, int64_t b, float &m) const
  {
    m = 0;
    for (int64_t i = a; i < b; i++)
      m += mass(i);
  }
  void acc_mass_type(int64_t a, int64_t b, float mt[TYPE_MAX]) const
  {
    for (int k = 0; k < TYPE_MAX; k++)
      mt[k] = 0;
    for (int64_t i = a; i < b; i++)
      mt[part(i).type] += mass(i);
  }
  // GetMassUpscaleFactor, :290-317
  void upscale_global(int64_t nlast, float mlast, int64_t mss, int64_t nuns)
  {
    reset();
    float munsampled = 0;
    for (int64_t i = 0; i < nuns; i++)
      munsampled += mass(i);
    float mtrue = mlast - munsampled;
    float msub = 0;
    for (int64_t i = nuns; i < nuns + mss; i++)
      msub += mass(i);
    for (int k = 0; k < TYPE_MAX; k++)
      upscale[k] = mtrue / msub;
  }
  // GetMassUpscaleFactorPerParticleType, :323-361
  void upscale_per_type(int64_t nlast, int64_t mss, int64_t nuns)
  {
    reset();
    float msub[TYPE_MAX], mtot[TYPE_MAX];
    acc_mass_type(nuns, nuns + mss, msub);
    acc_mass_type(nuns + mss, nlast, mtot);
    for (int k = 0; k < TYPE_MAX; k++)
      mtot[k] += msub[k];
    float f[TYPE_MAX];
    for (int k = 0; k < TYPE_MAX; k++)
      f[k] = msub[k] ? mtot[k] / msub[k] : 0.f;
    float ttrue = 0, tscaled = 0;
    for (int k = 0; k < TYPE_MAX; k++)
    {
      ttrue += mtot[k];
      tscaled += msub[k] * f[k];
    }
    for (int k = 0; k < TYPE_MAX; k++)
      upscale[k] = f[k] * (ttrue / tscaled);
  }
  // SetMassUpscaleFactor, :374-391
  int64_t set_upscale(int64_t nlast, float mlast, int64_t mss, int64_t nuns)
  {
    if ((mss == 0) || (nlast < (mss + nuns)))
      return nlast;
    if (c.p.dm_only || !c.p.upscale_per_type)
      upscale_global(nlast, mlast, mss, nuns);
    else
      upscale_per_type(nlast, mss, nuns);
    return mss + nuns;
  }
  // GetPotentialEnergy, :114-135
  float potential_energy(int64_t i, const xyz &rp, const xyz &rv) const
  {
    float E = elist[i].E;
    if (c.p.thermal)
      E -= part(i).u;
    const xyz &x = pos(i);
    const xyz &v = part(i).v;
    double dx[3], dv[3];
    for (int j = 0; j < 3; j++)
    {
      dx[j] = x[j] - rp[j];
      if (c.p.periodic)
        dx[j] = c.nearest(dx[j]);
      dx[j] *= c.p.scale_factor;
      dv[j] = v[j] - rv[j] + c.p.hz * dx[j];
      E -= 0.5 * dv[j] * dv[j];
    }
    return E;
  }
};

struct SubState
{
  std::vector<Particle> P;
  xyz ref_pos, ref_vel, mb_pos, mb_vel;
  int64_t nbound = 0;
  float mbound = 0;
  int64_t nbound_type[TYPE_MAX] = {0};
  float mbound_type[TYPE_MAX] = {0};
  int64_t tracer_index = 0;
  int64_t mostbound_id = -1;
  float spot = 0, skin = 0, sam[3] = {0, 0, 0};
  xyz core_pos, core_vel;
  float core_sr = 0, core_sv = 0;
  int iterations = 0;
  int disrupted = 0;
  std::vector<PE> elist;
  std::vector<float> pot_energy;
};

// Subhalo_t::CountParticleTypes, src/subhalo.cpp:311-392
void count_particle_types(const Cfg &c, SubState &S)
{
  if (c.p.dm_only)
  {
    S.tracer_index = 0;
    return;
  }
  int64_t first = -1;
  for (int k = 0; k < TYPE_MAX; k++)
  {
    S.nbound_type[k] = 0;
    S.mbound_type[k] = 0.f;
  }
  for (int64_t i = 0; i < S.nbound; i++)
  {
    const Particle &p = S.P[i];
    S.nbound_type[p.type]++;
    S.mbound_type[p.type] += p.m;
    if (first < 0 && is_tracer(c, p))
      first = i;
  }
  S.tracer_index = (first < 0) ? 0 : first;
}

// Subhalo_t::GetCorePhaseSpaceProperties, src/subhalo_merge.cpp:94-181 (+ GetCoreSize, src/subhalo.cpp:492-495)
void core_phase_space(const Cfg &c, int64_t nbound, const Particle *P, const xyz &mb_pos, const xyz &mb_vel,
                      xyz &cp, xyz &cv, float &sr, float &sv)
{
  if (nbound == 0)
  {
    cp = mb_pos;
    cv = mb_vel;
    sr = sv = 0.f;
    return;
  }
  int64_t num = 0;
  double msum = 0, pos[3] = {0, 0, 0}, pos2[3] = {0, 0, 0}, vel[3] = {0, 0, 0}, vel2[3] = {0, 0, 0};
  int64_t core = std::max((int64_t)(nbound * c.p.sub_core_size_factor), c.p.sub_core_size_min);
  double origin[3] = {0, 0, 0};
  if (c.p.periodic)
    for (int j = 0; j < 3; j++)
      origin[j] = P[0].x[j];
  for (int pass = 0; pass < 2; pass++)
  {
    for (int64_t i = 0; i < nbound; i++)
    {
      const int tr = is_tracer(c, P[i]);
      if ((tr && pass == 0) || (!tr && pass == 1))
      {
        num++;
        float m = P[i].m;
        msum += m;
        for (int d = 0; d < 3; d++)
        {
          double dx = P[i].x[d];
          if (c.p.periodic)
            dx = c.nearest(dx - origin[d]);
          pos[d] += m * dx;
          pos2[d] += m * dx * dx;
          double dv = P[i].v[d];
          vel[d] += m * dv;
          vel2[d] += m * dv * dv;
        }
      }
      if (num == core)
        break;
    }
    if (num == core)
      break;
  }
  for (int d = 0; d < 3; d++)
  {
    pos[d] /= msum;
    pos2[d] /= msum;
    cp[d] = pos[d];
    if (c.p.periodic)
      cp[d] += origin[d];
    pos2[d] -= pos[d] * pos[d];
    vel[d] /= msum;
    vel2[d] /= msum;
    cv[d] = vel[d];
    vel2[d] -= vel[d] * vel[d];
  }
  sr = std::sqrt(pos2[0] + pos2[1] + pos2[2]);
  sv = std::sqrt(vel2[0] + vel2[1] + vel2[2]);
}

// Subhalo_t::Unbind, src/subhalo_unbind.cpp:463-808.
// prop (may be null): warm-start proposal, one flag per particle.  The warm start
// has no reference counterpart; it is restated here with the library's rules (a
// first round with the proposed set as sources and every particle as target in
// the proposal's frame, acceptance test, else the plain iteration from every
// particle) so that the GPU path can be checked round for round.
void unbind(const Cfg &c, SubState &S, std::vector<int64_t> &final_order, bool want_pot,
            const uint8_t *prop = nullptr)
{
  std::vector<Particle> &P = S.P;
  const int64_t N = (int64_t)P.size();
  final_order.resize(N);
  for (int64_t i = 0; i < N; i++)
    final_order[i] = i;
  if (N == 0)
  {
    S.nbound = 0;
    S.mbound = 0;
    core_phase_space(c, 0, nullptr, S.mb_pos, S.mb_vel, S.core_pos, S.core_vel, S.core_sr, S.core_sv);
    return;
  }
  int64_t mss = c.p.max_sample_size;
  bool refine = (mss > 0 && c.p.refine_most_bound);
  float prec = c.p.bound_mass_precision;
  S.tracer_index = 0;
  xyz old_pos = {0, 0, 0}, old_vel = {0, 0, 0};
  xyz &ref_pos = S.ref_pos;
  xyz &ref_vel = S.ref_vel;
  int64_t nbound = N;
  Tree tree;
  // PrepareParticlesForSubsampling (:430-443): particles arrive partitioned and shuffled
  int64_t nuns = 0;
  if (mss > 0 && nbound > mss)
    while (nuns < N && do_not_subsample(c, P[nuns]))
      nuns++;
  std::vector<PE> &E = S.elist;
  E.assign(N, PE{0.f, 0});
  for (int64_t i = 0; i < N; i++)
    E[i].idx = i;
  ESnap es(E.data(), N, P, c);
  float mbound;
  if (c.p.dm_only)
    es.acc_mass(0, nbound, mbound);
  else
  {
    float mt[TYPE_MAX];
    es.acc_mass_type(0, nbound, mt);
    mbound = 0;
    for (int k = 0; k < TYPE_MAX; k++)
      mbound += mt[k];
  }
  int64_t nlast = 0;
  float mlast = 0;
  bool correction = false;
  int iters = 0;
  bool disrupted = false;
  // convergence (:687-758): bound part sorted, centre refinement, final permutation
  auto finalize = [&]() {
    std::sort(E.begin(), E.begin() + nbound, [](const PE &a, const PE &b) { return a.E < b.E; });
    if (refine && nbound > mss)
    {
      // RefineBindingEnergyOrder (:395-425)
      int64_t size = std::max(mss, (int64_t)(c.p.bound_fraction_center_refinement * nbound));
      tree.build(c, es, size);
      std::vector<PE> ein(size);
      for (int64_t i = 0; i < size; i++)
      {
        const Particle &q = P[E[i].idx];
        ein[i].idx = i;
        ein[i].E = tree.binding_energy(q.x, q.v, ref_pos, ref_vel, q.m);
      }
      std::sort(ein.begin(), ein.end(), [](const PE &a, const PE &b) { return a.E < b.E; });
      for (int64_t i = 0; i < size; i++)
        ein[i] = E[ein[i].idx];
      for (int64_t i = 0; i < size; i++)
        E[i] = ein[i];
    }
    std::vector<Particle> p2(N);
    for (int64_t i = 0; i < N; i++)
    {
      final_order[i] = E[i].idx;
      p2[i] = P[E[i].idx];
      E[i].idx = i;
    }
    P.swap(p2);
    S.mb_pos = P[0].x;
    S.mb_vel = P[0].v;
  };
  // warm start: eligible when the proposal holds at least 2 MinNumPartOfSub particles
  std::vector<PE> warm_src; // the correction sources after an accepted warm round
  bool warm_pending = false;
  bool done = false;
  if (prop && mss == 0)
  {
    int64_t npro = 0;
    for (int64_t i = 0; i < N; i++)
      npro += prop[i] ? 1 : 0;
    warm_pending = npro >= 2 * c.p.min_num_part;
  }
  if (warm_pending)
  {
    iters++;
    const float mtotal = mbound;
    const xyz in_pos = ref_pos, in_vel = ref_vel;
    std::vector<PE> EP;
    for (int64_t i = 0; i < N; i++)
      if (prop[i])
        EP.push_back(PE{0.f, i});
    const int64_t npro = (int64_t)EP.size();
    ESnap esp(EP.data(), npro, P, c);
    esp.com_frame(ref_pos, ref_vel, npro);
    tree.build(c, esp, npro);
    for (int64_t i = 0; i < N; i++)
    {
      const Particle &q = P[E[i].idx];
      E[i].E = tree.binding_energy(q.x, q.v, ref_pos, ref_vel, prop[E[i].idx] ? q.m : 0.f);
      if (c.p.thermal)
        E[i].E += q.u;
    }
    // every particle with E < 0 stays, proposed or not (np_out: bound, outside the proposal)
    int64_t np_out = 0;
    for (int64_t i = 0; i < N; i++)
      if (E[i].E < 0 && !prop[E[i].idx])
        np_out++;
    auto it = std::stable_partition(E.begin(), E.end(), [&](const PE &a) { return a.E < 0; });
    const int64_t nb = it - E.begin();
    const int64_t nin = nb - np_out;
    int64_t ntr = 0;
    for (int64_t i = 0; i < nb; i++)
      ntr += (c.p.dm_only || is_tracer(c, P[E[i].idx])) ? 1 : 0;
    const bool accept = nin >= 2 * c.p.min_num_part && ntr >= 2 * c.p.min_num_tracer && 2 * nin >= npro &&
                        np_out * 1000 <= nb;
    if (accept)
    {
      std::sort(E.begin() + nb, E.end(), [](const PE &a, const PE &b) { return a.E < b.E; });
      for (int64_t i = nb; i < N; i++)
        if (prop[E[i].idx])
          warm_src.push_back(E[i]);
      const int64_t nrem = npro - nin;
      nbound = nb;
      nlast = N;
      if (nrem < nb)
      {
        correction = true;
        old_pos = ref_pos;
        old_vel = ref_vel;
      }
      mbound = es.com_frame(ref_pos, ref_vel, nbound);
      if (correction && np_out > 0)
      {
        // the outside particles' potential is added to every candidate's energy
        // (one more round; frames and the correction state stay)
        iters++;
        std::vector<PE> EA;
        for (int64_t i = 0; i < nb; i++)
          if (!prop[E[i].idx])
            EA.push_back(E[i]);
        ESnap esa(EA.data(), (int64_t)EA.size(), P, c);
        tree.build(c, esa, (int64_t)EA.size());
        for (int64_t i = 0; i < nb; i++)
        {
          const Particle &q = P[E[i].idx];
          E[i].E = (float)((double)E[i].E + tree.potential(q.x, prop[E[i].idx] ? 0.f : q.m));
        }
        // any candidate the added potential left unbound joins the first correction's sources
        auto ia = std::stable_partition(E.begin(), E.begin() + nb, [](const PE &a) { return a.E < 0; });
        const int64_t nb2 = ia - E.begin();
        if (nb2 < nb)
        {
          std::sort(E.begin() + nb2, E.begin() + nb, [](const PE &a, const PE &b) { return a.E < b.E; });
          for (int64_t i = nb2; i < nb; i++)
            warm_src.push_back(E[i]);
          nbound = nb2;
          mbound = es.com_frame(ref_pos, ref_vel, nbound);
        }
      }
      // convergence against every particle: the warm round ends the iteration only
      // when nearly all particles are bound and no outside particle was added
      if (!(correction && np_out > 0) && nbound >= nlast * prec)
      {
        finalize();
        done = true;
      }
    }
    else
    {
      for (int64_t i = 0; i < N; i++)
        E[i] = PE{0.f, i};
      nbound = N;
      mbound = mtotal;
      ref_pos = in_pos;
      ref_vel = in_vel;
      warm_pending = false;
    }
  }
  bool after_warm = warm_pending; // the next correction's sources are the proposed, removed particles
  while (!done)
  {
    iters++;
    if (correction)
    {
      xyz dV;
      relative_velocity(c, old_pos, old_vel, ref_pos, ref_vel, dV);
      float dK = 0.5 * vecdot(dV, dV);
      const int64_t nsrc = after_warm ? (int64_t)warm_src.size() : nlast - nbound;
      ESnap esc(after_warm ? warm_src.data() : &E[nbound], nsrc, P, c);
      if (nsrc > 0) // (a plain correction always has sources; after a warm round there may be none)
        tree.build(c, esc, nsrc);
      for (int64_t i = 0; i < nbound; i++)
      {
        const Particle &q = P[E[i].idx];
        xyz ov;
        relative_velocity(c, q.x, q.v, old_pos, old_vel, ov);
        E[i].E += vecdot(ov, dV) + dK - (nsrc > 0 ? tree.potential(q.x, 0.f) : 0.0);
      }
      nlast = nbound;
    }
    else
    {
      nlast = nbound;
      mlast = mbound;
      int64_t ngs = es.set_upscale(nlast, mlast, mss, nuns);
      tree.build(c, es, ngs);
      for (int64_t i = 0; i < nlast; i++)
      {
        float pm = (i < ngs) ? es.mass(i) : 0.f;
        const Particle &q = P[E[i].idx];
        E[i].E = tree.binding_energy(q.x, q.v, ref_pos, ref_vel, pm);
        if (c.p.thermal)
          E[i].E += q.u;
      }
      es.reset();
    }
    after_warm = false;
    // RemoveUnboundParticles (:35-44)
    auto it = std::stable_partition(E.begin(), E.begin() + nlast, [](const PE &a) { return a.E < 0; });
    nbound = it - E.begin();
    // CountUnsampledParticles (:446-461)
    {
      int64_t nn = 0;
      for (int64_t i = 0; i < nuns; i++)
      {
        if (!do_not_subsample(c, P[E[i].idx]))
          break;
        nn++;
      }
      nuns = nn;
    }
    int64_t ntr;
    if (c.p.dm_only)
      ntr = nbound;
    else
    {
      ntr = 0;
      for (int64_t i = 0; i < nbound; i++)
      {
        if (is_tracer(c, P[E[i].idx]))
          ntr++;
        if (ntr >= c.p.min_num_tracer)
          break;
      }
    }
    if ((nbound < c.p.min_num_part) || (ntr < c.p.min_num_tracer))
    {
      ref_pos = S.mb_pos;
      ref_vel = S.mb_vel;
      nbound = 0;
      mbound = 0;
      disrupted = true;
      break;
    }
    std::sort(E.begin() + nbound, E.begin() + nlast, [](const PE &a, const PE &b) { return a.E < b.E; });
    int64_t ndiff = nlast - nbound;
    if (ndiff < nbound)
      if (mss <= 0 || ndiff < mss)
      {
        correction = true;
        old_pos = ref_pos;
        old_vel = ref_vel;
      }
    mbound = es.com_frame(ref_pos, ref_vel, nbound);
    if (nbound >= nlast * prec)
    {
      finalize();
      break;
    }
  }
  S.nbound = nbound;
  S.mbound = mbound;
  S.iterations = iters;
  S.disrupted = disrupted;
  es.average_kinematics(S.spot, S.skin, S.sam, nbound, ref_pos, ref_vel);
  count_particle_types(c, S);
  if (!disrupted)
    S.mostbound_id = P[S.tracer_index].id;
  core_phase_space(c, nbound, P.data(), S.mb_pos, S.mb_vel, S.core_pos, S.core_vel, S.core_sr, S.core_sv);
  if (want_pot)
  {
    S.pot_energy.resize(nbound);
    for (int64_t i = 0; i < nbound; i++)
      S.pot_energy[i] = es.potential_energy(i, ref_pos, ref_vel);
  }
}

struct ArraySources : public Sources
{
  const float *p;
  const float *m;
  mutable xyz tmp;
  std::vector<xyz> xs;
  std::vector<float> ms;
  ArraySources(int64_t n, const float *pos, const float *mass) : xs(n), ms(n)
  {
    for (int64_t i = 0; i < n; i++)
    {
      xs[i] = {pos[3 * i], pos[3 * i + 1], pos[3 * i + 2]};
      ms[i] = mass[i];
    }
  }
  const xyz &pos(int64_t i) const override
  {
    return xs[i];
  }
  float mass(int64_t i) const override
  {
    return ms[i];
  }
};

double kernel_fp64(double r2, double m, double h, double h_inv)
{
  double r = std::sqrt(r2);
  if (r >= h)
    return -m / r;
  double u = r * h_inv, wp;
  if (u < 0.5)
    wp = -2.8 + u * u * (5.333333333333 + u * u * (6.4 * u - 9.6));
  else
    wp = -3.2 + 0.066666666667 / u + u * u * (10.666666666667 + u * (-16.0 + u * (9.6 - 2.133333333333 * u)));
  return m * h_inv * wp;
}
} // namespace

static double unbind_batch_impl(const ounb_params *prm, int64_t n, const float *pos, const float *vel,
                                const float *mass, const float *u, const int32_t *type, const int64_t *id,
                                int64_t nsub, const int64_t *offsets, const float *ref_pos, const float *ref_vel,
                                const float *mb_pos, const float *mb_vel, const uint8_t *proposal, int nthreads,
                                int64_t *out_order, float *out_energy, float *out_potential, ounb_result *out_res)
{
  Cfg c(*prm);
  std::vector<SubState> subs(nsub);
  for (int64_t s = 0; s < nsub; s++)
  {
    SubState &S = subs[s];
    int64_t a = offsets[s], b = offsets[s + 1];
    S.P.resize(b - a);
    for (int64_t i = a; i < b; i++)
    {
      Particle &q = S.P[i - a];
      q.id = id ? id[i] : i - a;
      for (int j = 0; j < 3; j++)
      {
        q.x[j] = pos[3 * i + j];
        q.v[j] = vel[3 * i + j];
      }
      q.m = mass[i];
      q.u = u ? u[i] : 0.f;
      q.type = type ? type[i] : 1;
    }
    for (int j = 0; j < 3; j++)
    {
      S.ref_pos[j] = ref_pos[3 * s + j];
      S.ref_vel[j] = ref_vel[3 * s + j];
      S.mb_pos[j] = mb_pos ? mb_pos[3 * s + j] : S.ref_pos[j];
      S.mb_vel[j] = mb_vel ? mb_vel[3 * s + j] : S.ref_vel[j];
    }
  }
  std::vector<std::vector<int64_t>> orders(nsub);
  if (nthreads <= 0)
    nthreads = omp_get_max_threads();
  double t0 = omp_get_wtime();
#pragma omp parallel for schedule(dynamic, 1) num_threads(nthreads)
  for (int64_t s = 0; s < nsub; s++)
    unbind(c, subs[s], orders[s], out_potential != nullptr, proposal ? proposal + offsets[s] : nullptr);
  double t1 = omp_get_wtime();
  for (int64_t s = 0; s < nsub; s++)
  {
    SubState &S = subs[s];
    int64_t a = offsets[s], N = offsets[s + 1] - a;
    for (int64_t k = 0; k < N; k++)
    {
      out_order[a + k] = orders[s][k];
    }
    // energies in final particle order (E list is in final order after convergence;
    // a disrupted subhalo keeps its input order, so map through the E list indices)
    if (!S.disrupted)
      for (int64_t k = 0; k < N; k++)
        out_energy[a + k] = S.elist[k].E;
    else
      for (int64_t k = 0; k < N; k++)
        out_energy[a + S.elist[k].idx] = S.elist[k].E;
    if (out_potential)
      for (int64_t k = 0; k < N; k++)
        out_potential[a + k] = (k < S.nbound) ? S.pot_energy[k] : 0.f;
    ounb_result &R = out_res[s];
    std::memset(&R, 0, sizeof(R));
    R.nbound = S.nbound;
    R.mbound = S.mbound;
    for (int k = 0; k < TYPE_MAX; k++)
    {
      R.nbound_type[k] = S.nbound_type[k];
      R.mbound_type[k] = S.mbound_type[k];
    }
    for (int j = 0; j < 3; j++)
    {
      R.com_pos[j] = S.ref_pos[j];
      R.com_vel[j] = S.ref_vel[j];
      R.mostbound_pos[j] = S.mb_pos[j];
      R.mostbound_vel[j] = S.mb_vel[j];
      R.spec_am[j] = S.sam[j];
      R.core_pos[j] = S.core_pos[j];
      R.core_vel[j] = S.core_vel[j];
    }
    R.spec_pot = S.spot;
    R.spec_kin = S.skin;
    R.tracer_index = S.tracer_index;
    R.mostbound_id = S.mostbound_id;
    R.core_sigma_r = S.core_sr;
    R.core_sigma_v = S.core_sv;
    R.iterations = S.iterations;
    R.disrupted = S.disrupted;
  }
  return t1 - t0;
}

extern "C" double oracle_unbind_batch(const ounb_params *prm, int64_t n, const float *pos, const float *vel,
                                      const float *mass, const float *u, const int32_t *type, const int64_t *id,
                                      int64_t nsub, const int64_t *offsets, const float *ref_pos,
                                      const float *ref_vel, const float *mb_pos, const float *mb_vel, int nthreads,
                                      int64_t *out_order, float *out_energy, float *out_potential,
                                      ounb_result *out_res)
{
  return unbind_batch_impl(prm, n, pos, vel, mass, u, type, id, nsub, offsets, ref_pos, ref_vel, mb_pos, mb_vel,
                           nullptr, nthreads, out_order, out_energy, out_potential, out_res);
}

extern "C" double oracle_unbind_batch_warm(const ounb_params *prm, int64_t n, const float *pos, const float *vel,
                                           const float *mass, const float *u, const int32_t *type,
                                           const int64_t *id, int64_t nsub, const int64_t *offsets,
                                           const float *ref_pos, const float *ref_vel, const float *mb_pos,
                                           const float *mb_vel, const uint8_t *proposal, int nthreads,
                                           int64_t *out_order, float *out_energy, float *out_potential,
                                           ounb_result *out_res)
{
  return unbind_batch_impl(prm, n, pos, vel, mass, u, type, id, nsub, offsets, ref_pos, ref_vel, mb_pos, mb_vel,
                           proposal, nthreads, out_order, out_energy, out_potential, out_res);
}

extern "C" int64_t oracle_tree_potential(const ounb_params *prm, int64_t n_src, const float *src_pos,
                                         const float *src_mass, int64_t n_tgt, const float *tgt_pos,
                                         const float *tgt_mass, int nthreads, double *out_pot)
{
  Cfg c(*prm);
  ArraySources s(n_src, src_pos, src_mass);
  Tree t;
  int64_t ncell = t.build(c, s, n_src);
  if (nthreads <= 0)
    nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 256) num_threads(nthreads)
  for (int64_t i = 0; i < n_tgt; i++)
  {
    xyz x = {tgt_pos[3 * i], tgt_pos[3 * i + 1], tgt_pos[3 * i + 2]};
    out_pot[i] = t.potential(x, tgt_mass ? tgt_mass[i] : 0.f);
  }
  return ncell;
}

extern "C" void oracle_direct_potential(const ounb_params *prm, int64_t n_src, const float *src_pos,
                                        const float *src_mass, int64_t n_tgt, const float *tgt_pos,
                                        const float *tgt_mass, int nthreads, double *out_pot)
{
  Cfg c(*prm);
  double h = 2.8 * c.eps, h_inv = 1.0 / h;
  if (nthreads <= 0)
    nthreads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 64) num_threads(nthreads)
  for (int64_t i = 0; i < n_tgt; i++)
  {
    double px = tgt_pos[3 * i], py = tgt_pos[3 * i + 1], pz = tgt_pos[3 * i + 2];
    double pot = tgt_mass ? tgt_mass[i] / (double)c.eps : 0.0;
    for (int64_t j = 0; j < n_src; j++)
    {
      double dx = src_pos[3 * j] - px, dy = src_pos[3 * j + 1] - py, dz = src_pos[3 * j + 2] - pz;
      if (c.p.periodic)
      {
        dx = c.nearest(dx);
        dy = c.nearest(dy);
        dz = c.nearest(dz);
      }
      pot += kernel_fp64(dx * dx + dy * dy + dz * dz, src_mass[j], h, h_inv);
    }
    out_pot[i] = pot * c.p.G / c.p.scale_factor;
  }
}

extern "C" int64_t oracle_tree_dfs(const ounb_params *prm, int64_t n_src, const float *src_pos,
                                   const float *src_mass, int64_t cap, int32_t *kind, int64_t *pidx, float *mass,
                                   float *com, float *len, int64_t *skip, int32_t *depth)
{
  Cfg c(*prm);
  ArraySources s(n_src, src_pos, src_mass);
  Tree t;
  int64_t ncell = t.build(c, s, n_src);
  int64_t total = n_src + ncell;
  if (total > cap)
    return -total;
  // preorder position of every tree id (particles 0..np-1, cells np..)
  std::vector<int64_t> posof(total, -1);
  std::vector<int64_t> order;
  order.reserve(total);
  std::vector<int32_t> dep(total, 0);
  const Cell *nd = t.cells.data() - t.np;
  // explicit DFS following nextnode/next_from_particle, which visits in preorder
  int64_t no = t.np;
  // depth tracking: a cell's children have depth+1; use a stack of (sibling target, depth)
  std::vector<std::pair<int64_t, int32_t>> stack;
  int32_t d = 0;
  while (no >= 0)
  {
    // pop levels whose subtree ended (we arrived at that level's sibling)
    while (!stack.empty() && no == stack.back().first)
    {
      d = stack.back().second;
      stack.pop_back();
    }
    posof[no] = (int64_t)order.size();
    order.push_back(no);
    dep[no] = d;
    if (no < t.np)
      no = t.next_from_particle[no];
    else
    {
      stack.push_back({nd[no].way.sibling, d});
      d = d + 1;
      no = nd[no].way.nextnode;
    }
  }
  for (int64_t k = 0; k < (int64_t)order.size(); k++)
  {
    int64_t id = order[k];
    if (id < t.np)
    {
      kind[k] = 0;
      pidx[k] = id;
      mass[k] = s.ms[id];
      for (int j = 0; j < 3; j++)
        com[3 * k + j] = s.xs[id][j];
      len[k] = 0.f;
      int64_t nx = t.next_from_particle[id];
      skip[k] = (nx < 0) ? total : posof[nx];
    }
    else
    {
      const Cell &cl = nd[id];
      kind[k] = 1;
      pidx[k] = -1;
      mass[k] = cl.way.mass;
      for (int j = 0; j < 3; j++)
        com[3 * k + j] = cl.way.s[j];
      len[k] = cl.way.len;
      skip[k] = (cl.way.sibling < 0) ? total : posof[cl.way.sibling];
    }
    depth[k] = dep[id];
  }
  return (int64_t)order.size();
}

extern "C" void oracle_core_phase_space(const ounb_params *prm, int64_t nbound, const float *pos, const float *vel,
                                        const float *mass, const int32_t *type, float *core_pos, float *core_vel,
                                        float *sigma_r, float *sigma_v)
{
  Cfg c(*prm);
  std::vector<Particle> P(nbound);
  for (int64_t i = 0; i < nbound; i++)
  {
    for (int j = 0; j < 3; j++)
    {
      P[i].x[j] = pos[3 * i + j];
      P[i].v[j] = vel[3 * i + j];
    }
    P[i].m = mass[i];
    P[i].type = type ? type[i] : 1;
    P[i].id = i;
    P[i].u = 0;
  }
  xyz cp, cv, z = {0, 0, 0};
  float sr, sv;
  core_phase_space(c, nbound, P.data(), z, z, cp, cv, sr, sv);
  for (int j = 0; j < 3; j++)
  {
    core_pos[j] = cp[j];
    core_vel[j] = cv[j];
  }
  *sigma_r = sr;
  *sigma_v = sv;
}

// ---------------------------------------------------------------------------
// Subhalo_t::CalculateProfileProperties (src/subhalo.cpp:175-243) with
// Snapshot_t::SphericalOverdensitySize (src/snapshot.cpp:258-282), and
// Subhalo_t::CalculateShape (src/subhalo.cpp:245-309).
// ---------------------------------------------------------------------------
namespace
{
struct RadMassVel
{
  float r, m, v;
};

float periodic_distance(const ounb_params *prm, const float *x, const float *y)
{
  float dx[3];
  for (int k = 0; k < 3; k++)
  {
    dx[k] = x[k] - y[k];
    if (prm->periodic)
    {
      const float half = prm->box_size / 2.f;
      dx[k] = (dx[k] > half) ? dx[k] - prm->box_size : ((dx[k] < -half) ? dx[k] + prm->box_size : dx[k]);
    }
  }
  return std::sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]);
}

void profile_one(const ounb_params *prm, const float *pos, const float *mass, int64_t nb, float mbound,
                 ounb_profile *o)
{
  std::memset(o, 0, sizeof(*o));
  if (nb <= 1)
    return;
  const float G = (float)prm->G;
  const float vel_unit = G / prm->scale_factor;
  const float *cen = pos; // most bound particle (final slot 0)
  std::vector<RadMassVel> prof(nb);
  for (int64_t i = 0; i < nb; i++)
  {
    prof[i].r = periodic_distance(prm, cen, pos + 3 * i);
    prof[i].m = mass[i];
  }
  std::sort(prof.begin(), prof.end(), [](const RadMassVel &a, const RadMassVel &b) { return a.r < b.r; });
  double m_cum = 0.;
  for (auto &p : prof)
    p.m = (float)(m_cum += p.m);
  const float eps = (prm->max_phys_softening == -1.f)
                      ? prm->softening
                      : std::min(prm->softening, prm->max_phys_softening / prm->scale_factor);
  for (auto &p : prof)
  {
    if (p.r < eps)
      p.r = eps;
    p.v = p.m / p.r;
  }
  auto best = std::max_element(prof.begin(), prof.end(),
                               [](const RadMassVel &a, const RadMassVel &b) { return a.v < b.v; });
  o->rmax_comoving = best->r;
  o->vmax_physical = std::sqrt(best->v * vel_unit);
  o->r_half_comoving = prof[nb / 2].r;
  o->r_enclose_comoving = prof[nb - 1].r;
  const float rho = (float)((double)(200.f * prm->hz * prm->hz) / 2.0 / (double)G * (double)prm->scale_factor *
                            (double)prm->scale_factor * (double)prm->scale_factor);
  for (auto p = prof.rbegin(); p != prof.rend(); ++p)
    if (p->m > rho * p->r * p->r * p->r)
    {
      o->bound_m200_crit = p->m;
      o->bound_r200_crit_comoving = (float)std::pow(p->m / rho, 1.0 / 3);
      o->has_r200 = 1;
      break;
    }
  double I[6] = {0, 0, 0, 0, 0, 0}, W[6] = {0, 0, 0, 0, 0, 0};
  for (int64_t i = 1; i < nb; i++)
  {
    const float m = mass[i];
    float d[3];
    for (int k = 0; k < 3; k++)
    {
      d[k] = pos[3 * i + k] - cen[k];
      if (prm->periodic)
      {
        const float half = prm->box_size / 2.f;
        d[k] = (d[k] > half) ? d[k] - prm->box_size : ((d[k] < -half) ? d[k] + prm->box_size : d[k]);
      }
    }
    const float dx2 = d[0] * d[0], dy2 = d[1] * d[1], dz2 = d[2] * d[2];
    I[0] += dx2 * m;
    I[3] += dy2 * m;
    I[5] += dz2 * m;
    I[1] += d[0] * d[1] * m;
    I[2] += d[0] * d[2] * m;
    I[4] += d[1] * d[2] * m;
    float dr2 = dx2 + dy2 + dz2;
    dr2 /= m;
    W[0] += dx2 / dr2;
    W[3] += dy2 / dr2;
    W[5] += dz2 / dr2;
    W[1] += d[0] * d[1] / dr2;
    W[2] += d[0] * d[2] / dr2;
    W[4] += d[1] * d[2] / dr2;
  }
  for (int k = 0; k < 6; k++)
  {
    o->inertial_tensor[k] = (float)I[k] / mbound;
    o->inertial_tensor_weighted[k] = (float)W[k] / mbound;
  }
}
} // namespace

extern "C" void oracle_profiles(const ounb_params *prm, int64_t n, const float *pos, const float *mass, int64_t nsub,
                                const int64_t *offsets, const int64_t *nbound, const float *mbound, ounb_profile *out)
{
  (void)n;
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t s = 0; s < nsub; s++)
    profile_one(prm, pos + 3 * offsets[s], mass + offsets[s], nbound[s], mbound[s], out + s);
}

// ---------------------------------------------------------------------------
// SubhaloSnapshot_t::ReassignParticles (src/subhalo_reassign_gas.cpp:63-197) with
// GeoTree_t::NearestNeighbours / Search (src/geometric_tree.cpp:108-213) and
// NumNearestNeighboursCollector_t (src/datatypes.h:287-324).  The tree search is
// replaced by a scan over the group's tracers: it collects the same set (every
// tracer inside the per-axis and sphere tests).  The reference orders tracers at
// exactly equal float distances by its tree's traversal order and the heap's
// internals (and places tracers closer than TreeNodeResolution at random), so
// that order is not reproducible; here the heap orders by (distance, particle
// index), the rule the GPU implements too.  Parity unpinned for such ties.
// ---------------------------------------------------------------------------
namespace
{
struct Located
{
  int64_t index;
  float d2;
};
struct LocatedLess
{
  bool operator()(const Located &a, const Located &b) const
  {
    return a.d2 < b.d2 || (a.d2 == b.d2 && a.index < b.index);
  }
};
} // namespace

extern "C" void oracle_reassign_targets(const ounb_params *prm, int64_t n, const float *pos, const uint8_t *is_tracer,
                                        const int32_t *sub, int64_t ngroups, const int64_t *group_offsets, int32_t k,
                                        float rguess, int32_t *out_target)
{
  (void)n;
  const bool periodic = prm->periodic != 0;
  const double box = prm->box_size, half = prm->box_size / 2.0;
  auto nearest = [&](double d) { return (d > half) ? d - box : ((d < -half) ? d + box : d); };
#pragma omp parallel for schedule(dynamic, 1)
  for (int64_t g = 0; g < ngroups; g++)
  {
    const int64_t a = group_offsets[g], e = group_offsets[g + 1];
    std::vector<int64_t> tracers;
    for (int64_t i = a; i < e; i++)
      if (is_tracer[i])
        tracers.push_back(i);
    const int64_t nt = (int64_t)tracers.size();
    for (int64_t i = a; i < e; i++)
    {
      out_target[i] = -1;
      if (is_tracer[i] || nt == 0)
        continue;
      const int64_t nr_to_find = std::min<int64_t>(nt, k);
      std::priority_queue<Located, std::vector<Located>, LocatedLess> heap;
      const double x0 = pos[3 * i], y0 = pos[3 * i + 1], z0 = pos[3 * i + 2];
      float radius = rguess;
      auto search = [&](float r) {
        const double h2 = r * r; // float product, as in the reference
        for (int64_t t : tracers)
        {
          double dx = pos[3 * t] - x0, dy = pos[3 * t + 1] - y0, dz = pos[3 * t + 2] - z0;
          if (periodic)
          {
            dx = nearest(dx);
            dy = nearest(dy);
            dz = nearest(dz);
          }
          // per-axis box test before the sphere test (src/geometric_tree.cpp:160-176)
          if (dx > r || dx < -r || dy > r || dy < -r || dz > r || dz < -r)
            continue;
          const double r2 = dx * dx + dy * dy + dz * dz;
          if (r2 < h2)
          {
            const Located lp{t, (float)r2};
            if ((int64_t)heap.size() < k)
              heap.push(lp);
            else if (LocatedLess()(lp, heap.top()))
            {
              heap.push(lp);
              heap.pop();
            }
          }
        }
      };
      search(radius);
      while ((int64_t)heap.size() < nr_to_find)
      {
        radius = (float)(radius * 1.26);
        search(radius);
      }
      std::vector<int64_t> list;
      while (!heap.empty())
      {
        list.push_back(heap.top().index);
        heap.pop();
      }
      bool found = false;
      for (int64_t t : list)
        if (sub[t] == sub[i])
        {
          found = true;
          break;
        }
      if (!found)
      {
        const int32_t ns = sub[list.back()];
        if (ns >= 0 && ns != sub[i])
          out_target[i] = ns;
      }
    }
  }
}

// SetMassUpscaleFactor (src/subhalo_unbind.cpp:374-391, with GetMassUpscaleFactor
// :290-317 or GetMassUpscaleFactorPerParticleType :323-361) over particles in
// E-list order: returns the number of gravity sources and the per-type factors.
extern "C" int64_t oracle_mass_upscale(const ounb_params *prm, int64_t n, const float *mass, const int32_t *type,
                                       int64_t nlast, float mlast, int64_t mss, int64_t nuns, float *out_upscale)
{
  Cfg c(*prm);
  std::vector<Particle> P(n);
  std::vector<PE> E(n);
  for (int64_t i = 0; i < n; i++)
  {
    P[i] = Particle{i, {0.f, 0.f, 0.f}, {0.f, 0.f, 0.f}, mass[i], 0.f, type ? type[i] : 1};
    E[i] = PE{0.f, i};
  }
  ESnap es(E.data(), n, P, c);
  const int64_t ns = es.set_upscale(nlast, mlast, mss, nuns);
  for (int k = 0; k < TYPE_MAX; k++)
    out_upscale[k] = es.upscale[k];
  return ns;
}
