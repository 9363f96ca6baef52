// Shared device helpers for the hbtx kernels (gfx950, wave64).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#define HBTX_WAVE 64
#define HBTX_DEV __device__ __forceinline__
#define HBTX_HD __host__ __device__ __forceinline__

// Number of octree levels encoded in a 63-bit key (3 bits per level).
#define HBTX_KEY_LEVELS 21

#define GRID_STRIDE(i, n) for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < (n); i += (int64_t)gridDim.x * blockDim.x)
// Loop over the active-slot tiles assigned to this block (tile = int4: subhalo,
// first slot, end slot, first tile of the subhalo); the count lives on the device.
#define TILE_LOOP(tiles, ntiles_p) for (int t_ = blockIdx.x, nt_ = *(ntiles_p); t_ < nt_; t_ += gridDim.x)

namespace hbtx
{

// Subhalo status values
enum : int32_t
{
  ST_ACTIVE = 0,
  ST_CONVERGED = 1,
  ST_DISRUPTED = 2
};
// Per-round mode of a subhalo
enum : int32_t
{
  MODE_IDLE = 0,
  MODE_FULL = 1,  // all current candidates are targets, sources = candidates (or subsample)
  MODE_CORR = 2,  // correction: sources = particles removed in the previous round
  MODE_REFINE = 3, // centre refinement: self-binding energies of the most bound particles
  MODE_ADD = 4     // after an accepted warm round: the potential of the bound particles from
                   // outside the proposal is added to every candidate's energy
};
// Learned warm start, per subhalo
enum : int32_t
{
  WARM_OFF = 0,     // plain iteration
  WARM_PENDING = 1, // the warm round (proposed set as sources, every particle a target) comes next
  WARM_CORR = 2,    // accepted: the first correction (sources: proposed particles the warm round removed)
  WARM_DONE = 3,    // the correction sequence continues as in the plain iteration
  WARM_ADD = 4      // accepted with bound particles outside the proposal: a MODE_ADD round first
};

// Device copy of the unbinding parameters (derived values precomputed on the host
// exactly as the reference derives them, src/config_parser.cpp:200-212,641-649).
struct DevParams
{
  double G_over_a;     // PhysicalConst::G / ScaleFactor
  double res;          // TreeNodeResolution (0.1 SofteningHalo, float -> double)
  float eps;           // GetCurrentSoftening(a)
  float h;             // 2.8 eps
  float hinv;          // 1 / h
  float h2;            // h^2
  float theta2;        // TreeNodeOpenAngleSquare (float product)
  float a;             // ScaleFactor
  float hz;            // Hz
  float box;           // BoxSize
  float box_half;      // BoxHalf
  float prec;          // BoundMassPrecision
  float core_factor;   // SubCoreSizeFactor
  float refine_frac;   // BoundFractionCenterRefinement
  int32_t periodic;
  int32_t thermal;
  int32_t dm_only;
  int32_t tracer_mask;
  int32_t donot_mask;
  int32_t per_type;
  int32_t refine;
  int64_t mss;         // MaxSampleSizeOfPotentialEstimate
  int64_t min_part;
  int64_t min_tracer;
  int64_t core_min;
  int64_t direct_cutoff;
};

// Depth-first tree entries are kept as parallel arrays: float4 (x, y, z, m) -
// a cell's centre of mass and mass, a particle's position and (scaled) mass -,
// the depth (int8, -1 for a particle) and, for cells only, skip (int32: the
// first entry after the cell's subtree; a particle's is its own index + 1).
// The walk reads its own compact records, built alongside.

// Loads through the constant address space: a wave-uniform index then becomes a
// scalar (s_load) access.  Only for buffers that no kernel writes while it reads them.
HBTX_DEV float4 const_load(const float4 *p, int i)
{
  typedef const __attribute__((address_space(4))) float cf;
  const cf *q = (const cf *)(uintptr_t)(p + i);
  return make_float4(q[0], q[1], q[2], q[3]);
}
HBTX_DEV int2 const_load(const int2 *p, int i)
{
  typedef const __attribute__((address_space(4))) int ci;
  const ci *q = (const ci *)(uintptr_t)(p + i);
  return make_int2(q[0], q[1]);
}
// Walk records: the walk's own copy of every tree, in the entries' depth-first
// order but packed into 8-byte units.  A particle takes two units {x, y, z, m};
// a cell three {x, y, z, m with the sign bit set, crit2, rskip}, rskip in bytes
// from the tree's first record.  Source j's group (its cells, then the particle)
// starts at unit 2 j + 3 cscan[j].  A particle's skip is the next record and it
// is never opened.  Packing the particles into 16 bytes shrinks a tree's walk
// footprint by almost half against the 32-byte entries.

// a particle record (x, y, z, m) at an 8-byte aligned unit, through the scalar path
HBTX_DEV float4 const_load_f4u(const uint2 *p)
{
  typedef int v4i __attribute__((ext_vector_type(4)));
  typedef const __attribute__((address_space(4))) v4i c4;
  const v4i a = *(const c4 *)(uintptr_t)p;
  return make_float4(__int_as_float(a[0]), __int_as_float(a[1]), __int_as_float(a[2]), __int_as_float(a[3]));
}
// Longest walk-record span of one tree: the walk keeps its position as an int32
// byte offset and steps at most 24 bytes past the span's end.
#define WALK_MAX_SPAN_UNITS ((0x7fffffff - 64) / 8)
HBTX_HD int32_t wrec_unit(int32_t j, int32_t cscan_j)
{
  return 2 * j + 3 * cscan_j;
}


HBTX_DEV uint64_t ballot(bool p)
{
  return __ballot(p);
}

HBTX_DEV int lane_id()
{
  return __lane_id();
}

HBTX_DEV int uniform(int v)
{
  return __builtin_amdgcn_readfirstlane(v);
}

HBTX_DEV uint64_t lanemask_lt()
{
  const int l = lane_id();
  return (l == 0) ? 0ull : (~0ull >> (64 - l));
}

// float <-> uint32 with the float ordering preserved
HBTX_HD uint32_t float_to_ordered(float f)
{
  uint32_t u = __float_as_uint(f);
  return (u & 0x80000000u) ? ~u : (u | 0x80000000u);
}
HBTX_HD float ordered_to_float(uint32_t u)
{
  uint32_t v = (u & 0x80000000u) ? (u & 0x7fffffffu) : ~u;
  return __uint_as_float(v);
}

// NEAREST, src/config_parser.h:218-219
HBTX_DEV float nearest_f(float x, float box, float half)
{
  return (x > half) ? (x - box) : ((x < -half) ? (x + box) : x);
}
// NEAREST(a - b) rounded once, as the reference gets it from float operands in
// double arithmetic (src/gravity_tree.cpp:108-133).  The float difference s and
// its rounding error e (Knuth's two-sum: s + e = a - b exactly) are formed first;
// the box shift of s is exact (|s| lies within a factor of two of the box when it
// applies), and e is added last.  A plain float a - b loses up to half an ulp of
// the box size when the operands sit on opposite sides of the box edge.
HBTX_DEV float nearest_diff_f(float a, float b, float box, float half)
{
  const float s = a - b;
  const float bv = s - a;
  const float e = (a - (s - bv)) + (-b - bv);
  return nearest_f(s, box, half) + e;
}
HBTX_DEV double nearest_d(double x, double box, double half)
{
  return (x > half) ? (x - box) : ((x < -half) ? (x + box) : x);
}

HBTX_DEV int is_tracer(const DevParams &P, int type)
{
  return P.dm_only ? 1 : ((1 << type) & P.tracer_mask);
}
HBTX_DEV int donot_subsample(const DevParams &P, int type)
{
  return P.dm_only ? 0 : ((1 << type) & P.donot_mask);
}

// 63-bit octree key of a point in the cube (centre, lenhalf): the sequence of
// child indices (bit0 x, bit1 y, bit2 z) that the reference insertion computes
// level by level in double precision (src/oct_tree.tpp:377-410).
HBTX_HD uint64_t octree_key(float x, float y, float z, double cx, double cy, double cz, double lenhalf)
{
  uint64_t key = 0;
  double lh = lenhalf;
  const double px = x, py = y, pz = z;
#pragma unroll 3
  for (int l = 0; l < HBTX_KEY_LEVELS; l++)
  {
    lh *= 0.5;
    uint64_t sub = 0;
    if (px > cx) { cx += lh; sub |= 1; } else { cx -= lh; }
    if (py > cy) { cy += lh; sub |= 2; } else { cy -= lh; }
    if (pz > cz) { cz += lh; sub |= 4; } else { cz -= lh; }
    key = (key << 3) | sub;
  }
  return key;
}

// Depth below which the reference stops placing particles geometrically: a new
// cell of half-size lenhalf < TreeNodeResolutionHalf sends the displaced particle
// to a random child (src/oct_tree.tpp:426-438).  Returns the last geometric level.
HBTX_HD int resolution_depth(double lenhalf, double res_half)
{
  int d = 1;
  double lh = lenhalf * 0.5;
  while (d < HBTX_KEY_LEVELS && !(lh < res_half))
  {
    lh *= 0.5;
    d++;
  }
  return d;
}

// Levels deeper than the resolution depth take digits from a hash of the
// particle's index inside its subhalo: deterministic, and like the reference's
// random placement they separate coincident particles within a few levels.
HBTX_HD uint64_t hash_digits(uint64_t key, int dcut, uint32_t local_index)
{
  if (dcut >= HBTX_KEY_LEVELS)
    return key;
  uint64_t h = (uint64_t)local_index * 0x9E3779B97F4A7C15ull;
  h ^= h >> 29;
  h *= 0xBF58476D1CE4E5B9ull;
  h ^= h >> 32;
  const uint64_t mask = (1ull << (3 * (HBTX_KEY_LEVELS - dcut))) - 1ull;
  return (key & ~mask) | (h & mask);
}

// Number of leading octree levels two keys share (0..21).
HBTX_HD int common_levels(uint64_t a, uint64_t b)
{
  uint64_t x = a ^ b;
  if (x == 0)
    return HBTX_KEY_LEVELS;
#ifdef __HIP_DEVICE_COMPILE__
  int lz = __clzll((long long)x);
#else
  int lz = __builtin_clzll(x);
#endif
  return (lz - 1) / 3;
}

// Prefix of a key at depth l (l levels kept).
HBTX_HD uint64_t key_prefix(uint64_t k, int l)
{
  return (l == 0) ? 0ull : (k >> (3 * (HBTX_KEY_LEVELS - l)));
}

// Softened pair potential inside the kernel radius h, in units where the final
// sum is multiplied by G/a: the GADGET-2 cubic-spline polynomial of
// src/gravity_tree.cpp:150-162 (outside h the term is -m/r).  rinv = 1/sqrt(r2).
HBTX_DEV float spline_near(float r2, float rinv, float m, float hinv)
{
  const float r = (r2 > 0.f) ? r2 * rinv : 0.f;
  const float u = r * hinv;
  const float u2 = u * u;
  const float inner = -2.8f + u2 * (5.333333333333f + u2 * (6.4f * u - 9.6f));
  const float outer = -3.2f + 0.066666666667f * __builtin_amdgcn_rcpf(u) +
                      u2 * (10.666666666667f + u * (-16.0f + u * (9.6f - 2.133333333333f * u)));
  return m * hinv * ((u < 0.5f) ? inner : outer);
}

// Kahan-compensated accumulation of FP32 terms.
struct Kahan
{
  float s, c;
  HBTX_DEV void init()
  {
    s = 0.f;
    c = 0.f;
  }
  HBTX_DEV void add(float x)
  {
    float y = x - c;
    float t = s + y;
    c = (t - s) - y;
    s = t;
  }
};

} // namespace hbtx
