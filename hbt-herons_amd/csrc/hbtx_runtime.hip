// hbtx runtime: device buffers, the round pipeline, and the C-ABI entry points
// declared in include/hbtx.h.  All kernels live in the headers included below
// and are launched on the context's stream.
#include "hbtx.h"

#include "common.h"
#include "scan.h"
#include "sort.h"
#include "tree.h"
#include "walk.h"
#include "unbind.h"
#include "points.h"
#include "subsample.h"
#include "learn.h"
#include "tiles.h"
#include "profile.h"
#include "reassign.h"

#define WALK_MAX_CHUNKS 16
// entries whose walk records (at most 3 units each) keep unit offsets in int32
#define WREC_MAX_ENTRIES ((int64_t)0x7ffffff0 / 3 - 8)

#include <algorithm>
#include <atomic>
#include <cmath>
#include <dlfcn.h>
#include <fstream>
#include <sstream>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <numeric>
#include <stdexcept>
#include <string>
#include <vector>

using namespace hbtx;

namespace
{

struct Error : public std::runtime_error
{
  int code;
  Error(int c, const std::string &m) : std::runtime_error(m), code(c) {}
};

#define HIPCHK(x)                                                                                                    \
  do                                                                                                                 \
  {                                                                                                                  \
    hipError_t e_ = (x);                                                                                             \
    if (e_ != hipSuccess)                                                                                            \
      throw Error(HBTX_E_HIP, std::string(#x) + ": " + hipGetErrorString(e_) + " (" + __FILE__ + ":" +            \
                                  std::to_string(__LINE__) + ")");                                                   \
  } while (0)

#define REQUIRE(cond, code, msg)                                                                                     \
  do                                                                                                                 \
  {                                                                                                                  \
    if (!(cond))                                                                                                     \
      throw Error(code, msg);                                                                                        \
  } while (0)

// The array of hbtx_direct_potential_noalloc: packed source records, packed
// targets and the work counter, HBTX_NOALLOC_MAX_POINTS of each.  Every GPU holds
// its own copy at its own address until the process ends; all calls on one GPU
// share it (calls that overlap on one GPU race).
constexpr size_t kScratchTgt = 16 * (size_t)HBTX_NOALLOC_MAX_POINTS;
constexpr size_t kScratchCtr = 32 * (size_t)HBTX_NOALLOC_MAX_POINTS;
constexpr size_t kScratchBytes = kScratchCtr + 256;
__device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on the current device, looked up once per device
// (host code never uses the array's name as an address).
void *scratch_address(int device)
{
  static std::mutex mu;
  static void *addr[256] = {};
  REQUIRE(device >= 0 && device < 256, HBTX_E_INVALID, "device index beyond 255");
  std::lock_guard<std::mutex> lock(mu);
  if (!addr[device])
  {
    void *p = nullptr;
    HIPCHK(hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)));
    REQUIRE(p != nullptr, HBTX_E_HIP, "no device address for the scratch array");
    addr[device] = p;
  }
  return addr[device];
}

// Advanced by every device allocation and release: kernels captured in a graph
// hold buffer addresses, so a graph is valid only for the generation it was
// captured in.
std::atomic<uint64_t> g_alloc_gen{0};
std::atomic<int64_t> g_device_bytes{0}; // device memory the library's buffers hold, all contexts

template <class T>
struct DArr
{
  T *p = nullptr;
  size_t n = 0;
  void alloc(size_t count)
  {
    if (p && count <= n)
      return;
    release();
    g_alloc_gen++;
    size_t bytes = std::max<size_t>(count, 1) * sizeof(T);
    hipError_t e = hipMalloc((void **)&p, bytes);
    if (e != hipSuccess)
    {
      p = nullptr;
      n = 0;
      throw Error(HBTX_E_ALLOC, "hipMalloc of " + std::to_string(bytes) + " bytes failed: " + hipGetErrorString(e));
    }
    n = count;
    g_device_bytes += (int64_t)bytes;
  }
  void release()
  {
    if (p)
    {
      (void)hipFree(p);
      g_alloc_gen++;
      g_device_bytes -= (int64_t)(std::max<size_t>(n, 1) * sizeof(T));
    }
    p = nullptr;
    n = 0;
  }
  ~DArr()
  {
    release();
  }
  DArr() = default;
  DArr(const DArr &) = delete;
  DArr &operator=(const DArr &) = delete;
};

inline int blocks_for(int64_t n, int threads = 256, int cap = 16384)
{
  int64_t b = (n + threads - 1) / threads;
  return (int)std::max<int64_t>(1, std::min<int64_t>(b, cap));
}

} // namespace

// ---------------------------------------------------------------------------
// Batch state
// ---------------------------------------------------------------------------
struct Batch
{
  int64_t n = 0;
  int32_t nsub = 0;
  int32_t max_count = 0;
  bool has_type = false;
  bool has_u = false;
  bool has_proposal = false;
  std::vector<int64_t> h_off;
  std::vector<int64_t> h_id;
  std::vector<int32_t> h_count;
  // particles
  DArr<float4> posm, velu;
  DArr<int8_t> ptype;
  DArr<int32_t> sub_of, sub_zero, off, off_pairs, order, order_zero;
  DArr<int32_t> state, cur, cur2, src, src_sub, fvals, ford;
  DArr<float> energy, esl, selfm, fenergy, fpot, eref;
  DArr<uint8_t> proposal;
  DArr<int32_t> flag, crank, cnt, cscan;
  DArr<int64_t> flag64;
  DArr<uint64_t> keys;
  DArr<float4> ent;      // per entry (x, y, z, m): ent_cap + 1
  DArr<int32_t> eskip;   // per cell entry: skip
  DArr<int8_t> ent_depth;
  DArr<uint2> wt;        // walk records: 3 ent_cap + 4 units of 8 bytes
  DArr<int32_t> cell_u;  // per entry index: walk-record unit of a cell
  int32_t ent_cap = 0;
  int64_t ent_bound = 0; // entries of the largest trees the batch's subhaloes can make
  DArr<int2> items;
  DArr<int32_t> icnt, iscan;
  // subhaloes
  DArr<int32_t> status, corr, iters, mode, treemode, nlast, nbound, nrem, nuns, ntgt, nsrc, src_off, ent_begin,
    ent_end, nitems, bnd_cnt, bnd_tr, first_idx, seg_begin, seg_len, rbegin, rlen, type_cnt, first_tracer, aux32;
  DArr<float> mlast, mbound, ref, oldref, mb, upscale, ref_in, mb_in, minit;
  DArr<int32_t> warm, npro, nadd, reset_;
  DArr<float> feat, logit;
  // 32-particle groups of the subhaloes of at least egroups_min particles (-1: none built)
  DArr<int32_t> egroups;
  int64_t egroups_n = 0, egroups_min = -1;
  bool feat_full = false;  // feat holds every particle's features (hbtx_batch_features)
  // active-slot tiles of the current round and their prefix sums
  DArr<int4> stiles;
  DArr<int32_t> stile_n, stile_off, tsum32, tpre32, sub_first32;
  DArr<int64_t> tsum64, tpre64, sub_total64;
  // previous snapshot's outputs for the learned warm start (same slot = same particle)
  DArr<int32_t> prev_rank;
  DArr<uint8_t> prev_bound;
  DArr<float> prev_energy, prev_com, prev_mb, prev_rs;
  bool has_prev = false;
  DArr<double> cube, len0, red, red2;
  DArr<uint32_t> bbox;
  DArr<int32_t> misc; // [0] active, [1] (unused), [2] overflow, [3] entry total, [4] max depth
  DArr<int32_t> wctr; // walk work counters, 16 ints apart
  DArr<char2> lev;                // per dense source: the level pair of its tree
  DArr<int32_t> mom_slots;        // moment passes: per window / group, the cells that end beyond it
  DArr<hbtx_sub_result> dres;
  DArr<float> prof_r, prof_m; // profiles: sorted radii and cumulative masses
  DArr<hbtx_sub_profile> dprof;
  bool ran = false;
};

struct View
{
  int64_t n;
  int32_t nsub;
  SubArrays S;
  PartArrays Q;
  const int32_t *order;
  int64_t p0; // first particle of the view in the batch
  int32_t s0; // first subhalo of the view
  int32_t *ford;
  float *fenergy, *fpot;
  int32_t *fvals;
  hbtx_sub_result *dres;
  double *red, *red2;
  int32_t *type_cnt, *first_tracer;
};

struct hbtx_tree
{
  hbtx_ctx *ctx = nullptr;
  int64_t n = 0;
  int32_t ne = 0;
  int64_t ncells = 0;
  DevParams P;
  DArr<float4> ent;
  DArr<int32_t> eskip;
  DArr<int8_t> ent_depth;
  DArr<uint2> wt;   // walk records of the tree
  int32_t nw = 0;   // their units
  DArr<int32_t> src_of_entry;
  DArr<double> cube;
  int32_t treemode = 1;
};

struct hbtx_mlp
{
  DArr<float> master, m, v, wf, grad, partial;
  DArr<uint16_t> wb;
  DArr<double> lossp, loss;
  int64_t step = 0;
};

// What a captured round graph depends on besides buffer addresses.
struct GraphKey
{
  uint64_t gen = ~0ull;
  int64_t p0 = -1, n = -1;
  int32_t s0 = -1, nsub = -1;
  const void *ptype = nullptr, *proposal = nullptr;
  int32_t ent_cap = -1;
  DevParams P;
  bool operator==(const GraphKey &o) const
  {
    return gen == o.gen && p0 == o.p0 && n == o.n && s0 == o.s0 && nsub == o.nsub && ptype == o.ptype &&
           proposal == o.proposal && ent_cap == o.ent_cap && std::memcmp(&P, &o.P, sizeof(P)) == 0;
  }
};

struct hbtx_ctx
{
  int device = 0;
  hipStream_t stream = nullptr;
  std::string err;
  int num_cu = 256;
  int walk_chunks = 8;          // item chunks of the walk (one per XCD)
  int64_t direct_cutoff = 1024; // sources at or below use the exact sum when params ask for the tuned value
  int walk_waves = 32;          // resident walk waves per CU (grid of the persistent walk)
  void *scratch = nullptr;      // this device's address of the no-allocation entry point's array
  std::string tune_source = "built-in defaults";
  // scratch
  DArr<int32_t> part32;
  DArr<int64_t> part64;
  DArr<uint64_t> kalt;
  DArr<int32_t> valt, sort_ntile, sort_toff, ghist, gchunk;
  DArr<int4> sort_tiles;
  DArr<int32_t> red_ntile, red_toff;
  DArr<int4> red_tiles;
  DArr<double> red_partial;
  DArr<unsigned long long> stats;
  hipEvent_t ev0 = nullptr, ev1 = nullptr, wev0 = nullptr, wev1 = nullptr;
  int64_t walk_items_hint = INT64_MAX; // work items of the previous round of the running view
  double walk_ms = 0.0;
  bool count_work = false;
  double stage_ms[8] = {0};
  int rounds = 0;
  // round graphs: head (plan, sources, octree, walk items) and tail (partition,
  // decision, frame), replayed around the directly launched walk
  hipGraphExec_t g_head[3] = {nullptr, nullptr, nullptr}, g_tail = nullptr; // head per round level
  GraphKey g_key;
  int graph_rounds = 0, graph_captures = 0;
  int64_t tree_entries = 0, tree_entries_first = 0; // octree entries built, all rounds / first round
  Batch B;
  // particle reassignment buffers (grow-only, reused across calls)
  struct
  {
    DArr<float> pos, radii;
    DArr<uint8_t> tracer;
    DArr<int32_t> sub, goff, gid, tflag, cnt, loff, qtiles, tval, qval, seg, out, tot;
    DArr<RsGroup> grp;
    DArr<uint32_t> bbox;
    DArr<uint64_t> tkey, qkey;
    DArr<float4> trec, nodes;
    DArr<int64_t> lbase;
    DArr<int4> tiles;
  } rs;
};

namespace
{

// ---------------------------------------------------------------------------
// primitives
// ---------------------------------------------------------------------------
template <class T>
void scan_excl(hbtx_ctx *c, const T *in, T *out, int64_t n, DArr<T> &part, const int32_t *n_dev = nullptr)
{
  const int64_t nb = std::max<int64_t>(1, (n + SCAN_TILE - 1) / SCAN_TILE);
  part.alloc(nb);
  k_scan_reduce<T><<<(unsigned)nb, SCAN_THREADS, 0, c->stream>>>(in, n, n_dev, part.p);
  k_scan_partials<T><<<1, SCAN_THREADS, 0, c->stream>>>(part.p, nb);
  k_scan_down<T><<<(unsigned)nb, SCAN_THREADS, 0, c->stream>>>(in, out, n, n_dev, part.p);
  HIPCHK(hipGetLastError());
}

void scan32(hbtx_ctx *c, const int32_t *in, int32_t *out, int64_t n)
{
  scan_excl<int32_t>(c, in, out, n, c->part32);
}
// length min(n_max, *n_dev), known on the device only; out[length] = total
void scan32_dev(hbtx_ctx *c, const int32_t *in, int32_t *out, int64_t n_max, const int32_t *n_dev)
{
  scan_excl<int32_t>(c, in, out, n_max, c->part32, n_dev);
}
void scan64(hbtx_ctx *c, const int64_t *in, int64_t *out, int64_t n)
{
  scan_excl<int64_t>(c, in, out, n, c->part64);
}

// Radix tiles of a segmented sort over n elements in nseg segments: a segment
// longer than SORT_SMALL_MAX takes ceil(len / SORT_TILE) <= len / SORT_TILE + 1
// tiles, and at most n / (SORT_SMALL_MAX + 1) segments are that long.
int64_t sort_tile_bound(int64_t n, int64_t nseg)
{
  return n / SORT_TILE + std::min<int64_t>(nseg, n / (SORT_SMALL_MAX + 1)) + 2;
}

void reserve_sort(hbtx_ctx *c, int64_t n, int32_t nseg)
{
  c->kalt.alloc(n);
  c->valt.alloc(n);
  c->sort_ntile.alloc(nseg + 1);
  c->sort_toff.alloc(nseg + 1);
  const int64_t maxt = sort_tile_bound(n, nseg);
  c->sort_tiles.alloc(maxt);
  c->ghist.alloc(maxt * RADIX_MAX_BINS);
  c->gchunk.alloc(maxt * RADIX_MAX_BINS); // chunk totals and bases, stored at the chunks' head tiles
}

// Sort (keys, vals) within each segment [seg_begin, seg_begin+seg_len); keys use
// the low key_bits bits.
// small_only: every segment is known to hold at most SORT_SMALL_MAX elements.
void seg_sort(hbtx_ctx *c, uint64_t *keys, int32_t *vals, int32_t nseg, const int32_t *seg_begin,
              const int32_t *seg_len, int64_t n_cap, int key_bits, bool small_only = false)
{
  reserve_sort(c, n_cap, nseg);
  k_sort_wave<<<std::max(1, std::min((nseg + 3) / 4, 16384)), SORT_THREADS, 0, c->stream>>>(keys, vals, seg_begin,
                                                                                            seg_len, nseg);
  k_sort_wave4<<<std::max(1, std::min((nseg + 3) / 4, 16384)), SORT_THREADS, 0, c->stream>>>(keys, vals, seg_begin,
                                                                                            seg_len, nseg);
  k_sort_small<<<std::max(1, std::min(nseg, 16384)), SORT_THREADS, 0, c->stream>>>(keys, vals, seg_begin, seg_len,
                                                                                    nseg);
  if (small_only)
  {
    HIPCHK(hipGetLastError());
    return;
  }
  k_sort_ntiles<<<blocks_for(nseg, 256, 1 << 20), 256, 0, c->stream>>>(seg_len, nseg, c->sort_ntile.p);
  scan32(c, c->sort_ntile.p, c->sort_toff.p, nseg);
  k_sort_tiles<<<blocks_for(nseg, 256, 1 << 20), 256, 0, c->stream>>>(seg_begin, seg_len, c->sort_toff.p, nseg,
                                                                       c->sort_tiles.p);
  const int32_t *ntiles_p = c->sort_toff.p + nseg;
  const int64_t maxt = sort_tile_bound(n_cap, nseg);
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(maxt, 4 * c->num_cu));
  // an even number of passes leaves the result in place; 9-bit digits where they save passes
  auto even_passes = [&](int rb) {
    const int p = (key_bits + rb - 1) / rb;
    return p + (p & 1);
  };
  const int rb = (even_passes(9) < even_passes(8)) ? 9 : 8;
  const int passes = even_passes(rb);
  for (int pass = 0; pass < passes; pass++)
  {
    const int shift = pass * rb;
    const uint64_t *kin = (pass & 1) ? c->kalt.p : keys;
    const int32_t *vin = (pass & 1) ? c->valt.p : vals;
    uint64_t *kout = (pass & 1) ? keys : c->kalt.p;
    int32_t *vout = (pass & 1) ? vals : c->valt.p;
    if (rb == 9)
    {
      k_radix_hist<9><<<grid, SORT_THREADS, 0, c->stream>>>(kin, c->sort_tiles.p, ntiles_p, shift, c->ghist.p);
      k_radix_chunk_scan<9><<<grid, SORT_THREADS, 0, c->stream>>>(c->sort_tiles.p, ntiles_p, c->sort_toff.p,
                                                                  c->sort_ntile.p, c->ghist.p, c->gchunk.p);
      k_radix_seg_scan<9><<<grid, SORT_THREADS, 0, c->stream>>>(seg_begin, c->sort_tiles.p, ntiles_p,
                                                                c->sort_toff.p, c->sort_ntile.p, c->gchunk.p);
      k_radix_scatter<9><<<grid, SORT_THREADS, 0, c->stream>>>(kin, vin, kout, vout, c->sort_tiles.p, ntiles_p,
                                                               shift, c->ghist.p, c->sort_toff.p, c->gchunk.p);
      continue;
    }
    k_radix_hist<8><<<grid, SORT_THREADS, 0, c->stream>>>(kin, c->sort_tiles.p, ntiles_p, shift, c->ghist.p);
    k_radix_chunk_scan<8><<<grid, SORT_THREADS, 0, c->stream>>>(c->sort_tiles.p, ntiles_p, c->sort_toff.p,
                                                                c->sort_ntile.p, c->ghist.p, c->gchunk.p);
    k_radix_seg_scan<8><<<grid, SORT_THREADS, 0, c->stream>>>(seg_begin, c->sort_tiles.p, ntiles_p, c->sort_toff.p,
                                                              c->sort_ntile.p, c->gchunk.p);
    k_radix_scatter<8><<<grid, SORT_THREADS, 0, c->stream>>>(kin, vin, kout, vout, c->sort_tiles.p, ntiles_p, shift,
                                                          c->ghist.p, c->sort_toff.p, c->gchunk.p);
  }
  HIPCHK(hipGetLastError());
}

// Tiles of RED_TILE slots over [begin[s], begin[s]+len[s]); returns device pointer to tile count.
const int32_t *red_tiles(hbtx_ctx *c, const int32_t *begin, const int32_t *len, int32_t nsub, int64_t n_cap)
{
  c->red_ntile.alloc(nsub + 1);
  c->red_toff.alloc(nsub + 1);
  const int64_t maxt = n_cap / RED_TILE + nsub + 1;
  c->red_tiles.alloc(maxt);
  c->red_partial.alloc(maxt * RED_NV);
  k_red_ntiles<<<blocks_for(nsub, 256, 1 << 20), 256, 0, c->stream>>>(len, nsub, c->red_ntile.p);
  scan32(c, c->red_ntile.p, c->red_toff.p, nsub);
  k_red_tiles<<<blocks_for(nsub, 256, 1 << 20), 256, 0, c->stream>>>(begin, len, c->red_toff.p, nsub,
                                                                      c->red_tiles.p);
  return c->red_toff.p + nsub;
}

int red_grid(hbtx_ctx *c, int64_t n_cap, int32_t nsub)
{
  return (int)std::max<int64_t>(1, std::min<int64_t>(n_cap / RED_TILE + nsub + 1, 8 * c->num_cu));
}

DevParams make_dev_params(const hbtx_params &p);
const int32_t *red_tiles(hbtx_ctx *c, const int32_t *begin, const int32_t *len, int32_t nsub, int64_t n_cap);
int red_grid(hbtx_ctx *c, int64_t n_cap, int32_t nsub);

// Profiles and shapes of the bound sets described by A (positions, final order,
// results): sort by radius, tiled cumulative mass and candidates, per-subhalo
// combination into out (device).
void run_profiles(hbtx_ctx *c, const hbtx_params &prm, ProfArgs &A, int64_t n, int32_t nsub, const int32_t *sub_of,
                  int32_t *seg_begin, int32_t *seg_len, uint64_t *keys, int32_t *vals, hbtx_sub_profile *out)
{
  hbtx_params p = prm;
  if (p.direct_cutoff < 0)
    p.direct_cutoff = c->direct_cutoff;
  A.P = make_dev_params(p);
  A.keys = keys;
  A.vals = vals;
  // as the reference types them: G, a, H and the factor 200 are floats (src/snapshot.cpp:268-269)
  const float G = (float)p.G, a = p.scale_factor;
  A.rho = (float)((double)(200.f * p.hz * p.hz) / 2.0 / (double)G * (double)a * (double)a * (double)a);
  A.vel_unit = G / a;
  hipStream_t st = c->stream;
  const int gs = blocks_for(nsub, 256, 1 << 20);
  k_prof_segments<<<gs, 256, 0, st>>>(nsub, A.off, A.res, seg_begin, seg_len);
  k_prof_keys<<<blocks_for(n), 256, 0, st>>>(A, n, sub_of, keys, vals);
  seg_sort(c, keys, vals, nsub, seg_begin, seg_len, n, 32);
  const int32_t *nt = red_tiles(c, seg_begin, seg_len, nsub, n);
  A.part = c->red_partial.p;
  const int rg = red_grid(c, n, nsub);
  k_prof_tile_sum<<<rg, RED_THREADS, 0, st>>>(A, c->red_tiles.p, nt);
  k_prof_tile_prefix<<<gs, 256, 0, st>>>(nsub, c->red_toff.p, c->red_ntile.p, c->red_partial.p);
  k_prof_tile_main<<<rg, RED_THREADS, 0, st>>>(A, c->red_tiles.p, nt);
  k_prof_finish<<<gs, 256, 0, st>>>(A, nsub, c->red_toff.p, c->red_ntile.p, out);
  HIPCHK(hipGetLastError());
}

DevParams make_dev_params(const hbtx_params &p)
{
  REQUIRE(p.scale_factor > 0.f, HBTX_E_INVALID, "scale_factor must be positive");
  REQUIRE(p.softening > 0.f, HBTX_E_INVALID, "softening must be positive");
  REQUIRE(p.theta >= 0.f, HBTX_E_INVALID, "theta must be non-negative");
  REQUIRE(!p.periodic || p.box_size > 0.f, HBTX_E_INVALID, "box_size must be positive when periodic");
  DevParams P;
  std::memset(&P, 0, sizeof(P));
  const float eps = (p.max_phys_softening == -1.f) ? p.softening
                                                   : std::min(p.softening, p.max_phys_softening / p.scale_factor);
  P.G_over_a = p.G / (double)p.scale_factor;
  const float res = p.softening * 0.1f;
  P.res = res;
  P.eps = eps;
  P.h = (float)(2.8 * (double)eps);
  P.hinv = (float)(1.0 / (2.8 * (double)eps));
  P.h2 = P.h * P.h;
  P.theta2 = p.theta * p.theta;
  P.a = p.scale_factor;
  P.hz = p.hz;
  P.box = p.box_size;
  P.box_half = p.box_size / 2.f;
  P.prec = p.bound_mass_precision;
  P.core_factor = p.sub_core_size_factor;
  P.refine_frac = p.bound_fraction_center_refinement;
  P.periodic = p.periodic;
  P.thermal = p.dm_only ? 0 : p.thermal;
  P.dm_only = p.dm_only;
  P.tracer_mask = p.tracer_mask;
  P.donot_mask = p.donot_subsample_mask;
  P.per_type = p.upscale_per_type;
  P.refine = (p.max_sample_size > 0 && p.refine_most_bound) ? 1 : 0;
  P.mss = p.max_sample_size;
  P.min_part = p.min_num_part;
  P.min_tracer = p.min_num_tracer;
  P.core_min = p.sub_core_size_min;
  P.direct_cutoff = p.direct_cutoff;
  return P;
}

// default direct/tree cutoff when the caller asks for the tuned value
// ---------------------------------------------------------------------------
// Tuning: the direct/tree cutoff and the walk's item chunking.  Defaults are
// replaced by the autotuner's file (HBTX_TUNE_FILE, else hbtx_tune.json next to
// this library; written by hbt-herons_amd/autotune.py on the target GPU), then
// by the environment (HBTX_DIRECT_CUTOFF, HBTX_WALK_CHUNKS), then by
// hbtx_set_tuning.
// ---------------------------------------------------------------------------
bool json_int(const std::string &txt, const char *name, int64_t &out)
{
  const std::string key = std::string("\"") + name + "\"";
  const size_t k = txt.find(key);
  if (k == std::string::npos)
    return false;
  const size_t c = txt.find(':', k + key.size());
  if (c == std::string::npos)
    return false;
  char *end = nullptr;
  const long long v = strtoll(txt.c_str() + c + 1, &end, 10);
  if (end == txt.c_str() + c + 1)
    return false;
  out = v;
  return true;
}

std::string tune_file_path()
{
  if (const char *e = getenv("HBTX_TUNE_FILE"))
    return e;
  Dl_info info;
  if (dladdr((void *)&tune_file_path, &info) && info.dli_fname)
  {
    std::string lib = info.dli_fname;
    const size_t slash = lib.find_last_of('/');
    return (slash == std::string::npos ? std::string(".") : lib.substr(0, slash)) + "/hbtx_tune.json";
  }
  return "hbtx_tune.json";
}

// Most entries an octree over n sources can have: n particles, and at each of
// the key's depths 0..HBTX_KEY_LEVELS at most min(8^depth, n/2) cells (a cell
// holds >= 2 sources, and the cells of one depth are disjoint).  Every source
// set of a subhalo is a subset of its particles.
int64_t entry_bound(int64_t n)
{
  int64_t e = n, w = 1;
  if (n < 2)
    return e;
  for (int l = 0; l <= HBTX_KEY_LEVELS; l++)
  {
    e += std::min(w, n / 2);
    w = std::min<int64_t>(w * 8, (int64_t)1 << 40);
  }
  return e;
}

void batch_reserve(hbtx_ctx *c, Batch &B)
{
  const int64_t n = B.n;
  const int32_t ns = B.nsub;
  B.posm.alloc(n);
  B.velu.alloc(n);
  B.ptype.alloc(n);
  B.sub_of.alloc(n);
  B.sub_zero.alloc(std::max<int64_t>(B.max_count, 1));
  B.off.alloc(ns + 1);
  B.off_pairs.alloc(2 * (int64_t)ns + 2);
  B.order.alloc(ns + 1);
  B.order_zero.alloc(1);
  B.state.alloc(n);
  B.cur.alloc(n);
  B.cur2.alloc(n);
  B.src.alloc(n);
  B.src_sub.alloc(n);
  B.fvals.alloc(n);
  B.ford.alloc(n);
  B.energy.alloc(n);
  B.esl.alloc(n);
  B.selfm.alloc(n);
  B.fenergy.alloc(n);
  B.fpot.alloc(n);
  B.eref.alloc(n);
  B.proposal.alloc(n);
  B.flag.alloc(n + 1);
  B.crank.alloc(n + 1);
  B.cnt.alloc(n + 1);
  B.cscan.alloc(n + 1);
  B.flag64.alloc(n + 1);
  B.keys.alloc(n);
  B.lev.alloc(n + 1);
  B.ent_cap = (int32_t)std::min<int64_t>(3 * n + 64, WREC_MAX_ENTRIES);
  B.ent.alloc((int64_t)B.ent_cap + 1);
  B.eskip.alloc((int64_t)B.ent_cap + 1);
  B.ent_depth.alloc(((int64_t)B.ent_cap + 16) & ~(int64_t)15);
  B.wt.alloc(3 * (int64_t)B.ent_cap + 4);
  B.cell_u.alloc((int64_t)B.ent_cap + 1);
  B.items.alloc(n / HBTX_WAVE + ns + 1);
  B.icnt.alloc(ns + 1);
  B.iscan.alloc(ns + 1);
  DArr<int32_t> *i32s[] = {&B.status, &B.corr, &B.iters, &B.mode, &B.treemode, &B.nlast, &B.nbound, &B.nrem,
                           &B.nuns, &B.ntgt, &B.nsrc, &B.src_off, &B.ent_begin, &B.ent_end, &B.nitems, &B.bnd_cnt,
                           &B.bnd_tr, &B.first_idx, &B.seg_begin, &B.seg_len, &B.rbegin, &B.rlen, &B.first_tracer,
                           &B.aux32};
  for (auto *a : i32s)
    a->alloc(ns + 1);
  B.type_cnt.alloc(6 * (int64_t)ns + 6);
  DArr<int32_t> *w3[] = {&B.warm, &B.npro, &B.nadd, &B.reset_};
  for (auto *a : w3)
    a->alloc(ns + 1);
  DArr<float> *f1[] = {&B.mlast, &B.mbound, &B.minit};
  for (auto *a : f1)
    a->alloc(ns + 1);
  DArr<float> *f6[] = {&B.ref, &B.oldref, &B.mb, &B.upscale, &B.ref_in, &B.mb_in};
  for (auto *a : f6)
    a->alloc(6 * (int64_t)ns + 6);
  B.cube.alloc(4 * (int64_t)ns + 4);
  B.len0.alloc(ns + 1);
  B.red.alloc((int64_t)RED_NV * ns + RED_NV);
  B.red2.alloc((int64_t)RED_NV * ns + RED_NV);
  B.bbox.alloc(6 * (int64_t)ns + 6);
  B.misc.alloc(8);
  B.wctr.alloc(16 * WALK_MAX_CHUNKS);
  B.dres.alloc(ns + 1);
  c->stats.alloc(NSTATS);
  {
    const int64_t maxt = n / SLOT_TILE + ns + 2;
    B.stiles.alloc(maxt);
    B.stile_n.alloc(ns + 1);
    B.stile_off.alloc(ns + 1);
    B.tsum32.alloc(maxt + 1);
    B.tpre32.alloc(maxt + 1);
    B.tsum64.alloc(maxt + 1);
    B.tpre64.alloc(maxt + 1);
    B.sub_first32.alloc(ns + 1);
    B.sub_total64.alloc(ns + 1);
  }
  reserve_sort(c, n, ns);
}

View make_view(Batch &B, int64_t p0, int64_t n, int32_t s0, int32_t nsub, bool whole)
{
  View V;
  V.n = n;
  V.nsub = nsub;
  V.p0 = p0;
  V.s0 = s0;
  SubArrays &S = V.S;
  S.nsub = nsub;
  S.off = whole ? B.off.p : B.off_pairs.p + 2 * s0;
  S.sub_of = whole ? B.sub_of.p : B.sub_zero.p;
  S.status = B.status.p + s0;
  S.corr = B.corr.p + s0;
  S.iters = B.iters.p + s0;
  S.mode = B.mode.p + s0;
  S.treemode = B.treemode.p + s0;
  S.nlast = B.nlast.p + s0;
  S.nbound = B.nbound.p + s0;
  S.nrem = B.nrem.p + s0;
  S.nuns = B.nuns.p + s0;
  S.ntgt = B.ntgt.p + s0;
  S.nsrc = B.nsrc.p + s0;
  S.mlast = B.mlast.p + s0;
  S.mbound = B.mbound.p + s0;
  S.ref = B.ref.p + 6 * s0;
  S.oldref = B.oldref.p + 6 * s0;
  S.mb = B.mb.p + 6 * s0;
  S.upscale = B.upscale.p + 6 * s0;
  S.src_off = B.src_off.p + s0;
  S.ent_begin = B.ent_begin.p + s0;
  S.ent_end = B.ent_end.p + s0;
  S.nitems = B.nitems.p + s0;
  S.cube = B.cube.p + 4 * s0;
  S.len0 = B.len0.p + s0;
  S.bbox = B.bbox.p + 6 * s0;
  S.bnd_cnt = B.bnd_cnt.p + s0;
  S.bnd_tr = B.bnd_tr.p + s0;
  S.first_idx = B.first_idx.p + s0;
  S.active_count = B.misc.p + 0;
  S.max_active_n = B.misc.p + 6;
  S.warm = B.warm.p + s0;
  S.npro = B.npro.p + s0;
  S.nadd = B.nadd.p + s0;
  S.reset = B.reset_.p + s0;
  S.minit = B.minit.p + s0;
  PartArrays &Q = V.Q;
  Q.posm = B.posm.p + p0;
  Q.velu = B.velu.p + p0;
  Q.ptype = B.has_type ? B.ptype.p + p0 : nullptr;
  Q.state = B.state.p + p0;
  Q.energy = B.energy.p + p0;
  Q.esl = B.esl.p + p0;
  Q.selfm = B.selfm.p + p0;
  Q.cur = B.cur.p + p0;
  Q.cur2 = B.cur2.p + p0;
  Q.proposal = B.has_proposal ? B.proposal.p + p0 : nullptr;
  V.order = whole ? B.order.p : B.order_zero.p;
  V.ford = B.ford.p + p0;
  V.fenergy = B.fenergy.p + p0;
  V.fpot = B.fpot.p + p0;
  V.fvals = B.fvals.p + p0;
  V.dres = B.dres.p + s0;
  V.red = B.red.p + (int64_t)RED_NV * s0;
  V.red2 = B.red2.p + (int64_t)RED_NV * s0;
  V.type_cnt = B.type_cnt.p + 6 * s0;
  V.first_tracer = B.first_tracer.p + s0;
  return V;
}

TreeArgs tree_args(Batch &B, View &V, const DevParams &P)
{
  TreeArgs T;
  T.src = B.src.p;
  T.keys = B.keys.p;
  T.src_sub = B.src_sub.p;
  T.src_off = V.S.src_off;
  T.nsrc = V.S.nsrc;
  T.treemode = V.S.treemode;
  T.mode = V.S.mode;
  T.len0 = V.S.len0;
  T.posm = V.Q.posm;
  T.ptype = V.Q.ptype;
  T.upscale = V.S.upscale;
  T.cnt = B.cnt.p;
  T.cscan = B.cscan.p;
  T.s_total = B.tpre32.p + (V.n / SLOT_TILE + V.nsub + 1); // grand total of the source tile scan
  T.ent = B.ent.p;
  T.eskip = B.eskip.p;
  T.ent_depth = B.ent_depth.p;
  T.wt = B.wt.p;
  T.cell_u = B.cell_u.p;
  T.lev = B.lev.p;
  T.cap = B.ent_cap;
  T.overflow = B.misc.p + 2;
  T.P = P;
  return T;
}

void build_sources_tree(hbtx_ctx *c, Batch &B, View &V, const DevParams &P, const int32_t *nt, bool graph,
                        int level, bool tsum_ready);
// device round counter (1 in the first round), advanced by k_round_begin
int32_t *round_ptr(Batch &B)
{
  return B.misc.p + 5;
}

// entry buffers for at least `need` entries (slack: a quarter more, for the next rounds)
void grow_entries(Batch &B, int64_t need, bool slack = true)
{
  const int64_t cap = std::min<int64_t>(slack ? need + need / 4 + 64 : need, WREC_MAX_ENTRIES);
  REQUIRE(cap >= need, HBTX_E_CAPACITY, "tree entry count exceeds the walk-record range");
  B.ent.release();
  B.eskip.release();
  B.ent_depth.release();
  B.ent.alloc(cap + 1);
  B.eskip.alloc(cap + 1);
  B.ent_depth.alloc((cap + 16) & ~(int64_t)15);
  B.wt.release();
  B.cell_u.release();
  B.wt.alloc(3 * cap + 4);
  B.cell_u.alloc(cap + 1);
  B.ent_cap = (int32_t)cap;
}

int64_t slot_tile_cap(const View &V)
{
  return V.n / SLOT_TILE + V.nsub + 1; // (tree_args reads the source total at tpre32[this])
}
int tile_grid(hbtx_ctx *c, const View &V)
{
  return (int)std::max<int64_t>(1, std::min<int64_t>(slot_tile_cap(V), 16 * c->num_cu));
}

// Tiles over the slot ranges of the view's active subhaloes (mode set by the
// round's plan); returns the device-side tile count.
const int32_t *build_slot_tiles(hbtx_ctx *c, Batch &B, View &V)
{
  const int gs = blocks_for(V.nsub, 256, 1 << 20);
  k_slot_ntiles<<<gs, 256, 0, c->stream>>>(V.S, B.stile_n.p);
  scan32(c, B.stile_n.p, B.stile_off.p, V.nsub);
  k_slot_tiles<<<gs, 256, 0, c->stream>>>(V.S, B.stile_off.p, B.stiles.p);
  return B.stile_off.p + V.nsub;
}

// Sources of the round -> dense list, then (tree-mode subhaloes) the octree.
void round_sources_and_tree(hbtx_ctx *c, Batch &B, View &V, const DevParams &P, bool graph, int level)
{
  hipStream_t st = c->stream;
  const int64_t n = V.n;
  const int gb = blocks_for(n);
  const int gs = blocks_for(V.nsub, 256, 1 << 20);
  const int32_t *nt = build_slot_tiles(c, B, V);
  if (P.mss > 0)
  {
    // candidate ranks in E-list order and the mass upscale factors (:290-391)
    k_mark_candidates<<<gb, 256, 0, st>>>(V.S, V.Q, n, B.flag.p);
    scan32(c, B.flag.p, B.crank.p, n);
    k_upscale_ranges<<<gs, 256, 0, st>>>(V.S, B.rbegin.p, B.rlen.p);
    const int32_t *nt = red_tiles(c, B.rbegin.p, B.rlen.p, V.nsub, n);
    k_upscale_tiles<<<red_grid(c, n, V.nsub), 256, 0, st>>>(V.S, V.Q, P, c->red_tiles.p, nt, B.crank.p,
                                                            c->red_partial.p);
    k_red_finalize<<<red_finalize_blocks(V.nsub), 256, 0, st>>>(c->red_toff.p, c->red_ntile.p, V.nsub, c->red_partial.p, V.red2);
    k_upscale_finish<<<gs, 256, 0, st>>>(V.S, P, V.red2);
  }
  // (without subsampling the candidate rank is irrelevant: every candidate is a source)
  k_mark_sources<<<tile_grid(c, V), SCAN_THREADS, 0, st>>>(V.S, V.Q, P, B.stiles.p, nt, round_ptr(B),
                                                           P.mss > 0 ? B.crank.p : nullptr, B.flag.p, B.tsum32.p,
                                                           (int32_t)slot_tile_cap(V));
  build_sources_tree(c, B, V, P, nt, graph, level, true);
}

// Flags in B.flag (on the slots of the active tiles) -> dense source list, tree
// cube, keys, sort, octree entries.  tsum_ready: the flag producer already wrote
// the tile totals into B.tsum32.
void build_sources_tree(hbtx_ctx *c, Batch &B, View &V, const DevParams &P, const int32_t *nt, bool graph,
                        int level, bool tsum_ready)
{
  hipStream_t st = c->stream;
  const int64_t n = V.n;
  const int gb = blocks_for(n);
  const int gs = blocks_for(V.nsub, 256, 1 << 20);
  const int64_t maxt = slot_tile_cap(V);
  if (!tsum_ready)
    k_tile_sum<int32_t><<<tile_grid(c, V), SLOT_THREADS, 0, st>>>(B.stiles.p, nt, (int32_t)maxt, B.flag.p,
                                                                  B.tsum32.p);
  scan_excl<int32_t>(c, B.tsum32.p, B.tpre32.p, maxt, c->part32);
  const int32_t *s_total = B.tpre32.p + maxt;
  k_scatter_sources<<<tile_grid(c, V), SLOT_THREADS, 0, st>>>(V.S, B.stiles.p, nt, B.flag.p, B.tpre32.p, B.src.p,
                                                              B.src_sub.p, B.sub_first32.p);
  k_sub_src<<<gs, 256, 0, st>>>(V.S, B.sub_first32.p);
  k_tree_src_ranges<<<gs, 256, 0, st>>>(V.S, B.rbegin.p, B.rlen.p);
  const int32_t *nbt = red_tiles(c, B.rbegin.p, B.rlen.p, V.nsub, n);
  k_bbox_tiles<<<red_grid(c, n, V.nsub), 256, 0, st>>>(V.S, V.Q.posm, B.src.p, c->red_tiles.p, nbt);
  k_cube<<<gs, 256, 0, st>>>(V.S);
  k_keys<<<gb, 256, 0, st>>>(V.S, P, V.Q.posm, B.src.p, B.src_sub.p, s_total, B.keys.p);
  k_sort_segments<<<gs, 256, 0, st>>>(V.S, B.seg_begin.p, B.seg_len.p);
  seg_sort(c, B.keys.p, B.src.p, V.nsub, B.seg_begin.p, B.seg_len.p, n, 63, level >= 1);
  k_targets_from_sources<<<gb, 256, 0, st>>>(V.S, V.Q, B.src.p, B.src_sub.p, s_total);
  TreeArgs T = tree_args(B, V, P);
  k_tree_count<<<blocks_for(n, 256, 8 * 1024), 256, 0, st>>>(T, B.cnt.p);
  scan32_dev(c, B.cnt.p, B.cscan.p, n, s_total);
  if (!graph)
  {
    // exact entry count and depth of this round's trees; grow the entry buffers if needed
    k_entry_total<<<1, 1, 0, st>>>(s_total, B.cscan.p, B.misc.p + 3, 0x7fffffff, B.misc.p + 2);
    int32_t etot = 0;
    HIPCHK(hipMemcpyAsync(&etot, B.misc.p + 3, sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    REQUIRE(etot >= 0, HBTX_E_CAPACITY, "tree entry count overflows int32");
    if (etot > B.ent_cap)
    {
      grow_entries(B, etot);
      T = tree_args(B, V, P);
    }
  }
  // (a captured round cannot wait for the count: its buffers are sized for the
  // largest trees the batch can make)
  k_tree_emit<<<blocks_for(n, 256, 8 * 1024), 256, 0, st>>>(T);
  k_sub_entries<<<gs, 256, 0, st>>>(V.S, B.cscan.p, B.ent_cap, B.misc.p + 2);
  k_entry_total<<<1, 1, 0, st>>>(s_total, B.cscan.p, B.misc.p + 3, B.ent_cap, B.misc.p + 2);
  if (level >= 2)
  {
    HIPCHK(hipGetLastError());
    return; // direct sums only: no cells
  }
  // moments bottom-up: LDS-staged windows, then one pass per level of window
  // groups (levels sized for the entry capacity; each pass reads the entry total)
  std::vector<int64_t> units{(std::max<int64_t>(B.ent_cap, 1) + MOM_WIN - 1) / MOM_WIN};
  while (units.back() > 1)
    units.push_back((units.back() + MOM_GROUP - 1) / MOM_GROUP);
  int64_t nslots = 0;
  for (int64_t u : units)
    nslots += u;
  B.mom_slots.alloc(nslots * MOM_SLOTS);
  int32_t *sl = B.mom_slots.p;
  k_moments_windows<<<(int)std::min<int64_t>(units[0], 4096), MOM_THREADS, 0, st>>>(T, B.misc.p + 3, sl);
  int64_t span = MOM_WIN;
  for (size_t k = 1; k < units.size(); k++)
  {
    int32_t *up = sl + units[k - 1] * MOM_SLOTS;
    k_moments_groups<<<(int)std::min<int64_t>(units[k], 4096), MOM_GROUP, 0, st>>>(T, B.misc.p + 3, span, sl, up);
    sl = up;
    span *= MOM_GROUP;
  }
  HIPCHK(hipGetLastError());
}

WalkArgs walk_args(hbtx_ctx *c, Batch &B, View &V, const DevParams &P)
{
  WalkArgs W;
  W.items = B.items.p;
  W.nitems = B.iscan.p + V.nsub;
  W.counter = B.wctr.p;
  W.nchunks = c->walk_chunks;
  W.cur = V.Q.cur;
  W.off = V.S.off;
  W.ntgt = V.S.ntgt;
  W.mode = V.S.mode;
  W.treemode = V.S.treemode;
  W.ent_begin = V.S.ent_begin;
  W.ent_end = V.S.ent_end;
  W.wt = B.wt.p;
  W.posm = V.Q.posm;
  W.velu = V.Q.velu;
  W.selfm = V.Q.selfm;
  W.ref = V.S.ref;
  W.oldref = V.S.oldref;
  W.energy = V.Q.energy;
  W.esl = V.Q.esl;
  W.eref = B.eref.p + V.p0;
  W.phi_out = nullptr;
  W.stats = c->stats.p;
  W.overflow = B.misc.p + 2;
  W.P = P;
  return W;
}

// the walk's work items (largest subhaloes first) and its cleared chunk counters
void round_items(hbtx_ctx *c, Batch &B, View &V)
{
  hipStream_t st = c->stream;
  const int gs = blocks_for(V.nsub, 256, 1 << 20);
  k_items_count<<<gs, 256, 0, st>>>(V.S, V.order, B.icnt.p);
  scan32(c, B.icnt.p, B.iscan.p, V.nsub);
  k_items_emit<<<blocks_for((int64_t)V.nsub * HBTX_WAVE, 256, 1 << 24), 256, 0, st>>>(V.S, V.order, B.iscan.p,
                                                                                       B.items.p);
  k_fill_i32<<<1, 256, 0, st>>>(B.wctr.p, 16 * WALK_MAX_CHUNKS, 0);
  k_copy_i32<<<1, 64, 0, st>>>(B.iscan.p + V.nsub, B.misc.p + 7); // item count, read after the round
  HIPCHK(hipGetLastError());
}

// the persistent walk over the round's items, timed by the walk events
void round_walk_launch(hbtx_ctx *c, Batch &B, View &V, const DevParams &P)
{
  hipStream_t st = c->stream;
  WalkArgs W = walk_args(c, B, V, P);
  // (the persistent walk needs no more waves than items; a round has no more
  // items than the round before it, as every subhalo's target count only shrinks)
  const int64_t maxitems = std::min<int64_t>(V.n / HBTX_WAVE + V.nsub + 1, c->walk_items_hint);
  const int waves = (int)std::max<int64_t>(1, std::min<int64_t>(maxitems, (int64_t)c->walk_waves * c->num_cu));
  const int grid = (waves + HBTX_WALK_WG_WAVES - 1) / HBTX_WALK_WG_WAVES;
  HIPCHK(hipEventRecord(c->wev0, st));
  if (P.periodic)
  {
    if (c->count_work)
      k_walk_energy<true, true><<<grid, WALK_WG_THREADS, 0, st>>>(W);
    else
      k_walk_energy<true, false><<<grid, WALK_WG_THREADS, 0, st>>>(W);
  }
  else
  {
    if (c->count_work)
      k_walk_energy<false, true><<<grid, WALK_WG_THREADS, 0, st>>>(W);
    else
      k_walk_energy<false, false><<<grid, WALK_WG_THREADS, 0, st>>>(W);
  }
  HIPCHK(hipEventRecord(c->wev1, st));
  HIPCHK(hipGetLastError());
}

void round_walk(hbtx_ctx *c, Batch &B, View &V, const DevParams &P)
{
  round_items(c, B, V);
  round_walk_launch(c, B, V, P);
}

void round_partition(hbtx_ctx *c, Batch &B, View &V, const DevParams &P, bool warm_round)
{
  hipStream_t st = c->stream;
  const int64_t n = V.n;
  const int gb = blocks_for(n);
  const int gs = blocks_for(V.nsub, 256, 1 << 20);
  // the round's active tiles (built from the plan at the start of the round)
  const int32_t *nt = B.stile_off.p + V.nsub;
  const int tg = tile_grid(c, V);
  if (warm_round)
    k_fill_i32<<<gs, 256, 0, st>>>(B.aux32.p, V.nsub, 0);
  const int64_t maxt = slot_tile_cap(V);
  k_partition_flags<<<tg, SCAN_THREADS, 0, st>>>(V.S, V.Q, P, B.stiles.p, nt, B.flag64.p,
                                                 warm_round ? B.aux32.p : nullptr, B.tsum64.p, (int32_t)maxt);
  scan_excl<int64_t>(c, B.tsum64.p, B.tpre64.p, maxt, c->part64);
  k_partition_scatter<<<tg, SLOT_THREADS, 0, st>>>(V.S, V.Q, B.stiles.p, nt, round_ptr(B), B.flag64.p, B.tpre64.p,
                                                   B.sub_total64.p);
  k_copy_cur<<<tg, 256, 0, st>>>(V.S, V.Q, B.stiles.p, nt, B.sub_total64.p);
  if (P.mss > 0)
  {
    k_fill_i32<<<gs, 256, 0, st>>>(B.aux32.p, V.nsub, 0);
    k_count_unsampled<<<gb, 256, 0, st>>>(V.S, V.Q, P, n, B.aux32.p);
    k_set_unsampled<<<gs, 256, 0, st>>>(V.S, B.aux32.p);
  }
  k_round_decide<<<gs, 256, 0, st>>>(V.S, P, B.sub_total64.p, warm_round ? B.aux32.p : nullptr);
  k_com_ranges<<<gs, 256, 0, st>>>(V.S, B.rbegin.p, B.rlen.p);
  if (P.periodic)
  {
    k_fill_i32<<<gs, 256, 0, st>>>(V.S.first_idx, V.nsub, 0x7fffffff);
    k_first_index<<<gb, 256, 0, st>>>(V.S, V.Q, n);
  }
  const int32_t *ntiles = red_tiles(c, B.rbegin.p, B.rlen.p, V.nsub, n);
  k_red_tile<ComOp><<<red_grid(c, n, V.nsub), RED_THREADS, 0, st>>>(V.S, V.Q, P, c->red_tiles.p, ntiles,
                                                                     c->red_partial.p);
  k_red_finalize<<<red_finalize_blocks(V.nsub), 256, 0, st>>>(c->red_toff.p, c->red_ntile.p, V.nsub, c->red_partial.p, V.red);
  k_round_finish<<<gs, 256, 0, st>>>(V.S, P, V.red);
  if (P.periodic)
    k_round_finish_periodic<<<gs, 256, 0, st>>>(V.S, V.Q, P, V.red);
  if (warm_round)
  {
    k_warm_reset_particles<<<gb, 256, 0, st>>>(V.S, V.Q, n);
    k_warm_reset_subs<<<gs, 256, 0, st>>>(V.S, B.ref_in.p + 6 * V.s0);
  }
  HIPCHK(hipGetLastError());
}

__global__ void k_final_local(SubArrays S, int64_t n, const int32_t *vals, const float *energy, int32_t *ford,
                              float *fenergy)
{
  GRID_STRIDE(q, n)
  {
    const int s = S.sub_of[q];
    const int p = vals[q];
    ford[q] = p - S.off[s];
    fenergy[q] = energy[p];
  }
}

__global__ void k_final_sub(SubArrays S, PartArrays Q, DevParams P, const int32_t *vals, const double *kin,
                            const double *core, const int32_t *first_tracer, const float *ref_in,
                            hbtx_sub_result *out)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  hbtx_sub_result R = {};
  const int n = S.off[s + 1] - S.off[s];
  const int st = S.status[s];
  R.iterations = S.iters[s];
  R.mostbound_id = -1;
  const float *mb = S.mb + 6 * s;
  if (n == 0)
  {
    // orphan early exit (src/subhalo_unbind.cpp:472-489): frame unchanged, core = most-bound
    for (int k = 0; k < 3; k++)
    {
      R.com_pos[k] = ref_in[6 * s + k];
      R.com_vel[k] = ref_in[6 * s + 3 + k];
      R.mostbound_pos[k] = mb[k];
      R.mostbound_vel[k] = mb[3 + k];
      R.core_pos[k] = mb[k];
      R.core_vel[k] = mb[3 + k];
    }
    out[s] = R;
    return;
  }
  const float *rf = S.ref + 6 * s;
  for (int k = 0; k < 3; k++)
  {
    R.com_pos[k] = rf[k];
    R.com_vel[k] = rf[3 + k];
  }
  if (st != ST_CONVERGED)
  {
    R.disrupted = 1;
    for (int k = 0; k < 3; k++)
    {
      R.mostbound_pos[k] = mb[k];
      R.mostbound_vel[k] = mb[3 + k];
      R.core_pos[k] = mb[k];
      R.core_vel[k] = mb[3 + k];
    }
    out[s] = R;
    return;
  }
  const int nb = S.nbound[s];
  R.nbound = nb;
  R.mbound = S.mbound[s];
  const int p0 = vals[S.off[s]];
  const float4 x0 = Q.posm[p0];
  const float4 v0 = Q.velu[p0];
  R.mostbound_pos[0] = x0.x;
  R.mostbound_pos[1] = x0.y;
  R.mostbound_pos[2] = x0.z;
  R.mostbound_vel[0] = v0.x;
  R.mostbound_vel[1] = v0.y;
  R.mostbound_vel[2] = v0.z;
  const double *v = kin + (int64_t)RED_NV * s;
  if (nb > 1)
  {
    const double M = v[6];
    double E = v[0] - v[1];
    E /= M;
    double K = v[2] * (0.5 / M);
    R.spec_pot = (float)(E - K);
    R.spec_kin = (float)K;
    R.spec_am[0] = (float)(v[3] / M);
    R.spec_am[1] = (float)(v[4] / M);
    R.spec_am[2] = (float)(v[5] / M);
  }
  if (!P.dm_only)
  {
    for (int k = 0; k < 6; k++)
    {
      R.nbound_type[k] = (int64_t)v[13 + k];
      R.mbound_type[k] = (float)v[7 + k];
    }
    const int ft = first_tracer[s];
    R.tracer_index = (ft == 0x7fffffff) ? 0 : ft;
  }
  else
    R.tracer_index = 0;
  const double *c = core + (int64_t)RED_NV * s;
  const double msum = c[0];
  double s2r = 0., s2v = 0.;
  for (int d = 0; d < 3; d++)
  {
    const double pos = c[1 + d] / msum;
    const double pos2 = c[4 + d] / msum - pos * pos;
    const double vel = c[7 + d] / msum;
    const double vel2 = c[10 + d] / msum - vel * vel;
    float cp = (float)pos;
    if (P.periodic)
      cp = (float)(pos + (double)x0.x * (d == 0) + (double)x0.y * (d == 1) + (double)x0.z * (d == 2));
    R.core_pos[d] = cp;
    R.core_vel[d] = (float)vel;
    s2r += pos2;
    s2v += vel2;
  }
  R.core_sigma_r = (float)sqrt(s2r);
  R.core_sigma_v = (float)sqrt(s2v);
  out[s] = R;
}

void run_final(hbtx_ctx *c, Batch &B, View &V, const DevParams &P, bool want_pot, int rounds)
{
  hipStream_t st = c->stream;
  const int64_t n = V.n;
  const int gb = blocks_for(n);
  const int gs = blocks_for(V.nsub, 256, 1 << 20);
  k_final_keys<<<gb, 256, 0, st>>>(V.S, V.Q, n, rounds, B.keys.p, V.fvals);
  k_final_segments<<<gs, 256, 0, st>>>(V.S, B.seg_begin.p, B.seg_len.p);
  int gbits = 0;
  while ((1 << gbits) <= rounds)
    gbits++;
  seg_sort(c, B.keys.p, V.fvals, V.nsub, B.seg_begin.p, B.seg_len.p, n, 32 + gbits);
  if (P.refine)
  {
    // RefineBindingEnergyOrder (:395-425): tree over the most bound particles, true masses
    k_refine_plan<<<gs, 256, 0, st>>>(V.S, P);
    const int32_t *nt = build_slot_tiles(c, B, V);
    k_slot_of<<<gb, 256, 0, st>>>(V.S, n, V.fvals, V.Q.cur2);
    k_mark_refine<<<gb, 256, 0, st>>>(V.S, V.Q, n, V.Q.cur2, B.flag.p);
    build_sources_tree(c, B, V, P, nt, false, 0, false);
    c->walk_items_hint = INT64_MAX; // (other targets than the rounds')
    round_walk(c, B, V, P);
    k_refine_keys<<<gb, 256, 0, st>>>(V.S, n, V.fvals, B.eref.p + V.p0, B.keys.p, B.seg_begin.p, B.seg_len.p);
    k_refine_segments<<<gs, 256, 0, st>>>(V.S, B.seg_begin.p, B.seg_len.p);
    seg_sort(c, B.keys.p, V.fvals, V.nsub, B.seg_begin.p, B.seg_len.p, n, 32);
  }
  k_final_local<<<gb, 256, 0, st>>>(V.S, n, V.fvals, V.Q.energy, V.ford, V.fenergy);
  // kinematics / types over the bound particles (input order)
  k_final_ranges<<<gs, 256, 0, st>>>(V.S, B.rbegin.p, B.rlen.p);
  const int32_t *ntiles = red_tiles(c, B.rbegin.p, B.rlen.p, V.nsub, n);
  if (V.Q.ptype)
    k_final_tile<true><<<red_grid(c, n, V.nsub), RED_THREADS, 0, st>>>(V.S, V.Q, P, c->red_tiles.p, ntiles,
                                                                      c->red_partial.p);
  else
    k_final_tile<false><<<red_grid(c, n, V.nsub), RED_THREADS, 0, st>>>(V.S, V.Q, P, c->red_tiles.p, ntiles,
                                                                       c->red_partial.p);
  k_red_finalize<<<red_finalize_blocks(V.nsub), 256, 0, st>>>(c->red_toff.p, c->red_ntile.p, V.nsub, c->red_partial.p, V.red);
  const int gw = blocks_for((int64_t)V.nsub * HBTX_WAVE, 256, 1 << 24);
  k_first_tracer<<<gw, 256, 0, st>>>(V.S, V.Q, P, V.fvals, V.first_tracer);
  k_core_wave<<<gw, 256, 0, st>>>(V.S, V.Q, P, V.fvals, V.red2);
  k_final_sub<<<gs, 256, 0, st>>>(V.S, V.Q, P, V.fvals, V.red, V.red2, V.first_tracer, B.ref_in.p + 6 * V.s0,
                                  V.dres);
  if (want_pot)
    k_potential_energy<<<gb, 256, 0, st>>>(V.S, V.Q, P, n, V.fvals, V.fpot);
  HIPCHK(hipGetLastError());
}

__global__ void k_all_ranges(SubArrays S, int32_t *begin, int32_t *len)
{
  const int s = blockIdx.x * blockDim.x + threadIdx.x;
  if (s >= S.nsub)
    return;
  begin[s] = S.off[s];
  len[s] = S.off[s + 1] - S.off[s];
}

void k_final_ranges_all(hbtx_ctx *c, Batch &B, View &V)
{
  k_all_ranges<<<blocks_for(V.nsub, 256, 1 << 20), 256, 0, c->stream>>>(V.S, B.rbegin.p, B.rlen.p);
}

// Iterate one view to convergence. Returns rounds used.
// Per-view state at the start of the unbinding: particles, frames, the initial
// Mbound (:537-541) and the leading do-not-subsample run (:522-524).
void init_view(hbtx_ctx *c, Batch &B, View &V, const DevParams &P, bool warm)
{
  hipStream_t st = c->stream;
  const int gb = blocks_for(V.n);
  const int gs = blocks_for(V.nsub, 256, 1 << 20);
  k_init_particles<<<gb, 256, 0, st>>>(V.S, V.Q, V.n);
  k_init_subs<<<gs, 256, 0, st>>>(V.S, B.ref_in.p + 6 * V.s0, B.mb_in.p + 6 * V.s0, nullptr);
  k_fill_i32<<<1, 64, 0, st>>>(B.misc.p, 8, 0);
  k_final_ranges_all(c, B, V);
  const int32_t *nt = red_tiles(c, B.rbegin.p, B.rlen.p, V.nsub, V.n);
  k_fill_i32<<<gs, 256, 0, st>>>(B.aux32.p, V.nsub, 0x7fffffff);
  k_fill_i32<<<gs, 256, 0, st>>>(V.S.first_idx, V.nsub, 0x7fffffff);
  if (warm && P.periodic)
    k_first_proposed<<<gb, 256, 0, st>>>(V.S, V.Q, V.n);
  if (P.periodic)
    k_init_mass_tiles<true><<<red_grid(c, V.n, V.nsub), 256, 0, st>>>(V.S, V.Q, P, c->red_tiles.p, nt, c->red_partial.p,
                                                                B.aux32.p);
  else
    k_init_mass_tiles<false><<<red_grid(c, V.n, V.nsub), 256, 0, st>>>(V.S, V.Q, P, c->red_tiles.p, nt, c->red_partial.p,
                                                                B.aux32.p);
  k_red_finalize<<<red_finalize_blocks(V.nsub), 256, 0, st>>>(c->red_toff.p, c->red_ntile.p, V.nsub, c->red_partial.p, V.red);
  k_init_mass_finish<<<gs, 256, 0, st>>>(V.S, V.Q, P, V.red, B.aux32.p, warm ? 1 : 0);
  HIPCHK(hipGetLastError());
}

// First part of a round, up to the walk: plan, sources, octrees, work items.
// graph: no host synchronisation (the part is being captured).  level (a bound
// the host knows from the previous round): 0 any subhalo; 1 every active
// subhalo fits the LDS sort (no radix passes); 2 every active subhalo also takes
// the direct sum (no tree cells, no moment passes).
void round_head(hbtx_ctx *c, Batch &B, View &V, const DevParams &P, bool graph, bool warm_round, int level = 0)
{
  hipStream_t st = c->stream;
  const int gs = blocks_for(V.nsub, 256, 1 << 20);
  k_round_begin<<<1, 64, 0, st>>>(V.S.active_count, round_ptr(B), V.S.max_active_n);
  k_plan<<<gs, 256, 0, st>>>(V.S, P, round_ptr(B));
  round_sources_and_tree(c, B, V, P, graph, level);
  if (warm_round)
  {
    // warm round: every target, in Hilbert order of its cell in the proposal's tree cube
    k_warm_keys<<<blocks_for(V.n), 256, 0, st>>>(V.S, V.Q, V.n, B.keys.p, B.seg_begin.p, B.seg_len.p);
    k_warm_segments<<<gs, 256, 0, st>>>(V.S, B.seg_begin.p, B.seg_len.p);
    seg_sort(c, B.keys.p, V.Q.cur, V.nsub, B.seg_begin.p, B.seg_len.p, V.n, 30);
  }
  round_items(c, B, V);
}

void drop_graphs(hbtx_ctx *c)
{
  for (auto &g : c->g_head)
  {
    if (g)
      (void)hipGraphExecDestroy(g);
    g = nullptr;
  }
  if (c->g_tail)
    (void)hipGraphExecDestroy(c->g_tail);
  c->g_tail = nullptr;
  c->g_key = GraphKey();
}

template <class F>
hipGraphExec_t capture(hbtx_ctx *c, F &&record)
{
  hipStream_t st = c->stream;
  HIPCHK(hipStreamBeginCapture(st, hipStreamCaptureModeThreadLocal));
  try
  {
    record();
  }
  catch (...)
  {
    hipGraph_t g = nullptr;
    (void)hipStreamEndCapture(st, &g);
    if (g)
      (void)hipGraphDestroy(g);
    (void)hipGetLastError();
    throw;
  }
  hipGraph_t g = nullptr;
  HIPCHK(hipStreamEndCapture(st, &g));
  hipGraphExec_t ex = nullptr;
  const hipError_t e = hipGraphInstantiate(&ex, g, nullptr, nullptr, 0);
  (void)hipGraphDestroy(g);
  HIPCHK(e);
  return ex;
}

// Round graphs for this view, captured again when a buffer moved or the view or
// the parameters changed.  Capture only records: every buffer the round touches
// already has its size from the first (direct) round.  Returns false (rounds
// then run directly) if the capture could not be made.
bool ensure_round_graphs(hbtx_ctx *c, Batch &B, View &V, const DevParams &P, int level)
{
  GraphKey k;
  k.p0 = V.p0;
  k.n = V.n;
  k.s0 = V.s0;
  k.nsub = V.nsub;
  k.ptype = V.Q.ptype;
  k.proposal = V.Q.proposal;
  k.ent_cap = B.ent_cap;
  k.P = P;
  k.gen = g_alloc_gen.load();
  if (c->g_head[level] && c->g_tail && k == c->g_key)
    return true;
  if (!(k == c->g_key))
    drop_graphs(c);
  try
  {
    c->g_head[level] = capture(c, [&] { round_head(c, B, V, P, true, false, level); });
    if (!c->g_tail)
      c->g_tail = capture(c, [&] { round_partition(c, B, V, P, false); });
  }
  catch (const Error &)
  {
    drop_graphs(c);
    return false;
  }
  if (g_alloc_gen.load() != k.gen)
  {
    drop_graphs(c);
    return false;
  }
  c->g_key = k;
  c->graph_captures++;
  return true;
}

// The overflow word of a round: bit 1 the entry buffers were too small, bit 2 a
// tree's records exceed the walk's int32 byte offsets, bit 4 a walk stack was full
// (not reachable for trees of at most 22 cell levels).
void check_overflow(int32_t ov)
{
  REQUIRE(!(ov & WALK_STACK_OVERFLOW), HBTX_E_CAPACITY, "tree walk stack capacity exceeded");
  REQUIRE(!(ov & 2), HBTX_E_CAPACITY,
          "a subhalo's octree exceeds the walk's 2 GiB record span (int32 byte offsets)");
  REQUIRE(ov == 0, HBTX_E_CAPACITY, "tree entry capacity exceeded");
}

// Iterate one view to convergence. Returns rounds used.  graphs: rounds after
// the first replay the captured head and tail around the walk.
int run_view(hbtx_ctx *c, Batch &B, View &V, const DevParams &P, bool want_pot, bool warm, bool graphs)
{
  hipStream_t st = c->stream;
  if (graphs)
  {
    // replayed rounds need entry buffers for the largest trees the view can make
    if (B.ent_bound > WREC_MAX_ENTRIES)
      graphs = false;
    else if (B.ent_cap < B.ent_bound)
      grow_entries(B, B.ent_bound, false);
  }
  init_view(c, B, V, P, warm);
  c->walk_items_hint = INT64_MAX;
  int round = 0, level = 0;
  const int max_rounds = 2000;
  while (true)
  {
    REQUIRE(round < max_rounds, HBTX_E_INVALID, "unbinding did not converge within the round limit");
    round++;
    const bool replay = graphs && round >= 2 && ensure_round_graphs(c, B, V, P, level);
    if (replay)
    {
      HIPCHK(hipGraphLaunch(c->g_head[level], st));
      round_walk_launch(c, B, V, P);
      HIPCHK(hipGraphLaunch(c->g_tail, st));
      c->graph_rounds++;
    }
    else
    {
      round_head(c, B, V, P, false, warm && round == 1, round == 1 ? 0 : level);
      round_walk_launch(c, B, V, P);
      round_partition(c, B, V, P, warm && round == 1);
    }
    int32_t h_misc[8];
    HIPCHK(hipMemcpyAsync(h_misc, B.misc.p, sizeof(h_misc), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    float wms = 0.f;
    HIPCHK(hipEventElapsedTime(&wms, c->wev0, c->wev1));
    c->walk_ms += wms;
    c->walk_items_hint = std::max(h_misc[7], 1);
    c->tree_entries += h_misc[3];
    if (round == 1)
      c->tree_entries_first += h_misc[3];
    // the next round's level: subhaloes only leave the active set
    const int64_t maxn = h_misc[6];
    level = (maxn <= P.direct_cutoff) ? 2 : (maxn <= SORT_SMALL_MAX ? 1 : 0);
    check_overflow(h_misc[2]);
    if (h_misc[0] == 0)
      break;
  }
  run_final(c, B, V, P, want_pot, round);
  return round;
}

// The batch layout limits, checked before anything touches the device:
// particle slots and subhalo indices are int32, and the walk addresses one
// subhalo's records by int32 byte offsets, so a subhalo's tree (16 bytes a
// particle, 24 a cell) must stay below 2 GiB.  Particles alone fix the bound
// below; the cells are counted on the device once the tree exists (the span
// check of k_sub_entries).
void check_layout(int64_t n, int64_t nsub, const int64_t *offsets)
{
  REQUIRE(n >= 0 && nsub >= 1, HBTX_E_INVALID, "need n >= 0 and nsub >= 1");
  REQUIRE(n < (int64_t)0x20000000, HBTX_E_INVALID, "at most 2^29 particles per batch");
  REQUIRE(nsub < (int64_t)0x7fffffff, HBTX_E_INVALID, "too many subhaloes");
  REQUIRE(offsets, HBTX_E_INVALID, "null offsets");
  REQUIRE(offsets[0] == 0 && offsets[nsub] == n, HBTX_E_INVALID, "offsets must start at 0 and end at n");
  for (int64_t s = 0; s < nsub; s++)
  {
    REQUIRE(offsets[s + 1] >= offsets[s], HBTX_E_INVALID, "offsets must be non-decreasing");
    REQUIRE(offsets[s + 1] - offsets[s] <= HBTX_MAX_SUBHALO_PARTICLES, HBTX_E_CAPACITY,
            "subhalo " + std::to_string(s) + " has " + std::to_string(offsets[s + 1] - offsets[s]) +
              " particles: its octree would exceed the walk's 2 GiB record span (at most " +
              std::to_string((long long)HBTX_MAX_SUBHALO_PARTICLES) + " particles per subhalo)");
  }
}

void upload_batch(hbtx_ctx *c, Batch &B, int64_t n, const float *pos, const float *vel, const float *mass,
                  const float *u, const int32_t *type, const int64_t *id, int64_t nsub, const int64_t *offsets,
                  const float *ref_pos, const float *ref_vel, const float *mb_pos, const float *mb_vel)
{
  check_layout(n, nsub, offsets);
  REQUIRE(pos && vel && mass && ref_pos && ref_vel, HBTX_E_INVALID, "null input array");
  B.n = n;
  B.nsub = (int32_t)nsub;
  B.has_type = type != nullptr;
  B.has_u = u != nullptr;
  B.has_proposal = false;
  B.has_prev = false;
  B.egroups_min = -1;
  B.feat_full = false;
  B.h_off.assign(offsets, offsets + nsub + 1);
  B.h_count.resize(nsub);
  B.max_count = 0;
  for (int64_t s = 0; s < nsub; s++)
  {
    B.h_count[s] = (int32_t)(offsets[s + 1] - offsets[s]);
    B.max_count = std::max(B.max_count, B.h_count[s]);
  }
  B.ent_bound = 0;
  for (int64_t s = 0; s < nsub; s++)
    B.ent_bound += entry_bound(B.h_count[s]);
  if (id)
    B.h_id.assign(id, id + n);
  else
  {
    B.h_id.resize(n);
    for (int64_t s = 0; s < nsub; s++)
      for (int64_t i = offsets[s]; i < offsets[s + 1]; i++)
        B.h_id[i] = i - offsets[s];
  }
  batch_reserve(c, B);
  std::vector<float4> hp(n), hv(n);
  std::vector<int32_t> hsub(n);
  for (int64_t i = 0; i < n; i++)
  {
    // (the walk records mark cells by the sign bit of the mass)
    REQUIRE(mass[i] >= 0.f, HBTX_E_INVALID, "particle masses must be non-negative");
    hp[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], mass[i]);
    hv[i] = make_float4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], u ? u[i] : 0.f);
  }
  for (int64_t s = 0; s < nsub; s++)
    for (int64_t i = offsets[s]; i < offsets[s + 1]; i++)
      hsub[i] = (int32_t)s;
  std::vector<int32_t> hoff(nsub + 1), hpairs(2 * nsub + 2, 0), horder(nsub);
  for (int64_t s = 0; s <= nsub; s++)
    hoff[s] = (int32_t)offsets[s];
  for (int64_t s = 0; s < nsub; s++)
    hpairs[2 * s + 1] = B.h_count[s];
  std::iota(horder.begin(), horder.end(), 0);
  std::stable_sort(horder.begin(), horder.end(),
                   [&](int32_t a, int32_t b) { return B.h_count[a] > B.h_count[b]; });
  std::vector<float> hr(6 * nsub), hm(6 * nsub);
  for (int64_t s = 0; s < nsub; s++)
    for (int k = 0; k < 3; k++)
    {
      hr[6 * s + k] = ref_pos[3 * s + k];
      hr[6 * s + 3 + k] = ref_vel[3 * s + k];
      hm[6 * s + k] = mb_pos ? mb_pos[3 * s + k] : ref_pos[3 * s + k];
      hm[6 * s + 3 + k] = mb_vel ? mb_vel[3 * s + k] : ref_vel[3 * s + k];
    }
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(B.posm.p, hp.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(B.velu.p, hv.data(), n * sizeof(float4), hipMemcpyHostToDevice, st));
  if (type)
  {
    std::vector<int8_t> ht(n);
    for (int64_t i = 0; i < n; i++)
    {
      REQUIRE(type[i] >= 0 && type[i] < 6, HBTX_E_INVALID, "particle type out of range [0,6)");
      ht[i] = (int8_t)type[i];
    }
    HIPCHK(hipMemcpyAsync(B.ptype.p, ht.data(), n, hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
  }
  HIPCHK(hipMemcpyAsync(B.sub_of.p, hsub.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(B.sub_zero.p, 0, std::max<int64_t>(B.max_count, 1) * sizeof(int32_t), st));
  HIPCHK(hipMemsetAsync(B.order_zero.p, 0, sizeof(int32_t), st));
  HIPCHK(hipMemcpyAsync(B.off.p, hoff.data(), (nsub + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(B.off_pairs.p, hpairs.data(), (2 * nsub + 2) * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(B.order.p, horder.data(), nsub * sizeof(int32_t), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(B.ref_in.p, hr.data(), 6 * nsub * sizeof(float), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(B.mb_in.p, hm.data(), 6 * nsub * sizeof(float), hipMemcpyHostToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  B.ran = false;
}

void run_batch(hbtx_ctx *c, const hbtx_params &prm, uint32_t flags, double *elapsed_ms)
{
  Batch &B = c->B;
  REQUIRE(B.n > 0 || B.nsub > 0, HBTX_E_INVALID, "no batch loaded");
  hbtx_params p = prm;
  if (p.direct_cutoff < 0)
    p.direct_cutoff = c->direct_cutoff;
  REQUIRE(p.max_sample_size >= 0, HBTX_E_INVALID, "MaxSampleSizeOfPotentialEstimate must be >= 0");
  const bool warm = (flags & HBTX_WARMSTART) != 0;
  REQUIRE(!warm || B.has_proposal, HBTX_E_INVALID, "warm start needs a proposal (hbtx_batch_set_proposal / propose)");
  REQUIRE(!warm || p.max_sample_size == 0, HBTX_E_INVALID, "warm start is exclusive with subsampling");
  const DevParams P = make_dev_params(p);
  hipStream_t st = c->stream;
  HIPCHK(hipMemsetAsync(c->stats.p, 0, NSTATS * sizeof(unsigned long long), st));
  c->walk_ms = 0.0;
  c->rounds = 0;
  c->graph_rounds = 0;
  c->graph_captures = 0;
  c->tree_entries = c->tree_entries_first = 0;
  HIPCHK(hipEventRecord(c->ev0, st));
  const bool want_pot = (flags & HBTX_SAVE_POTENTIAL) != 0;
  c->count_work = (flags & HBTX_COUNT) != 0;
  if (flags & HBTX_PLAIN)
  {
    for (int32_t s = 0; s < B.nsub; s++)
    {
      View V = make_view(B, B.h_off[s], B.h_count[s], s, 1, false);
      c->rounds += run_view(c, B, V, P, want_pot, warm, false);
    }
  }
  else
  {
    View V = make_view(B, 0, B.n, 0, B.nsub, true);
    c->rounds = run_view(c, B, V, P, want_pot, warm, (flags & HBTX_NO_GRAPH) == 0);
  }
  HIPCHK(hipEventRecord(c->ev1, st));
  HIPCHK(hipEventSynchronize(c->ev1));
  float ms = 0.f;
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  if (elapsed_ms)
    *elapsed_ms = ms;
  B.ran = true;
}

void fetch_batch(hbtx_ctx *c, int64_t *out_order, float *out_energy, float *out_potential,
                 hbtx_sub_result *out_res)
{
  Batch &B = c->B;
  REQUIRE(B.ran, HBTX_E_INVALID, "batch has not been run");
  hipStream_t st = c->stream;
  const int64_t n = B.n;
  std::vector<int32_t> ford(n);
  HIPCHK(hipMemcpyAsync(ford.data(), B.ford.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
  if (out_energy)
    HIPCHK(hipMemcpyAsync(out_energy, B.fenergy.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
  if (out_potential)
    HIPCHK(hipMemcpyAsync(out_potential, B.fpot.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
  std::vector<hbtx_sub_result> res(B.nsub);
  HIPCHK(hipMemcpyAsync(res.data(), B.dres.p, B.nsub * sizeof(hbtx_sub_result), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  for (int64_t s = 0; s < B.nsub; s++)
  {
    const int64_t a = B.h_off[s];
    for (int64_t k = 0; k < B.h_count[s]; k++)
      if (out_order)
        out_order[a + k] = ford[a + k];
    hbtx_sub_result &R = res[s];
    if (R.nbound > 0 && !R.disrupted)
    {
      const int64_t ti = R.tracer_index;
      R.mostbound_id = B.h_id[a + ford[a + ti]];
    }
    else
      R.mostbound_id = -1;
    if (out_res)
      out_res[s] = R;
  }
}

// ---------------------------------------------------------------------------
// tree API
// ---------------------------------------------------------------------------
void build_tree(hbtx_ctx *c, const hbtx_params &prm, int64_t n, const float *pos, const float *mass, hbtx_tree *T,
                int treemode)
{
  REQUIRE(n >= 1, HBTX_E_INVALID, "a tree needs at least one source");
  Batch &B = c->B;
  std::vector<float> vel(3 * n, 0.f);
  const int64_t off[2] = {0, n};
  float zero3[3] = {0.f, 0.f, 0.f};
  upload_batch(c, B, n, pos, vel.data(), mass, nullptr, nullptr, nullptr, 1, off, zero3, zero3, nullptr, nullptr);
  hbtx_params p = prm;
  p.direct_cutoff = treemode ? 0 : ((int64_t)1 << 40);
  p.max_sample_size = 0;
  const DevParams P = make_dev_params(p);
  View V = make_view(B, 0, n, 0, 1, true);
  hipStream_t st = c->stream;
  k_init_particles<<<blocks_for(n), 256, 0, st>>>(V.S, V.Q, n);
  k_init_subs<<<1, 64, 0, st>>>(V.S, B.ref_in.p, B.mb_in.p, nullptr);
  k_fill_i32<<<1, 64, 0, st>>>(B.misc.p, 8, 0);
  k_round_begin<<<1, 64, 0, st>>>(V.S.active_count, round_ptr(B), V.S.max_active_n);
  k_plan<<<1, 64, 0, st>>>(V.S, P, round_ptr(B));
  round_sources_and_tree(c, B, V, P, false, 0);
  int32_t h_misc[4];
  HIPCHK(hipMemcpyAsync(h_misc, B.misc.p, sizeof(h_misc), hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  check_overflow(h_misc[2]);
  const int32_t ne = h_misc[3];
  T->ctx = c;
  T->n = n;
  T->ne = ne;
  T->ncells = ne - n;
  T->P = P;
  T->treemode = treemode;
  T->ent.alloc((int64_t)ne + 1);
  T->eskip.alloc((int64_t)ne + 1);
  T->ent_depth.alloc(ne);
  T->src_of_entry.alloc(n);
  T->cube.alloc(4);
  // walk records: 2 units per particle, 3 per cell, and two spare units
  T->nw = (int32_t)(2 * n + 3 * ((int64_t)ne - n));
  T->wt.alloc((int64_t)T->nw + 2);
  HIPCHK(hipMemcpyAsync(T->wt.p, B.wt.p, ((int64_t)T->nw + 2) * sizeof(uint2), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(T->ent.p, B.ent.p, ((int64_t)ne + 1) * sizeof(float4), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(T->eskip.p, B.eskip.p, ((int64_t)ne + 1) * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(T->ent_depth.p, B.ent_depth.p, ne, hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(T->src_of_entry.p, B.src.p, n * sizeof(int32_t), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipMemcpyAsync(T->cube.p, B.cube.p, 4 * sizeof(double), hipMemcpyDeviceToDevice, st));
  HIPCHK(hipStreamSynchronize(st));
  B.ran = false;
}

void tree_points(hbtx_ctx *c, hbtx_tree *T, int64_t nt, const float *tpos, const float *tmass, double *out)
{
  REQUIRE(nt >= 0 && nt < (int64_t)0x20000000, HBTX_E_INVALID, "bad target count");
  if (nt == 0)
    return;
  hipStream_t st = c->stream;
  DArr<float4> dt;
  DArr<float> phi;
  DArr<double> phid;
  DArr<uint64_t> keys;
  DArr<int32_t> vals, segb, segl, counter;
  dt.alloc(nt);
  phi.alloc(nt);
  phid.alloc(nt);
  keys.alloc(nt);
  vals.alloc(nt);
  segb.alloc(1);
  segl.alloc(1);
  counter.alloc(2); // [0] work counter, [1] overflow word
  std::vector<float4> h(nt);
  for (int64_t i = 0; i < nt; i++)
    h[i] = make_float4(tpos[3 * i], tpos[3 * i + 1], tpos[3 * i + 2], tmass ? tmass[i] : 0.f);
  HIPCHK(hipMemcpyAsync(dt.p, h.data(), nt * sizeof(float4), hipMemcpyHostToDevice, st));
  // walk order: targets sorted by their key in the tree cube
  k_point_keys<<<blocks_for(nt), 256, 0, st>>>(dt.p, (int32_t)nt, T->cube.p, keys.p, vals.p);
  int32_t hb = 0, hl = (int32_t)nt;
  HIPCHK(hipMemcpyAsync(segb.p, &hb, 4, hipMemcpyHostToDevice, st));
  HIPCHK(hipMemcpyAsync(segl.p, &hl, 4, hipMemcpyHostToDevice, st));
  if (T->treemode)
    seg_sort(c, keys.p, vals.p, 1, segb.p, segl.p, nt, 63);
  HIPCHK(hipMemsetAsync(counter.p, 0, 8, st));
  PointArgs A;
  A.tgt = dt.p;
  A.tgt_order = vals.p;
  A.ntgt = (int32_t)nt;
  A.eb = 0;
  A.ee = T->nw;
  A.treemode = T->treemode;
  A.wt = T->wt.p;
  A.phi = phi.p;
  A.phi_d = phid.p;
  A.counter = counter.p;
  A.stats = c->stats.p;
  A.overflow = counter.p + 1;
  A.P = T->P;
  HIPCHK(hipMemsetAsync(c->stats.p, 0, NSTATS * sizeof(unsigned long long), st));
  const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(nt / HBTX_WAVE + 1, 32 * c->num_cu) /
                                                 HBTX_WALK_WG_WAVES);
  HIPCHK(hipEventRecord(c->wev0, st));
  if (T->P.periodic)
  {
    if (c->count_work)
      k_points_potential<true, true><<<grid, WALK_WG_THREADS, 0, st>>>(A);
    else
      k_points_potential<true, false><<<grid, WALK_WG_THREADS, 0, st>>>(A);
  }
  else
  {
    if (c->count_work)
      k_points_potential<false, true><<<grid, WALK_WG_THREADS, 0, st>>>(A);
    else
      k_points_potential<false, false><<<grid, WALK_WG_THREADS, 0, st>>>(A);
  }
  HIPCHK(hipEventRecord(c->wev1, st));
  HIPCHK(hipGetLastError());
  HIPCHK(hipMemcpyAsync(out, phid.p, nt * sizeof(double), hipMemcpyDeviceToHost, st));
  int32_t ov = 0;
  HIPCHK(hipMemcpyAsync(&ov, counter.p + 1, 4, hipMemcpyDeviceToHost, st));
  HIPCHK(hipStreamSynchronize(st));
  REQUIRE(ov == 0, HBTX_E_CAPACITY, "tree walk stack capacity exceeded");
  float wms = 0.f;
  HIPCHK(hipEventElapsedTime(&wms, c->wev0, c->wev1));
  c->walk_ms = wms;
}

int fail(hbtx_ctx *c, int code, const std::string &msg)
{
  if (c)
    c->err = msg;
  return code;
}

} // namespace

#define API_GUARD(ctx, ...)                                                                                          \
  try                                                                                                                \
  {                                                                                                                  \
    __VA_ARGS__;                                                                                                     \
    return HBTX_OK;                                                                                                  \
  }                                                                                                                  \
  catch (const Error &e)                                                                                             \
  {                                                                                                                  \
    return fail(ctx, e.code, e.what());                                                                              \
  }                                                                                                                  \
  catch (const std::exception &e)                                                                                    \
  {                                                                                                                  \
    return fail(ctx, HBTX_E_INVALID, e.what());                                                                      \
  }

extern "C" {

void hbtx_default_params(hbtx_params *p)
{
  if (!p)
    return;
  std::memset(p, 0, sizeof(*p));
  p->G = 43.0071;
  p->scale_factor = 1.f;
  p->hz = 0.f;
  p->softening = 0.001f;
  p->max_phys_softening = -1.f;
  p->theta = 0.45f;
  p->bound_mass_precision = 0.995f;
  p->max_sample_size = 1000;
  p->refine_most_bound = 1;
  p->bound_fraction_center_refinement = 0.1f;
  p->min_num_part = 20;
  p->min_num_tracer = 10;
  p->tracer_mask = (1 << 1) | (1 << 4);
  p->donot_subsample_mask = (1 << 5);
  p->periodic = 1;
  p->box_size = 0.f;
  p->thermal = 0;
  p->dm_only = 1;
  p->upscale_per_type = 1;
  p->sub_core_size_factor = 0.f;
  p->sub_core_size_min = 20;
  p->direct_cutoff = -1;
}

} // extern "C"

void load_tuning(hbtx_ctx *c)
{
  const std::string path = tune_file_path();
  std::ifstream f(path);
  if (f)
  {
    std::stringstream ss;
    ss << f.rdbuf();
    const std::string txt = ss.str();
    int64_t v;
    if (json_int(txt, "direct_cutoff", v) && v >= 0)
      c->direct_cutoff = v;
    if (json_int(txt, "walk_chunks", v) && v >= 1)
      c->walk_chunks = (int)std::min<int64_t>(WALK_MAX_CHUNKS, v);
    if (json_int(txt, "walk_waves", v) && v >= 1)
      c->walk_waves = (int)std::min<int64_t>(64, v);
    c->tune_source = path;
  }
  if (const char *e = getenv("HBTX_DIRECT_CUTOFF"))
  {
    c->direct_cutoff = std::max<int64_t>(0, atoll(e));
    c->tune_source = "environment";
  }
  if (const char *e = getenv("HBTX_WALK_CHUNKS"))
  {
    c->walk_chunks = std::min(WALK_MAX_CHUNKS, std::max(1, atoi(e)));
    c->tune_source = "environment";
  }
  if (const char *e = getenv("HBTX_WALK_WAVES"))
  {
    c->walk_waves = std::min(64, std::max(1, atoi(e)));
    c->tune_source = "environment";
  }
}

extern "C" {

int hbtx_create(hbtx_ctx **out, int device)
{
  if (!out)
    return HBTX_E_INVALID;
  *out = nullptr;
  hbtx_ctx *c = new (std::nothrow) hbtx_ctx();
  if (!c)
    return HBTX_E_ALLOC;
  try
  {
    int ndev = 0;
    HIPCHK(hipGetDeviceCount(&ndev));
    REQUIRE(device >= 0 && device < ndev, HBTX_E_INVALID, "no such HIP device");
    c->device = device;
    HIPCHK(hipSetDevice(device));
    hipDeviceProp_t prop;
    HIPCHK(hipGetDeviceProperties(&prop, device));
    c->num_cu = std::max(1, prop.multiProcessorCount);
    c->scratch = scratch_address(device);
    load_tuning(c);
    HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
    HIPCHK(hipEventCreate(&c->ev0));
    HIPCHK(hipEventCreate(&c->ev1));
    HIPCHK(hipEventCreate(&c->wev0));
    HIPCHK(hipEventCreate(&c->wev1));
    c->stats.alloc(NSTATS);
  }
  catch (const Error &e)
  {
    int code = e.code;
    delete c;
    return code;
  }
  *out = c;
  return HBTX_OK;
}

void hbtx_destroy(hbtx_ctx *c)
{
  if (!c)
    return;
  (void)hipSetDevice(c->device);
  if (c->stream)
    (void)hipStreamSynchronize(c->stream);
  if (c->ev0)
    (void)hipEventDestroy(c->ev0);
  if (c->ev1)
    (void)hipEventDestroy(c->ev1);
  if (c->wev0)
    (void)hipEventDestroy(c->wev0);
  if (c->wev1)
    (void)hipEventDestroy(c->wev1);
  drop_graphs(c);
  hipStream_t st = c->stream;
  delete c;
  if (st)
    (void)hipStreamDestroy(st);
}

const char *hbtx_last_error(const hbtx_ctx *c)
{
  return c ? c->err.c_str() : "null context";
}

int hbtx_device_sync(hbtx_ctx *c)
{
  if (!c)
    return HBTX_E_INVALID;
  API_GUARD(c, HIPCHK(hipStreamSynchronize(c->stream)));
}

int hbtx_check_layout(int64_t n, int64_t nsub, const int64_t *offsets, char *msg, int32_t msg_len)
{
  std::string m;
  int rc = HBTX_OK;
  try
  {
    check_layout(n, nsub, offsets);
  }
  catch (const Error &e)
  {
    rc = e.code;
    m = e.what();
  }
  if (msg && msg_len > 0)
  {
    const size_t k = std::min<size_t>(m.size(), (size_t)msg_len - 1);
    std::memcpy(msg, m.data(), k);
    msg[k] = 0;
  }
  return rc;
}

int hbtx_batch_load(hbtx_ctx *c, int64_t n, const float *pos, const float *vel, const float *mass, const float *u,
                    const int32_t *type, const int64_t *id, int64_t nsub, const int64_t *offsets,
                    const float *ref_pos, const float *ref_vel, const float *mb_pos, const float *mb_vel)
{
  if (!c)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    upload_batch(c, c->B, n, pos, vel, mass, u, type, id, nsub, offsets, ref_pos, ref_vel, mb_pos, mb_vel);
  });
}

int hbtx_batch_set_proposal(hbtx_ctx *c, const uint8_t *proposal)
{
  if (!c)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    REQUIRE(proposal, HBTX_E_INVALID, "null proposal");
    HIPCHK(hipMemcpyAsync(c->B.proposal.p, proposal, c->B.n, hipMemcpyHostToDevice, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    c->B.has_proposal = true;
  });
}

int hbtx_batch_run(hbtx_ctx *c, const hbtx_params *prm, uint32_t flags, double *elapsed_ms)
{
  if (!c || !prm)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    run_batch(c, *prm, flags, elapsed_ms);
  });
}

int hbtx_batch_fetch(hbtx_ctx *c, int64_t *out_order, float *out_energy, float *out_potential,
                     hbtx_sub_result *out_res)
{
  if (!c)
    return HBTX_E_INVALID;
  API_GUARD(c, fetch_batch(c, out_order, out_energy, out_potential, out_res));
}

static const char *kStatNames[NSTATS] = {"tree_wave_steps", "tree_lane_visits", "tree_lane_accepts",
                                         "direct_lane_pairs", "targets", "work_items", "near_interactions",
                                         "tree_entries", "tree_entries_first", "unused9", "unused10", "unused11",
                                         "walk_ms",
                                         "rounds", "graph_rounds", "graph_captures"};

int hbtx_batch_stats(hbtx_ctx *c, double *stats, int nstats)
{
  if (!c || !stats)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    unsigned long long h[NSTATS];
    HIPCHK(hipMemcpy(h, c->stats.p, sizeof(h), hipMemcpyDeviceToHost));
    for (int i = 0; i < nstats && i < NSTATS; i++)
      stats[i] = (double)h[i];
    if (nstats > 7)
      stats[7] = (double)c->tree_entries;
    if (nstats > 8)
      stats[8] = (double)c->tree_entries_first;
    if (nstats > 12)
      stats[12] = c->walk_ms;
    if (nstats > 13)
      stats[13] = c->rounds;
    if (nstats > 14)
      stats[14] = c->graph_rounds;
    if (nstats > 15)
      stats[15] = c->graph_captures;
  });
}

int hbtx_device_bytes(int64_t *bytes)
{
  if (!bytes)
    return HBTX_E_INVALID;
  *bytes = g_device_bytes.load();
  return 0;
}

const char *hbtx_stat_name(int i)
{
  return (i >= 0 && i < NSTATS) ? kStatNames[i] : "";
}

int hbtx_unbind_batch(hbtx_ctx *c, const hbtx_params *prm, int64_t n, const float *pos, const float *vel,
                      const float *mass, const float *u, const int32_t *type, const int64_t *id, int64_t nsub,
                      const int64_t *offsets, const float *ref_pos, const float *ref_vel, const float *mb_pos,
                      const float *mb_vel, uint32_t flags, int64_t *out_order, float *out_energy,
                      float *out_potential, hbtx_sub_result *out_res)
{
  if (!c || !prm)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    upload_batch(c, c->B, n, pos, vel, mass, u, type, id, nsub, offsets, ref_pos, ref_vel, mb_pos, mb_vel);
    run_batch(c, *prm, flags, nullptr);
    fetch_batch(c, out_order, out_energy, out_potential, out_res);
  });
}

int hbtx_tree_build(hbtx_ctx *c, const hbtx_params *prm, int64_t n, const float *pos, const float *mass,
                    hbtx_tree **tree, int64_t *ncells)
{
  if (!c || !prm || !tree || !pos || !mass)
    return HBTX_E_INVALID;
  *tree = nullptr;
  hbtx_tree *T = new (std::nothrow) hbtx_tree();
  if (!T)
    return HBTX_E_ALLOC;
  try
  {
    HIPCHK(hipSetDevice(c->device));
    build_tree(c, *prm, n, pos, mass, T, 1);
  }
  catch (const Error &e)
  {
    delete T;
    return fail(c, e.code, e.what());
  }
  *tree = T;
  if (ncells)
    *ncells = T->ncells;
  return HBTX_OK;
}

int hbtx_tree_potential(hbtx_ctx *c, hbtx_tree *T, int64_t n_tgt, const float *tgt_pos, const float *tgt_mass,
                        double *out_pot)
{
  if (!c || !T || (n_tgt > 0 && (!tgt_pos || !out_pot)))
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    tree_points(c, T, n_tgt, tgt_pos, tgt_mass, out_pot);
  });
}

int hbtx_tree_binding_energy(hbtx_ctx *c, hbtx_tree *T, int64_t n_tgt, const float *tgt_pos, const float *tgt_vel,
                             const float *ref_pos, const float *ref_vel, const float *tgt_mass, double *out_energy)
{
  if (!c || !T || (n_tgt > 0 && (!tgt_pos || !tgt_vel || !ref_pos || !ref_vel || !out_energy)))
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    tree_points(c, T, n_tgt, tgt_pos, tgt_mass, out_energy);
    const float ha = T->P.hz * T->P.a;
    for (int64_t i = 0; i < n_tgt; i++)
    {
      // Snapshot_t::RelativeVelocity in float, then VecNorm * 0.5 + pot
      float k2 = 0.f;
      for (int j = 0; j < 3; j++)
      {
        float dx = tgt_pos[3 * i + j] - ref_pos[j];
        if (T->P.periodic)
          dx = (dx > T->P.box_half) ? dx - T->P.box : ((dx < -T->P.box_half) ? dx + T->P.box : dx);
        float dv = tgt_vel[3 * i + j] - ref_vel[j];
        dv += ha * dx;
        k2 += dv * dv;
      }
      out_energy[i] = (double)k2 * 0.5 + out_energy[i];
    }
  });
}

int hbtx_tree_export(hbtx_ctx *c, hbtx_tree *T, int64_t cap, int32_t *kind, int64_t *pidx, float *mass, float *com,
                     float *len, int64_t *skip, int32_t *depth, int64_t *nentries)
{
  if (!c || !T)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    const int32_t ne = T->ne;
    if (nentries)
      *nentries = ne;
    REQUIRE(cap >= ne, HBTX_E_CAPACITY, "export capacity too small");
    std::vector<float4> en(ne);
    std::vector<int32_t> esk(ne);
    std::vector<int8_t> dep(ne);
    std::vector<int32_t> src(T->n);
    double cube[4];
    HIPCHK(hipMemcpy(en.data(), T->ent.p, ne * sizeof(float4), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(esk.data(), T->eskip.p, ne * sizeof(int32_t), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(cube, T->cube.p, sizeof(cube), hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(dep.data(), T->ent_depth.p, ne, hipMemcpyDeviceToHost));
    HIPCHK(hipMemcpy(src.data(), T->src_of_entry.p, T->n * sizeof(int32_t), hipMemcpyDeviceToHost));
    int64_t pi = 0;
    for (int32_t q = 0; q < ne; q++)
    {
      const bool part = dep[q] < 0;
      if (kind)
        kind[q] = part ? 0 : 1;
      if (pidx)
        pidx[q] = part ? (int64_t)src[pi] : -1;
      if (part)
        pi++;
      if (mass)
        mass[q] = en[q].w;
      if (com)
      {
        com[3 * q] = en[q].x;
        com[3 * q + 1] = en[q].y;
        com[3 * q + 2] = en[q].z;
      }
      if (len)
      {
        // the cell's side: the root cube's, halved per level while not below the
        // resolution (the same rule and float rounding as the build)
        double L = 2.0 * cube[3];
        for (int l = 0; l < dep[q]; l++)
          L = (L >= T->P.res) ? L * 0.5 : L;
        const float lf = (float)L;
        len[q] = part ? 0.f : std::sqrt(lf * lf);
      }
      if (skip)
        skip[q] = part ? q + 1 : esk[q];
      if (depth)
        depth[q] = dep[q];
    }
  });
}

void hbtx_tree_free(hbtx_tree *T)
{
  delete T;
}

int hbtx_direct_potential(hbtx_ctx *c, const hbtx_params *prm, int64_t n_src, const float *src_pos,
                          const float *src_mass, int64_t n_tgt, const float *tgt_pos, const float *tgt_mass,
                          double *out_pot)
{
  if (!c || !prm)
    return HBTX_E_INVALID;
  hbtx_tree T;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    build_tree(c, *prm, n_src, src_pos, src_mass, &T, 0);
    tree_points(c, &T, n_tgt, tgt_pos, tgt_mass, out_pot);
  });
}

int hbtx_direct_potential_noalloc(hbtx_ctx *c, const hbtx_params *prm, int64_t n_src, const float *d_src_pos,
                                  const float *d_src_mass, int64_t n_tgt, const float *d_tgt_pos,
                                  const float *d_tgt_mass, double *d_out_pot)
{
  if (!c || !prm)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    REQUIRE(n_src >= 1 && n_tgt >= 0, HBTX_E_INVALID, "need at least one source and n_tgt >= 0");
    REQUIRE(n_src <= HBTX_NOALLOC_MAX_POINTS && n_tgt <= HBTX_NOALLOC_MAX_POINTS, HBTX_E_CAPACITY,
            "hbtx_direct_potential_noalloc takes at most HBTX_NOALLOC_MAX_POINTS (" +
              std::to_string(HBTX_NOALLOC_MAX_POINTS) + ") sources and targets; got " + std::to_string(n_src) +
              " and " + std::to_string(n_tgt));
    REQUIRE(d_src_pos && d_src_mass, HBTX_E_INVALID, "null source array");
    REQUIRE(n_tgt == 0 || (d_tgt_pos && d_out_pot), HBTX_E_INVALID, "null target or output array");
    REQUIRE(c->scratch, HBTX_E_HIP, "the scratch array has no device address");
    HIPCHK(hipSetDevice(c->device));
    hbtx_params p = *prm;
    p.direct_cutoff = (int64_t)1 << 40;
    p.max_sample_size = 0;
    const DevParams P = make_dev_params(p);
    hipStream_t st = c->stream;
    unsigned char *base = static_cast<unsigned char *>(c->scratch);
    uint2 *wt = reinterpret_cast<uint2 *>(base);
    float4 *tgt = reinterpret_cast<float4 *>(base + kScratchTgt);
    int32_t *counter = reinterpret_cast<int32_t *>(base + kScratchCtr);
    k_pack_records<<<blocks_for(n_src), 256, 0, st>>>(d_src_pos, d_src_mass, (int32_t)n_src, wt);
    if (n_tgt > 0)
    {
      k_pack_points<<<blocks_for(n_tgt), 256, 0, st>>>(d_tgt_pos, d_tgt_mass, (int32_t)n_tgt, tgt);
      k_fill_i32<<<1, 64, 0, st>>>(counter, 1, 0); // the previous call's count stays in the array
      PointArgs A;
      A.tgt = tgt;
      A.tgt_order = nullptr;
      A.ntgt = (int32_t)n_tgt;
      A.eb = 0;
      A.ee = (int32_t)(2 * n_src);
      A.treemode = 0;
      A.wt = wt;
      A.phi = nullptr;
      A.phi_d = d_out_pot;
      A.counter = counter;
      A.stats = nullptr;
      A.overflow = nullptr; // direct sums only
      A.P = P;
      const int grid = (int)std::max<int64_t>(1, std::min<int64_t>(n_tgt / HBTX_WAVE + 1, 32 * c->num_cu) /
                                                     HBTX_WALK_WG_WAVES);
      if (P.periodic)
        k_points_potential<true, false><<<grid, WALK_WG_THREADS, 0, st>>>(A);
      else
        k_points_potential<false, false><<<grid, WALK_WG_THREADS, 0, st>>>(A);
    }
    HIPCHK(hipGetLastError());
  });
}

int hbtx_batch_profiles(hbtx_ctx *c, const hbtx_params *prm, hbtx_sub_profile *out)
{
  if (!c || !prm || !out)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    Batch &B = c->B;
    REQUIRE(B.ran, HBTX_E_INVALID, "batch has not been run");
    B.prof_r.alloc(std::max<int64_t>(B.n, 1));
    B.prof_m.alloc(std::max<int64_t>(B.n, 1));
    B.dprof.alloc(std::max<int32_t>(B.nsub, 1));
    ProfArgs A;
    A.off = B.off.p;
    A.ford = B.ford.p;
    A.posm = B.posm.p;
    A.res = B.dres.p;
    A.seg_len = B.seg_len.p;
    A.rc = B.prof_r.p;
    A.mc = B.prof_m.p;
    run_profiles(c, *prm, A, B.n, B.nsub, B.sub_of.p, B.seg_begin.p, B.seg_len.p, B.keys.p, B.src.p, B.dprof.p);
    HIPCHK(hipMemcpyAsync(out, B.dprof.p, B.nsub * sizeof(hbtx_sub_profile), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
  });
}

int hbtx_profiles(hbtx_ctx *c, const hbtx_params *prm, int64_t n, const float *pos, const float *mass, int64_t nsub,
                  const int64_t *offsets, const int64_t *nbound, const float *mbound, hbtx_sub_profile *out)
{
  if (!c || !prm || !out || (n > 0 && (!pos || !mass)) || !offsets || !nbound || !mbound || nsub < 1)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    REQUIRE(n < (int64_t)0x20000000 && nsub < (int64_t)0x7fffffff, HBTX_E_INVALID, "too large");
    REQUIRE(offsets[0] == 0 && offsets[nsub] == n, HBTX_E_INVALID, "offsets must start at 0 and end at n");
    hipStream_t st = c->stream;
    std::vector<float4> hp(std::max<int64_t>(n, 1));
    std::vector<int32_t> hoff(nsub + 1), hsub(std::max<int64_t>(n, 1)), hord(std::max<int64_t>(n, 1));
    std::vector<hbtx_sub_result> hres(nsub);
    for (int64_t i = 0; i < n; i++)
      hp[i] = make_float4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], mass[i]);
    for (int64_t s = 0; s < nsub; s++)
    {
      const int64_t a = offsets[s], e = offsets[s + 1];
      REQUIRE(a <= e && e <= n && nbound[s] >= 0 && nbound[s] <= e - a, HBTX_E_INVALID,
              "bound sets must lie inside ascending subhalo ranges");
      hoff[s] = (int32_t)a;
      hbtx_sub_result &R = hres[s];
      std::memset(&R, 0, sizeof(R));
      R.nbound = nbound[s];
      R.mbound = mbound[s];
      if (nbound[s] > 0)
        for (int k = 0; k < 3; k++)
          R.mostbound_pos[k] = pos[3 * a + k];
      for (int64_t i = a; i < e; i++)
      {
        hsub[i] = (int32_t)s;
        hord[i] = (int32_t)(i - a);
      }
    }
    hoff[nsub] = (int32_t)n;
    DArr<float4> dp;
    DArr<int32_t> doff, dsub, dord, sb, sl, vals;
    DArr<uint64_t> keys;
    DArr<hbtx_sub_result> dres;
    DArr<float> rc, mc;
    DArr<hbtx_sub_profile> dout;
    dp.alloc(hp.size());
    doff.alloc(nsub + 1);
    dsub.alloc(hsub.size());
    dord.alloc(hord.size());
    sb.alloc(nsub + 1);
    sl.alloc(nsub + 1);
    vals.alloc(hsub.size());
    keys.alloc(hsub.size());
    dres.alloc(nsub);
    rc.alloc(hsub.size());
    mc.alloc(hsub.size());
    dout.alloc(nsub);
    HIPCHK(hipMemcpyAsync(dp.p, hp.data(), hp.size() * sizeof(float4), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(doff.p, hoff.data(), (nsub + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dsub.p, hsub.data(), hsub.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dord.p, hord.data(), hord.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dres.p, hres.data(), nsub * sizeof(hbtx_sub_result), hipMemcpyHostToDevice, st));
    ProfArgs A;
    A.off = doff.p;
    A.ford = dord.p;
    A.posm = dp.p;
    A.res = dres.p;
    A.seg_len = sl.p;
    A.rc = rc.p;
    A.mc = mc.p;
    run_profiles(c, *prm, A, n, (int32_t)nsub, dsub.p, sb.p, sl.p, keys.p, vals.p, dout.p);
    HIPCHK(hipMemcpyAsync(out, dout.p, nsub * sizeof(hbtx_sub_profile), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  });
}

// ---------------------------------------------------------------------------
// particle reassignment
// ---------------------------------------------------------------------------
int hbtx_reassign_targets(hbtx_ctx *c, const hbtx_params *prm, int64_t n, const float *pos, const uint8_t *is_tracer,
                          const int32_t *sub, int64_t ngroups, const int64_t *group_offsets, int32_t k, float rguess,
                          int32_t *out_target, double *elapsed_ms)
{
  if (!c || !prm || !group_offsets || ngroups < 1 || (n > 0 && (!pos || !is_tracer || !sub || !out_target)))
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    REQUIRE(n < (int64_t)0x20000000 && ngroups < (int64_t)0x20000000, HBTX_E_INVALID, "too large");
    REQUIRE(k >= 1 && k <= 32, HBTX_E_INVALID, "k (NumNeighboursForReassignment) must lie in 1..32");
    REQUIRE(group_offsets[0] == 0 && group_offsets[ngroups] == n, HBTX_E_INVALID,
            "group offsets must start at 0 and end at n");
    for (int64_t g = 0; g < ngroups; g++)
      REQUIRE(group_offsets[g] <= group_offsets[g + 1], HBTX_E_INVALID, "group offsets must ascend");
    // GeoTree_t::NearestNeighbours(centre, k, SofteningHalo * 0.1) (src/subhalo_reassign_gas.cpp:144)
    const float r0 = (rguess > 0.f) ? rguess : (float)(prm->softening * 0.1);
    REQUIRE(r0 > 0.f && std::isfinite(r0), HBTX_E_INVALID, "the first search radius must be positive");
    std::vector<float> rad(RS_RADII);
    rad[0] = r0;
    for (int j = 1; j < RS_RADII; j++)
      rad[j] = (float)((double)rad[j - 1] * 1.26); // rguess *= 1.26 (src/geometric_tree.cpp:127)
    REQUIRE(rad[RS_RADII - 1] >= 1e30f, HBTX_E_INVALID, "the first search radius is too small");
    if (elapsed_ms)
      *elapsed_ms = 0.0;
    if (n == 0)
      return HBTX_OK;
    auto &R = c->rs;
    const int32_t G = (int32_t)ngroups;
    std::vector<int32_t> hoff(G + 1);
    for (int32_t g = 0; g <= G; g++)
      hoff[g] = (int32_t)group_offsets[g];
    R.pos.alloc(3 * n);
    R.radii.alloc(RS_RADII);
    R.tracer.alloc(n);
    R.sub.alloc(n);
    R.goff.alloc(G + 1);
    R.gid.alloc(n);
    R.tflag.alloc(n + 1);
    R.cnt.alloc((int64_t)RS_LEVELS * G);
    R.loff.alloc((int64_t)RS_LEVELS * (G + 1));
    R.qtiles.alloc(G + 1);
    R.tval.alloc(n);
    R.qval.alloc(n);
    R.tkey.alloc(n);
    R.qkey.alloc(n);
    R.seg.alloc(4 * (int64_t)G);
    R.out.alloc(n);
    R.tot.alloc(RS_LEVELS + 2);
    R.grp.alloc(G);
    R.bbox.alloc(7 * (int64_t)G);
    R.trec.alloc(n);
    R.lbase.alloc(RS_LEVELS);
    hipStream_t st = c->stream;
    HIPCHK(hipMemcpyAsync(R.pos.p, pos, 3 * n * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(R.tracer.p, is_tracer, n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(R.sub.p, sub, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(R.goff.p, hoff.data(), (G + 1) * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(R.radii.p, rad.data(), RS_RADII * sizeof(float), hipMemcpyHostToDevice, st));
    RsArgs A;
    A.n = (int32_t)n;
    A.ngroups = G;
    A.pos = R.pos.p;
    A.tracer = R.tracer.p;
    A.sub = R.sub.p;
    A.goff = R.goff.p;
    A.box = prm->box_size;
    A.half = prm->box_size / 2;
    A.periodic = prm->periodic ? 1 : 0;
    A.k = k;
    A.gid = R.gid.p;
    A.tflag = R.tflag.p;
    A.grp = R.grp.p;
    A.bbox = R.bbox.p;
    A.cnt = R.cnt.p;
    A.loff = R.loff.p;
    A.qtiles_n = R.qtiles.p;
    A.tkey = R.tkey.p;
    A.qkey = R.qkey.p;
    A.tval = R.tval.p;
    A.qval = R.qval.p;
    A.seg = R.seg.p;
    A.trec = R.trec.p;
    A.nodes = nullptr;
    A.lbase = R.lbase.p;
    A.out = R.out.p;
    HIPCHK(hipEventRecord(c->ev0, st));
    k_rs_flags<<<blocks_for(n), 256, 0, st>>>(A);
    scan32(c, R.tflag.p, R.tflag.p, n);
    k_rs_groups<<<blocks_for(G, 256, 1 << 20), 256, 0, st>>>(A);
    k_rs_scatter<<<blocks_for(n, RS_CHUNK), RS_THREADS, 0, st>>>(A);
    for (int l = 0; l < RS_LEVELS; l++)
      scan32(c, R.cnt.p + (int64_t)l * G, R.loff.p + (int64_t)l * (G + 1), G);
    scan32(c, R.qtiles.p, R.qtiles.p, G);
    k_rs_totals<<<1, 64, 0, st>>>(A, R.tot.p);
    int32_t tot[RS_LEVELS + 2];
    HIPCHK(hipMemcpyAsync(tot, R.tot.p, sizeof(tot), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    const int32_t NT = tot[RS_LEVELS], ntiles = tot[RS_LEVELS + 1];
    int64_t lb[RS_LEVELS], nodes = 0;
    for (int l = 0; l < RS_LEVELS; l++)
    {
      lb[l] = nodes;
      nodes += tot[l];
    }
    R.nodes.alloc(2 * std::max<int64_t>(nodes, 1));
    R.tiles.alloc(std::max<int32_t>(ntiles, 1));
    A.nodes = R.nodes.p;
    HIPCHK(hipMemcpyAsync(R.lbase.p, lb, sizeof(lb), hipMemcpyHostToDevice, st));
    k_rs_keys<<<blocks_for(n), 256, 0, st>>>(A, NT);
    k_rs_group_finish<<<blocks_for(G, 256, 1 << 20), 256, 0, st>>>(A);
    seg_sort(c, R.tkey.p, R.tval.p, G, R.seg.p, R.seg.p + G, std::max<int32_t>(NT, 1), 3 * RS_KEY_BITS);
    seg_sort(c, R.qkey.p, R.qval.p, G, R.seg.p + 2 * (int64_t)G, R.seg.p + 3 * (int64_t)G,
             std::max<int64_t>(n - NT, 1), 3 * RS_KEY_BITS);
    if (NT > 0)
      k_rs_records<<<blocks_for(NT), 256, 0, st>>>(A, NT);
    for (int l = 0; l < RS_LEVELS; l++)
      if (tot[l] > 0)
        k_rs_level<<<blocks_for(tot[l]), 256, 0, st>>>(A, l, tot[l]);
    if (ntiles > 0)
    {
      k_rs_tiles<<<blocks_for(ntiles), 256, 0, st>>>(A, R.tiles.p, ntiles);
      const int wpb = RS_THREADS / HBTX_WAVE;
      const int grid = (int)std::min<int64_t>((ntiles + wpb - 1) / wpb, 1 << 20);
      if (k <= 8)
        k_rs_knn<8><<<grid, RS_THREADS, 0, st>>>(A, R.tiles.p, ntiles, R.radii.p);
      else if (k <= 10) // the reference's default NumNeighboursForReassignment
        k_rs_knn<10><<<grid, RS_THREADS, 0, st>>>(A, R.tiles.p, ntiles, R.radii.p);
      else if (k <= 16)
        k_rs_knn<16><<<grid, RS_THREADS, 0, st>>>(A, R.tiles.p, ntiles, R.radii.p);
      else
        k_rs_knn<32><<<grid, RS_THREADS, 0, st>>>(A, R.tiles.p, ntiles, R.radii.p);
    }
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventRecord(c->ev1, st));
    HIPCHK(hipMemcpyAsync(out_target, R.out.p, n * sizeof(int32_t), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (elapsed_ms)
    {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
      *elapsed_ms = ms;
    }
  });
}

int hbtx_segmented_sort(hbtx_ctx *c, int64_t n, uint64_t *keys, int32_t *vals, int64_t nseg,
                        const int64_t *seg_offsets)
{
  if (!c || (n > 0 && (!keys || !vals)) || !seg_offsets)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    REQUIRE(n < (int64_t)0x40000000 && nseg < (int64_t)0x7fffffff, HBTX_E_INVALID, "too large");
    hipStream_t st = c->stream;
    DArr<uint64_t> dk;
    DArr<int32_t> dv, sb, sl;
    dk.alloc(n);
    dv.alloc(n);
    sb.alloc(nseg);
    sl.alloc(nseg);
    std::vector<int32_t> hb(nseg), hl(nseg);
    for (int64_t s = 0; s < nseg; s++)
    {
      REQUIRE(seg_offsets[s + 1] >= seg_offsets[s], HBTX_E_INVALID, "bad segments");
      hb[s] = (int32_t)seg_offsets[s];
      hl[s] = (int32_t)(seg_offsets[s + 1] - seg_offsets[s]);
    }
    HIPCHK(hipMemcpyAsync(dk.p, keys, n * 8, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dv.p, vals, n * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(sb.p, hb.data(), nseg * 4, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(sl.p, hl.data(), nseg * 4, hipMemcpyHostToDevice, st));
    // digit passes only over the bits the keys use
    uint64_t any = 0;
    for (int64_t i = 0; i < n; i++)
      any |= keys[i];
    int bits = 1;
    while (bits < 64 && (any >> bits) != 0)
      bits++;
    seg_sort(c, dk.p, dv.p, (int32_t)nseg, sb.p, sl.p, n, bits);
    HIPCHK(hipMemcpyAsync(keys, dk.p, n * 8, hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(vals, dv.p, n * 4, hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  });
}

} // extern "C"

namespace
{
uint64_t splitmix(uint64_t &x)
{
  uint64_t z = (x += 0x9E3779B97F4A7C15ull);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
  return z ^ (z >> 31);
}
double urand(uint64_t &x)
{
  return (double)(splitmix(x) >> 11) * (1.0 / 9007199254740992.0);
}

void mlp_alloc(hbtx_mlp *M)
{
  M->master.alloc(MLP_NPAR);
  M->m.alloc(MLP_NPAR);
  M->v.alloc(MLP_NPAR);
  M->wf.alloc(MLP_NPAR);
  M->grad.alloc(MLP_NPAR);
  M->wb.alloc(MLP_NPAR);
  M->loss.alloc(1);
}

void mlp_upload(hbtx_ctx *c, hbtx_mlp *M, const float *master)
{
  hipStream_t st = c->stream;
  HIPCHK(hipMemcpyAsync(M->master.p, master, MLP_NPAR * sizeof(float), hipMemcpyHostToDevice, st));
  HIPCHK(hipMemsetAsync(M->m.p, 0, MLP_NPAR * sizeof(float), st));
  HIPCHK(hipMemsetAsync(M->v.p, 0, MLP_NPAR * sizeof(float), st));
  k_mlp_round<<<(MLP_NPAR + 255) / 256, 256, 0, st>>>(M->master.p, M->wf.p, M->wb.p);
  HIPCHK(hipGetLastError());
  HIPCHK(hipStreamSynchronize(st));
  M->step = 0;
}

// forward+backward of nb samples (rows idx[0..nb) of dX, or rows 0..nb) -> M->grad, M->loss
void mlp_grad_device(hbtx_ctx *c, hbtx_mlp *M, const float *dX, const uint8_t *dy, const int32_t *idx, int64_t nb)
{
  hipStream_t st = c->stream;
  const int nblk = (int)((nb + MLP_BLOCK - 1) / MLP_BLOCK);
  M->partial.alloc((int64_t)nblk * MLP_NPAR);
  M->lossp.alloc(nblk);
  k_mlp_grad<<<nblk, MLP_BLOCK, 0, st>>>(M->wf.p, dX, dy, idx, nb, 1.f / (float)nb, M->partial.p, M->lossp.p);
  k_mlp_reduce<<<(MLP_NPAR + 255) / 256, 256, 0, st>>>(M->partial.p, nblk, M->grad.p, M->lossp.p, M->loss.p);
  HIPCHK(hipGetLastError());
}
} // namespace

extern "C" {

int hbtx_mlp_create(hbtx_ctx *c, uint64_t seed, hbtx_mlp **out)
{
  if (!c || !out)
    return HBTX_E_INVALID;
  *out = nullptr;
  hbtx_mlp *M = new (std::nothrow) hbtx_mlp();
  if (!M)
    return HBTX_E_ALLOC;
  try
  {
    HIPCHK(hipSetDevice(c->device));
    mlp_alloc(M);
    std::vector<float> w(MLP_NPAR, 0.f);
    uint64_t x = seed;
    const double l1 = std::sqrt(6.0 / MLP_IN), l2 = std::sqrt(6.0 / MLP_H), l3 = std::sqrt(3.0 / MLP_H);
    for (int i = 0; i < MLP_H * MLP_IN; i++)
      w[MLP_OFF_W1 + i] = (float)((2.0 * urand(x) - 1.0) * l1);
    for (int i = 0; i < MLP_H * MLP_H; i++)
      w[MLP_OFF_W2 + i] = (float)((2.0 * urand(x) - 1.0) * l2);
    for (int i = 0; i < MLP_H; i++)
      w[MLP_OFF_W3 + i] = (float)((2.0 * urand(x) - 1.0) * l3);
    mlp_upload(c, M, w.data());
  }
  catch (const Error &e)
  {
    delete M;
    return fail(c, e.code, e.what());
  }
  *out = M;
  return HBTX_OK;
}

void hbtx_mlp_free(hbtx_mlp *M)
{
  delete M;
}

int hbtx_mlp_get_weights(hbtx_ctx *c, hbtx_mlp *M, float *master, uint16_t *bf16)
{
  if (!c || !M)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    HIPCHK(hipStreamSynchronize(c->stream));
    if (master)
      HIPCHK(hipMemcpy(master, M->master.p, MLP_NPAR * sizeof(float), hipMemcpyDeviceToHost));
    if (bf16)
      HIPCHK(hipMemcpy(bf16, M->wb.p, MLP_NPAR * sizeof(uint16_t), hipMemcpyDeviceToHost));
  });
}

int hbtx_mlp_set_weights(hbtx_ctx *c, hbtx_mlp *M, const float *master)
{
  if (!c || !M || !master)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    mlp_upload(c, M, master);
  });
}

int hbtx_mlp_gradient(hbtx_ctx *c, hbtx_mlp *M, int64_t n, const float *X, const uint8_t *y, float *grad,
                      double *loss)
{
  if (!c || !M || n <= 0 || !X || !y)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    hipStream_t st = c->stream;
    DArr<float> dX;
    DArr<uint8_t> dy;
    dX.alloc(n * MLP_IN);
    dy.alloc(n);
    HIPCHK(hipMemcpyAsync(dX.p, X, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dy.p, y, n, hipMemcpyHostToDevice, st));
    mlp_grad_device(c, M, dX.p, dy.p, nullptr, n);
    double l = 0.0;
    if (grad)
      HIPCHK(hipMemcpyAsync(grad, M->grad.p, MLP_NPAR * sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHK(hipMemcpyAsync(&l, M->loss.p, sizeof(double), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
    if (loss)
      *loss = l / (double)n;
  });
}

int hbtx_mlp_train(hbtx_ctx *c, hbtx_mlp *M, int64_t n, const float *X, const uint8_t *y, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, double *mean_loss)
{
  if (!c || !M || n <= 0 || !X || !y || epochs <= 0 || batch <= 0)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    REQUIRE(n < (int64_t)0x7fffffff, HBTX_E_INVALID, "too many training samples");
    hipStream_t st = c->stream;
    DArr<float> dX;
    DArr<uint8_t> dy;
    DArr<int32_t> didx;
    dX.alloc(n * MLP_IN);
    dy.alloc(n);
    didx.alloc(n);
    HIPCHK(hipMemcpyAsync(dX.p, X, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(dy.p, y, n, hipMemcpyHostToDevice, st));
    std::vector<int32_t> perm(n);
    uint64_t rs = seed;
    const float b1 = 0.9f, b2 = 0.999f;
    double last = 0.0;
    std::vector<double> lbatch;
    for (int ep = 0; ep < epochs; ep++)
    {
      std::iota(perm.begin(), perm.end(), 0);
      for (int64_t i = n - 1; i > 0; i--)
      {
        const int64_t j = (int64_t)(splitmix(rs) % (uint64_t)(i + 1));
        std::swap(perm[i], perm[j]);
      }
      HIPCHK(hipMemcpyAsync(didx.p, perm.data(), n * sizeof(int32_t), hipMemcpyHostToDevice, st));
      DArr<double> losses;
      const int64_t nsteps = (n + batch - 1) / batch;
      losses.alloc(nsteps);
      for (int64_t k = 0; k < nsteps; k++)
      {
        const int64_t b0 = k * batch, nb = std::min<int64_t>(batch, n - b0);
        mlp_grad_device(c, M, dX.p, dy.p, didx.p + b0, nb);
        M->step++;
        const float c1 = 1.f - std::pow(b1, (float)M->step), c2 = 1.f - std::pow(b2, (float)M->step);
        k_mlp_adam<<<(MLP_NPAR + 255) / 256, 256, 0, st>>>(M->master.p, M->m.p, M->v.p, M->grad.p, M->wf.p,
                                                            M->wb.p, lr, b1, b2, 1e-8f, c1, c2);
        HIPCHK(hipMemcpyAsync(losses.p + k, M->loss.p, sizeof(double), hipMemcpyDeviceToDevice, st));
      }
      lbatch.resize(nsteps);
      HIPCHK(hipMemcpyAsync(lbatch.data(), losses.p, nsteps * sizeof(double), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
      double tot = 0.0;
      for (double v : lbatch)
        tot += v;
      last = tot / (double)n;
    }
    if (mean_loss)
      *mean_loss = last;
  });
}

int hbtx_mlp_predict(hbtx_ctx *c, hbtx_mlp *M, int64_t n, const float *X, float *logit)
{
  if (!c || !M || n < 0 || (n > 0 && (!X || !logit)))
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    if (n == 0)
      return HBTX_OK;
    hipStream_t st = c->stream;
    DArr<float> dX, dl;
    dX.alloc(n * MLP_IN);
    dl.alloc(n);
    HIPCHK(hipMemcpyAsync(dX.p, X, n * MLP_IN * sizeof(float), hipMemcpyHostToDevice, st));
    k_mlp_forward<<<mlp_forward_blocks(n), 256, 0, st>>>(M->wf.p, M->wb.p, dX.p, n, dl.p, nullptr, 0);
    HIPCHK(hipGetLastError());
    HIPCHK(hipMemcpyAsync(logit, dl.p, n * sizeof(float), hipMemcpyDeviceToHost, st));
    HIPCHK(hipStreamSynchronize(st));
  });
}

int hbtx_batch_set_previous(hbtx_ctx *c, const int32_t *prev_rank, const uint8_t *prev_bound,
                            const float *prev_energy, const float *prev_com_pos, const float *prev_com_vel,
                            const float *prev_mbound, const float *prev_rscale)
{
  if (!c || !prev_rank || !prev_bound || !prev_energy || !prev_com_pos || !prev_com_vel || !prev_mbound ||
      !prev_rscale)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    Batch &B = c->B;
    REQUIRE(B.n > 0, HBTX_E_INVALID, "no batch loaded");
    hipStream_t st = c->stream;
    const int64_t n = B.n, ns = B.nsub;
    B.prev_rank.alloc(n);
    B.prev_bound.alloc(n);
    B.prev_energy.alloc(n);
    B.prev_com.alloc(6 * ns);
    B.prev_mb.alloc(ns);
    B.prev_rs.alloc(ns);
    std::vector<float> com(6 * ns);
    for (int64_t s = 0; s < ns; s++)
      for (int k = 0; k < 3; k++)
      {
        com[6 * s + k] = prev_com_pos[3 * s + k];
        com[6 * s + 3 + k] = prev_com_vel[3 * s + k];
      }
    HIPCHK(hipMemcpyAsync(B.prev_rank.p, prev_rank, n * sizeof(int32_t), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.prev_bound.p, prev_bound, n, hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.prev_energy.p, prev_energy, n * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.prev_com.p, com.data(), 6 * ns * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.prev_mb.p, prev_mbound, ns * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHK(hipMemcpyAsync(B.prev_rs.p, prev_rscale, ns * sizeof(float), hipMemcpyHostToDevice, st));
    HIPCHK(hipStreamSynchronize(st));
    B.has_prev = true;
    B.feat_full = false;  // the features depend on the previous outputs
  });
}

int hbtx_batch_features(hbtx_ctx *c, const hbtx_params *prm, float *out_features)
{
  if (!c || !prm)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    Batch &B = c->B;
    REQUIRE(B.n > 0 && B.has_prev, HBTX_E_INVALID, "load a batch and its previous-snapshot outputs first");
    const DevParams P = make_dev_params(*prm);
    hipStream_t st = c->stream;
    const int64_t n = B.n;
    B.feat.alloc(n * MLP_IN);
    View V = make_view(B, 0, n, 0, B.nsub, true);
    PrevArrays A{B.prev_rank.p, B.prev_bound.p, B.prev_energy.p, B.prev_com.p, B.prev_mb.p, B.prev_rs.p};
    k_features<<<blocks_for(n), 256, 0, st>>>(V.S, V.Q, P, A, n, B.feat.p, nullptr, 0);
    HIPCHK(hipGetLastError());
    B.feat_full = true;
    if (out_features)
    {
      HIPCHK(hipMemcpyAsync(out_features, B.feat.p, n * MLP_IN * sizeof(float), hipMemcpyDeviceToHost, st));
      HIPCHK(hipStreamSynchronize(st));
    }
  });
}

int hbtx_batch_get_proposal(hbtx_ctx *c, uint8_t *proposal)
{
  if (!c || !proposal)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    Batch &B = c->B;
    REQUIRE(B.n > 0 && B.has_proposal, HBTX_E_INVALID, "no proposal on the device");
    HIPCHK(hipStreamSynchronize(c->stream));
    HIPCHK(hipMemcpy(proposal, B.proposal.p, B.n, hipMemcpyDeviceToHost));
  });
}

int hbtx_batch_propose(hbtx_ctx *c, hbtx_mlp *M, float threshold, int64_t min_particles, int64_t *nproposed,
                       double *elapsed_ms)
{
  if (!c || !M || !(threshold > 0.f && threshold < 1.f) || min_particles < 0)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    Batch &B = c->B;
    REQUIRE(B.n > 0 && B.feat.p && B.feat_full, HBTX_E_INVALID, "compute the batch features first");
    hipStream_t st = c->stream;
    const int64_t n = B.n;
    B.logit.alloc(n);
    HIPCHK(hipEventRecord(c->wev0, st));
    k_mlp_forward<<<mlp_forward_blocks(n), 256, 0, st>>>(M->wf.p, M->wb.p, B.feat.p, n, B.logit.p, nullptr, 0);
    k_proposal_from_logit<<<blocks_for(n), 256, 0, st>>>(B.logit.p, n, std::log(threshold / (1.f - threshold)),
                                                          B.sub_of.p, B.off.p, min_particles, B.proposal.p);
    HIPCHK(hipEventRecord(c->wev1, st));
    HIPCHK(hipGetLastError());
    HIPCHK(hipEventSynchronize(c->wev1));
    float ms = 0.f;
    HIPCHK(hipEventElapsedTime(&ms, c->wev0, c->wev1));
    if (elapsed_ms)
      *elapsed_ms = ms;
    B.has_proposal = true;
    if (nproposed)
    {
      std::vector<uint8_t> h(n);
      HIPCHK(hipMemcpy(h.data(), B.proposal.p, n, hipMemcpyDeviceToHost));
      int64_t cnt = 0;
      for (int64_t i = 0; i < n; i++)
        cnt += h[i];
      *nproposed = cnt;
    }
  });
}

int hbtx_batch_warm_proposal(hbtx_ctx *c, const hbtx_params *prm, hbtx_mlp *M, float threshold,
                             int64_t min_particles, int64_t *nproposed, double *elapsed_ms)
{
  if (!c || !prm || !M || !(threshold > 0.f && threshold < 1.f) || min_particles < 0)
    return HBTX_E_INVALID;
  API_GUARD(c, {
    HIPCHK(hipSetDevice(c->device));
    Batch &B = c->B;
    REQUIRE(B.n > 0 && B.has_prev, HBTX_E_INVALID, "load a batch and its previous-snapshot outputs first");
    const DevParams P = make_dev_params(*prm);
    hipStream_t st = c->stream;
    const int64_t n = B.n;
    if (B.egroups_min != min_particles)
    {
      // the 32-particle groups that hold a particle of a subhalo of at least
      // min_particles particles (subhaloes are contiguous, so the list ascends)
      std::vector<int32_t> g;
      int64_t last = -1;
      for (int32_t s = 0; s < B.nsub; s++)
      {
        const int64_t a = B.h_off[s], e = B.h_off[s + 1];
        if (e - a < std::max<int64_t>(min_particles, 1))
          continue;
        for (int64_t q = std::max<int64_t>(a / 32, last + 1); q <= (e - 1) / 32; q++)
          g.push_back((int32_t)q);
        last = (e - 1) / 32;
      }
      B.egroups.alloc(std::max<size_t>(g.size(), 1));
      if (!g.empty())
        HIPCHK(hipMemcpyAsync(B.egroups.p, g.data(), g.size() * sizeof(int32_t), hipMemcpyHostToDevice, st));
      HIPCHK(hipStreamSynchronize(st));
      B.egroups_n = (int64_t)g.size();
      B.egroups_min = min_particles;
    }
    B.feat.alloc(n * MLP_IN);
    B.logit.alloc(n);
    B.feat_full = false;  // only the listed groups' features are written
    View V = make_view(B, 0, n, 0, B.nsub, true);
    PrevArrays A{B.prev_rank.p, B.prev_bound.p, B.prev_energy.p, B.prev_com.p, B.prev_mb.p, B.prev_rs.p};
    HIPCHK(hipEventRecord(c->wev0, st));
    if (B.egroups_n > 0)
    {
      k_features<<<blocks_for(32 * B.egroups_n), 256, 0, st>>>(V.S, V.Q, P, A, n, B.feat.p, B.egroups.p,
                                                               B.egroups_n);
      k_mlp_forward<<<mlp_forward_blocks(32 * B.egroups_n), 256, 0, st>>>(M->wf.p, M->wb.p, B.feat.p, n,
                                                                          B.logit.p, B.egroups.p, B.egroups_n);
    }
    k_proposal_from_logit<<<blocks_for(n), 256, 0, st>>>(B.logit.p, n, std::log(threshold / (1.f - threshold)),
                                                          B.sub_of.p, B.off.p, std::max<int64_t>(min_particles, 1),
                                                          B.proposal.p);
    HIPCHK(hipEventRecord(c->wev1, st));
    HIPCHK(hipGetLastError());
    B.has_proposal = true;
    if (elapsed_ms || nproposed)
      HIPCHK(hipEventSynchronize(c->wev1));
    if (elapsed_ms)
    {
      float ms = 0.f;
      HIPCHK(hipEventElapsedTime(&ms, c->wev0, c->wev1));
      *elapsed_ms = ms;
    }
    if (nproposed)
    {
      std::vector<uint8_t> h(n);
      HIPCHK(hipMemcpy(h.data(), B.proposal.p, n, hipMemcpyDeviceToHost));
      int64_t cnt = 0;
      for (int64_t i = 0; i < n; i++)
        cnt += h[i];
      *nproposed = cnt;
    }
  });
}

} // extern "C"

extern "C" {

int hbtx_get_tuning(hbtx_ctx *c, int64_t *direct_cutoff, int32_t *walk_chunks)
{
  if (!c)
    return HBTX_E_INVALID;
  if (direct_cutoff)
    *direct_cutoff = c->direct_cutoff;
  if (walk_chunks)
    *walk_chunks = c->walk_chunks;
  return HBTX_OK;
}

int hbtx_set_tuning(hbtx_ctx *c, int64_t direct_cutoff, int32_t walk_chunks)
{
  if (!c || direct_cutoff < 0 || walk_chunks < 1 || walk_chunks > WALK_MAX_CHUNKS)
    return HBTX_E_INVALID;
  c->direct_cutoff = direct_cutoff;
  c->walk_chunks = walk_chunks;
  c->tune_source = "hbtx_set_tuning";
  return HBTX_OK;
}

} // extern "C"
