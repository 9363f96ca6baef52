/* hbtx — MI355X-native gravitational unbinding: the C-ABI drop-in boundary.
 *
 * Every entry point takes plain pointers and sizes; the library never aborts and
 * reports failures as negative status codes with a message in hbtx_last_error().
 * Host arrays are owned by the caller; a context owns device buffers, its HIP
 * stream and its cached graphs.  One context per host thread.
 *
 * Reference interfaces replaced (paths relative to the reference repository):
 *   hbtx_unbind_batch        Subhalo_t::Unbind(const Snapshot_t&)        src/subhalo.h:152, src/subhalo_unbind.cpp:463-808
 *                            (per subhalo, HBTX_PLAIN) and the batch loop of
 *                            SubhaloSnapshot_t::RefineParticles           src/subhalo_unbind.cpp:855-919 (default)
 *   hbtx_tree_build          GravityTree_t::Reserve + OctTree_t::Build     src/oct_tree.tpp:11-156, src/gravity_tree.cpp:50-80
 *   hbtx_tree_potential      GravityTree_t::EvaluatePotential              src/gravity_tree.cpp:82-166
 *   hbtx_tree_binding_energy GravityTree_t::BindingEnergy                  src/gravity_tree.cpp:168-178
 *   hbtx_tree_free           OctTree_t::Clear                              src/oct_tree.tpp:158-164
 *   hbtx_params              HBTConfig fields used by unbinding            src/config_parser.h:98-172, :200-212, :641-649
 *   hbtx_reassign_targets    SubhaloSnapshot_t::ReassignParticles (the     src/subhalo_reassign_gas.cpp:65-197,
 *                            choice of target) with GeoTree_t's kNN        src/geometric_tree.cpp:108-210
 */
#ifndef HBTX_H
#define HBTX_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define HBTX_OK 0
#define HBTX_E_INVALID (-1)
#define HBTX_E_ALLOC (-2)
#define HBTX_E_HIP (-3)
#define HBTX_E_CAPACITY (-4)
#define HBTX_E_NOT_IMPLEMENTED (-5)

/* Largest subhalo a batch accepts: the walk addresses one subhalo's octree
 * records (16 bytes a particle, 24 a cell) by int32 byte offsets, so a tree must
 * stay under 2 GiB.  Larger subhaloes are rejected with HBTX_E_CAPACITY before
 * upload; a tree whose cells push it over the span is reported as HBTX_E_CAPACITY
 * by the round that builds it. */
#define HBTX_MAX_SUBHALO_PARTICLES ((int64_t)(1 << 27) - 64)

/* Flags for hbtx_unbind_batch / hbtx_batch_run */
#define HBTX_BATCHED 0u          /* all subhaloes advance together, device-side control (default) */
#define HBTX_PLAIN 1u            /* one subhalo at a time, host-driven iteration (Subhalo_t::Unbind loop) */
#define HBTX_WARMSTART 2u        /* start from the proposed bound set (hbtx_batch_set_proposal) */
#define HBTX_SAVE_POTENTIAL 4u   /* fill out_potential (ParticlePotentialEnergies) */
#define HBTX_NO_GRAPH 8u         /* batched rounds after the first replay two captured HIP graphs (plan, sources,
                                    octrees, work items | partition, decision, frames) around the persistent
                                    walk launch; this flag launches every kernel directly (same results) */
#define HBTX_COUNT 16u           /* count walk work (entries visited, interactions) for hbtx_batch_stats;
                                    results are identical, timing is not representative */

typedef struct hbtx_ctx hbtx_ctx;
typedef struct hbtx_tree hbtx_tree;

/* Unbinding parameters: the same names, meaning and defaults as the reference's
 * HBTConfig (src/config_parser.h:98-172), plus the epoch and two tuning knobs. */
typedef struct
{
  double G;                       /* PhysicalConst::G = 43.0071 (Msun/h 1e10, Mpc/h, km/s) */
  float scale_factor;             /* Cosmology.ScaleFactor */
  float hz;                       /* Cosmology.Hz, internal units */
  float softening;                /* SofteningHalo (comoving) */
  float max_phys_softening;       /* MaxPhysicalSofteningHalo, -1 = unset */
  float theta;                    /* TreeNodeOpenAngle */
  float bound_mass_precision;     /* BoundMassPrecision */
  int64_t max_sample_size;        /* MaxSampleSizeOfPotentialEstimate, 0 disables subsampling */
  int32_t refine_most_bound;      /* RefineMostBoundParticle */
  float bound_fraction_center_refinement; /* BoundFractionCenterRefinement */
  int64_t min_num_part;           /* MinNumPartOfSub */
  int64_t min_num_tracer;         /* MinNumTracerPartOfSub */
  int32_t tracer_mask;            /* TracerParticleBitMask */
  int32_t donot_subsample_mask;   /* DoNotSubsampleParticleBitMask */
  int32_t periodic;               /* PeriodicBoundaryOn */
  float box_size;                 /* BoxSize */
  int32_t thermal;                /* UNBIND_WITH_THERMAL_ENERGY semantics */
  int32_t dm_only;                /* DM_ONLY semantics (no types; every particle a tracer) */
  int32_t upscale_per_type;       /* PotentialEstimateUpscaleMassesPerType */
  float sub_core_size_factor;     /* SubCoreSizeFactor */
  int64_t sub_core_size_min;      /* SubCoreSizeMin */
  int64_t direct_cutoff;          /* sources <= cutoff use the exact O(N^2) sum; 0 = tree everywhere,
                                     -1 = autotuned value */
  int32_t reserved[2];
} hbtx_params;

/* Per-subhalo outputs of Subhalo_t::Unbind (src/subhalo.h:26-216 fields). */
typedef struct
{
  int64_t nbound;
  float mbound;
  int64_t nbound_type[6];
  float mbound_type[6];
  float com_pos[3];      /* ComovingAveragePosition */
  float com_vel[3];      /* PhysicalAverageVelocity */
  float mostbound_pos[3];/* ComovingMostBoundPosition */
  float mostbound_vel[3];/* PhysicalMostBoundVelocity */
  float spec_pot;        /* SpecificSelfPotentialEnergy */
  float spec_kin;        /* SpecificSelfKineticEnergy */
  float spec_am[3];      /* SpecificAngularMomentum */
  int64_t tracer_index;  /* TracerIndex */
  int64_t mostbound_id;  /* MostBoundParticleId (-1 when disrupted) */
  float core_pos[3], core_vel[3];       /* CoreComovingPosition, CorePhysicalVelocity */
  float core_sigma_r, core_sigma_v;     /* CoreComovingSigmaR, CorePhysicalSigmaV */
  int32_t iterations;    /* NumberUnbindingIterations */
  int32_t disrupted;     /* 1 if Nbound fell below MinNumPartOfSub / MinNumTracerPartOfSub */
} hbtx_sub_result;

void hbtx_default_params(hbtx_params *p);

int hbtx_create(hbtx_ctx **ctx, int device);
void hbtx_destroy(hbtx_ctx *ctx);
const char *hbtx_last_error(const hbtx_ctx *ctx);
int hbtx_device_sync(hbtx_ctx *ctx);

/* One-call unbinding of nsub subhaloes stored back to back (CSR offsets[nsub+1]).
 * Inputs are structure-of-arrays host buffers: pos/vel float[n][3], mass float[n],
 * u float[n] or NULL (specific internal energy), type int32[n] or NULL (all DM),
 * id int64[n] or NULL, ref_pos/ref_vel float[nsub][3] (the pre-set frame,
 * src/subhalo_unbind.cpp:464), mb_pos/mb_vel float[nsub][3] or NULL (most-bound
 * position/velocity used when a subhalo disrupts, :660-661).
 * With max_sample_size > 0 each subhalo must arrive partitioned and shuffled as
 * PrepareParticlesForSubsampling leaves it (:430-443).
 * Outputs: out_order int64[n] (local input index of the particle in each final
 * slot: bound sorted by energy, then unbound batches, latest first), out_energy
 * float[n] (energy of the particle in each slot), out_potential float[n] or NULL
 * (ParticlePotentialEnergies of the bound part, needs HBTX_SAVE_POTENTIAL),
 * out_res[nsub]. */
int hbtx_unbind_batch(hbtx_ctx *ctx, const hbtx_params *prm, int64_t n, const float *pos, const float *vel,
                      const float *mass, const float *u, const int32_t *type, const int64_t *id, int64_t nsub,
                      const int64_t *offsets, const float *ref_pos, const float *ref_vel, const float *mb_pos,
                      const float *mb_vel, uint32_t flags, int64_t *out_order, float *out_energy,
                      float *out_potential, hbtx_sub_result *out_res);

/* Host-only check of a batch layout (no device needed), the checks
 * hbtx_batch_load and hbtx_unbind_batch make before touching the device:
 * offsets[0] = 0, non-decreasing, offsets[nsub] = n, n < 2^29 and every subhalo
 * at most HBTX_MAX_SUBHALO_PARTICLES.  Returns HBTX_OK, HBTX_E_INVALID or
 * HBTX_E_CAPACITY; msg (may be NULL) receives the reason.  No reference
 * counterpart: the reference's Unbind has no size limit. */
int hbtx_check_layout(int64_t n, int64_t nsub, const int64_t *offsets, char *msg, int32_t msg_len);

/* Device-resident batch: upload once, run (timed on the context stream with HIP
 * events, inputs already in HBM), fetch.  Same array meanings as above. */
int hbtx_batch_load(hbtx_ctx *ctx, int64_t n, const float *pos, const float *vel, const float *mass,
                    const float *u, const int32_t *type, const int64_t *id, int64_t nsub, const int64_t *offsets,
                    const float *ref_pos, const float *ref_vel, const float *mb_pos, const float *mb_vel);
/* Optional warm start: proposed bound flag per particle (uint8[n], 1 = proposed bound). */
int hbtx_batch_set_proposal(hbtx_ctx *ctx, const uint8_t *proposal);
int hbtx_batch_run(hbtx_ctx *ctx, const hbtx_params *prm, uint32_t flags, double *elapsed_ms);
int hbtx_batch_fetch(hbtx_ctx *ctx, int64_t *out_order, float *out_energy, float *out_potential,
                     hbtx_sub_result *out_res);
/* Post-unbinding properties of the last run's subhaloes, centred on the most bound
 * particle: Subhalo_t::CalculateProfileProperties (src/subhalo.cpp:175-243, with
 * Snapshot_t::SphericalOverdensitySize, src/snapshot.cpp:258-282) and
 * Subhalo_t::CalculateShape (src/subhalo.cpp:245-309).  prm supplies G,
 * scale_factor, hz, the softening and the periodic box.  All zero when Nbound <= 1. */
typedef struct
{
  float rmax_comoving;            /* RmaxComoving */
  float vmax_physical;            /* VmaxPhysical */
  float r_half_comoving;          /* RHalfComoving */
  float r_enclose_comoving;       /* REncloseComoving */
  float bound_r200_crit_comoving; /* BoundR200CritComoving (when has_r200) */
  float bound_m200_crit;          /* BoundM200Crit (when has_r200) */
  float inertial_tensor[6];       /* InertialTensor {xx, xy, xz, yy, yz, zz} */
  float inertial_tensor_weighted[6]; /* InertialTensorWeighted */
  int32_t has_r200;               /* 0: no radius met the overdensity; the reference keeps the old values */
  int32_t reserved;
} hbtx_sub_profile;
int hbtx_batch_profiles(hbtx_ctx *ctx, const hbtx_params *prm, hbtx_sub_profile *out);
/* The same for bound sets given by the caller: subhalo s owns particles
 * [offsets[s], offsets[s+1]) of pos float[n][3] / mass float[n] in final order,
 * the first nbound[s] of them bound (the first is the centre), Mbound mbound[s]. */
int hbtx_profiles(hbtx_ctx *ctx, const hbtx_params *prm, int64_t n, const float *pos, const float *mass, int64_t nsub,
                  const int64_t *offsets, const int64_t *nbound, const float *mbound, hbtx_sub_profile *out);

/* Particle reassignment, the choice made by SubhaloSnapshot_t::ReassignParticles
 * (src/subhalo_reassign_gas.cpp:65-197): for every non-tracer particle, the k
 * nearest tracers of its host halo as GeoTree_t::NearestNeighbours finds them
 * (src/geometric_tree.cpp:108-141: a sphere of radius rguess grown by 1.26 until
 * min(tracers, k) are collected, the tracers of every sphere collected again into
 * one bounded heap); the particle moves to the subhalo of the nearest tracer
 * unless one of those neighbours belongs to its own subhalo.
 * Particles come host halo by host halo: halo g owns [group_offsets[g],
 * group_offsets[g+1]) (ngroups+1 ascending values, 0 .. n), in the order the
 * reference visits them (the subhaloes of MemberTable.SubGroups[g] in order, each
 * one's Particles).  pos float[n][3] comoving, is_tracer uint8[n]
 * (Particle_t::IsTracer), sub int32[n] the particle's subhalo.
 * out_target int32[n]: the subhalo to append the particle to, or -1 (tracers,
 * particles that stay).  k = NumNeighboursForReassignment, 1..32; rguess <= 0
 * selects SofteningHalo * 0.1 (prm->softening).  prm supplies the periodic box.
 * *elapsed_ms (optional): device time from the uploaded inputs to the targets. */
int hbtx_reassign_targets(hbtx_ctx *ctx, const hbtx_params *prm, int64_t n, const float *pos,
                          const uint8_t *is_tracer, const int32_t *sub, int64_t ngroups,
                          const int64_t *group_offsets, int32_t k, float rguess, int32_t *out_target,
                          double *elapsed_ms);

/* Kernel-level counters of the last run (work counts need HBTX_COUNT): stats[0..15], see hbtx_stat_name;
 * 12 walk ms, 13 rounds, 14 rounds replayed from graphs, 15 graph captures. */
int hbtx_batch_stats(hbtx_ctx *ctx, double *stats, int nstats);
const char *hbtx_stat_name(int i);
/* Device memory the library's buffers hold now, over all contexts of the process
 * (batch arrays, tree entries and records, sort and reduction scratch, trees,
 * networks); no counterpart in the reference (a footprint diagnostic). */
int hbtx_device_bytes(int64_t *bytes);

/* GravityTree_t replacement.  Build over n sources (pos float[n][3], mass float[n]);
 * *ncells receives the number of tree cells (OctTree_t::Build's return value). */
int hbtx_tree_build(hbtx_ctx *ctx, const hbtx_params *prm, int64_t n, const float *pos, const float *mass,
                    hbtx_tree **tree, int64_t *ncells);
/* Specific physical potential at n_tgt points (targetMass = tgt_mass[i], or 0
 * when tgt_mass is NULL, removing the self term of a target that is a source). */
int hbtx_tree_potential(hbtx_ctx *ctx, hbtx_tree *tree, int64_t n_tgt, const float *tgt_pos,
                        const float *tgt_mass, double *out_pot);
/* Potential + 0.5 |v - V + H a (x - X)|^2 at n_tgt points, one reference frame. */
int hbtx_tree_binding_energy(hbtx_ctx *ctx, hbtx_tree *tree, int64_t n_tgt, const float *tgt_pos,
                             const float *tgt_vel, const float *ref_pos, const float *ref_vel,
                             const float *tgt_mass, double *out_energy);
/* Depth-first preorder export of the tree (tests): per entry kind (0 particle,
 * 1 cell), source index or -1, mass, centre of mass [3], len, skip, depth. */
int hbtx_tree_export(hbtx_ctx *ctx, hbtx_tree *tree, int64_t cap, int32_t *kind, int64_t *pidx, float *mass,
                     float *com, float *len, int64_t *skip, int32_t *depth, int64_t *nentries);
void hbtx_tree_free(hbtx_tree *tree);

/* Exact softened potential by direct summation on the GPU (sources = targets allowed). */
int hbtx_direct_potential(hbtx_ctx *ctx, const hbtx_params *prm, int64_t n_src, const float *src_pos,
                          const float *src_mass, int64_t n_tgt, const float *tgt_pos, const float *tgt_mass,
                          double *out_pot);

/* Exact direct-sum potential without allocation: the same results as
 * hbtx_direct_potential, for callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  Inputs and the output are
 * DEVICE arrays of the context's device: d_src_pos float[n_src][3], d_src_mass
 * float[n_src] (non-negative), d_tgt_pos float[n_tgt][3], d_tgt_mass float[n_tgt]
 * or NULL (targetMass: its self term is removed), d_out_pot double[n_tgt].
 * Staging (the packed sources and targets, the work counter) lives in one
 * file-scope device array of the library (about 32 MiB a GPU, held until the
 * process ends), so n_src and n_tgt are at most HBTX_NOALLOC_MAX_POINTS
 * (HBTX_E_CAPACITY beyond).  The call only enqueues work on the context's
 * stream: synchronize it (hbtx_device_sync) before reading d_out_pot.  Every
 * call on one GPU shares that array: calls that overlap on one GPU (several
 * contexts, several host threads) race. */
#define HBTX_NOALLOC_MAX_POINTS (1 << 20)
int hbtx_direct_potential_noalloc(hbtx_ctx *ctx, const hbtx_params *prm, int64_t n_src, const float *d_src_pos,
                                  const float *d_src_mass, int64_t n_tgt, const float *d_tgt_pos,
                                  const float *d_tgt_mass, double *d_out_pot);

/* ---------------------------------------------------------------------------
 * Learned warm start (no counterpart in the reference; it only proposes the
 * starting set: the proposal is verified by an exact potential computation and
 * falls back to the plain iteration, src/subhalo_unbind.cpp:555-758).
 * Per-particle classifier 8 -> 64 -> 64 -> 1 (ReLU, logistic output), FP32 master
 * weights, BF16 weights in every forward/backward pass, FP32 accumulation.
 * ------------------------------------------------------------------------- */
#define HBTX_MLP_NIN 8
#define HBTX_MLP_NPAR 4801 /* W1[64][8] | b1[64] | W2[64][64] | b2[64] | w3[64] | b3 */
typedef struct hbtx_mlp hbtx_mlp;
int hbtx_mlp_create(hbtx_ctx *ctx, uint64_t seed, hbtx_mlp **mlp);
void hbtx_mlp_free(hbtx_mlp *mlp);
/* master FP32 weights and their BF16 working copy (either may be NULL) */
int hbtx_mlp_get_weights(hbtx_ctx *ctx, hbtx_mlp *mlp, float *master, uint16_t *bf16);
int hbtx_mlp_set_weights(hbtx_ctx *ctx, hbtx_mlp *mlp, const float *master);
/* mean binary cross-entropy and its gradient over n samples X[n][8], labels y[n] */
int hbtx_mlp_gradient(hbtx_ctx *ctx, hbtx_mlp *mlp, int64_t n, const float *X, const uint8_t *y, float *grad,
                      double *loss);
/* Adam training on the GPU, shuffled minibatches; *mean_loss = mean loss of the last epoch */
int hbtx_mlp_train(hbtx_ctx *ctx, hbtx_mlp *mlp, int64_t n, const float *X, const uint8_t *y, int32_t epochs,
                   int32_t batch, float lr, uint64_t seed, double *mean_loss);
int hbtx_mlp_predict(hbtx_ctx *ctx, hbtx_mlp *mlp, int64_t n, const float *X, float *logit);
/* Previous snapshot's outputs for the loaded batch (same slot = same particle):
 * prev_rank int32[n] (slot in the previous final order, -1 if absent), prev_bound
 * uint8[n], prev_energy float[n]; per subhalo prev_com_pos, prev_com_vel float[nsub][3],
 * prev_mbound and prev_rscale float[nsub] (rms radius of the previous bound set). */
int hbtx_batch_set_previous(hbtx_ctx *ctx, const int32_t *prev_rank, const uint8_t *prev_bound,
                            const float *prev_energy, const float *prev_com_pos, const float *prev_com_vel,
                            const float *prev_mbound, const float *prev_rscale);
/* Per-particle features float[n][8] on the device (out_features may be NULL). */
int hbtx_batch_features(hbtx_ctx *ctx, const hbtx_params *prm, float *out_features);
/* Inference on the batch features: proposal = P(bound) > threshold, kept on the device
 * for HBTX_WARMSTART; subhaloes of fewer than min_particles particles get no proposal
 * (they run the plain iteration); *elapsed_ms times the inference kernels. Needs
 * hbtx_batch_features after the last hbtx_batch_set_previous (HBTX_E_INVALID otherwise). */
int hbtx_batch_propose(hbtx_ctx *ctx, hbtx_mlp *mlp, float threshold, int64_t min_particles, int64_t *nproposed,
                       double *elapsed_ms);
/* Features and inference for the subhaloes of at least min_particles particles
 * only, in one call (the others get no proposal and run the plain iteration):
 * the same proposal as hbtx_batch_features followed by hbtx_batch_propose, with
 * the features and the classifier evaluated on those subhaloes' particles alone.
 * *elapsed_ms times the feature and inference kernels. Leaves the other particles'
 * features unset, so hbtx_batch_propose needs hbtx_batch_features again. */
int hbtx_batch_warm_proposal(hbtx_ctx *ctx, const hbtx_params *prm, hbtx_mlp *mlp, float threshold,
                             int64_t min_particles, int64_t *nproposed, double *elapsed_ms);
/* Copy of the current proposal (uint8[n]). */
int hbtx_batch_get_proposal(hbtx_ctx *ctx, uint8_t *proposal);

/* Tuning of the batched path (no counterpart in the reference): the source count
 * at or below which a subhalo takes the exact O(N^2) sum (used when
 * hbtx_params.direct_cutoff is -1) and the number of contiguous item chunks of the
 * walk's work distribution (one per XCD).  Loaded at hbtx_create from the
 * autotuner's file (HBTX_TUNE_FILE, else hbtx_tune.json beside the library), then
 * HBTX_DIRECT_CUTOFF / HBTX_WALK_CHUNKS. */
int hbtx_get_tuning(hbtx_ctx *ctx, int64_t *direct_cutoff, int32_t *walk_chunks);
int hbtx_set_tuning(hbtx_ctx *ctx, int64_t direct_cutoff, int32_t walk_chunks);

/* Segmented sort of uint64 keys with int32 payloads (used by the tree build and
 * the final binding-energy order; exported for tests). */
int hbtx_segmented_sort(hbtx_ctx *ctx, int64_t n, uint64_t *keys, int32_t *vals, int64_t nseg,
                        const int64_t *seg_offsets);

#ifdef __cplusplus
}
#endif
#endif
