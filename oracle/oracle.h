/* CPU oracle for the gravitational-unbinding path.  TEST INFRASTRUCTURE ONLY:
 * only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg may load
 * it, as the checker or as the reported CPU baseline.  The product library
 * never links or calls it.
 *
 * It restates, in plain C++ (float storage, double walk and accumulators, the
 * same operation order), the reference's
 *   Subhalo_t::Unbind            src/subhalo_unbind.cpp:463-808
 *   GravityTree_t (Build / UpdateInternalNodes / EvaluatePotential / BindingEnergy)
 *                                src/gravity_tree.cpp:10-178, src/oct_tree.tpp:4-164
 * plus an FP64 direct-sum potential with the same spline kernel.
 *
 * Deliberate, documented differences from the reference:
 *   - near-coincident particle placement (src/oct_tree.tpp:426-438) draws from a
 *     per-tree xorshift generator instead of the global drand48() stream;
 *   - subsampling expects particles already partitioned/shuffled by the caller
 *     (src/subhalo_unbind.cpp:430-443 uses the global rand() stream).
 */
#ifndef HBTX_ORACLE_H
#define HBTX_ORACLE_H
#include <stdint.h>
#ifdef __cplusplus
extern "C" {
#endif

typedef struct
{
  double G;                       /* PhysicalConst::G, src/config_parser.cpp:200 */
  float scale_factor;             /* Cosmology.ScaleFactor */
  float hz;                       /* Cosmology.Hz */
  float softening;                /* SofteningHalo (comoving) */
  float max_phys_softening;       /* MaxPhysicalSofteningHalo, -1 = unset */
  float theta;                    /* TreeNodeOpenAngle */
  float bound_mass_precision;     /* BoundMassPrecision */
  int64_t max_sample_size;        /* MaxSampleSizeOfPotentialEstimate */
  int32_t refine_most_bound;      /* RefineMostBoundParticle */
  float bound_fraction_center_refinement;
  int64_t min_num_part;           /* MinNumPartOfSub */
  int64_t min_num_tracer;         /* MinNumTracerPartOfSub */
  int32_t tracer_mask;            /* TracerParticleBitMask */
  int32_t donot_subsample_mask;   /* DoNotSubsampleParticleBitMask */
  int32_t periodic;               /* PeriodicBoundaryOn */
  float box_size;                 /* BoxSize */
  int32_t thermal;                /* UNBIND_WITH_THERMAL_ENERGY */
  int32_t dm_only;                /* DM_ONLY build */
  int32_t upscale_per_type;       /* PotentialEstimateUpscaleMassesPerType */
  float sub_core_size_factor;     /* SubCoreSizeFactor */
  int64_t sub_core_size_min;      /* SubCoreSizeMin */
} ounb_params;

typedef struct
{
  int64_t nbound;
  float mbound;
  int64_t nbound_type[6];
  float mbound_type[6];
  float com_pos[3];      /* ComovingAveragePosition after unbinding */
  float com_vel[3];      /* PhysicalAverageVelocity after unbinding */
  float mostbound_pos[3];
  float mostbound_vel[3];
  float spec_pot;        /* SpecificSelfPotentialEnergy */
  float spec_kin;        /* SpecificSelfKineticEnergy */
  float spec_am[3];      /* SpecificAngularMomentum */
  int64_t tracer_index;
  int64_t mostbound_id;
  float core_pos[3], core_vel[3];
  float core_sigma_r, core_sigma_v;
  int32_t iterations;
  int32_t disrupted;
} ounb_result;

/* Unbind every subhalo (CSR offsets) independently, OpenMP dynamic schedule over
 * subhaloes with a serial inner loop (src/subhalo_unbind.cpp:871-903 with
 * omp_set_max_active_levels(1)).  Outputs per particle, in each subhalo's final
 * particle order: out_order[k] = local input index of the particle now at slot k,
 * out_energy[k] = its last Elist energy, out_potential[k] (bound part only, may be
 * NULL) = ParticlePotentialEnergies.  Returns wall seconds of the unbinding loop. */
double oracle_unbind_batch(const ounb_params *prm, int64_t n, const float *pos, const float *vel,
                           const float *mass, const float *u, const int32_t *type, const int64_t *id,
                           int64_t nsub, const int64_t *offsets, const float *ref_pos, const float *ref_vel,
                           const float *mb_pos, const float *mb_vel, int nthreads, int64_t *out_order,
                           float *out_energy, float *out_potential, ounb_result *out_res);

/* The same with the library's learned warm start restated (no reference
 * counterpart): proposal uint8[n], 1 = proposed bound.  Requires max_sample_size 0. */
double oracle_unbind_batch_warm(const ounb_params *prm, int64_t n, const float *pos, const float *vel,
                                const float *mass, const float *u, const int32_t *type, const int64_t *id,
                                int64_t nsub, const int64_t *offsets, const float *ref_pos, const float *ref_vel,
                                const float *mb_pos, const float *mb_vel, const uint8_t *proposal, int nthreads,
                                int64_t *out_order, float *out_energy, float *out_potential, ounb_result *out_res);

/* GravityTree_t::Build over n_src sources followed by EvaluatePotential at every
 * target (targetMass = tgt_mass[i], or 0 when tgt_mass is NULL).  Double result.
 * Returns the number of tree cells. */
int64_t oracle_tree_potential(const ounb_params *prm, int64_t n_src, const float *src_pos, const float *src_mass,
                              int64_t n_tgt, const float *tgt_pos, const float *tgt_mass, int nthreads,
                              double *out_pot);

/* SetMassUpscaleFactor over n particles in E-list order (mass, type or NULL):
 * returns the number of gravity sources; out_upscale[6] = per-type factors. */
int64_t oracle_mass_upscale(const ounb_params *prm, int64_t n, const float *mass, const int32_t *type,
                            int64_t nlast, float mlast, int64_t mss, int64_t nuns, float *out_upscale);

/* FP64 all-pairs sum with the reference spline kernel (theta -> 0 limit of the walk). */
void oracle_direct_potential(const ounb_params *prm, int64_t n_src, const float *src_pos, const float *src_mass,
                             int64_t n_tgt, const float *tgt_pos, const float *tgt_mass, int nthreads,
                             double *out_pot);

/* The tree of oracle_tree_potential flattened into depth-first preorder (the
 * order in which a walk that opens every cell visits entries).  Per entry:
 * kind (0 particle, 1 cell), particle index or -1, cell mass/centre of mass/len,
 * and the preorder index of the entry that follows the subtree ("skip").
 * Returns the entry count (n_src + cells); arrays must hold cap entries. */
int64_t oracle_tree_dfs(const ounb_params *prm, int64_t n_src, const float *src_pos, const float *src_mass,
                        int64_t cap, int32_t *kind, int64_t *pidx, float *mass, float *com, float *len,
                        int64_t *skip, int32_t *depth);

/* Subhalo_t::GetCorePhaseSpaceProperties on an already ordered bound list. */
void oracle_core_phase_space(const ounb_params *prm, int64_t nbound, const float *pos, const float *vel,
                             const float *mass, const int32_t *type, float *core_pos, float *core_vel,
                             float *sigma_r, float *sigma_v);

/* Subhalo_t::CalculateProfileProperties and CalculateShape (src/subhalo.cpp:175-309). */
typedef struct
{
  float rmax_comoving;            /* RmaxComoving */
  float vmax_physical;            /* VmaxPhysical */
  float r_half_comoving;          /* RHalfComoving */
  float r_enclose_comoving;       /* REncloseComoving */
  float bound_r200_crit_comoving; /* BoundR200CritComoving */
  float bound_m200_crit;          /* BoundM200Crit */
  float inertial_tensor[6];       /* InertialTensor {xx, xy, xz, yy, yz, zz} */
  float inertial_tensor_weighted[6];
  int32_t has_r200;               /* 0: no radius met the overdensity (the reference leaves the fields as they were) */
  int32_t reserved;
} ounb_profile;

/* Per subhalo: its bound particles [offsets[s], offsets[s] + nbound[s]) in final
 * order (the first is the most bound, the centre, src/subhalo_unbind.cpp:754),
 * mbound[s] its Mbound. */
void oracle_profiles(const ounb_params *prm, int64_t n, const float *pos, const float *mass, int64_t nsub,
                     const int64_t *offsets, const int64_t *nbound, const float *mbound, ounb_profile *out);

/* SubhaloSnapshot_t::ReassignParticles (src/subhalo_reassign_gas.cpp:63-197) for
 * the particle selection: groups are contiguous ranges [group_offsets[g],
 * group_offsets[g+1]) of the particle arrays, in the reference's order (the
 * group's subhaloes in turn, each subhalo's particles).  Every non-tracer whose
 * k nearest tracers of its group, found by GeoTree_t::NearestNeighbours
 * (src/geometric_tree.cpp:108-141: all tracers within a radius that starts at
 * rguess and grows by 1.26 until enough were collected, collected again at every
 * radius), include none of its own subhalo gets out_target = the subhalo of the
 * nearest tracer; others get -1.  A group without tracers moves nothing. */
void oracle_reassign_targets(const ounb_params *prm, int64_t n, const float *pos, const uint8_t *is_tracer,
                             const int32_t *sub, int64_t ngroups, const int64_t *group_offsets, int32_t k,
                             float rguess, int32_t *out_target);

#ifdef __cplusplus
}
#endif
#endif
