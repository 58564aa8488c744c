/*
 * plan_hip.h — C ABI of the MI355X (gfx950) clump-finding library.
 *
 * The reference (astroboylrx/PLAN) exposes this path as public C++ members
 * of its BHtree / ParticleSet / PlanetesimalList templates; there is no
 * plugin registry or FFI.  Each entry point below replaces one of those
 * members for a caller that hands over plain arrays.  All functions return
 * 0 on success and a nonzero PLAN_E_* code on failure, with the message in
 * plan_last_error(); none of them exits the process (the reference calls
 * exit(2..5) on fatal errors, src/global.cpp:469-512, src/tree.hpp:1341).
 *
 * One plan_ctx owns one HIP stream on one device and all device memory of
 * one snapshot; it is not thread-safe, separate contexts may run
 * concurrently.  Device buffers are reused across calls.
 */
#ifndef PLAN_HIP_H
#define PLAN_HIP_H

#include <stdint.h>

#include "plan_params.h"

#ifdef __cplusplus
extern "C" {
#endif

enum {
    PLAN_OK = 0,
    PLAN_E_ARG = 1,      /* invalid argument (reference: exit(4), src/global.cpp:204,215) */
    PLAN_E_STATE = 2,    /* call out of order (no particles / no tree yet) */
    PLAN_E_DEVICE = 3,   /* HIP runtime error or out of device memory */
    PLAN_E_INTERNAL = 4, /* unexpected condition inside the pipeline */
    PLAN_E_IO = 5        /* a file cannot be opened or is shorter than its header states
                            (reference: exit(3), src/tree.hpp:1341, 1354, 1529) */
};

typedef struct plan_ctx plan_ctx;

/* Tree summary after plan_build_tree (BHtree members, src/tree.hpp:2461-2531). */
typedef struct plan_tree_info {
    uint32_t num_particles;   /* after sink merging */
    uint32_t num_nodes;
    uint32_t num_leaf_nodes;
    uint32_t num_sink_particles;
    uint32_t depth;           /* deepest node level (root = 1) */
    uint32_t hop_auto;        /* 1: FindPlanetesimals raises H to 64 for this snapshot by the
                                 automatic choice (src/tree.hpp:3583-3586), which the reference
                                 keeps in its parameters for the snapshots after */
    double   root_center[3];
    double   half_width;
    int32_t  level_count[32];
} plan_tree_info;

/* Create a context on HIP device `device`. */
int plan_ctx_create(plan_ctx **out, int device);
void plan_ctx_destroy(plan_ctx *ctx);
const char *plan_last_error(const plan_ctx *ctx);

/* Load the particle set of one snapshot, already in the reference's
 * ReadLisFile order (sorted by cpu_id, id_in_run; id = index):
 * replaces ParticleSet<3>::ReadLisFile's result (src/tree.hpp:1540-1621).
 * pos, vel: N x 3 row-major; density: the .lis grid density (ath_density);
 * property_index: species; id: permanent particle id.  params carry the
 * header box/time and the parameter-file keys (plan_params.h). */
int plan_set_particles(plan_ctx *ctx, const plan_params *params, uint32_t N, const double *pos, const double *vel,
                       const double *density, const int32_t *property_index, const uint32_t *id);
/* Same from the .lis-native float32 fields (converted to binary64 on the
 * device exactly as the reference's static_cast<double>, src/tree.hpp:1397-1404). */
int plan_set_particles_f32(plan_ctx *ctx, const plan_params *params, uint32_t N, const float *pos, const float *vel,
                           const float *density, const int32_t *property_index, const uint32_t *id);

/* ParticleSet<3>::ReadLisFile (src/tree.hpp:1540-1583) for one snapshot:
 * one combined file (num_paths = 1, ReadSingleLisFile, src/tree.hpp:1434-1529)
 * or one file per process of the run, in process order
 * (ReadMultipleLisFile, src/tree.hpp:1302-1430).  The records travel to the
 * device unchanged and are put in (cpu_id, id_in_run) order there, id =
 * index (src/tree.hpp:1552-1560); records with equal (cpu_id, id_in_run),
 * which the reference's std::sort leaves in unspecified order, keep their
 * file order.  params: in, the parameter-file keys; out, box_min/box_max
 * (the header's domain limits), time, num_types and type_info from the first
 * file's header.  dt (may be NULL) receives the header's dt, N_out the
 * particle count.  A missing file or a short header or body is PLAN_E_IO;
 * more than 16 species or a species index outside them is PLAN_E_ARG. */
int plan_read_lis(plan_ctx *ctx, plan_params *params, const char *const *paths, uint32_t num_paths, double *dt,
                  uint32_t *N_out);
/* The same from records already in host memory: N records of 44 bytes in
 * .lis layout (float pos[3], float vel[3], float density, int
 * property_index, long id_in_run, int cpu_id), in file order; params
 * complete, as for plan_set_particles. */
int plan_set_particles_lis(plan_ctx *ctx, const plan_params *params, uint32_t N, const void *records);

/* ParticleSet<3>::MakeGhostParticles (src/tree.hpp:1626-1754), on the device.
 * T_out receives N + number of ghosts.  With params->no_ghost set it is a no-op. */
int plan_make_ghost_particles(plan_ctx *ctx, uint32_t *T_out);

/* BHtree<3>::BuildTree(paras, particle_set, quiet, check_pos) (src/tree.hpp:2759-2908).
 * Consumes the particle set like the reference (src/tree.hpp:2820). */
int plan_build_tree(plan_ctx *ctx, int check_pos, plan_tree_info *info);

/* Step 1 of FindPlanetesimals (src/tree.hpp:3664-3689): new_density of every
 * tree particle from its K neighbours (K in 1..64 here and in every KNN entry
 * point below; H likewise).  kernel: 0 median spherical
 * (src/tree.hpp:3473-3491), 1 pseudo quadratic spline (src/tree.hpp:3427-3434),
 * -1 the reference's automatic choice (src/tree.hpp:3576-3612). */
int plan_knn_density(plan_ctx *ctx, uint32_t K, int kernel);

/* Steps 2-3 (src/tree.hpp:3711-3809): densest of the H nearest neighbours
 * for every particle with new_density >= threshold, then each one's peak.
 * Returns the number of such particles and of naive groups. */
int plan_hop(plan_ctx *ctx, uint32_t H, double threshold, uint32_t *num_above, uint32_t *num_groups);

/* BHtree<3>::FindPlanetesimals(ds, loop_count) (src/tree.hpp:3576-4339) on a
 * built tree: densities, HOP, merging, unbinding, rebinding, filters,
 * angular momentum.  *out is allocated by the library (plan_catalog_free). */
int plan_find_planetesimals(plan_ctx *ctx, plan_catalog **out);
void plan_catalog_free(plan_catalog *cat);

/* ---- catalog files ---------------------------------------------------------- */

/* The files PlanetesimalList<3> writes after FindPlanetesimals, byte for byte,
 * for a catalog of this context's current tree (before the next
 * plan_build_tree or particle load).  File errors are PLAN_E_IO.
 * plan_write_peaks: the peaks table of OutputPlanetesimalsInfo
 * (src/tree.hpp:4958-5150): clumps by ascending mass, Npar with sinks
 * expanded, mass, Hill radius, centre, velocity, J, and the geometric-mean
 * and median member offsets after rotating J onto z.  The reference names it
 * peaks_at_<time as %07.3f>.txt in the output directory (src/tree.hpp:4960-4968). */
int plan_write_peaks(plan_ctx *ctx, const plan_catalog *cat, const char *path);
/* OutputSinglePlanetesimal (src/tree.hpp:5191-5222) for every clump:
 * dir/<peak_index>.txt, every sampling_rate-th member in distance order
 * (original id, position, velocity, mass).  The reference's dir is
 * ParList.<NNNN> (src/tree.hpp:4976-4985); the caller creates it. */
int plan_write_clump_particles(plan_ctx *ctx, const plan_catalog *cat, const char *dir, uint32_t sampling_rate);
/* WriteBasicResults (src/tree.hpp:4935-4953): one line (time, number of
 * clumps, largest and total clump mass in code units, as fractions of the
 * total, in Ceres masses) appended to path, after the column header when
 * header != 0 (the reference writes it for the first snapshot). */
int plan_write_basic_results(plan_ctx *ctx, const plan_catalog *cat, const char *path, int header);

/* ---- analyses on the tree --------------------------------------------------- */

/* BasicAnalysesWithTree's density against scale (src/analyses.hpp:196-333)
 * on the built tree: for radii r_i = (cell/2) 2^i, i < num_levels, the
 * largest sheared-periodic ball count (RecursiveBallSearchCountWithShear,
 * src/tree.hpp:3383-3402) over the cell-centre grid from z_min to the top
 * particle (levels 0-2 also at half-cell offsets, level 0 at quarter-cell
 * offsets), as max_rhop[i] = count x m_particle / (4/3 pi r_i^3).  sums[i]
 * is the sum of the counts over the level's searches (each distinct search
 * once).  Writes at most cap entries of each array (any may be NULL).  With
 * path non-NULL the reference's max_rhop_vs_scale lines are appended there
 * (the radii header line first when header != 0). */
int plan_density_vs_scale(plan_ctx *ctx, uint32_t cap, double *radii, double *max_rhop, uint64_t *sums,
                          uint32_t *num_levels, const char *path, int header);

/* ---- inspection (tests, tools) ------------------------------------------------ */

/* the loaded particle set (ParticleSet<3>::particles, src/tree.hpp:1200-1298),
 * with ghosts after plan_make_ghost_particles: the first min(cap, count)
 * particles; *count receives the set's size.  Any pointer may be NULL.
 * pos/vel: n x 3; density: the .lis grid density. */
int plan_get_particle_set(plan_ctx *ctx, uint32_t cap, double *pos, double *vel, double *density, int32_t *prop,
                          uint32_t *id, uint32_t *count);
/* tree particles in tree order; any pointer may be NULL.  pos/vel: n x 3. */
int plan_get_tree_particles(plan_ctx *ctx, double *pos, double *vel, double *mass, double *ath_density,
                            double *new_density, uint32_t *original_id, uint32_t *sink_count);
/* sorted unique Morton keys as words (bits 0-31, 32-63, 64-95) */
int plan_get_keys(plan_ctx *ctx, uint32_t *w0, uint32_t *w1, uint32_t *w2);
/* node arrays (num_nodes); center: num_nodes x 3 */
int plan_get_nodes(plan_ctx *ctx, double *center, double *half_width, uint32_t *begin, uint32_t *end,
                   uint32_t *parent, uint32_t *first_daughter, uint8_t *num_daughters, uint8_t *level,
                   uint16_t *orthant, int32_t *node2leaf, int32_t *leaf_nodes);
/* densest neighbour and peak per tree particle after plan_hop (peak = n for
 * particles below the threshold) */
int plan_get_hop(plan_ctx *ctx, uint32_t *densest, uint32_t *peak);
/* K-nearest-neighbour queries with the reference's candidate-set rule
 * (KNN_Search_OpenMP, src/tree.hpp:3198-3251).  Either particle queries
 * (particle_idx non-NULL, start radius from the tree, est may be NULL) or
 * point queries (pts: nq x 3 with est[nq] > 0).  Outputs nq x K indices and
 * squared distances in ascending order, and the final search radius. */
int plan_knn_query(plan_ctx *ctx, uint32_t nq, const uint32_t *particle_idx, const double *pts, const double *est,
                   uint32_t K, uint32_t *idx, double *d2, double *rstop);
/* plan_knn_query for callers that cannot allocate device memory (code under
 * HIP graph capture, tight serving loops, harnesses that forbid hipMalloc);
 * it has no counterpart in the reference.  Same queries and results as
 * plan_knn_query, but every pointer is DEVICE memory on the context's device:
 * particle_idx[nq] or pts[nq x 3] with est[nq], and the outputs idx[nq x K],
 * d2[nq x K], rstop[nq] (any output may be NULL).  It calls no hipMalloc or
 * hipFree: its working buffers live in a fixed device array of
 * PLAN_NOALLOC_SCRATCH_BYTES per GPU, held until the process ends, so nq is
 * limited to PLAN_NOALLOC_MAX_QUERIES (larger is PLAN_E_ARG) and K to 1..64.
 * The array is shared by every call on one GPU: calls that overlap
 * on one GPU race.  Returns after the results are written (the context's
 * stream is synchronized). */
#define PLAN_NOALLOC_MAX_QUERIES 16384u
#define PLAN_NOALLOC_SCRATCH_BYTES (16u << 20)
int plan_knn_query_noalloc(plan_ctx *ctx, uint32_t nq, const uint32_t *particle_idx, const double *pts,
                           const double *est, uint32_t K, uint32_t *idx, double *d2, double *rstop);
/* Densities at points (IsSaddlePointDeepEnough's query, src/tree.hpp:4880-4892). */
int plan_density_at(plan_ctx *ctx, uint32_t nq, const double *pts, const double *est, uint32_t K, int kernel,
                    double *rho);
/* BallSearch (src/tree.hpp:3297-3304): indices in traversal order; returns
 * the count in *count (writes at most cap). */
int plan_ball_query(plan_ctx *ctx, const double *center, double radius, uint32_t *out, uint32_t cap,
                    uint32_t *count);
/* Keep a device copy of the loaded N-particle set / restore it (so a caller
 * can rerun ghosts + BuildTree + FindPlanetesimals on resident input). */
int plan_keep_input(plan_ctx *ctx);
int plan_restore_input(plan_ctx *ctx);
/* device bytes held by the context */
int plan_device_memory(plan_ctx *ctx, uint64_t *bytes);
/* wall-clock milliseconds of the last pipeline stages: ghosts, build, density,
 * hop, clumps, total (hipEvent timing on the context stream; clumps = host). */
int plan_stage_times(plan_ctx *ctx, double *ms, int n);
/* counters of the last FindPlanetesimals, up to n of these 48 values:
 *  0 particles above threshold, 1 HOP groups, 2 naive clumps, 3 clumps after
 *  merging, 4 merge passes, 5 pair candidates, 6 saddle queries, 7 rebinding
 *  ball searches, 8 member-table copy ms, 9 rebinding reruns,
 *  10-19 host ms of: sort, merge A, merge B, re-key, unbind, rebind, filters,
 *  rip, J, ranking,
 *  20 rebinding iterations, 21-23 iterations, members and ms of the slowest
 *  clump's rebinding,
 *  24-26 merge ms in clump-tree searches, saddle densities, applying merges,
 *  27-29 rebinding ms in superset searches, parallel runs, in-order validation,
 *  30-31 waves of the density and densest-neighbour passes whose node list
 *  overflowed the on-chip capacity (redone in smaller waves),
 *  32-34 the slowest clump's rebinding: exact device searches, their ms, ms
 *  re-sorting members,
 *  35-37 all rebinding runs: summed ms, summed ms in exact device searches
 *  (waits for the device included), schedule slot of the slowest run,
 *  38-40 the run that ended last: its end (ms after the first run started),
 *  schedule slot and ms,
 *  41-42 queries of the density and densest-neighbour passes that walked in
 *  several list fills (one query per wave),
 *  43 clumps keyed by tree index 0 that a merge loop's zero-filled delete
 *  list erased (src/tree.hpp:3874 -> 4000, 4019 -> 4146),
 *  44 densities of the density pass whose x^1.5 lay so near a rounding
 *  midpoint (about 0.013 ulp) that they were checked on the host against the C library's
 *  pow, which the reference calls (src/tree.hpp:3489), 45 those of them
 *  where that pow rounds the other way and which were redone with it,
 *  46 host ms building the plan_catalog, 47 host ms of the whole clump logic
 *  (merging to ranking),
 *  48-51 the slowest clump's rebinding: ms scanning its live candidate
 *  records, ms evaluating candidate energies and claiming, live records at
 *  its start, candidate energies evaluated,
 *  52-55 all rebinding runs: summed ms of live-record scans, of energy
 *  evaluation, of member re-sorts and of building each run's live set and
 *  member arrays */
int plan_clump_stats(plan_ctx *ctx, double *values, int n);
/* device milliseconds of the last density pass's gather kernel (the K-nearest
 * selection and density, one launch) and of its radius-schedule kernel */
int plan_last_density_kernel_ms(plan_ctx *ctx, double *ms);
int plan_last_density_radius_ms(plan_ctx *ctx, double *ms);
/* Test hook, no reference counterpart: cap the KNN include list of a wave
 * at max_entries (clamped to 8 .. the on-chip capacity; 0 restores the
 * capacity) and start the level-parallel ball searches with a frontier of
 * 8 * max_entries, so small inputs exercise the overflow paths (16-query
 * waves, resumable multi-fill walks, frontier regrowth).  Results do not
 * change, only speed. */
int plan_set_knn_list_limit(plan_ctx *ctx, uint32_t max_entries);
/* Test hook, no C API counterpart in the reference: the merge passes' clump
 * tree (BuildClumpTree's BuildTree over n points with the given root box,
 * src/tree.hpp:5251-5284) built on the host, and RecursiveBallSearch
 * (src/tree.hpp:3275-3293) for nq queries (x, y, z, radius).  pts4 holds
 * (x, y, z, payload) per point; query q's results are the payloads of the
 * particles found, in the walk's order, at out[offsets[q] .. offsets[q+1])
 * (offsets has nq + 1 entries).  Returns PLAN_E_ARG when the results exceed
 * cap (offsets[nq] then holds the count needed).  No device is used. */
int plan_host_tree_balls(const plan_params *params, uint32_t n, const double *pts4, const double *root_center,
                         double half_width, const double *queries4, uint32_t nq, uint32_t *offsets, double *out,
                         uint32_t cap);
/* num_hop for the next plan_find_planetesimals of the loaded snapshot (1..64).
 * The reference keeps H in its global parameters and its automatic choice
 * overwrites it (src/tree.hpp:3583-3586), so in a series the H a snapshot
 * starts from is the previous one's: the params' num_hop, or 64 once an
 * earlier snapshot's plan_tree_info.hop_auto was 1.  No reference function:
 * it replaces that carried parameter. */
int plan_set_num_hop(plan_ctx *ctx, uint32_t num_hop);

/* ---- pairwise-potential boundness (extension: no reference counterpart) ----
 *
 * The reference decides boundness by a two-body test against the rest of the
 * clump (RemoveUnboundParticles, src/tree.hpp:4615-4674).  These entry points
 * evaluate it with the full pairwise potential instead, per clump with bound
 * set B (initially all members), iterating until no member leaves:
 *   phi_i = -G sum_{j in B, x_j != x_i} m_j / |x_i - x_j|
 *   dv_i = v_i - V + (0, q Omega (X_x - x_i,x), 0)   (V, X: mass-weighted over B)
 *   e_i = |dv_i|^2 / 2 + phi_i, members with e_i > 0 leave B;
 *   E_kin = sum_B m |dv|^2 / 2, E_pot = sum_B m phi / 2.
 * Forms: PLAN_BOUND_FP64 (binary64 pair sums, one clump at a time) and
 * PLAN_BOUND_FP32 (binary32 pair sums with compensated accumulation, all
 * clumps per launch, iterations replayed from a HIP graph, binary64 totals,
 * binary64 re-checks of members within 1e-5 and clumps within 1e-3 of zero
 * energy). */
enum { PLAN_BOUND_FP64 = 0, PLAN_BOUND_FP32 = 1 };

typedef struct plan_boundness_in {
    uint32_t num_clumps;
    const uint32_t *offsets;   /* num_clumps + 1 member offsets */
    const double *pos;         /* members x 3 */
    const double *vel;         /* members x 3 */
    const double *mass;        /* members */
    double G, Omega, q;
} plan_boundness_in;

typedef struct plan_boundness_clump {
    uint32_t n_in, n_bound, iterations, reserved;
    double mass, center_of_mass[3], vel_com[3];
    double E_kin, E_pot;
} plan_boundness_clump;

typedef struct plan_boundness_stats {
    double ms;                 /* device time of the unbinding, inputs resident in HBM */
    double pair_evaluations;   /* pair terms evaluated */
    uint32_t refit_clumps;     /* clumps redone in binary64 (near-zero energy) */
    uint32_t graph_replays;
    int32_t tile_shape;        /* chosen pair-kernel shape (FP32 form), -1 otherwise */
    int32_t reserved;
    double tune_ms[4];         /* autotuner timing of each candidate shape */
} plan_boundness_stats;

/* stage a batch of clumps in HBM (replaces any previous batch) */
int plan_boundness_load(plan_ctx *ctx, const plan_boundness_in *in);
/* stage the clumps of a catalog from the last plan_find_planetesimals, with
 * member positions, velocities and masses taken from the device tree */
int plan_boundness_load_catalog(plan_ctx *ctx, const plan_catalog *cat);
/* run one form on the staged batch; max_iter bounds the unbinding iterations */
int plan_boundness_run(plan_ctx *ctx, int form, uint32_t max_iter, plan_boundness_stats *stats);
/* results of the last run: bound flag per member (may be NULL), per clump */
int plan_boundness_result(plan_ctx *ctx, uint8_t *bound, plan_boundness_clump *clumps);

#ifdef __cplusplus
}
#endif

#endif /* PLAN_HIP_H */
