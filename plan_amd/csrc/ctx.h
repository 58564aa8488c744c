// ctx.h — device-resident state of one clump-finding context.
#pragma once

#include "devmem.h"
#include "../../include/plan_params.h"

#include <vector>

namespace plan {

/* An octree in the reference's level-major layout (src/tree.hpp:2419-2455,
 * 2849-2886), stored structure-of-arrays in HBM.  Particles are in Morton
 * order; nodes are level-major, key-ordered within a level. */
struct DevTree {
    uint32_t n = 0;               /* tree particles after sink merging */
    uint32_t num_nodes = 0;
    uint32_t num_leaves = 0;
    uint32_t num_sinks = 0;
    bool mass_uniform = false;    /* every tree particle has species 0's mass */
    int depth = 0;
    double root_center[3] = {0, 0, 0};
    double half_width = 0;
    double epsilon = 0;
    double boxmin[3] = {0, 0, 0}, scale[3] = {1, 1, 1};
    int level_count[32];
    int level_ptr[32];

    double4 *pm = nullptr;        /* x, y, z, mass            */
    double4 *va = nullptr;        /* vx, vy, vz, ath_density  */
    double *rho = nullptr;        /* new_density              */
    uint32_t *oid = nullptr;      /* original_id              */
    uint32_t *sinkcnt = nullptr;  /* members merged into this particle (1 = not a sink) */
    uint32_t *sinkoff = nullptr;  /* first member in sorted order (for sink expansion) */
    uint4 *skeys = nullptr;       /* all T keys in sorted order, payload = particle-set index (sink members) */
    uint4 *keys = nullptr;        /* sorted unique keys {w0, w1, w2, first sorted slot} */
    size_t node_cap = 0;          /* capacity of the node arrays (grow-only) */
    uint32_t list_limit = 0xffffffffu; /* KNN include-list entries per wave (test hook) */
    NodeGeo *ngeo = nullptr;
    NodeTop *ntop = nullptr;
    uint32_t *ninfo = nullptr;    /* num_daughters | level << 8 | orthant << 16 */
    uint32_t *skip = nullptr;     /* next node in pre-order after the subtree */
    uint32_t *leaf_of = nullptr;  /* leaf node holding each particle */
    int32_t *node2leaf = nullptr;
    int32_t *leaf_nodes = nullptr;
    Arena arena;
    HostPinned host_stage; /* pinned host staging of the density re-check */
};

/* Particle set before the tree (ParticleSet<3>, src/tree.hpp:1200-1298). */
struct DevParticles {
    uint32_t N = 0;       /* real particles */
    uint32_t T = 0;       /* with ghosts */
    uint32_t cap = 0;
    double4 *pos = nullptr;   /* x, y, z, density (.lis grid density) */
    double4 *vel = nullptr;   /* vx, vy, vz, 0 */
    int32_t *prop = nullptr;
    uint32_t *id = nullptr;
    Arena arena;
};

/* derived parameters: the reference's NumericalParameters after ReadLisFile */
void derive_params(const plan_params *p, uint32_t N, plan_derived *d);

/* lis.hip: ParticleSet<3>::ReadLisFile with the particle work on the device */
struct LisHeader {
    double coor_lim[12] = {};   /* grid lo/hi x3, then domain lo/hi x3 */
    uint32_t num_types = 0;
    double type_info[PLAN_MAX_TYPES] = {};
    double time = 0, dt = 0;
    uint64_t body_offset = 0;   /* first record's byte offset in every file */
    uint32_t num_particles = 0; /* over all files */
};
/* headers of the files (error message, or "" on success; io: the message is
 * about a file that cannot be opened or read, not about its contents) */
std::string lis_read_headers(const char *const *paths, uint32_t num_paths, LisHeader &h,
                             std::vector<uint64_t> &counts, bool &io);
/* all files' records, in file order, to consecutive device bytes at dst
 * (error message, always about reading a file, or "" on success) */
std::string lis_upload_records(const char *const *paths, uint32_t num_paths, const LisHeader &h,
                               const std::vector<uint64_t> &counts, uint8_t *dst, hipStream_t s);
/* n device records -> ps (capacity >= n) in (cpu_id, id_in_run) order with
 * id = index; false if a species lies outside [0, num_types) */
bool dev_lis_to_particles(Arena &a, const uint32_t *rec, uint32_t n, uint32_t num_types, DevParticles &ps,
                          hipStream_t s);

/* writers.hip: the reference's catalog files */
struct CatalogMembers {
    std::vector<double4> pm, va;  /* tree particle of each catalog member: pos + mass, vel + ath_density */
    std::vector<uint32_t> off;    /* member i's particles (sinks expanded) are exp[off[i], off[i+1]) */
    std::vector<double4> exp;     /* their positions, sink members in sorted order */
};
void catalog_members(const DevTree &t, const DevParticles &ps, const plan_catalog &cat, Arena &a, hipStream_t s,
                     CatalogMembers &out);
/* the peaks table of OutputPlanetesimalsInfo ("" or an error message) */
std::string write_peaks(const plan_catalog &cat, const CatalogMembers &mem, const char *path);
/* OutputSinglePlanetesimal for every clump: dir/<peak_index>.txt */
std::string write_clump_particles(const plan_catalog &cat, const CatalogMembers &mem, const char *dir,
                                  uint32_t rate);
/* WriteBasicResults: one line (after the header line if header) appended to path */
std::string write_basic_results(const plan_catalog &cat, double mass_total, double G_tilde, double time,
                                const char *path, bool header);

/* build.hip */
void dev_make_ghosts(DevParticles &ps, const plan_params &pp, const plan_derived &d, hipStream_t s);
void dev_build_tree(DevTree &t, const DevParticles &ps, const plan_derived &d, bool check_pos,
                    const double *root_center_in, double half_width_in, hipStream_t s);

/* query.hip */
/* the tree as the traversal kernels see it */
struct TreeView {
    const double4 *pm;
    const double *rho;
    const NodeGeo *ngeo;
    const NodeTop *ntop;
    const uint32_t *ninfo;
    const uint32_t *skip;
    const uint32_t *leaf_of;
    uint32_t n;
    uint32_t nn;
    /* neighbours per query (1..64); the kernels are built for 8, 16, 32 and
     * 64 and keep the kq nearest of the next of those */
    uint32_t kq = 0;
    double to_diag;
    double cell_len[3], half_cell[3], cell_sq[3], three_half_cell[3];
    double cell_volume;
    /* bound on how far a particle can sit outside its node's box (Morton
     * quantisation of the key grid), with a wide allowance for rounding */
    double margin_abs;
    uint32_t list_limit; /* include-list entries per wave, at most the kernel's capacity */
    /* msum[k]: k masses m0 summed one by one from zero (k <= 64), or null */
    const double *msum = nullptr;
    double m0 = 0;
    bool mass_uniform = false; /* every tree particle's mass is m0 */
    /* densities to recompute with the C library's pow: count, capacity, and
     * per entry the output index, x and the mass sum (null: none kept) */
    uint32_t *pw_count = nullptr;
    uint32_t pw_cap = 0;
    uint32_t *pw_idx = nullptr;
    double *pw_x = nullptr, *pw_num = nullptr;
};
TreeView tree_view(const DevTree &t, const plan_derived &d);

/* analyses.hip: BasicAnalysesWithTree's density against scale */
struct ScaleResult {
    std::vector<double> radii, max_rhop; /* per level */
    std::vector<uint64_t> sums;          /* sum of the counts of every search of the level */
    double z_min = 0, z_max = 0;
    uint64_t points = 0;                 /* cell centres searched */
};
void dev_density_vs_scale(const DevTree &t, const plan_params &pp, const plan_derived &d, Arena &a, hipStream_t s,
                          ScaleResult &res);

struct KnnOut {
    uint32_t *idx = nullptr;   /* nq * K, ascending d^2 (optional) */
    double *d2 = nullptr;      /* nq * K (optional) */
    double *rho = nullptr;     /* nq densities (optional) */
    uint32_t *densest = nullptr; /* nq densest-neighbour ids (hop mode) */
    double *rstop = nullptr;   /* nq final radius (optional) */
    hipEvent_t *ev = nullptr;  /* 3 events: before the radius pass, around the first select launch */
    uint32_t *novf = nullptr;  /* host counter: waves whose include list overflowed */
    uint32_t *nres = nullptr;  /* host counter: queries walked in several fills (one per wave) */
    uint32_t *nchk = nullptr;  /* host counter: densities checked against the C library's pow */
    uint32_t *npow = nullptr;  /* host counter: densities redone with the C library's pow */
};
/* neighbour counts the KNN kernels take (a wave's lanes hold one list) */
inline bool knn_k_supported(uint32_t K) { return K >= 1 && K <= 64; }
/* mode: 0 density, 1 hop (densest neighbour), 2 lists only.
 * K must satisfy knn_k_supported(K).
 * Queries are tree particles qidx[i] (or i when qidx is null) or explicit
 * points qpts with start radii est. kernel: 0 median spherical, 1 QS.
 * Working buffers come from scratch when given, else from the tree's arena. */
void dev_knn(const DevTree &t, uint32_t K, int mode, int kernel, uint32_t nq, const uint32_t *qidx,
             const double4 *qpts, const double *est, const double *rho_in, const plan_derived &d, KnnOut out,
             hipStream_t s, Arena *scratch = nullptr);
/* ball search for nq centres/radii: two-pass (count, fill) into a CSR list,
 * optionally keeping only particles with rho >= rho_min (rho_min_on). */
void dev_ball(const DevTree &t, uint32_t nq, const double4 *centre_r, uint32_t *counts, const uint32_t *offsets,
              uint32_t *out, const double *rho, double rho_min, int filter, hipStream_t s);
/* the same searches expanded level by level over all queries at once; results
 * per query in traversal order (ascending tree index) */
void dev_ball_lists(const DevTree &t, Arena &a, const std::vector<double4> &queries, const double *rho,
                    double rho_min, int filter, std::vector<std::vector<uint32_t>> &out, hipStream_t s);

}  // namespace plan
