// clumps.h — clump post-processing after HOP (host control over device queries).
#pragma once

#include "hop.h"
#include "../../include/plan_params.h"

#include <functional>
#include <map>
#include <vector>

namespace plan {

/* Device services the clump logic calls; implemented by the context. */
struct ClumpServices {
    virtual ~ClumpServices() {}
    /* density at points {x, y, z, start radius} with the run's K and kernel */
    virtual void saddle_density(const std::vector<double4> &pts, std::vector<double> &rho) = 0;
    /* build the clump tree over pts {com, peak id} and ball-search it;
     * returns, per query {x, y, z, r}, the peak ids found in traversal order */
    virtual void clump_tree_balls(const std::vector<double4> &com_peak, const std::vector<double4> &vel,
                                  const std::vector<double4> &queries,
                                  std::vector<std::vector<uint32_t>> &peaks) = 0;
    /* ball search on the particle tree keeping rho >= rho_min; tree indices */
    virtual void particle_balls(const std::vector<double4> &queries, double rho_min,
                                std::vector<std::vector<uint32_t>> &out) = 0;
};

struct ClumpSettings {
    unsigned K = 32;
    int kernel = 0;
    int enough_resolution = 1;
    double density_threshold = 0;
    double peak_density_threshold = 0;
    double clump_mass_threshold = 0;
    double diffuse = 0.55;
    double hill_frac = 0.25;
    double G = 0;
    double Omega = 1;
    double q = 1.5;
    double box_min[3], box_max[3];
    uint32_t num_peaks = 0;
    /* distance band around a ball's surface inside which a plain distance
     * test may disagree with the tree's box tests (particles sit up to the
     * Morton quantisation outside their leaf box) */
    double ball_margin_abs = 0;
};

/* per-clump state (Planetesimal<3>, src/tree.hpp:4351-4427); members are
 * handles into the host member table (records of particles above threshold) */
struct HClump {
    uint32_t peak_index = 0;
    std::vector<uint32_t> subpeaks;
    std::vector<uint32_t> indices;
    bool mask = true;
    double total_mass = 0;
    double com[3] = {0, 0, 0}, vel[3] = {0, 0, 0};
    std::vector<std::pair<uint32_t, double>> particles;
    double outer = 0, inner = 0, one10th = 0, two_sigma = 0, hill = 0;
    double J[3] = {0, 0, 0}, Jq[3] = {0, 0, 0};
};

/* a view of n elements in memory owned elsewhere (pinned host staging) */
template <class T> struct HostSpan {
    T *p = nullptr;
    size_t n = 0;
    T &operator[](size_t i) const { return p[i]; }
    size_t size() const { return n; }
    T *data() const { return p; }
    T *begin() const { return p; }
    T *end() const { return p + n; }
};

struct HostMembers {
    HostSpan<MemberRec> rec;            /* grouped order */
    std::vector<uint32_t> cstart;       /* naive clump starts, plus na at the end */
    std::vector<uint32_t> cpeak;        /* naive clump peak (tree index), ascending */
    HostSpan<uint32_t> sorted_tidx;     /* tree indices ascending */
    HostSpan<uint32_t> rec_of_sorted;   /* record of sorted_tidx[i] */
    HostSpan<uint32_t> sinkcnt;         /* per record: particles merged into it */
    /* optional dense tree index -> record table (entries of tree indices
     * that are not records hold anything; rec_of_tidx checks the record's
     * own tree index) */
    const uint32_t *tidx_map = nullptr;
    size_t tidx_map_n = 0;
    uint32_t rec_of_tidx(uint32_t tidx) const;
    /* when set, records may still be arriving: wait_records(e) returns once
     * records [0, e) are in rec */
    std::function<void(uint32_t)> wait_records;
};

struct ClumpStats {
    uint32_t naive = 0;
    uint32_t after_merge = 0;
    uint32_t merge_passes = 0;
    uint32_t saddle_queries = 0;
    uint32_t rebind_searches = 0;
    uint32_t pair_candidates = 0;
    uint32_t rebind_reruns = 0;
    /* clumps keyed by tree index 0 erased after a merge loop */
    uint32_t key0_erased = 0;
    uint32_t rebind_iters = 0, rebind_max_iters = 0, rebind_max_n = 0, rebind_max_exact = 0;
    double rebind_max_ms = 0, rebind_max_exact_ms = 0, rebind_max_sort_ms = 0;
    /* the longest run's live-record scans, energy blocks, initial live
     * records and energy evaluations; scans, energies and re-sorts summed
     * over all runs */
    double rebind_max_scan_ms = 0, rebind_max_energy_ms = 0, rebind_max_live = 0, rebind_max_evals = 0;
    double rebind_scan_total_ms = 0, rebind_energy_total_ms = 0, rebind_sort_total_ms = 0;
    /* all rebinding runs: summed ms building the live set and member arrays */
    double rebind_setup_total_ms = 0;
    /* rebinding runs: summed time, summed exact-search time (waits included),
     * schedule position of the longest run */
    double rebind_work_ms = 0, rebind_exact_total_ms = 0;
    uint32_t rebind_max_slot = 0;
    /* the run that ended last: its end (ms after the first run started),
     * schedule slot and length */
    double rebind_last_end_ms = 0, rebind_last_ms = 0;
    uint32_t rebind_last_slot = 0;
    /* host milliseconds: sort, merge A, merge B, re-key, unbind, rebind,
     * filters, rip, J, ranking */
    double ms[10] = {0, 0, 0, 0, 0, 0, 0, 0, 0, 0};
    /* merge loops: clump-tree searches, saddle densities, applying merges;
     * rebinding: superset searches, parallel runs, in-order validation */
    double mparts[3] = {0, 0, 0};
    double rparts[3] = {0, 0, 0};
};

void run_clump_pipeline(HostMembers &hm, const ClumpSettings &cs, ClumpServices &svc,
                        std::map<uint32_t, HClump> &out, double &Mp_max, double &Mp_tot, ClumpStats &st);

}  // namespace plan
