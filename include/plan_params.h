/*
 * plan_params.h — plain-old-data parameter block shared by the C ABI
 * (plan_hip.h) and its callers.
 *
 * Every field corresponds to a value the reference reads from the snapshot
 * header or from the key = value parameter file:
 *   box_min / box_max / time : .lis header domain limits and time
 *                              (src/tree.hpp:1565-1566, 1557-1560 of ReadSingleLisFile)
 *   solid_to_gas_ratio       : key "Z"             (src/global.cpp:110-113)
 *   box_resolution           : keys "Nx1..Nx3"     (src/global.cpp:115-122)
 *   four_pi_G                : key "four_pi_G_par" (src/global.cpp:123-126)
 *   num_knn                  : keys "num_neighbors_in_knn_search", "num_knn", "knn" (src/global.cpp:156-170),
 *                              1..64 (one wave's lanes hold a query's list; larger is PLAN_E_ARG)
 *   num_hop / fixed_num_hop  : keys "num_neighbors_to_hop", "num_hop", "hop"          (src/global.cpp:171-188),
 *                              1..64 likewise (the automatic choice is 64 at most, src/tree.hpp:3583-3589)
 *   num_peaks                : keys "num_peaks", "num_peak" (src/global.cpp:146-155)
 *   clump_diffuse_threshold  : key "clump_diffuse_threshold" (src/global.cpp:199-208), <0 = automatic
 *   Hill_fraction_for_merge  : key "Hill_fraction_for_merge" (src/global.cpp:210-219), <0 = automatic
 *   no_ghost                 : CLI flag --No_Ghost (src/tree.hpp:1627-1629)
 *   type_info                : .lis header per-species mass weights (src/tree.hpp:1602-1611)
 * Quantities the reference derives from these (cell sizes, particle masses,
 * G, thresholds) are recomputed by each consumer with the reference's
 * operation order (src/global.cpp:28-47, src/tree.hpp:1597-1612).
 */
#ifndef PLAN_PARAMS_H
#define PLAN_PARAMS_H

#include <stdint.h>

#define PLAN_MAX_TYPES 16

typedef struct plan_params {
    double   box_min[3];
    double   box_max[3];
    double   time;
    double   solid_to_gas_ratio;
    double   four_pi_G;
    double   clump_diffuse_threshold;
    double   Hill_fraction_for_merge;
    double   type_info[PLAN_MAX_TYPES];
    uint32_t box_resolution[3];
    uint32_t num_types;
    uint32_t num_knn;
    uint32_t num_hop;
    int32_t  fixed_num_hop;
    uint32_t num_peaks;
    int32_t  no_ghost;
    /* 0: this snapshot starts a series, and box, species weights and particle
     * masses come from it.  Otherwise the particle count of the series' first
     * snapshot, whose box and species weights these params carry: the
     * reference derives them at its first loop only (ReadLisFile at
     * loop_begin, src/tree.hpp:1563-1581), so later snapshots keep the first
     * one's particle masses (M_tot / N_first) and plan_read_lis keeps these
     * params' box and weights instead of the file header's. */
    uint32_t series_first_n;
} plan_params;

/* Values of the reference's NumericalParameters after ReadLisFile
 * (src/global.hpp:678-874 defaults; src/global.cpp:28-47; src/tree.hpp:1597-1612). */
typedef struct plan_derived {
    double box_center[3];
    double box_length[3];
    double box_half_width[3];
    double max_half_width;
    double ghost_zone_width[3];
    double max_ghost_zone_width;
    double cell_length[3];
    double cell_volume;
    double mass_total_code_units;
    double grav_constant;
    double shear_speed;
    double q;
    double Omega;
    double min_trusted_mass_code_unit;
    double mass_per_particle[PLAN_MAX_TYPES];
} plan_derived;

/* One clump of the final catalog: the fields of Planetesimal<3>
 * (src/tree.hpp:4357-4427) that FindPlanetesimals leaves behind. */
typedef struct plan_clump {
    uint32_t peak_index;          /* map key: tree index of the peak particle */
    uint32_t num_members;         /* tree particles in the clump */
    uint32_t num_particles;       /* members with sink particles expanded (src/tree.hpp:5036-5043) */
    uint32_t member_offset;       /* first member in plan_catalog.members */
    double   total_mass;
    double   center_of_mass[3];
    double   vel_com[3];
    double   Hill_radius;
    double   J[3];                /* in units of M R_H^2 Omega */
    double   J_quarter_Hill[3];   /* accumulated within R_H / 4 */
    double   one10th_radius;
    double   two_sigma_mass_radius;
    double   max_distance;        /* particles.back().second */
    double   peak_density;        /* new_density of the peak particle */
} plan_clump;

/* The clump catalog in ascending peak-index (std::map) order.  members[]
 * holds tree indices in each clump's final distance order;
 * member_original_ids[] the matching permanent particle ids. */
typedef struct plan_catalog {
    uint32_t    num_clumps;
    uint32_t    num_members_total;
    plan_clump *clumps;
    uint32_t   *members;
    uint32_t   *member_original_ids;
    double      density_threshold;
    double      peak_density_threshold;
    double      clump_mass_threshold;
    double      Mp_max;
    double      Mp_tot;
} plan_catalog;

#endif /* PLAN_PARAMS_H */
