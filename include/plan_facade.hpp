// plan_facade.hpp — C++ host side of the clump-finding path, with the
// reference's object model over the C ABI (plan_hip.h).
//
// PLAN exposes the path as public members of ParticleSet<3>, BHtree<3>,
// PlanetesimalList<3> and DataSet<T,3> (src/tree.hpp:1200-1298, 2345-2531,
// 4815-4860, 5298-5320) and calls them from src/main.cpp:72-79:
//
//     ds.particle_set.MakeGhostParticles(progIO->numerical_parameters);
//     ds.tree.BuildTree(progIO->numerical_parameters, ds.particle_set);
//     ds.tree.FindPlanetesimals(ds, loop_count);
//
// The classes below keep those names, argument meanings and effects so that
// the same three lines run on an MI355X:
// - ParticleSet<D>::particles holds the snapshot in ReadLisFile order;
//   MakeGhostParticles generates the ghosts on the device and updates
//   num_ghost_particles / num_total_particles;
// - BHtree<D>::BuildTree consumes the particle set (the reference frees the
//   caller's array, src/tree.hpp:2820) and fills the tree summary;
// - BHtree<D>::FindPlanetesimals(ds, loop_count) fills
//   ds.planetesimal_list.planetesimals (map keyed by peak tree index) and
//   peaks_and_masses, as the reference does (src/tree.hpp:3576-4339).
// The numerical parameters the reference reads from progIO are passed as a
// plan_params (plan_params.h).  Errors throw plan_amd::Error carrying the C
// status and message; the reference's exit codes for them are in
// exit_code_of().
#ifndef PLAN_FACADE_HPP
#define PLAN_FACADE_HPP

#include <algorithm>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <map>
#include <memory>
#include <stdexcept>
#include <string>
#include <utility>
#include <vector>

#include "plan_hip.h"

namespace plan_amd {

class Error : public std::runtime_error {
public:
    Error(int code, const std::string &msg) : std::runtime_error(msg), code(code) {}
    int code;
};

/* the reference's exit status for each failure class (src/global.cpp:204,215,
 * 469-512, src/tree.hpp:1341) */
inline int exit_code_of(int status)
{
    switch (status) {
    case PLAN_E_ARG: return 4;
    case PLAN_E_STATE: return 3;
    case PLAN_E_DEVICE: return 5;
    case PLAN_E_IO: return 3;
    default: return 2;
    }
}

/* one device context (one HIP stream, all device memory of a snapshot) */
class Context {
public:
    explicit Context(int device = 0)
    {
        plan_ctx *c = nullptr;
        int rc = plan_ctx_create(&c, device);
        if (rc != PLAN_OK) throw Error(rc, "plan_ctx_create failed");
        ctx_.reset(c);
    }
    plan_ctx *get() const { return ctx_.get(); }
    void check(int rc, const char *what) const
    {
        if (rc != PLAN_OK) throw Error(rc, std::string(what) + ": " + plan_last_error(ctx_.get()));
    }

private:
    struct Del {
        void operator()(plan_ctx *c) const { plan_ctx_destroy(c); }
    };
    std::unique_ptr<plan_ctx, Del> ctx_;
};

/* Particle<D> (src/tree.hpp:1154-1198) */
template <int D> struct Particle {
    double pos[D], vel[D];
    double density;
    uint32_t id_in_run;
    uint32_t id;
    int property_index;
    uint16_t cpu_id;
};

/* ParticleSet<D> (src/tree.hpp:1200-1298) */
template <int D> class ParticleSet {
public:
    uint32_t num_particles = 0;
    uint32_t num_ghost_particles = 0;
    uint32_t num_total_particles = 0;
    unsigned int num_types = 0;
    std::vector<Particle<D>> particles;
    std::shared_ptr<Context> ctx;

    explicit ParticleSet(std::shared_ptr<Context> c = nullptr) : ctx(c ? c : std::make_shared<Context>(0)) {}

    void AllocateSpace(uint32_t N)
    {
        particles.assign(N, Particle<D>{});
        num_particles = num_total_particles = N;
    }

    double time = 0, dt = 0;

    /* ReadLisFile (src/tree.hpp:1540-1583) straight into device memory: one
     * combined file or one file per process, in process order.  paras: in,
     * the parameter-file keys; out, the header's box limits, time and species
     * weights.  The particles stay on the device (`particles` is left empty)
     * until BuildTree consumes them. */
    void ReadLisFile(const std::vector<std::string> &paths, plan_params &paras)
    {
        std::vector<const char *> p;
        for (const auto &f : paths) p.push_back(f.c_str());
        uint32_t N = 0;
        ctx->check(plan_read_lis(ctx->get(), &paras, p.data(), (uint32_t)p.size(), &dt, &N), "ReadLisFile");
        particles.clear();
        num_particles = num_total_particles = N;
        num_ghost_particles = 0;
        num_types = paras.num_types;
        time = paras.time;
        params_ = paras;
        uploaded_ = from_lis_ = true;
    }

    /* MakeGhostParticles (src/tree.hpp:1626-1754): uploads the particles
     * (unless ReadLisFile put them on the device) and generates the ghost
     * copies on the device */
    void MakeGhostParticles(const plan_params &paras)
    {
        if (!from_lis_) upload(paras);
        uint32_t T = 0;
        ctx->check(plan_make_ghost_particles(ctx->get(), &T), "MakeGhostParticles");
        num_total_particles = T;
        num_ghost_particles = T - num_particles;
    }

    void upload(const plan_params &paras)
    {
        const uint32_t N = num_particles;
        std::vector<double> pos(3 * (size_t)N), vel(3 * (size_t)N), dens(N);
        std::vector<int32_t> prop(N);
        std::vector<uint32_t> id(N);
        for (uint32_t i = 0; i < N; ++i) {
            for (int d = 0; d < 3; ++d) {
                pos[3 * (size_t)i + d] = particles[i].pos[d];
                vel[3 * (size_t)i + d] = particles[i].vel[d];
            }
            dens[i] = particles[i].density;
            prop[i] = particles[i].property_index;
            id[i] = particles[i].id;
        }
        ctx->check(plan_set_particles(ctx->get(), &paras, N, pos.data(), vel.data(), dens.data(), prop.data(),
                                      id.data()),
                   "set particles");
        params_ = paras;
        uploaded_ = true;
        from_lis_ = false;
    }
    bool uploaded() const { return uploaded_; }
    const plan_params &params() const { return params_; }

private:
    plan_params params_{};
    bool uploaded_ = false;
    bool from_lis_ = false;
};

/* InternalParticle (src/tree.hpp:2367-2410), fetched on demand */
template <int D> struct InternalParticle {
    double pos[D], vel[D];
    double mass;
    double ath_density;
    double new_density;
    uint32_t original_id;
};

/* Planetesimal<D> (src/tree.hpp:4351-4427): the catalog fields */
template <int D> struct Planetesimal {
    uint32_t peak_index = 0;
    std::vector<uint32_t> indices;      /* tree indices, sorted by distance from the COM */
    std::vector<uint32_t> original_ids;
    uint32_t num_particles = 0;         /* with sink members expanded */
    double total_mass = 0;
    double center_of_mass[D] = {0};
    double vel_com[D] = {0};
    double Hill_radius = 0;
    double J[D] = {0};
    double accumulated_J_in_quarter_Hill_radius[D] = {0};
    double one10th_radius = 0;
    double two_sigma_mass_radius = 0;
    double max_distance = 0;
    double peak_density = 0;
};

/* PlanetesimalList<D> (src/tree.hpp:4815-4860) */
template <int D> class PlanetesimalList {
public:
    uint32_t num_planetesimals = 0;
    double density_threshold = 0;
    double clump_mass_threshold = 0;
    double peak_density_threshold = 0;
    double Mp_max = 0, Mp_tot = 0;
    std::vector<std::pair<uint32_t, double>> peaks_and_masses;
    std::map<uint32_t, Planetesimal<D>> planetesimals;
    std::shared_ptr<Context> ctx; /* the context whose tree the indices refer to */

    /* WriteBasicResults (src/tree.hpp:4935-4953): the summary line appended
     * to path, after the column header for the first snapshot */
    void WriteBasicResults(const std::string &path, bool first_snapshot) const
    {
        Catalog c(*this);
        ctx->check(plan_write_basic_results(ctx->get(), &c.cat, path.c_str(), first_snapshot ? 1 : 0),
                   "WriteBasicResults");
    }

    /* OutputPlanetesimalsInfo (src/tree.hpp:4958-5150): peaks_at_<time>.txt
     * in dir and, with parlist_dir non-empty (save_clumps), one particle list
     * per clump there (the caller creates ParList.<NNNN>) */
    void OutputPlanetesimalsInfo(const std::string &dir, double time, const std::string &parlist_dir = "",
                                 uint32_t sampling_rate = 1) const
    {
        Catalog c(*this);
        char name[64];
        std::snprintf(name, sizeof(name), "peaks_at_%07.3f.txt", time);
        const std::string path = dir.empty() ? std::string(name) : dir + "/" + name;
        ctx->check(plan_write_peaks(ctx->get(), &c.cat, path.c_str()), "OutputPlanetesimalsInfo");
        if (!parlist_dir.empty())
            ctx->check(plan_write_clump_particles(ctx->get(), &c.cat, parlist_dir.c_str(), sampling_rate),
                       "OutputSinglePlanetesimal");
    }

private:
    /* a plan_catalog view of the list, clumps in map order */
    struct Catalog {
        std::vector<plan_clump> clumps;
        std::vector<uint32_t> members, oids;
        plan_catalog cat{};
        explicit Catalog(const PlanetesimalList &l)
        {
            for (const auto &it : l.planetesimals) {
                const Planetesimal<D> &p = it.second;
                plan_clump k{};
                k.peak_index = p.peak_index;
                k.num_members = (uint32_t)p.indices.size();
                k.num_particles = p.num_particles;
                k.member_offset = (uint32_t)members.size();
                k.total_mass = p.total_mass;
                for (int d = 0; d < D; ++d) {
                    k.center_of_mass[d] = p.center_of_mass[d];
                    k.vel_com[d] = p.vel_com[d];
                    k.J[d] = p.J[d];
                    k.J_quarter_Hill[d] = p.accumulated_J_in_quarter_Hill_radius[d];
                }
                k.Hill_radius = p.Hill_radius;
                k.one10th_radius = p.one10th_radius;
                k.two_sigma_mass_radius = p.two_sigma_mass_radius;
                k.max_distance = p.max_distance;
                k.peak_density = p.peak_density;
                clumps.push_back(k);
                members.insert(members.end(), p.indices.begin(), p.indices.end());
                oids.insert(oids.end(), p.original_ids.begin(), p.original_ids.end());
            }
            cat.num_clumps = (uint32_t)clumps.size();
            cat.num_members_total = (uint32_t)members.size();
            cat.clumps = clumps.data();
            cat.members = members.data();
            cat.member_original_ids = oids.data();
            cat.density_threshold = l.density_threshold;
            cat.peak_density_threshold = l.peak_density_threshold;
            cat.clump_mass_threshold = l.clump_mass_threshold;
            cat.Mp_max = l.Mp_max;
            cat.Mp_tot = l.Mp_tot;
        }
    };
};

template <int D> class BHtree;

/* DataSet<T,D> (src/tree.hpp:5298-5320) */
template <typename T, int D> class DataSet {
public:
    ParticleSet<D> particle_set;
    BHtree<D> tree;
    PlanetesimalList<D> planetesimal_list;
    explicit DataSet(std::shared_ptr<Context> c = nullptr) : particle_set(c), tree(particle_set.ctx) {}
};

/* BHtree<D> (src/tree.hpp:2345-2531) */
template <int D> class BHtree {
public:
    uint32_t num_particles = 0;
    uint32_t num_nodes = 0;
    uint32_t num_leaf_nodes = 0;
    uint32_t num_sink_particles = 0;
    int max_level = 0;
    double root_center[D] = {0};
    double half_width = 0;
    std::shared_ptr<Context> ctx;

    explicit BHtree(std::shared_ptr<Context> c) : ctx(std::move(c)) {}

    /* BuildTree(paras, particle_set, quiet, check_pos) (src/tree.hpp:2759-2908) */
    void BuildTree(const plan_params &paras, ParticleSet<D> &particle_set, bool quiet = false, bool check_pos = true)
    {
        (void)quiet;
        if (!particle_set.uploaded()) {
            plan_params p = paras;
            p.no_ghost = 1;
            particle_set.upload(p);
            uint32_t T = 0;
            ctx->check(plan_make_ghost_particles(ctx->get(), &T), "upload");
        }
        plan_tree_info info{};
        ctx->check(plan_build_tree(ctx->get(), check_pos ? 1 : 0, &info), "BuildTree");
        num_particles = info.num_particles;
        num_nodes = info.num_nodes;
        num_leaf_nodes = info.num_leaf_nodes;
        num_sink_particles = info.num_sink_particles;
        max_level = (int)info.depth;
        for (int d = 0; d < D; ++d) root_center[d] = info.root_center[d];
        half_width = info.half_width;
        /* the reference consumes the particle set (src/tree.hpp:2820) */
        std::vector<Particle<D>>().swap(particle_set.particles);
    }

    /* FindPlanetesimals(ds, loop_count) (src/tree.hpp:3576-4339) */
    template <typename T> void FindPlanetesimals(DataSet<T, D> &ds, int loop_count)
    {
        (void)loop_count;
        plan_catalog *cat = nullptr;
        ctx->check(plan_find_planetesimals(ctx->get(), &cat), "FindPlanetesimals");
        std::unique_ptr<plan_catalog, void (*)(plan_catalog *)> hold(cat, plan_catalog_free);
        PlanetesimalList<D> &pl = ds.planetesimal_list;
        pl.ctx = ctx;
        pl.planetesimals.clear();
        pl.peaks_and_masses.clear();
        pl.density_threshold = cat->density_threshold;
        pl.peak_density_threshold = cat->peak_density_threshold;
        pl.clump_mass_threshold = cat->clump_mass_threshold;
        pl.Mp_max = cat->Mp_max;
        pl.Mp_tot = cat->Mp_tot;
        for (uint32_t i = 0; i < cat->num_clumps; ++i) {
            const plan_clump &c = cat->clumps[i];
            Planetesimal<D> p;
            p.peak_index = c.peak_index;
            p.indices.assign(cat->members + c.member_offset, cat->members + c.member_offset + c.num_members);
            p.original_ids.assign(cat->member_original_ids + c.member_offset,
                                  cat->member_original_ids + c.member_offset + c.num_members);
            p.num_particles = c.num_particles;
            p.total_mass = c.total_mass;
            for (int d = 0; d < D; ++d) {
                p.center_of_mass[d] = c.center_of_mass[d];
                p.vel_com[d] = c.vel_com[d];
                p.J[d] = c.J[d];
                p.accumulated_J_in_quarter_Hill_radius[d] = c.J_quarter_Hill[d];
            }
            p.Hill_radius = c.Hill_radius;
            p.one10th_radius = c.one10th_radius;
            p.two_sigma_mass_radius = c.two_sigma_mass_radius;
            p.max_distance = c.max_distance;
            p.peak_density = c.peak_density;
            pl.peaks_and_masses.emplace_back(c.peak_index, c.total_mass);
            pl.planetesimals.emplace(c.peak_index, std::move(p));
        }
        /* peaks_and_masses in the reference's ranking order: (mass, peak) */
        std::sort(pl.peaks_and_masses.begin(), pl.peaks_and_masses.end(),
                  [](const std::pair<uint32_t, double> &a, const std::pair<uint32_t, double> &b) {
                      if (a.second == b.second) return a.first < b.first;
                      return a.second < b.second;
                  });
        pl.num_planetesimals = (uint32_t)pl.planetesimals.size();
    }

    /* the tree's particle_list (src/tree.hpp:2474-2476) in tree order */
    std::vector<InternalParticle<D>> particle_list() const
    {
        const uint32_t n = num_particles;
        std::vector<double> pos(3 * (size_t)n), vel(3 * (size_t)n), m(n), ath(n), rho(n);
        std::vector<uint32_t> oid(n);
        ctx->check(plan_get_tree_particles(ctx->get(), pos.data(), vel.data(), m.data(), ath.data(), rho.data(),
                                           oid.data(), nullptr),
                   "particle_list");
        std::vector<InternalParticle<D>> out(n);
        for (uint32_t i = 0; i < n; ++i) {
            for (int d = 0; d < D; ++d) {
                out[i].pos[d] = pos[3 * (size_t)i + d];
                out[i].vel[d] = vel[3 * (size_t)i + d];
            }
            out[i].mass = m[i];
            out[i].ath_density = ath[i];
            out[i].new_density = rho[i];
            out[i].original_id = oid[i];
        }
        return out;
    }
};

}  // namespace plan_amd

#endif /* PLAN_FACADE_HPP */
