// api.cpp — the C ABI (include/plan_hip.h): context, particle upload, stage
// orchestration of FindPlanetesimals and the device services of the host
// clump logic.
#include "../../include/plan_hip.h"
#include "clumps.h"
#include "ctx.h"
#include "hop.h"
#include "boundness.h"
#include "host_tree.h"

#include <chrono>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <map>
#include <string>
#include <atomic>
#include <condition_variable>
#include <memory>
#include <thread>
#include <exception>
#include <mutex>
#include <vector>

using namespace plan;

struct plan_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    plan_params pp{};
    plan_derived d{};
    bool have_particles = false;
    bool have_tree = false;
    bool have_density = false;
    bool have_hop = false;
    DevParticles ps;
    DevParticles saved;      /* optional pristine copy of the N input particles */
    BoundnessEngine bnd;     /* staged pairwise-boundness batch */
    bool have_bnd_result = false;
    DevTree tree;
    DevTree ctree;
    DevParticles cps;
    Arena work;
    HopResult hop;
    uint32_t K = 0, H = 0;
    int kernel = 0;
    double times[6] = {0, 0, 0, 0, 0, 0};
    ClumpStats cstats;
    uint32_t hop_na = 0, hop_groups = 0;
    double d2h_ms = 0;
    double clump_logic_ms = 0; /* run_clump_pipeline of the last FindPlanetesimals */
    double catalog_ms = 0;     /* building its plan_catalog */
    HostPinned hpin;
    std::vector<uint32_t> tidx_map; /* tree index -> member record (grow-only) */
    double density_kernel_ms = 0;
    double density_radius_ms = 0;
    uint32_t density_ovf_waves = 0;
    uint32_t density_res_queries = 0; /* queries walked in several fills */
    uint32_t density_pow_checks = 0; /* densities checked against the C library's pow */
    uint32_t density_pow_fixes = 0;  /* densities redone with the C library's pow */
    hipEvent_t ev[8];
    /* the member records reach the host in chunks, one event each */
    hipEvent_t rec_ev[8];
};

namespace plan {

/* clump trees of at most this many centres are built and searched on the
 * host (HostTree), larger ones on the device */
constexpr uint32_t HOST_CLUMP_TREE_MAX = 65536;

/* NumericalParameters after ReadLisFile (src/tree.hpp:1597-1612) and
 * CalculateNewParameters (src/global.cpp:28-47). */
void derive_params(const plan_params *p, uint32_t N, plan_derived *d)
{
    std::memset(d, 0, sizeof(*d));
    const double PI = 3.14159265358979323846;
    d->q = 1.5;
    d->Omega = 1.0;
    for (int k = 0; k < 3; ++k) {
        d->box_center[k] = (p->box_min[k] + p->box_max[k]) / 2.0;
        d->box_length[k] = p->box_max[k] - p->box_min[k];
    }
    d->max_half_width = 0.0;
    d->max_ghost_zone_width = 0.025;
    for (int k = 0; k < 3; ++k) {
        d->box_half_width[k] = d->box_length[k] / 2.0;
        d->ghost_zone_width[k] = 0.025;
        d->max_half_width = std::max(d->max_half_width, d->box_half_width[k]);
        d->max_ghost_zone_width = std::max(d->max_ghost_zone_width, d->ghost_zone_width[k]);
    }
    for (int k = 0; k < 3; ++k) d->cell_length[k] = d->box_length[k] / (double)p->box_resolution[k];
    d->cell_volume = ((1.0 * d->cell_length[0]) * d->cell_length[1]) * d->cell_length[2];
    d->grav_constant = p->four_pi_G / (4. * PI);
    d->shear_speed = (d->q * d->Omega) * d->box_length[0];
    d->min_trusted_mass_code_unit =
        ((((3 * std::pow(d->cell_length[0], 3)) * d->Omega) * d->Omega) / 1.0) / d->grav_constant;
    uint32_t nt = p->num_types > PLAN_MAX_TYPES ? PLAN_MAX_TYPES : p->num_types;
    double frac[PLAN_MAX_TYPES];
    double sum = 0;
    for (uint32_t i = 0; i < nt; ++i) {
        frac[i] = p->type_info[i];
        sum += p->type_info[i];
    }
    d->mass_total_code_units = ((std::sqrt(2 * PI) * d->box_length[0]) * d->box_length[1]) * p->solid_to_gas_ratio;
    for (uint32_t i = 0; i < nt; ++i) {
        frac[i] /= sum;
        d->mass_total_code_units = ((p->solid_to_gas_ratio * d->box_length[0]) * d->box_length[1]) * std::sqrt(2. * PI);
        d->mass_per_particle[i] = (frac[i] * d->mass_total_code_units) / (double)(p->series_first_n ? p->series_first_n : N);
    }
}

}  // namespace plan

namespace {

int fail(plan_ctx *c, int code, const std::string &msg)
{
    if (c) c->err = msg;
    return code;
}

template <class F> int guarded(plan_ctx *c, F f)
{
    if (!c) return PLAN_E_ARG;
    try {
        HIPCHK(hipSetDevice(c->device));
        return f();
    } catch (const HipError &e) {
        return fail(c, PLAN_E_DEVICE, e.what());
    } catch (const std::exception &e) {
        return fail(c, PLAN_E_INTERNAL, e.what());
    }
}

void free_tree_nodes(DevTree &t)
{
    if (t.ngeo) (void)hipFree(t.ngeo);
    if (t.ntop) (void)hipFree(t.ntop);
    if (t.ninfo) (void)hipFree(t.ninfo);
    t.ngeo = nullptr;
    t.ntop = nullptr;
    t.ninfo = nullptr;
    t.node_cap = 0;
}

void free_particles(DevParticles &p)
{
    if (p.pos) (void)hipFree(p.pos);
    if (p.vel) (void)hipFree(p.vel);
    if (p.prop) (void)hipFree(p.prop);
    if (p.id) (void)hipFree(p.id);
    p.pos = nullptr;
    p.vel = nullptr;
    p.prop = nullptr;
    p.id = nullptr;
    p.cap = 0;
}

void ensure_particles(DevParticles &p, uint32_t N)
{
    if (p.cap >= N && p.pos) return;
    free_particles(p);
    size_t cap = (size_t)N + N / 2 + 64;
    HIPCHK(hipMalloc(&p.pos, cap * sizeof(double4)));
    HIPCHK(hipMalloc(&p.vel, cap * sizeof(double4)));
    HIPCHK(hipMalloc(&p.prop, cap * sizeof(int32_t)));
    HIPCHK(hipMalloc(&p.id, cap * sizeof(uint32_t)));
    p.cap = (uint32_t)cap;
}

__global__ void k_f32_to_set(const float *__restrict__ pos, const float *__restrict__ vel,
                             const float *__restrict__ dens, uint32_t N, double4 *__restrict__ P,
                             double4 *__restrict__ V)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    P[i] = make_double4((double)pos[3 * i], (double)pos[3 * i + 1], (double)pos[3 * i + 2], (double)dens[i]);
    V[i] = make_double4((double)vel[3 * i], (double)vel[3 * i + 1], (double)vel[3 * i + 2], 0.0);
}

__global__ void k_f64_to_set(const double *__restrict__ pos, const double *__restrict__ vel,
                             const double *__restrict__ dens, uint32_t N, double4 *__restrict__ P,
                             double4 *__restrict__ V)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= N) return;
    P[i] = make_double4(pos[3 * i], pos[3 * i + 1], pos[3 * i + 2], dens[i]);
    V[i] = make_double4(vel[3 * i], vel[3 * i + 1], vel[3 * i + 2], 0.0);
}

__global__ void k_rec_of_apos(const uint4 *__restrict__ order, uint32_t na, uint32_t *__restrict__ rec_of_apos)
{
    uint32_t j = blockIdx.x * blockDim.x + threadIdx.x;
    if (j >= na) return;
    rec_of_apos[order[j].w] = j;
}

__global__ void k_clump_peaks(const uint4 *__restrict__ order, const uint32_t *__restrict__ cstart, uint32_t nc,
                              const uint32_t *__restrict__ a_idx, uint32_t *__restrict__ cpeak)
{
    uint32_t c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nc) return;
    cpeak[c] = a_idx[order[cstart[c]].z];
}

__global__ void k_hop_full(const uint32_t *__restrict__ apos, uint32_t n, const uint32_t *__restrict__ densest,
                           const uint32_t *__restrict__ root, const uint32_t *__restrict__ a_idx,
                           uint32_t *__restrict__ dn_out, uint32_t *__restrict__ pk_out)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    uint32_t a = apos[i];
    if (a == PLAN_NONE) {
        dn_out[i] = i;
        pk_out[i] = n;
    } else {
        dn_out[i] = densest[a];
        pk_out[i] = a_idx[root[a]];
    }
}

__global__ void k_split_pts(const double *__restrict__ pts, const double *__restrict__ est, uint32_t nq,
                            double4 *__restrict__ q4, double *__restrict__ e)
{
    uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    q4[i] = make_double4(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], 0.0);
    e[i] = est[i];
}

/* device services for the host clump logic */
struct Services : ClumpServices {
    plan_ctx *c;
    /* ball searches asked for by the parallel rebinding runs, waiting for the
     * device (one search batch on the stream at a time) */
    struct BallRequest {
        const std::vector<double4> *queries;
        double rho_min;
        std::vector<std::vector<uint32_t>> *out;
        bool done;
        std::exception_ptr err;
    };
    std::mutex mu;
    std::condition_variable cv;
    std::vector<BallRequest *> pending;
    bool device_busy = false;
    explicit Services(plan_ctx *cc) : c(cc) {}

    void saddle_density(const std::vector<double4> &pts, std::vector<double> &rho) override
    {
        hipStream_t s = c->stream;
        uint32_t nq = (uint32_t)pts.size();
        double4 *q = c->work.get_t<double4>("sd_q", nq);
        double *e = c->work.get_t<double>("sd_e", nq);
        double *r = c->work.get_t<double>("sd_r", nq);
        std::vector<double4> qh(nq);
        std::vector<double> eh(nq);
        for (uint32_t i = 0; i < nq; ++i) {
            qh[i] = make_double4(pts[i].x, pts[i].y, pts[i].z, 0.0);
            eh[i] = pts[i].w;
        }
        HIPCHK(hipMemcpyAsync(q, qh.data(), nq * sizeof(double4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(e, eh.data(), nq * sizeof(double), hipMemcpyHostToDevice, s));
        KnnOut ko;
        ko.rho = r;
        dev_knn(c->tree, c->K, 0, c->kernel, nq, nullptr, q, e, c->tree.rho, c->d, ko, s);
        rho.resize(nq);
        HIPCHK(hipMemcpyAsync(rho.data(), r, nq * sizeof(double), hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
    }

    void run_balls(const DevTree &t, const std::vector<double4> &queries, const double *rho, double rho_min,
                   int filter, std::vector<std::vector<uint32_t>> &out)
    {
        dev_ball_lists(t, c->work, queries, rho, rho_min, filter, out, c->stream);
    }

    void clump_tree_balls(const std::vector<double4> &com_peak, const std::vector<double4> &vel,
                          const std::vector<double4> &queries, std::vector<std::vector<uint32_t>> &peaks) override
    {
        hipStream_t s = c->stream;
        uint32_t n = (uint32_t)com_peak.size();
        peaks.assign(queries.size(), {});
        if (n == 0) return;
        if (n <= HOST_CLUMP_TREE_MAX) {
            /* a few thousand clump centres: the same tree and searches on the
             * host, without the device round trips */
            HostTree ht;
            ht.build(com_peak, c->tree.root_center, c->tree.half_width, c->d);
            const long nq = (long)queries.size();
#pragma omp parallel
            {
                std::vector<uint32_t> l;
#pragma omp for schedule(dynamic, 16)
                for (long qi = 0; qi < nq; ++qi) {
                    const double4 &q = queries[qi];
                    ht.ball(q.x, q.y, q.z, q.w, l);
                    peaks[qi].reserve(l.size());
                    for (uint32_t j : l) peaks[qi].push_back((uint32_t)ht.payload(j));
                }
            }
            return;
        }
        ensure_particles(c->cps, n);
        std::vector<int32_t> prop(n, 0);
        std::vector<uint32_t> ids(n);
        for (uint32_t i = 0; i < n; ++i) ids[i] = i;
        HIPCHK(hipMemcpyAsync(c->cps.pos, com_peak.data(), n * sizeof(double4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->cps.vel, vel.data(), n * sizeof(double4), hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->cps.prop, prop.data(), n * 4, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->cps.id, ids.data(), n * 4, hipMemcpyHostToDevice, s));
        c->cps.N = c->cps.T = n;
        dev_build_tree(c->ctree, c->cps, c->d, false, c->tree.root_center, c->tree.half_width, s);
        std::vector<std::vector<uint32_t>> idx;
        run_balls(c->ctree, queries, nullptr, 0.0, 0, idx);
        std::vector<double4> va(c->ctree.n);
        HIPCHK(hipMemcpyAsync(va.data(), c->ctree.va, c->ctree.n * sizeof(double4), hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        for (size_t qi = 0; qi < queries.size(); ++qi) {
            peaks[qi].reserve(idx[qi].size());
            for (uint32_t j : idx[qi]) peaks[qi].push_back((uint32_t)va[j].w);
        }
    }

    /* Called from the parallel rebinding runs.  A search costs about the same
     * for one query as for many (a level-synchronous walk), so the requests
     * that wait while the device is busy go to it together: the caller that
     * finds the device free takes every waiting request with its threshold,
     * runs them as one batch and hands each caller its own lists (each
     * query's list does not depend on the others in the batch). */
    void particle_balls(const std::vector<double4> &queries, double rho_min,
                        std::vector<std::vector<uint32_t>> &out) override
    {
        BallRequest me{&queries, rho_min, &out, false, nullptr};
        std::unique_lock<std::mutex> lk(mu);
        pending.push_back(&me);
        while (!me.done) {
            if (device_busy) {
                cv.wait(lk);
                continue;
            }
            device_busy = true;
            std::vector<BallRequest *> batch, rest;
            const double thr = pending.front()->rho_min;
            for (BallRequest *r : pending) (r->rho_min == thr ? batch : rest).push_back(r);
            pending.swap(rest);
            lk.unlock();
            std::exception_ptr err = nullptr;
            try {
                std::vector<double4> q;
                for (BallRequest *r : batch) q.insert(q.end(), r->queries->begin(), r->queries->end());
                std::vector<std::vector<uint32_t>> lists;
                run_balls(c->tree, q, c->tree.rho, thr, 1, lists);
                size_t k = 0;
                for (BallRequest *r : batch) {
                    r->out->resize(r->queries->size());
                    for (auto &l : *r->out) l.swap(lists[k++]);
                }
            } catch (...) {
                err = std::current_exception();
            }
            lk.lock();
            for (BallRequest *r : batch) {
                r->err = err;
                r->done = true;
            }
            device_busy = false;
            cv.notify_all();
        }
        lk.unlock();
        if (me.err) std::rethrow_exception(me.err);
    }
};

double ms_between(hipEvent_t a, hipEvent_t b)
{
    float ms = 0;
    (void)hipEventElapsedTime(&ms, a, b);
    return ms;
}

/* kernel choice and H (src/tree.hpp:3576-3612) */
void choose_kernel(plan_ctx *c, uint32_t N, int &kernel, uint32_t &H, int &enough)
{
    const plan_params &pp = c->pp;
    const plan_derived &d = c->d;
    uint32_t hres = pp.box_resolution[0] * pp.box_resolution[1];
    if (c->tree.half_width < d.max_half_width + d.max_ghost_zone_width - 1e-16)
        hres = (uint32_t)std::pow(c->tree.half_width / d.cell_length[0] *
                                      ((c->tree.half_width - d.max_ghost_zone_width) / c->tree.half_width),
                                  2.0);
    H = pp.num_hop;
    if (N >= 4 * hres) {
        if (!pp.fixed_num_hop) {
            if (N >= 1.67e7) H = 64;
            else if (N >= 2.68e8) H = 128;
        }
        enough = 1;
        kernel = 0;
    } else {
        enough = 0;
        kernel = 1;
    }
}

double density_threshold(const plan_ctx *c, uint32_t N)
{
    const plan_derived &d = c->d;
    double thr = std::pow(d.Omega, 2.) / 3.14159265358979323846 / d.grav_constant * 2.;
    if (d.mass_total_code_units / N / d.cell_volume > thr / 4) thr *= 9. / 8.;
    return thr;
}

}  // namespace

extern "C" {

namespace {
int scratch_init(int device);
}

int plan_ctx_create(plan_ctx **out, int device)
{
    if (!out) return PLAN_E_ARG;
    *out = nullptr;
    plan_ctx *c = new plan_ctx();
    c->device = device;
    try {
        HIPCHK(hipSetDevice(device));
        HIPCHK(hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking));
        for (auto &e : c->ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventBlockingSync));
        for (auto &e : c->rec_ev) HIPCHK(hipEventCreateWithFlags(&e, hipEventDisableTiming | hipEventBlockingSync));
        if (scratch_init(device) != PLAN_OK) throw std::runtime_error("scratch symbol address");
    } catch (const std::exception &e) {
        delete c;
        return PLAN_E_DEVICE;
    }
    *out = c;
    return PLAN_OK;
}

void plan_ctx_destroy(plan_ctx *c)
{
    if (!c) return;
    (void)hipSetDevice(c->device);
    if (c->stream) (void)plan::stream_wait(c->stream);
    free_tree_nodes(c->tree);
    free_tree_nodes(c->ctree);
    free_particles(c->ps);
    free_particles(c->cps);
    free_particles(c->saved);
    for (auto &e : c->ev) (void)hipEventDestroy(e);
    for (auto &e : c->rec_ev) (void)hipEventDestroy(e);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *plan_last_error(const plan_ctx *c) { return c ? c->err.c_str() : "null context"; }

static int set_common(plan_ctx *c, const plan_params *params, uint32_t N)
{
    if (!params) return fail(c, PLAN_E_ARG, "params is NULL");
    if (N == 0) return fail(c, PLAN_E_ARG, "no particles");
    if (params->num_types == 0 || params->num_types > PLAN_MAX_TYPES)
        return fail(c, PLAN_E_ARG, "num_types must be 1..16");
    for (int k = 0; k < 3; ++k) {
        if (params->box_resolution[k] == 0) return fail(c, PLAN_E_ARG, "box_resolution must be positive");
        if (!(params->box_max[k] > params->box_min[k])) return fail(c, PLAN_E_ARG, "empty box");
    }
    if (params->clump_diffuse_threshold == 0)
        return fail(c, PLAN_E_ARG, "clump_diffuse_threshold must be larger than 0 (or negative for automatic)");
    if (!knn_k_supported(params->num_knn))
        return fail(c, PLAN_E_ARG, "num_knn must be in 1..64 (neighbour lists hold at most 64 per query)");
    if (!knn_k_supported(params->num_hop))
        return fail(c, PLAN_E_ARG, "num_hop must be in 1..64 (neighbour lists hold at most 64 per query)");
    c->pp = *params;
    derive_params(params, N, &c->d);
    c->ps.N = c->ps.T = N;
    c->have_particles = true;
    c->have_tree = c->have_density = c->have_hop = false;
    return PLAN_OK;
}

int plan_set_particles(plan_ctx *c, const plan_params *params, uint32_t N, const double *pos, const double *vel,
                       const double *density, const int32_t *prop, const uint32_t *id)
{
    return guarded(c, [&]() -> int {
        if (!pos || !vel || !density || !prop || !id) return fail(c, PLAN_E_ARG, "NULL particle array");
        int r = set_common(c, params, N);
        if (r) return r;
        for (uint32_t i = 0; i < N; ++i)
            if (prop[i] < 0 || (uint32_t)prop[i] >= params->num_types)
                return fail(c, PLAN_E_ARG, "property_index outside the species table");
        ensure_particles(c->ps, N);
        hipStream_t s = c->stream;
        double *dp = c->work.get_t<double>("up_pos", (size_t)3 * N);
        double *dv = c->work.get_t<double>("up_vel", (size_t)3 * N);
        double *dd = c->work.get_t<double>("up_den", N);
        HIPCHK(hipMemcpyAsync(dp, pos, (size_t)24 * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dv, vel, (size_t)24 * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dd, density, (size_t)8 * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->ps.prop, prop, (size_t)4 * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->ps.id, id, (size_t)4 * N, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_f64_to_set, dim3(ceil_div(N, 256)), dim3(256), 0, s, dp, dv, dd, N, c->ps.pos, c->ps.vel);
        LAUNCHCHK();
        HIPCHK(plan::stream_wait(s));
        return PLAN_OK;
    });
}

int plan_set_particles_f32(plan_ctx *c, const plan_params *params, uint32_t N, const float *pos, const float *vel,
                           const float *density, const int32_t *prop, const uint32_t *id)
{
    return guarded(c, [&]() -> int {
        if (!pos || !vel || !density || !prop || !id) return fail(c, PLAN_E_ARG, "NULL particle array");
        int r = set_common(c, params, N);
        if (r) return r;
        for (uint32_t i = 0; i < N; ++i)
            if (prop[i] < 0 || (uint32_t)prop[i] >= params->num_types)
                return fail(c, PLAN_E_ARG, "property_index outside the species table");
        ensure_particles(c->ps, N);
        hipStream_t s = c->stream;
        float *dp = c->work.get_t<float>("upf_pos", (size_t)3 * N);
        float *dv = c->work.get_t<float>("upf_vel", (size_t)3 * N);
        float *dd = c->work.get_t<float>("upf_den", N);
        HIPCHK(hipMemcpyAsync(dp, pos, (size_t)12 * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dv, vel, (size_t)12 * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(dd, density, (size_t)4 * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->ps.prop, prop, (size_t)4 * N, hipMemcpyHostToDevice, s));
        HIPCHK(hipMemcpyAsync(c->ps.id, id, (size_t)4 * N, hipMemcpyHostToDevice, s));
        hipLaunchKernelGGL(k_f32_to_set, dim3(ceil_div(N, 256)), dim3(256), 0, s, dp, dv, dd, N, c->ps.pos, c->ps.vel);
        LAUNCHCHK();
        HIPCHK(plan::stream_wait(s));
        return PLAN_OK;
    });
}

/* the header's fields of plan_params (ReadLisFile, src/tree.hpp:1564-1583):
 * domain limits, time and species weights */
static void params_from_lis(const LisHeader &h, plan_params &p)
{
    for (int k = 0; k < 3; ++k) {
        p.box_min[k] = h.coor_lim[6 + 2 * k];
        p.box_max[k] = h.coor_lim[7 + 2 * k];
    }
    p.time = h.time;
    p.num_types = h.num_types;
    for (int k = 0; k < PLAN_MAX_TYPES; ++k) p.type_info[k] = k < (int)h.num_types ? h.type_info[k] : 0.0;
}

static int lis_finish(plan_ctx *c, const uint32_t *rec, uint32_t N, uint32_t num_types)
{
    if (!dev_lis_to_particles(c->work, rec, N, num_types, c->ps, c->stream)) {
        c->have_particles = false;
        return fail(c, PLAN_E_ARG, "property_index outside the species table");
    }
    return PLAN_OK;
}

int plan_read_lis(plan_ctx *c, plan_params *params, const char *const *paths, uint32_t num_paths, double *dt,
                  uint32_t *N_out)
{
    return guarded(c, [&]() -> int {
        if (!params) return fail(c, PLAN_E_ARG, "params is NULL");
        LisHeader h;
        std::vector<uint64_t> counts;
        bool io = false;
        std::string e = lis_read_headers(paths, num_paths, h, counts, io);
        if (!e.empty()) return fail(c, io ? PLAN_E_IO : PLAN_E_ARG, e);
        plan_params p = *params;
        if (p.series_first_n) p.time = h.time; /* a later snapshot of a series: box and weights as given */
        else params_from_lis(h, p);
        const uint32_t N = h.num_particles;
        int r = set_common(c, &p, N);
        if (r) return r;
        c->have_particles = false;
        ensure_particles(c->ps, N);
        DevBuf rec((size_t)N * 44);
        e = lis_upload_records(paths, num_paths, h, counts, static_cast<uint8_t *>(rec.p), c->stream);
        if (!e.empty()) return fail(c, PLAN_E_IO, e);
        r = lis_finish(c, static_cast<const uint32_t *>(rec.p), N, p.num_types);
        if (r) return r;
        c->have_particles = true;
        *params = p;
        if (dt) *dt = h.dt;
        if (N_out) *N_out = N;
        return PLAN_OK;
    });
}

int plan_set_particles_lis(plan_ctx *c, const plan_params *params, uint32_t N, const void *records)
{
    return guarded(c, [&]() -> int {
        if (!records) return fail(c, PLAN_E_ARG, "NULL records");
        int r = set_common(c, params, N);
        if (r) return r;
        c->have_particles = false;
        ensure_particles(c->ps, N);
        DevBuf rec((size_t)N * 44);
        HIPCHK(hipMemcpyAsync(rec.p, records, (size_t)N * 44, hipMemcpyHostToDevice, c->stream));
        r = lis_finish(c, static_cast<const uint32_t *>(rec.p), N, params->num_types);
        if (r) return r;
        c->have_particles = true;
        return PLAN_OK;
    });
}

int plan_get_particle_set(plan_ctx *c, uint32_t cap, double *pos, double *vel, double *density, int32_t *prop,
                          uint32_t *id, uint32_t *count)
{
    return guarded(c, [&]() -> int {
        if (!c->have_particles) return fail(c, PLAN_E_STATE, "no particle set loaded");
        const uint32_t T = c->ps.T, n = cap < T ? cap : T;
        if (count) *count = T;
        if (n == 0) return PLAN_OK;
        hipStream_t s = c->stream;
        std::vector<double4> P(n), V(n);
        HIPCHK(hipMemcpyAsync(P.data(), c->ps.pos, (size_t)n * 32, hipMemcpyDeviceToHost, s));
        HIPCHK(hipMemcpyAsync(V.data(), c->ps.vel, (size_t)n * 32, hipMemcpyDeviceToHost, s));
        if (prop) HIPCHK(hipMemcpyAsync(prop, c->ps.prop, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        if (id) HIPCHK(hipMemcpyAsync(id, c->ps.id, (size_t)n * 4, hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        for (uint32_t i = 0; i < n; ++i) {
            if (pos) {
                pos[3 * i] = P[i].x;
                pos[3 * i + 1] = P[i].y;
                pos[3 * i + 2] = P[i].z;
            }
            if (vel) {
                vel[3 * i] = V[i].x;
                vel[3 * i + 1] = V[i].y;
                vel[3 * i + 2] = V[i].z;
            }
            if (density) density[i] = P[i].w;
        }
        return PLAN_OK;
    });
}

int plan_make_ghost_particles(plan_ctx *c, uint32_t *T_out)
{
    return guarded(c, [&]() -> int {
        if (!c->have_particles) return fail(c, PLAN_E_STATE, "no particle set loaded");
        HIPCHK(hipEventRecord(c->ev[0], c->stream));
        if (!c->pp.no_ghost) dev_make_ghosts(c->ps, c->pp, c->d, c->stream);
        HIPCHK(hipEventRecord(c->ev[1], c->stream));
        HIPCHK(hipEventSynchronize(c->ev[1]));
        c->times[0] = ms_between(c->ev[0], c->ev[1]);
        if (T_out) *T_out = c->ps.T;
        return PLAN_OK;
    });
}

int plan_build_tree(plan_ctx *c, int check_pos, plan_tree_info *info)
{
    return guarded(c, [&]() -> int {
        if (!c->have_particles) return fail(c, PLAN_E_STATE, "no particle set loaded");
        HIPCHK(hipEventRecord(c->ev[0], c->stream));
        double zc[3] = {0, 0, 0};
        dev_build_tree(c->tree, c->ps, c->d, check_pos != 0, zc, 0.0, c->stream);
        HIPCHK(hipEventRecord(c->ev[1], c->stream));
        HIPCHK(hipEventSynchronize(c->ev[1]));
        c->times[1] = ms_between(c->ev[0], c->ev[1]);
        c->have_tree = true;
        c->have_density = c->have_hop = false;
        if (info) {
            std::memset(info, 0, sizeof(*info));
            info->num_particles = c->tree.n;
            info->num_nodes = c->tree.num_nodes;
            info->num_leaf_nodes = c->tree.num_leaves;
            info->num_sink_particles = c->tree.num_sinks;
            info->depth = (uint32_t)c->tree.depth;
            for (int k = 0; k < 3; ++k) info->root_center[k] = c->tree.root_center[k];
            info->half_width = c->tree.half_width;
            for (int l = 0; l < 32; ++l) info->level_count[l] = c->tree.level_count[l];
            int kernel, enough;
            uint32_t H;
            choose_kernel(c, c->ps.N, kernel, H, enough);
            info->hop_auto = (enough && !c->pp.fixed_num_hop && c->ps.N >= 1.67e7) ? 1u : 0u;
        }
        return PLAN_OK;
    });
}

static int run_density(plan_ctx *c, uint32_t K, int kernel)
{
    if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
    if (!knn_k_supported(K) || K > c->tree.n) return fail(c, PLAN_E_ARG, "K must be in 1..min(64, num_particles)");
    if (kernel < 0) {
        uint32_t H;
        int enough;
        choose_kernel(c, c->ps.N, kernel, H, enough);
    }
    c->K = K;
    c->kernel = kernel;
    KnnOut ko;
    ko.rho = c->tree.rho;
    ko.ev = &c->ev[5];
    c->density_ovf_waves = 0;
    c->density_res_queries = 0;
    c->density_pow_checks = 0;
    c->density_pow_fixes = 0;
    ko.novf = &c->density_ovf_waves;
    ko.nres = &c->density_res_queries;
    ko.nchk = &c->density_pow_checks;
    ko.npow = &c->density_pow_fixes;
    HIPCHK(hipEventRecord(c->ev[2], c->stream));
    dev_knn(c->tree, K, 0, kernel, c->tree.n, nullptr, nullptr, nullptr, nullptr, c->d, ko, c->stream);
    HIPCHK(hipEventRecord(c->ev[3], c->stream));
    HIPCHK(hipEventSynchronize(c->ev[3]));
    c->density_kernel_ms = ms_between(c->ev[6], c->ev[7]);
    c->density_radius_ms = ms_between(c->ev[5], c->ev[6]);
    c->times[2] = ms_between(c->ev[2], c->ev[3]);
    c->have_density = true;
    return PLAN_OK;
}

int plan_knn_density(plan_ctx *c, uint32_t K, int kernel)
{
    return guarded(c, [&]() -> int { return run_density(c, K, kernel); });
}

static int run_hop(plan_ctx *c, uint32_t H, double thr)
{
    if (!c->have_density) return fail(c, PLAN_E_STATE, "densities not computed");
    if (!knn_k_supported(H) || H > c->tree.n) return fail(c, PLAN_E_ARG, "H must be in 1..min(64, num_particles)");
    HIPCHK(hipEventRecord(c->ev[0], c->stream));
    dev_hop(c->tree, c->work, c->d, thr, H, c->hop, c->stream);
    HIPCHK(hipEventRecord(c->ev[4], c->stream));
    HIPCHK(hipEventSynchronize(c->ev[4]));
    c->times[3] = ms_between(c->ev[0], c->ev[4]);
    c->H = H;
    c->have_hop = true;
    return PLAN_OK;
}

int plan_hop(plan_ctx *c, uint32_t H, double threshold, uint32_t *num_above, uint32_t *num_groups)
{
    return guarded(c, [&]() -> int {
        int r = run_hop(c, H, threshold);
        if (r) return r;
        if (num_above) *num_above = c->hop.na;
        if (num_groups) *num_groups = c->hop.nclumps;
        return PLAN_OK;
    });
}

static std::mutex &host_phase_mutex()
{
    static std::mutex m;
    return m;
}

int plan_find_planetesimals(plan_ctx *c, plan_catalog **out)
{
    return guarded(c, [&]() -> int {
        if (!out) return fail(c, PLAN_E_ARG, "out is NULL");
        *out = nullptr;
        if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
        auto t0 = std::chrono::steady_clock::now();
        const uint32_t N = c->ps.N;
        int kernel, enough;
        uint32_t H;
        choose_kernel(c, N, kernel, H, enough);
        uint32_t K = c->pp.num_knn;
        if (K == 0 || K > c->tree.n || H > c->tree.n) return fail(c, PLAN_E_ARG, "neighbour count exceeds particles");
        const double thr = density_threshold(c, N);
        int r = run_density(c, K, kernel);
        if (r) return r;
        r = run_hop(c, H, thr);
        if (r) return r;
        /* the host part (member table, clump logic, catalog) uses every host
         * core: contexts take turns at it (two at once would oversubscribe
         * the cores), while another context's kernels keep the GPU busy */
        std::unique_lock<std::mutex> host_turn(host_phase_mutex());
        auto t1 = std::chrono::steady_clock::now();
        /* member table to the host */
        hipStream_t s = c->stream;
        HostMembers hm;
        std::thread rec_waiter;
        struct Join {
            std::thread &t;
            ~Join()
            {
                if (t.joinable()) t.join();
            }
        } join_waiter{rec_waiter};
        const uint32_t na = c->hop.na, nc = c->hop.nclumps;
        {
            /* pinned, grow-only staging for the member table */
            const size_t rb = ((size_t)na * sizeof(MemberRec) + 255) & ~(size_t)255;
            const size_t ub = ((size_t)na * 4 + 255) & ~(size_t)255;
            char *base = (char *)c->hpin.get(rb + 3 * ub + 256);
            hm.rec = HostSpan<MemberRec>{(MemberRec *)base, na};
            hm.sinkcnt = HostSpan<uint32_t>{(uint32_t *)(base + rb), na};
            hm.sorted_tidx = HostSpan<uint32_t>{(uint32_t *)(base + rb + ub), na};
            hm.rec_of_sorted = HostSpan<uint32_t>{(uint32_t *)(base + rb + 2 * ub), na};
        }
        hm.cstart.resize(nc + 1);
        hm.cpeak.resize(nc);
        if (na) {
            uint32_t *roa = c->work.get_t<uint32_t>("rec_of_apos", na);
            uint32_t *cpk = c->work.get_t<uint32_t>("cpeak", nc + 1);
            hipLaunchKernelGGL(k_rec_of_apos, dim3(ceil_div(na, 256)), dim3(256), 0, s, c->hop.order, na, roa);
            LAUNCHCHK();
            hipLaunchKernelGGL(k_clump_peaks, dim3(ceil_div(nc, 256)), dim3(256), 0, s, c->hop.order, c->hop.cstart,
                               nc, c->hop.a_idx, cpk);
            LAUNCHCHK();
            HIPCHK(hipMemcpyAsync(hm.sinkcnt.data(), c->hop.msink, (size_t)na * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(hm.cstart.data(), c->hop.cstart, (size_t)nc * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(hm.cpeak.data(), cpk, (size_t)nc * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(hm.sorted_tidx.data(), c->hop.a_idx, (size_t)na * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(hipMemcpyAsync(hm.rec_of_sorted.data(), roa, (size_t)na * 4, hipMemcpyDeviceToHost, s));
            HIPCHK(plan::stream_wait(s));
            /* the records in chunks: the naive clumps' first pass starts on
             * the clumps whose records have arrived while the rest stream */
            constexpr uint32_t NCH = sizeof(c->rec_ev) / sizeof(c->rec_ev[0]);
            const uint32_t step = (na + NCH - 1) / NCH;
            std::vector<uint32_t> ends;
            for (uint32_t k = 0; k < NCH && k * step < na; ++k) {
                const uint32_t b = k * step, e = std::min(na, b + step);
                HIPCHK(hipMemcpyAsync(hm.rec.data() + b, c->hop.members + b, (size_t)(e - b) * sizeof(MemberRec),
                                      hipMemcpyDeviceToHost, s));
                HIPCHK(hipEventRecord(c->rec_ev[k], s));
                ends.push_back(e);
            }
            /* one thread waits for the chunks in turn and publishes how many
             * records have arrived; the clump workers poll that count */
            auto ready = std::make_shared<std::atomic<uint32_t>>(0u);
            auto failed = std::make_shared<std::atomic<bool>>(false);
            rec_waiter = std::thread([c, ends, ready, failed]() {
                for (size_t k = 0; k < ends.size(); ++k) {
                    if (hipEventSynchronize(c->rec_ev[k]) != hipSuccess) {
                        failed->store(true);
                        return;
                    }
                    ready->store(ends[k], std::memory_order_release);
                }
            });
            hm.wait_records = [ready, failed](uint32_t upto) {
                while (ready->load(std::memory_order_acquire) < upto) {
                    if (failed->load()) throw std::runtime_error("member record copy to the host failed");
                    std::this_thread::yield();
                }
            };
        }
        hm.cstart[nc] = na;
        if (c->tidx_map.size() < c->tree.n) c->tidx_map.resize(c->tree.n);
        {
            uint32_t *map = c->tidx_map.data();
            const uint32_t *st = hm.sorted_tidx.data(), *ro = hm.rec_of_sorted.data();
#pragma omp parallel for schedule(static)
            for (long k = 0; k < (long)na; ++k) map[st[k]] = ro[k];
            hm.tidx_map = map;
            hm.tidx_map_n = c->tree.n;
        }
        ClumpSettings cs;
        cs.K = K;
        cs.kernel = kernel;
        cs.enough_resolution = enough;
        cs.density_threshold = thr;
        cs.peak_density_threshold = 3 * thr;
        cs.clump_mass_threshold = c->d.min_trusted_mass_code_unit;
        {
            double mn = 1e300;
            for (int k = 0; k < 3; ++k) mn = std::min(mn, (double)c->pp.box_resolution[k] / c->d.box_length[k]);
            unsigned hrh = (unsigned)mn;
            cs.diffuse = 0.55;
            cs.hill_frac = 0.25;
            if (hrh <= 1024) {
                cs.diffuse = 0.35;
                cs.hill_frac = 0.75;
            } else if (hrh <= 1536) {
                cs.diffuse = 0.4;
                cs.hill_frac = 0.5;
            } else if (hrh <= 2048) {
                cs.diffuse = 0.5;
                cs.hill_frac = 0.35;
            }
            if (c->pp.clump_diffuse_threshold > 0) cs.diffuse = c->pp.clump_diffuse_threshold;
            if (c->pp.Hill_fraction_for_merge > 0) cs.hill_frac = c->pp.Hill_fraction_for_merge;
        }
        cs.G = c->d.grav_constant;
        cs.Omega = c->d.Omega;
        cs.q = c->d.q;
        for (int k = 0; k < 3; ++k) {
            cs.box_min[k] = c->pp.box_min[k];
            cs.box_max[k] = c->pp.box_max[k];
        }
        cs.num_peaks = c->pp.num_peaks;
        cs.ball_margin_abs = 2.5 * c->tree.epsilon;
        Services svc(c);
        std::map<uint32_t, HClump> cl;
        double Mp_max = 0, Mp_tot = 0;
        ClumpStats st;
        c->d2h_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t1).count();
        const auto tc0 = std::chrono::steady_clock::now();
        run_clump_pipeline(hm, cs, svc, cl, Mp_max, Mp_tot, st);
        const auto tc1 = std::chrono::steady_clock::now();
        c->clump_logic_ms = std::chrono::duration<double, std::milli>(tc1 - tc0).count();
        c->cstats = st;
        c->hop_na = na;
        c->hop_groups = nc;
        /* catalog */
        plan_catalog *cat = (plan_catalog *)std::calloc(1, sizeof(plan_catalog));
        uint32_t tot = 0;
        for (auto &it : cl) tot += (uint32_t)it.second.particles.size();
        cat->num_clumps = (uint32_t)cl.size();
        cat->num_members_total = tot;
        cat->clumps = (plan_clump *)std::calloc(cl.size() ? cl.size() : 1, sizeof(plan_clump));
        cat->members = (uint32_t *)std::calloc(tot ? tot : 1, 4);
        cat->member_original_ids = (uint32_t *)std::calloc(tot ? tot : 1, 4);
        cat->density_threshold = thr;
        cat->peak_density_threshold = cs.peak_density_threshold;
        cat->clump_mass_threshold = cs.clump_mass_threshold;
        cat->Mp_max = Mp_max;
        cat->Mp_tot = Mp_tot;
        /* clump fields and member offsets in key order, then every clump's
         * members (records scattered in distance order) in parallel */
        std::vector<const HClump *> qs;
        uint32_t off = 0, i = 0;
        for (auto &it : cl) {
            const HClump &q = it.second;
            plan_clump &o = cat->clumps[i++];
            qs.push_back(&q);
            o.peak_index = it.first;
            o.num_members = (uint32_t)q.particles.size();
            o.member_offset = off;
            o.total_mass = q.total_mass;
            for (int k = 0; k < 3; ++k) {
                o.center_of_mass[k] = q.com[k];
                o.vel_com[k] = q.vel[k];
                o.J[k] = q.J[k];
                o.J_quarter_Hill[k] = q.Jq[k];
            }
            o.Hill_radius = q.hill;
            o.one10th_radius = q.one10th;
            o.two_sigma_mass_radius = q.two_sigma;
            o.max_distance = q.particles.empty() ? 0.0 : q.particles.back().second;
            uint32_t kr = hm.rec_of_tidx(it.first);
            o.peak_density = kr == PLAN_NONE ? 0.0 : hm.rec[kr].rho;
            off += o.num_members;
        }
        const long ncl = (long)qs.size();
#pragma omp parallel for schedule(dynamic, 1)
        for (long k = 0; k < ncl; ++k) {
            const HClump &q = *qs[k];
            plan_clump &o = cat->clumps[k];
            const size_t nm = q.particles.size();
            uint32_t np = 0, w = o.member_offset;
            for (size_t j = 0; j < nm; ++j) {
                if (j + 12 < nm) {
                    __builtin_prefetch(&hm.rec[q.particles[j + 12].first].rho);
                    __builtin_prefetch(&hm.sinkcnt[q.particles[j + 12].first]);
                }
                const uint32_t h = q.particles[j].first;
                const MemberRec &r = hm.rec[h];
                cat->members[w] = r.tidx;
                cat->member_original_ids[w] = r.oid;
                np += hm.sinkcnt[h];
                ++w;
            }
            o.num_particles = np;
        }
        auto t2 = std::chrono::steady_clock::now();
        c->catalog_ms = std::chrono::duration<double, std::milli>(t2 - tc1).count();
        c->times[4] = std::chrono::duration<double, std::milli>(t2 - t1).count();
        c->times[5] = std::chrono::duration<double, std::milli>(t2 - t0).count();
        *out = cat;
        return PLAN_OK;
    });
}

/* members of cat on the device tree they index (the context's current one) */
static int load_members(plan_ctx *c, const plan_catalog *cat, CatalogMembers &mem)
{
    if (!cat) return fail(c, PLAN_E_ARG, "catalog is NULL");
    if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
    catalog_members(c->tree, c->ps, *cat, c->work, c->stream, mem);
    return PLAN_OK;
}

int plan_write_peaks(plan_ctx *c, const plan_catalog *cat, const char *path)
{
    return guarded(c, [&]() -> int {
        if (!path) return fail(c, PLAN_E_ARG, "path is NULL");
        CatalogMembers mem;
        int r = load_members(c, cat, mem);
        if (r) return r;
        const std::string e = write_peaks(*cat, mem, path);
        return e.empty() ? PLAN_OK : fail(c, PLAN_E_IO, e);
    });
}

int plan_write_clump_particles(plan_ctx *c, const plan_catalog *cat, const char *dir, uint32_t sampling_rate)
{
    return guarded(c, [&]() -> int {
        if (!dir) return fail(c, PLAN_E_ARG, "dir is NULL");
        CatalogMembers mem;
        int r = load_members(c, cat, mem);
        if (r) return r;
        const std::string e = write_clump_particles(*cat, mem, dir, sampling_rate);
        return e.empty() ? PLAN_OK : fail(c, PLAN_E_IO, e);
    });
}

int plan_write_basic_results(plan_ctx *c, const plan_catalog *cat, const char *path, int header)
{
    return guarded(c, [&]() -> int {
        if (!cat || !path) return fail(c, PLAN_E_ARG, "catalog or path is NULL");
        if (!c->have_particles) return fail(c, PLAN_E_STATE, "no particle set loaded");
        const double G_tilde = c->pp.four_pi_G / c->d.Omega / c->d.Omega; /* src/global.cpp:40 */
        const std::string e =
            write_basic_results(*cat, c->d.mass_total_code_units, G_tilde, c->pp.time, path, header != 0);
        return e.empty() ? PLAN_OK : fail(c, PLAN_E_IO, e);
    });
}

int plan_density_vs_scale(plan_ctx *c, uint32_t cap, double *radii, double *max_rhop, uint64_t *sums,
                          uint32_t *num_levels, const char *path, int header)
{
    return guarded(c, [&]() -> int {
        if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
        ScaleResult r;
        dev_density_vs_scale(c->tree, c->pp, c->d, c->work, c->stream, r);
        const uint32_t L = (uint32_t)r.radii.size();
        if (num_levels) *num_levels = L;
        for (uint32_t i = 0; i < L && i < cap; ++i) {
            if (radii) radii[i] = r.radii[i];
            if (max_rhop) max_rhop[i] = r.max_rhop[i];
            if (sums) sums[i] = r.sums[i];
        }
        if (!path) return PLAN_OK;
        /* the max_rhop_vs_scale file (src/analyses.hpp:255-264, 326-331):
         * width 15, scientific, default precision */
        std::FILE *f = std::fopen(path, "a");
        if (!f) return fail(c, PLAN_E_IO, std::string("Failed to open file ") + path);
        if (header) {
            std::fprintf(f, "%15s", "#time");
            for (double x : r.radii) std::fprintf(f, "%15.6e", x);
            std::fprintf(f, "\n");
        }
        std::fprintf(f, "%15.6e", c->pp.time);
        for (double x : r.max_rhop) std::fprintf(f, "%15.6e", x);
        std::fprintf(f, "\n");
        const bool ok = std::fclose(f) == 0;
        return ok ? PLAN_OK : fail(c, PLAN_E_IO, std::string("Failed to write file ") + path);
    });
}

void plan_catalog_free(plan_catalog *cat)
{
    if (!cat) return;
    std::free(cat->clumps);
    std::free(cat->members);
    std::free(cat->member_original_ids);
    std::free(cat);
}

int plan_get_tree_particles(plan_ctx *c, double *pos, double *vel, double *mass, double *ath, double *rho,
                            uint32_t *oid, uint32_t *sinkcnt)
{
    return guarded(c, [&]() -> int {
        if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
        const uint32_t n = c->tree.n;
        std::vector<double4> pm(n), va(n);
        HIPCHK(hipMemcpy(pm.data(), c->tree.pm, (size_t)n * 32, hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(va.data(), c->tree.va, (size_t)n * 32, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) {
            if (pos) {
                pos[3 * i] = pm[i].x;
                pos[3 * i + 1] = pm[i].y;
                pos[3 * i + 2] = pm[i].z;
            }
            if (vel) {
                vel[3 * i] = va[i].x;
                vel[3 * i + 1] = va[i].y;
                vel[3 * i + 2] = va[i].z;
            }
            if (mass) mass[i] = pm[i].w;
            if (ath) ath[i] = va[i].w;
        }
        if (rho) HIPCHK(hipMemcpy(rho, c->tree.rho, (size_t)n * 8, hipMemcpyDeviceToHost));
        if (oid) HIPCHK(hipMemcpy(oid, c->tree.oid, (size_t)n * 4, hipMemcpyDeviceToHost));
        if (sinkcnt) HIPCHK(hipMemcpy(sinkcnt, c->tree.sinkcnt, (size_t)n * 4, hipMemcpyDeviceToHost));
        return PLAN_OK;
    });
}

int plan_get_keys(plan_ctx *c, uint32_t *w0, uint32_t *w1, uint32_t *w2)
{
    return guarded(c, [&]() -> int {
        if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
        const uint32_t n = c->tree.n;
        std::vector<uint4> k(n);
        HIPCHK(hipMemcpy(k.data(), c->tree.keys, (size_t)n * 16, hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) {
            if (w0) w0[i] = k[i].x;
            if (w1) w1[i] = k[i].y;
            if (w2) w2[i] = k[i].z;
        }
        return PLAN_OK;
    });
}

int plan_get_nodes(plan_ctx *c, double *center, double *hw, uint32_t *begin, uint32_t *end, uint32_t *parent,
                   uint32_t *first_daughter, uint8_t *num_daughters, uint8_t *level, uint16_t *orthant,
                   int32_t *node2leaf, int32_t *leaf_nodes)
{
    return guarded(c, [&]() -> int {
        if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
        const uint32_t nn = c->tree.num_nodes;
        std::vector<NodeGeo> g(nn);
        std::vector<NodeTop> t(nn);
        std::vector<uint32_t> info(nn);
        HIPCHK(hipMemcpy(g.data(), c->tree.ngeo, (size_t)nn * sizeof(NodeGeo), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(t.data(), c->tree.ntop, (size_t)nn * sizeof(NodeTop), hipMemcpyDeviceToHost));
        HIPCHK(hipMemcpy(info.data(), c->tree.ninfo, (size_t)nn * 4, hipMemcpyDeviceToHost));
        for (uint32_t v = 0; v < nn; ++v) {
            if (center) {
                center[3 * v] = g[v].cx;
                center[3 * v + 1] = g[v].cy;
                center[3 * v + 2] = g[v].cz;
            }
            if (hw) hw[v] = g[v].hw;
            if (begin) begin[v] = t[v].begin;
            if (end) end[v] = t[v].end;
            if (parent) parent[v] = t[v].parent;
            if (first_daughter) first_daughter[v] = t[v].first_daughter;
            if (num_daughters) num_daughters[v] = (uint8_t)(info[v] & 0xff);
            if (level) level[v] = (uint8_t)((info[v] >> 8) & 0xff);
            if (orthant) orthant[v] = (uint16_t)(info[v] >> 16);
        }
        if (node2leaf) HIPCHK(hipMemcpy(node2leaf, c->tree.node2leaf, (size_t)nn * 4, hipMemcpyDeviceToHost));
        if (leaf_nodes)
            HIPCHK(hipMemcpy(leaf_nodes, c->tree.leaf_nodes, (size_t)c->tree.num_leaves * 4, hipMemcpyDeviceToHost));
        return PLAN_OK;
    });
}

int plan_get_hop(plan_ctx *c, uint32_t *densest, uint32_t *peak)
{
    return guarded(c, [&]() -> int {
        if (!c->have_hop) return fail(c, PLAN_E_STATE, "hop not run");
        const uint32_t n = c->tree.n;
        uint32_t *dn = c->work.get_t<uint32_t>("gh_dn", n);
        uint32_t *pk = c->work.get_t<uint32_t>("gh_pk", n);
        hipLaunchKernelGGL(k_hop_full, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, c->hop.apos, n,
                           c->hop.densest, c->hop.root, c->hop.a_idx, dn, pk);
        LAUNCHCHK();
        HIPCHK(plan::stream_wait(c->stream));
        if (densest) HIPCHK(hipMemcpy(densest, dn, (size_t)n * 4, hipMemcpyDeviceToHost));
        if (peak) HIPCHK(hipMemcpy(peak, pk, (size_t)n * 4, hipMemcpyDeviceToHost));
        return PLAN_OK;
    });
}

int plan_knn_query(plan_ctx *c, uint32_t nq, const uint32_t *pidx, const double *pts, const double *est, uint32_t K,
                   uint32_t *idx, double *d2, double *rstop)
{
    return guarded(c, [&]() -> int {
        if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
        if (!pidx && (!pts || !est)) return fail(c, PLAN_E_ARG, "point queries need pts and est");
        if (!knn_k_supported(K) || K > c->tree.n) return fail(c, PLAN_E_ARG, "K must be in 1..min(64, num_particles)");
        if (nq == 0) return PLAN_OK;
        hipStream_t s = c->stream;
        uint32_t *di = c->work.get_t<uint32_t>("kq_idx", (size_t)nq * K);
        double *dd = c->work.get_t<double>("kq_d2", (size_t)nq * K);
        double *dr = c->work.get_t<double>("kq_r", nq);
        KnnOut ko;
        ko.idx = di;
        ko.d2 = dd;
        ko.rstop = dr;
        if (pidx) {
            for (uint32_t i = 0; i < nq; ++i)
                if (pidx[i] >= c->tree.n) return fail(c, PLAN_E_ARG, "particle index out of range");
            uint32_t *dq = c->work.get_t<uint32_t>("kq_q", nq);
            HIPCHK(hipMemcpyAsync(dq, pidx, (size_t)nq * 4, hipMemcpyHostToDevice, s));
            double *de = nullptr;
            if (est) {
                de = c->work.get_t<double>("kq_e", nq);
                HIPCHK(hipMemcpyAsync(de, est, (size_t)nq * 8, hipMemcpyHostToDevice, s));
            }
            dev_knn(c->tree, K, 2, 0, nq, dq, nullptr, de, c->tree.rho, c->d, ko, s);
        } else {
            for (uint32_t i = 0; i < nq; ++i)
                if (!(est[i] > 0)) return fail(c, PLAN_E_ARG, "point queries need a positive start radius");
            double *dp = c->work.get_t<double>("kq_p", (size_t)nq * 3);
            double *de = c->work.get_t<double>("kq_e", nq);
            double4 *q4 = c->work.get_t<double4>("kq_q4", nq);
            double *e2 = c->work.get_t<double>("kq_e2", nq);
            HIPCHK(hipMemcpyAsync(dp, pts, (size_t)nq * 24, hipMemcpyHostToDevice, s));
            HIPCHK(hipMemcpyAsync(de, est, (size_t)nq * 8, hipMemcpyHostToDevice, s));
            hipLaunchKernelGGL(k_split_pts, dim3(ceil_div(nq, 256)), dim3(256), 0, s, dp, de, nq, q4, e2);
            LAUNCHCHK();
            dev_knn(c->tree, K, 2, 0, nq, nullptr, q4, e2, c->tree.rho, c->d, ko, s);
        }
        HIPCHK(plan::stream_wait(s));
        if (idx) HIPCHK(hipMemcpy(idx, di, (size_t)nq * K * 4, hipMemcpyDeviceToHost));
        if (d2) HIPCHK(hipMemcpy(d2, dd, (size_t)nq * K * 8, hipMemcpyDeviceToHost));
        if (rstop) HIPCHK(hipMemcpy(rstop, dr, (size_t)nq * 8, hipMemcpyDeviceToHost));
        return PLAN_OK;
    });
}

/* ---- the KNN query entry point that allocates nothing ----------------------
 *
 * Its working buffers (radii, counts, overflow flags and scans, the best
 * lists of queries walked in several fills, the packed query points) are
 * carved from `scratch`, a fixed array in device memory that each GPU holds
 * until the process ends.  A kernel needs its device address, which
 * hipGetSymbolAddress returns once per device (the array's name in host code
 * is only a host-side stand-in); plan_ctx_create does that with the device
 * current.  Every call on one device shares the array: calls that overlap on
 * one GPU race. */
static __device__ __align__(256) unsigned char scratch[PLAN_NOALLOC_SCRATCH_BYTES];

namespace {
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
void *g_scratch_ptr[kMaxDevices] = {};

/* device address of this device's scratch array (the device is current) */
int scratch_init(int device)
{
    if (device < 0 || device >= kMaxDevices) return PLAN_E_ARG;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    if (g_scratch_ptr[device]) return PLAN_OK;
    void *p = nullptr;
    if (hipGetSymbolAddress(&p, HIP_SYMBOL(scratch)) != hipSuccess || !p) return PLAN_E_DEVICE;
    g_scratch_ptr[device] = p;
    return PLAN_OK;
}

void *scratch_ptr(int device)
{
    if (device < 0 || device >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> lk(g_scratch_mu);
    return g_scratch_ptr[device];
}
}  // namespace

/* bad[0] += queries with a particle index out of range or, for point
 * queries, a start radius that is not positive */
__global__ void k_check_queries(const uint32_t *__restrict__ pidx, const double *__restrict__ est, uint32_t nq,
                                uint32_t n, uint32_t *__restrict__ bad)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nq) return;
    const bool wrong = pidx ? pidx[i] >= n : !(est[i] > 0);
    if (wrong) atomicAdd(bad, 1u);
}

int plan_knn_query_noalloc(plan_ctx *c, uint32_t nq, const uint32_t *pidx, const double *pts, const double *est,
                           uint32_t K, uint32_t *idx, double *d2, double *rstop)
{
    return guarded(c, [&]() -> int {
        if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
        if (!pidx && (!pts || !est)) return fail(c, PLAN_E_ARG, "point queries need pts and est");
        if (!knn_k_supported(K) || K > c->tree.n) return fail(c, PLAN_E_ARG, "K must be in 1..min(64, num_particles)");
        if (nq > PLAN_NOALLOC_MAX_QUERIES)
            return fail(c, PLAN_E_ARG, "plan_knn_query_noalloc: at most PLAN_NOALLOC_MAX_QUERIES queries per call");
        if (nq == 0) return PLAN_OK;
        void *base = scratch_ptr(c->device);
        if (!base) return fail(c, PLAN_E_STATE, "scratch array not set up on this device");
        hipStream_t s = c->stream;
        Arena a(base, PLAN_NOALLOC_SCRATCH_BYTES);
        uint32_t *bad = a.get_t<uint32_t>("kq_bad", 4);
        HIPCHK(hipMemsetAsync(bad, 0, 4, s));
        hipLaunchKernelGGL(k_check_queries, dim3(ceil_div(nq, 256)), dim3(256), 0, s, pidx, est, nq, c->tree.n, bad);
        LAUNCHCHK();
        uint32_t nbad = 0;
        HIPCHK(hipMemcpyAsync(&nbad, bad, 4, hipMemcpyDeviceToHost, s));
        HIPCHK(plan::stream_wait(s));
        if (nbad)
            return fail(c, PLAN_E_ARG, pidx ? "particle index out of range"
                                            : "point queries need a positive start radius");
        KnnOut ko;
        ko.idx = idx;
        ko.d2 = d2;
        ko.rstop = rstop;
        if (pidx) {
            dev_knn(c->tree, K, 2, 0, nq, pidx, nullptr, est, c->tree.rho, c->d, ko, s, &a);
        } else {
            double4 *q4 = a.get_t<double4>("kq_q4", nq);
            double *e2 = a.get_t<double>("kq_e2", nq);
            hipLaunchKernelGGL(k_split_pts, dim3(ceil_div(nq, 256)), dim3(256), 0, s, pts, est, nq, q4, e2);
            LAUNCHCHK();
            dev_knn(c->tree, K, 2, 0, nq, nullptr, q4, e2, c->tree.rho, c->d, ko, s, &a);
        }
        HIPCHK(plan::stream_wait(s));
        return PLAN_OK;
    });
}

int plan_density_at(plan_ctx *c, uint32_t nq, const double *pts, const double *est, uint32_t K, int kernel,
                    double *rho)
{
    return guarded(c, [&]() -> int {
        if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
        if (!pts || !est || !rho) return fail(c, PLAN_E_ARG, "NULL argument");
        if (!knn_k_supported(K) || K > c->tree.n) return fail(c, PLAN_E_ARG, "K must be in 1..min(64, num_particles)");
        std::vector<double4> q(nq);
        for (uint32_t i = 0; i < nq; ++i) {
            if (!(est[i] > 0)) return fail(c, PLAN_E_ARG, "point queries need a positive start radius");
            q[i] = make_double4(pts[3 * i], pts[3 * i + 1], pts[3 * i + 2], est[i]);
        }
        uint32_t oldK = c->K;
        int oldk = c->kernel;
        c->K = K;
        c->kernel = kernel;
        std::vector<double> r;
        Services svc(c);
        svc.saddle_density(q, r);
        c->K = oldK;
        c->kernel = oldk;
        std::memcpy(rho, r.data(), (size_t)nq * 8);
        return PLAN_OK;
    });
}

int plan_ball_query(plan_ctx *c, const double *center, double radius, uint32_t *out, uint32_t cap, uint32_t *count)
{
    return guarded(c, [&]() -> int {
        if (!c->have_tree) return fail(c, PLAN_E_STATE, "no tree built");
        if (!center || !count) return fail(c, PLAN_E_ARG, "NULL argument");
        Services svc(c);
        std::vector<std::vector<uint32_t>> l;
        svc.run_balls(c->tree, {make_double4(center[0], center[1], center[2], radius)}, nullptr, 0.0, 0, l);
        *count = (uint32_t)l[0].size();
        if (out)
            for (uint32_t i = 0; i < *count && i < cap; ++i) out[i] = l[0][i];
        return PLAN_OK;
    });
}

int plan_keep_input(plan_ctx *c)
{
    return guarded(c, [&]() -> int {
        if (!c->have_particles || c->ps.T != c->ps.N) return fail(c, PLAN_E_STATE, "load particles before ghosts");
        const uint32_t N = c->ps.N;
        ensure_particles(c->saved, N);
        hipStream_t s = c->stream;
        HIPCHK(hipMemcpyAsync(c->saved.pos, c->ps.pos, (size_t)N * 32, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(c->saved.vel, c->ps.vel, (size_t)N * 32, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(c->saved.prop, c->ps.prop, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(c->saved.id, c->ps.id, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
        HIPCHK(plan::stream_wait(s));
        c->saved.N = c->saved.T = N;
        return PLAN_OK;
    });
}

int plan_restore_input(plan_ctx *c)
{
    return guarded(c, [&]() -> int {
        if (!c->saved.pos || c->saved.N == 0) return fail(c, PLAN_E_STATE, "no kept input");
        const uint32_t N = c->saved.N;
        ensure_particles(c->ps, N);
        hipStream_t s = c->stream;
        HIPCHK(hipMemcpyAsync(c->ps.pos, c->saved.pos, (size_t)N * 32, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(c->ps.vel, c->saved.vel, (size_t)N * 32, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(c->ps.prop, c->saved.prop, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
        HIPCHK(hipMemcpyAsync(c->ps.id, c->saved.id, (size_t)N * 4, hipMemcpyDeviceToDevice, s));
        HIPCHK(plan::stream_wait(s));
        c->ps.N = c->ps.T = N;
        c->have_particles = true;
        c->have_tree = c->have_density = c->have_hop = false;
        return PLAN_OK;
    });
}

int plan_device_memory(plan_ctx *c, uint64_t *bytes)
{
    if (!c || !bytes) return PLAN_E_ARG;
    *bytes = c->ps.arena.bytes() + c->tree.arena.bytes() + c->work.bytes() + c->ctree.arena.bytes() + (uint64_t)c->ps.cap * 72 + (uint64_t)c->saved.cap * 72 + (uint64_t)c->tree.num_nodes * 56;
    return PLAN_OK;
}

int plan_stage_times(plan_ctx *c, double *ms, int n)
{
    if (!c || !ms) return PLAN_E_ARG;
    for (int i = 0; i < n && i < 6; ++i) ms[i] = c->times[i];
    return PLAN_OK;
}

int plan_clump_stats(plan_ctx *c, double *v, int n)
{
    if (!c || !v) return PLAN_E_ARG;
    double a[56] = {(double)c->hop_na, (double)c->hop_groups, (double)c->cstats.naive, (double)c->cstats.after_merge,
                    (double)c->cstats.merge_passes, (double)c->cstats.pair_candidates,
                    (double)c->cstats.saddle_queries, (double)c->cstats.rebind_searches, c->d2h_ms,
                    (double)c->cstats.rebind_reruns};
    for (int k = 0; k < 10; ++k) a[10 + k] = c->cstats.ms[k];
    a[20] = c->cstats.rebind_iters;
    a[21] = c->cstats.rebind_max_iters;
    a[22] = c->cstats.rebind_max_n;
    a[23] = c->cstats.rebind_max_ms;
    for (int k = 0; k < 3; ++k) a[24 + k] = c->cstats.mparts[k];
    for (int k = 0; k < 3; ++k) a[27 + k] = c->cstats.rparts[k];
    a[30] = c->density_ovf_waves;
    a[31] = c->hop.knn_overflow_waves;
    a[32] = c->cstats.rebind_max_exact;
    a[33] = c->cstats.rebind_max_exact_ms;
    a[34] = c->cstats.rebind_max_sort_ms;
    a[35] = c->cstats.rebind_work_ms;
    a[36] = c->cstats.rebind_exact_total_ms;
    a[37] = c->cstats.rebind_max_slot;
    a[38] = c->cstats.rebind_last_end_ms;
    a[39] = c->cstats.rebind_last_slot;
    a[40] = c->cstats.rebind_last_ms;
    a[41] = c->density_res_queries;
    a[42] = c->hop.knn_res_queries;
    a[43] = c->cstats.key0_erased;
    a[44] = c->density_pow_checks;
    a[45] = c->density_pow_fixes;
    a[46] = c->catalog_ms;
    a[47] = c->clump_logic_ms;
    a[48] = c->cstats.rebind_max_scan_ms;
    a[49] = c->cstats.rebind_max_energy_ms;
    a[50] = c->cstats.rebind_max_live;
    a[51] = c->cstats.rebind_max_evals;
    a[52] = c->cstats.rebind_scan_total_ms;
    a[53] = c->cstats.rebind_energy_total_ms;
    a[54] = c->cstats.rebind_sort_total_ms;
    a[55] = c->cstats.rebind_setup_total_ms;
    for (int i = 0; i < n && i < 56; ++i) v[i] = a[i];
    return PLAN_OK;
}

int plan_last_density_kernel_ms(plan_ctx *c, double *ms)
{
    if (!c || !ms) return PLAN_E_ARG;
    ms[0] = c->density_kernel_ms;
    return PLAN_OK;
}

int plan_set_num_hop(plan_ctx *c, uint32_t num_hop)
{
    if (!c) return PLAN_E_ARG;
    if (!c->have_particles) return fail(c, PLAN_E_STATE, "no particle set loaded");
    if (!knn_k_supported(num_hop)) return fail(c, PLAN_E_ARG, "num_hop must be in 1..64");
    c->pp.num_hop = num_hop;
    return PLAN_OK;
}

int plan_host_tree_balls(const plan_params *params, uint32_t n, const double *pts4, const double *root_center,
                         double half_width, const double *queries4, uint32_t nq, uint32_t *offsets, double *out,
                         uint32_t cap)
{
    if (!params || (n && !pts4) || !root_center || (nq && !queries4) || !offsets) return PLAN_E_ARG;
    try {
        plan_derived d;
        derive_params(params, n, &d);
        std::vector<double4> pts(n);
        for (uint32_t i = 0; i < n; ++i)
            pts[i] = make_double4(pts4[4 * i], pts4[4 * i + 1], pts4[4 * i + 2], pts4[4 * i + 3]);
        HostTree ht;
        ht.build(pts, root_center, half_width, d);
        std::vector<uint32_t> l;
        uint32_t k = 0;
        for (uint32_t q = 0; q < nq; ++q) {
            offsets[q] = k;
            if (n) ht.ball(queries4[4 * q], queries4[4 * q + 1], queries4[4 * q + 2], queries4[4 * q + 3], l);
            for (uint32_t j : l) {
                if (k < cap && out) out[k] = ht.payload(j);
                ++k;
            }
        }
        offsets[nq] = k;
        return k > cap ? PLAN_E_ARG : PLAN_OK;
    } catch (...) {
        return PLAN_E_INTERNAL;
    }
}

int plan_set_knn_list_limit(plan_ctx *c, uint32_t max_entries)
{
    if (!c) return PLAN_E_ARG;
    c->tree.list_limit = max_entries == 0 ? 0xffffffffu : (max_entries < 8 ? 8u : max_entries);
    c->ctree.list_limit = c->tree.list_limit;
    return PLAN_OK;
}

int plan_last_density_radius_ms(plan_ctx *c, double *ms)
{
    if (!c || !ms) return PLAN_E_ARG;
    ms[0] = c->density_radius_ms;
    return PLAN_OK;
}

int plan_boundness_load(plan_ctx *c, const plan_boundness_in *in)
{
    return guarded(c, [&]() -> int {
        if (!in || !in->offsets) return fail(c, PLAN_E_ARG, "boundness: no input");
        const uint32_t nm = in->offsets[in->num_clumps];
        if (in->offsets[0] != 0) return fail(c, PLAN_E_ARG, "boundness: offsets[0] must be 0");
        for (uint32_t k = 0; k < in->num_clumps; ++k)
            if (in->offsets[k + 1] < in->offsets[k]) return fail(c, PLAN_E_ARG, "boundness: offsets must not decrease");
        if (nm && (!in->pos || !in->vel || !in->mass)) return fail(c, PLAN_E_ARG, "boundness: missing member arrays");
        if (!(in->G > 0)) return fail(c, PLAN_E_ARG, "boundness: G must be positive");
        c->bnd.upload(*in, c->stream);
        c->have_bnd_result = false;
        return PLAN_OK;
    });
}

int plan_boundness_load_catalog(plan_ctx *c, const plan_catalog *cat)
{
    return guarded(c, [&]() -> int {
        if (!cat) return fail(c, PLAN_E_ARG, "boundness: no catalog");
        if (!c->have_tree) return fail(c, PLAN_E_STATE, "boundness: no tree");
        std::vector<uint32_t> off(cat->num_clumps + 1, 0), mem;
        for (uint32_t k = 0; k < cat->num_clumps; ++k) {
            const plan_clump &q = cat->clumps[k];
            off[k + 1] = off[k] + q.num_members;
            for (uint32_t j = 0; j < q.num_members; ++j) {
                const uint32_t t = cat->members[q.member_offset + j];
                if (t >= c->tree.n) return fail(c, PLAN_E_ARG, "boundness: catalog member outside the tree");
                mem.push_back(t);
            }
        }
        c->bnd.upload_tree(cat->num_clumps, off.data(), mem.data(), c->tree.pm, c->tree.va, c->d.grav_constant,
                           c->d.Omega, c->d.q, c->stream);
        c->have_bnd_result = false;
        return PLAN_OK;
    });
}

int plan_boundness_run(plan_ctx *c, int form, uint32_t max_iter, plan_boundness_stats *st)
{
    return guarded(c, [&]() -> int {
        if (!c->bnd.loaded()) return fail(c, PLAN_E_STATE, "boundness: no batch loaded");
        if (form != PLAN_BOUND_FP64 && form != PLAN_BOUND_FP32) return fail(c, PLAN_E_ARG, "boundness: unknown form");
        if (max_iter == 0) return fail(c, PLAN_E_ARG, "boundness: max_iter must be positive");
        c->bnd.run(form, max_iter, c->stream, st);
        c->have_bnd_result = true;
        return PLAN_OK;
    });
}

int plan_boundness_result(plan_ctx *c, uint8_t *bound, plan_boundness_clump *clumps)
{
    return guarded(c, [&]() -> int {
        if (!c->have_bnd_result) return fail(c, PLAN_E_STATE, "boundness: no result");
        c->bnd.download(bound, clumps, c->stream);
        return PLAN_OK;
    });
}

}  // extern "C"
