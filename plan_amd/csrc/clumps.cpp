// clumps.cpp — clump post-processing after HOP: per-clump kinematics, the two
// greedy merge loops, real-peak re-keying, unbinding, rebinding, envelope
// ripping, filters, angular momentum and ranking.
//
// Restated from BHtree::FindPlanetesimals (src/tree.hpp:3810-4339) and
// Planetesimal / PlanetesimalList (src/tree.hpp:4351-4909).  The control flow
// is inherently sequential (map-ordered greedy merges, key-ordered rebinding
// claims), so it runs on the host over the table of particles above the
// density threshold, and every tree query it needs goes to the device in
// batches: the clump tree and its ball searches for a whole merge pass, all
// saddle-point densities of a pass, and the Hill-sphere ball searches of the
// rebinding step.  Floating-point expressions keep the reference's operation
// order; the library is built without FMA contraction.
#include "clumps.h"

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <exception>
#include <limits>
#include <stdexcept>

namespace plan {

typedef std::pair<uint32_t, double> PD;
static bool pd_less(const PD &a, const PD &b) { return a.second < b.second; }

/* OpenMP loop over [0, n), one index at a time and in index order (callers
 * put the longest work first; a nonmonotonic dynamic schedule may hand a
 * thread a block of indices up front, which starts a long one late); the
 * first exception thrown by a body is rethrown on the calling thread after
 * the loop. */
template <class F> static void parallel_for(long n, F f)
{
    std::exception_ptr err = nullptr;
#pragma omp parallel for schedule(monotonic : dynamic, 1)
    for (long i = 0; i < n; ++i) {
        try {
            f(i);
        } catch (...) {
#pragma omp critical(plan_clump_error)
            if (!err) err = std::current_exception();
        }
    }
    if (err) std::rethrow_exception(err);
}

uint32_t HostMembers::rec_of_tidx(uint32_t tidx) const
{
    if (tidx_map) {
        if (tidx >= tidx_map_n) return PLAN_NONE;
        const uint32_t r = tidx_map[tidx];
        return (r < rec.size() && rec[r].tidx == tidx) ? r : PLAN_NONE;
    }
    auto it = std::lower_bound(sorted_tidx.begin(), sorted_tidx.end(), tidx);
    if (it == sorted_tidx.end() || *it != tidx) return PLAN_NONE;
    return rec_of_sorted[it - sorted_tidx.begin()];
}

/* std::sort(v) by distance for non-negative distances, in linear time: the
 * entries become (fixed-point distance << 32 | position) items, sorted by
 * two LSD radix passes (22-bit keys and 11-bit digits; 16-bit keys and 8-bit
 * digits below 4,096 entries, where the histograms' fixed cost would
 * dominate); the key is monotone in the
 * distance and the passes are stable, so only runs of equal keys can be out
 * of distance order, and an insertion pass over those runs finishes the
 * sort.  Without distance ties the sorted order is unique, so this is
 * std::sort's result; with a tie (or a distance that is not a number)
 * std::sort runs on the original sequence, which fixes libstdc++'s tie
 * order.  tmp, a, b: scratch. */
static void sort_by_distance_keys(std::vector<PD> &v, std::vector<PD> &tmp, std::vector<uint64_t> &a,
                                  std::vector<uint64_t> &b)
{
    const size_t n = v.size();
    double dmax = 0;
    for (const PD &p : v) dmax = std::max(dmax, p.second);
    if (n < 64 || n >= (1ull << 32) || !(dmax > 0) || !std::isfinite(dmax)) {
        std::sort(v.begin(), v.end(), pd_less);
        return;
    }
    const uint32_t DB = n < 4096 ? 8u : 11u, DIG = 1u << DB, KMAX = (1u << (2 * DB)) - 1;
    const double scale = (double)KMAX / dmax;
    a.resize(n);
    b.resize(n);
    uint32_t h0[2048], h1[2048];
    std::memset(h0, 0, DIG * sizeof(uint32_t));
    std::memset(h1, 0, DIG * sizeof(uint32_t));
    bool bad = false;
    for (size_t i = 0; i < n; ++i) {
        const double dd = v[i].second;
        bad |= !(dd >= 0);
        const double f = dd * scale;
        const uint32_t k = f >= (double)KMAX ? KMAX : f > 0 ? (uint32_t)f : 0u;
        b[i] = (uint64_t)k << 32 | i;
        h0[k & (DIG - 1)]++;
        h1[k >> DB]++;
    }
    if (bad) {
        std::sort(v.begin(), v.end(), pd_less);
        return;
    }
    for (uint32_t d = 0, s0 = 0, s1 = 0; d < DIG; ++d) {
        const uint32_t t0 = h0[d], t1 = h1[d];
        h0[d] = s0;
        h1[d] = s1;
        s0 += t0;
        s1 += t1;
    }
    for (size_t i = 0; i < n; ++i) {
        const uint64_t x = b[i];
        a[h0[(x >> 32) & (DIG - 1)]++] = x;
    }
    for (size_t i = 0; i < n; ++i) {
        const uint64_t x = a[i];
        b[h1[x >> (32 + DB)]++] = x;
    }
    uint64_t *q = b.data();
    auto dist = [&](uint64_t x) { return v[(uint32_t)x].second; };
    for (size_t r0 = 0; r0 < n;) {
        size_t r1 = r0 + 1;
        while (r1 < n && (q[r1] >> 32) == (q[r0] >> 32)) ++r1;
        if (r1 - r0 > 32) {
            std::sort(q + r0, q + r1, [&](uint64_t x, uint64_t y) { return dist(x) < dist(y); });
        } else {
            for (size_t i = r0 + 1; i < r1; ++i) {
                const uint64_t x = q[i];
                const double xd = dist(x);
                size_t j = i;
                while (j > r0 && xd < dist(q[j - 1])) {
                    q[j] = q[j - 1];
                    --j;
                }
                q[j] = x;
            }
        }
        r0 = r1;
    }
    for (size_t i = 1; i < n; ++i)
        if ((q[i] >> 32) == (q[i - 1] >> 32) && dist(q[i]) == dist(q[i - 1])) {
            std::sort(v.begin(), v.end(), pd_less);
            return;
        }
    tmp.resize(n);
    for (size_t i = 0; i < n; ++i) tmp[i] = v[(uint32_t)q[i]];
    v.swap(tmp);
}

/* The distance order of the positions seq[0, n) (distances D[p] >= 0, no
 * pairs built): the same two LSD radix passes and run repair as
 * sort_by_distance_keys on items (key << 32 | position), written to out.
 * Without distance ties the order is unique, so it is std::sort's on the
 * sequence of (position, distance) pairs; false (out unspecified) when two
 * distances tie, a distance is not a non-negative number, or n is out of the
 * radix range — the caller then sorts the pairs. */
static bool radix_order(const double *D, const uint32_t *seq, size_t n, std::vector<uint64_t> &a,
                        std::vector<uint64_t> &b, uint32_t *out)
{
    if (n < 64 || n >= (1ull << 32)) return false;
    double dmax = 0;
    bool bad = false;
    for (size_t i = 0; i < n; ++i) {
        const double dd = D[seq[i]];
        bad |= !(dd >= 0);
        dmax = std::max(dmax, dd);
    }
    if (bad || !(dmax > 0) || !std::isfinite(dmax)) return false;
    const uint32_t DB = n < 4096 ? 8u : 11u, DIG = 1u << DB, KMAX = (1u << (2 * DB)) - 1;
    const double scale = (double)KMAX / dmax;
    a.resize(n);
    b.resize(n);
    uint32_t h0[2048], h1[2048];
    std::memset(h0, 0, DIG * sizeof(uint32_t));
    std::memset(h1, 0, DIG * sizeof(uint32_t));
    for (size_t i = 0; i < n; ++i) {
        const uint32_t p = seq[i];
        const double f = D[p] * scale;
        const uint32_t k = f >= (double)KMAX ? KMAX : f > 0 ? (uint32_t)f : 0u;
        b[i] = (uint64_t)k << 32 | p;
        h0[k & (DIG - 1)]++;
        h1[k >> DB]++;
    }
    for (uint32_t d = 0, s0 = 0, s1 = 0; d < DIG; ++d) {
        const uint32_t t0 = h0[d], t1 = h1[d];
        h0[d] = s0;
        h1[d] = s1;
        s0 += t0;
        s1 += t1;
    }
    for (size_t i = 0; i < n; ++i) {
        const uint64_t x = b[i];
        a[h0[(x >> 32) & (DIG - 1)]++] = x;
    }
    for (size_t i = 0; i < n; ++i) {
        const uint64_t x = a[i];
        b[h1[x >> (32 + DB)]++] = x;
    }
    uint64_t *q = b.data();
    auto dist = [&](uint64_t x) { return D[(uint32_t)x]; };
    for (size_t r0 = 0; r0 < n;) {
        size_t r1 = r0 + 1;
        while (r1 < n && (q[r1] >> 32) == (q[r0] >> 32)) ++r1;
        if (r1 - r0 > 32) {
            std::sort(q + r0, q + r1, [&](uint64_t x, uint64_t y) { return dist(x) < dist(y); });
        } else {
            for (size_t i = r0 + 1; i < r1; ++i) {
                const uint64_t x = q[i];
                const double xd = dist(x);
                size_t j = i;
                while (j > r0 && xd < dist(q[j - 1])) {
                    q[j] = q[j - 1];
                    --j;
                }
                q[j] = x;
            }
        }
        r0 = r1;
    }
    for (size_t i = 1; i < n; ++i)
        if ((q[i] >> 32) == (q[i - 1] >> 32) && dist(q[i]) == dist(q[i - 1])) return false;
    for (size_t i = 0; i < n; ++i) out[i] = (uint32_t)q[i];
    return true;
}

namespace {

/* Loops over a clump's members visit the 80-byte member records in
 * distance order, scattered over the clump's record ranges: the record
 * PF_AHEAD steps ahead is prefetched (both cache lines it may span). */
constexpr size_t PF_AHEAD = 12;
static inline void prefetch_rec(const MemberRec *P, uint32_t h)
{
    const char *a = reinterpret_cast<const char *>(P + h);
    __builtin_prefetch(a);
    __builtin_prefetch(a + 64);
}

struct Ops {
    const MemberRec *P;
    const ClumpSettings &cs;

    double dist_to(uint32_t h, const double *c) const
    {
        const MemberRec &r = P[h];
        double dx = r.x - c[0], dy = r.y - c[1], dz = r.z - c[2];
        double s = 0;
        s += dx * dx;
        s += dy * dy;
        s += dz * dz;
        return std::sqrt(s);
    }
    static double dist_pts(const double *a, const double *b)
    {
        double s = 0;
        for (int k = 0; k < 3; ++k) {
            double t = a[k] - b[k];
            s += t * t;
        }
        return std::sqrt(s);
    }
    static const double *pos(const MemberRec &r) { return &r.x; }

    /* CalculateKinematicProperties (src/tree.hpp:4431-4442) */
    void kinematics(HClump &c) const
    {
        c.total_mass = 0;
        for (int k = 0; k < 3; ++k) c.com[k] = c.vel[k] = 0;
        const size_t n = c.indices.size();
        for (size_t i = 0; i < n; ++i) {
            if (i + PF_AHEAD < n) prefetch_rec(P, c.indices[i + PF_AHEAD]);
            const MemberRec &r = P[c.indices[i]];
            c.total_mass += r.m;
            c.com[0] += r.x * r.m;
            c.com[1] += r.y * r.m;
            c.com[2] += r.z * r.m;
            c.vel[0] += r.vx * r.m;
            c.vel[1] += r.vy * r.m;
            c.vel[2] += r.vz * r.m;
        }
        for (int k = 0; k < 3; ++k) c.com[k] /= c.total_mass;
        for (int k = 0; k < 3; ++k) c.vel[k] /= c.total_mass;
    }
    /* radii of SortParticles / MergeAnotherPlanetesimal (src/tree.hpp:4461-4511) */
    void radii(HClump &c) const
    {
        double tenth = P[c.indices[c.indices.size() / 2]].rho / 10.;
        auto it = c.particles.begin();
        for (; it != c.particles.end(); ++it)
            if (P[it->first].rho < tenth) break;
        c.inner = (it != c.particles.end()) ? it->second : c.particles.back().second;
        auto rit = c.particles.rbegin();
        for (; rit != c.particles.rend(); ++rit)
            if (P[rit->first].rho > tenth) break;
        c.outer = rit->second;
        c.one10th = (c.inner + c.outer) / 2.;
        rit = c.particles.rbegin();
        double tm = 0., lim = c.total_mass * (1.0 - 0.95);
        for (; rit != c.particles.rend(); ++rit) {
            tm += P[rit->first].m;
            if (tm > lim) break;
        }
        c.two_sigma = rit->second;
        if (c.inner > 0 && c.outer / c.inner > 2.5) c.one10th = std::min(c.outer, c.two_sigma);
        else c.one10th = std::min(c.one10th, c.two_sigma);
    }
    /* v[i].second = distance of member v[i].first to c */
    void set_distances(std::vector<PD> &v, const double *c) const
    {
        const size_t n = v.size();
        for (size_t i = 0; i < n; ++i) {
            if (i + PF_AHEAD < n) prefetch_rec(P, v[i + PF_AHEAD].first);
            v[i].second = dist_to(v[i].first, c);
        }
    }
    /* SortParticles (src/tree.hpp:4446-4512) */
    void sort_particles(HClump &c, bool nearly_sorted = false) const
    {
        kinematics(c);
        c.particles.reserve(c.particles.size() + c.indices.size());
        const size_t n = c.indices.size();
        for (size_t i = 0; i < n; ++i) {
            if (i + PF_AHEAD < n) prefetch_rec(P, c.indices[i + PF_AHEAD]);
            const uint32_t h = c.indices[i];
            c.particles.push_back(PD(h, dist_to(h, c.com)));
        }
        sort_by_distance(c.particles, nearly_sorted);
        for (size_t i = 0; i < c.particles.size(); ++i) c.indices[i] = c.particles[i].first;
        radii(c);
    }
    /* Sort by distance.  Without exact distance ties the sorted order is
     * unique, so an insertion sort of an almost sorted list (members re-sorted
     * around a slightly moved centre) or the distance-bucket sort gives
     * std::sort's result in O(n); on a tie std::sort runs on the original
     * order so libstdc++'s tie order is kept exactly. */
    static void sort_by_distance(std::vector<PD> &v, bool nearly_sorted)
    {
        static thread_local std::vector<PD> tmp;
        static thread_local std::vector<uint64_t> ka, kb;
        if (!nearly_sorted || v.size() < 32) {
            sort_by_distance_keys(v, tmp, ka, kb);
            return;
        }
        static thread_local std::vector<PD> orig;
        orig.assign(v.begin(), v.end());
        size_t budget = 16 * v.size() + 1024, moves = 0;
        bool ok = true;
        for (size_t i = 1; i < v.size() && ok; ++i) {
            PD x = v[i];
            size_t j = i;
            while (j > 0 && x.second < v[j - 1].second) {
                v[j] = v[j - 1];
                --j;
                if (++moves > budget) {
                    ok = false;
                    break;
                }
            }
            v[j] = x;
        }
        if (ok)
            for (size_t i = 1; i < v.size(); ++i)
                if (v[i].second == v[i - 1].second) {
                    ok = false;
                    break;
                }
        if (!ok) {
            v.swap(orig);
            sort_by_distance_keys(v, tmp, ka, kb);
        }
    }
    /* std::sort of v = A ++ B by distance, where A (the first na entries) is
     * usually nearly sorted: A by insertion (by the distance-bucket sort when
     * that takes too many moves), B by the bucket sort, then a linear merge.
     * Without distance
     * ties the sorted order is unique and equals std::sort's; on a tie
     * std::sort runs on the original v, or, when strict, v is left unsorted
     * and false is returned. */
    static bool sort_union(std::vector<PD> &v, size_t na, bool strict = false)
    {
        if (!strict && (v.size() < 64 || na < 32)) {
            std::sort(v.begin(), v.end(), pd_less);
            return true;
        }
        static thread_local std::vector<PD> a, b, m, tmp;
        static thread_local std::vector<uint64_t> ka, kb;
        a.assign(v.begin(), v.begin() + na);
        b.assign(v.begin() + na, v.end());
        size_t budget = 16 * a.size() + 1024, moves = 0;
        bool far = false;
        for (size_t i = 1; i < a.size() && !far; ++i) {
            PD x = a[i];
            size_t j = i;
            while (j > 0 && x.second < a[j - 1].second) {
                a[j] = a[j - 1];
                --j;
                if (++moves > budget) {
                    far = true;
                    break;
                }
            }
            a[j] = x;
        }
        if (far) sort_by_distance_keys(a, tmp, ka, kb);
        bool ok = true;
        {
            sort_by_distance_keys(b, tmp, ka, kb);
            m.resize(v.size());
            std::merge(a.begin(), a.end(), b.begin(), b.end(), m.begin(), pd_less);
            for (size_t i = 1; i < m.size(); ++i)
                if (m[i].second == m[i - 1].second) {
                    ok = false;
                    break;
                }
            if (ok) {
                std::copy(m.begin(), m.end(), v.begin());
                return true;
            }
        }
        if (strict) return false;
        std::sort(v.begin(), v.end(), pd_less);
        return true;
    }
    /* IsPositionDispersion2Large (src/tree.hpp:4516-4528) */
    static bool dispersion_too_large(const HClump &c, double tol = 0.55)
    {
        size_t n = c.particles.size();
        double s = 0.0;
        for (size_t i = 0; i < n; ++i) s = s + c.particles[i].second * c.particles[i].second;
        return s / n > std::pow(tol * c.particles[n - 1].second, 2);
    }
    /* MergeAnotherPlanetesimal (src/tree.hpp:4532-4611) */
    void merge(HClump &a, const HClump &b, double rho_a_peak, double rho_b_peak) const
    {
        if (rho_a_peak < rho_b_peak) {
            a.subpeaks.push_back(a.peak_index);
            a.peak_index = b.peak_index;
        } else {
            a.subpeaks.push_back(b.peak_index);
        }
        double tot = a.total_mass + b.total_mass;
        for (int k = 0; k < 3; ++k) a.com[k] = (a.com[k] * a.total_mass + b.com[k] * b.total_mass) / tot;
        for (int k = 0; k < 3; ++k) a.vel[k] = (a.vel[k] * a.total_mass + b.vel[k] * b.total_mass) / tot;
        a.total_mass = tot;
        const size_t na = a.particles.size();
        a.particles.insert(a.particles.end(), b.particles.begin(), b.particles.end());
        set_distances(a.particles, a.com);
        sort_union(a.particles, na);
        a.indices.resize(a.particles.size());
        for (size_t i = 0; i < a.particles.size(); ++i) a.indices[i] = a.particles[i].first;
        radii(a);
    }
    /* Several MergeAnotherPlanetesimal calls into one predator in a row
     * (bs in order).  Mass, centre, velocity and peaks are updated merge by
     * merge as the reference does; the particle list is re-sorted once, by
     * distance to the final centre, which is the last merge's result whenever
     * the final distances have no ties (the sorted order is then unique).
     * With a tie the merges are replayed one by one. */
    void merge_batch(HClump &a, const std::vector<const HClump *> &bs, const HostMembers &hm, bool hill) const
    {
        if (bs.size() == 1) {
            merge(a, *bs[0], rho_of(hm, a.peak_index), rho_of(hm, bs[0]->peak_index));
            if (hill) hill_radius(a);
            return;
        }
        HClump a0 = a;
        const size_t na = a.particles.size();
        for (const HClump *b : bs) {
            if (rho_of(hm, a.peak_index) < rho_of(hm, b->peak_index)) {
                a.subpeaks.push_back(a.peak_index);
                a.peak_index = b->peak_index;
            } else {
                a.subpeaks.push_back(b->peak_index);
            }
            double tot = a.total_mass + b->total_mass;
            for (int k = 0; k < 3; ++k) a.com[k] = (a.com[k] * a.total_mass + b->com[k] * b->total_mass) / tot;
            for (int k = 0; k < 3; ++k) a.vel[k] = (a.vel[k] * a.total_mass + b->vel[k] * b->total_mass) / tot;
            a.total_mass = tot;
            a.particles.insert(a.particles.end(), b->particles.begin(), b->particles.end());
        }
        set_distances(a.particles, a.com);
        if (!sort_union(a.particles, na, true)) {
            a = a0;
            for (const HClump *b : bs) {
                merge(a, *b, rho_of(hm, a.peak_index), rho_of(hm, b->peak_index));
                if (hill) hill_radius(a);
            }
            return;
        }
        a.indices.resize(a.particles.size());
        for (size_t i = 0; i < a.particles.size(); ++i) a.indices[i] = a.particles[i].first;
        radii(a);
        if (hill) hill_radius(a);
    }
    static double rho_of(const HostMembers &hm, uint32_t tidx)
    {
        uint32_t r = hm.rec_of_tidx(tidx);
        if (r == PLAN_NONE) throw std::runtime_error("clump peak is not a particle above the density threshold");
        return hm.rec[r].rho;
    }
    /* RemoveUnboundParticles (src/tree.hpp:4615-4674) */
    void remove_unbound(HClump &c) const
    {
        const double G = cs.G, sh = cs.q * cs.Omega;
        size_t i = c.particles.size() - 1;
        while (i != 0) {
            if (i > PF_AHEAD) prefetch_rec(P, c.particles[i - PF_AHEAD].first);
            const MemberRec &q = P[c.particles[i].first];
            double tm = c.total_mass - q.m;
            double tc[3], tv[3];
            const double *qp = &q.x, *qv = &q.vx;
            for (int k = 0; k < 3; ++k) tc[k] = (c.com[k] * c.total_mass - qp[k] * q.m) / tm;
            for (int k = 0; k < 3; ++k) tv[k] = (c.vel[k] * c.total_mass - qv[k] * q.m) / tm;
            double a = tc[0] - qp[0];
            double sv[3] = {0.0 * a, sh * a, 0.0 * a};
            double n2 = 0;
            for (int k = 0; k < 3; ++k) {
                double dv = (tv[k] - sv[k]) - qv[k];
                n2 += dv * dv;
            }
            double e = (0.5 / c.total_mass) * n2 - G / dist_pts(tc, qp);
            if (e > 0) {
                std::swap(c.indices[i], c.indices.back());
                c.indices.pop_back();
                std::swap(c.particles[i], c.particles.back());
                c.particles.pop_back();
                for (int k = 0; k < 3; ++k) {
                    c.vel[k] = tv[k];
                    c.com[k] = tc[k];
                }
                c.total_mass = tm;
            }
            i--;
        }
        std::vector<PD>().swap(c.particles);
        sort_particles(c, true);
    }
    /* CalculateHillRadius (src/tree.hpp:4715-4717) */
    void hill_radius(HClump &c) const
    {
        c.hill = std::pow((((c.total_mass * cs.G) / 3.) / cs.Omega) / cs.Omega, 1. / 3.);
    }
    /* CalculateAngularMomentum (src/tree.hpp:4721-4747) */
    void angular_momentum(HClump &c) const
    {
        double quarter = 0.25 * c.hill;
        double hu = ((c.total_mass * c.hill) * c.hill) * cs.Omega;
        double s1 = cs.q * cs.Omega;
        const size_t n = c.particles.size();
        for (size_t i = 0; i < n; ++i) {
            const auto &it = c.particles[i];
            if (i + PF_AHEAD < n) prefetch_rec(P, c.particles[i + PF_AHEAD].first);
            const MemberRec &q = P[it.first];
            const double *qp = &q.x, *qv = &q.vx;
            double dr[3], dv[3], j[3];
            for (int k = 0; k < 3; ++k) dr[k] = qp[k] - c.com[k];
            double sp[3] = {0. * qp[0], s1 * qp[0], 0. * qp[0]};
            double sc[3] = {0. * c.com[0], s1 * c.com[0], 0. * c.com[0]};
            for (int k = 0; k < 3; ++k) dv[k] = (qv[k] - sp[k]) - (c.vel[k] - sc[k]);
            j[0] = dr[1] * dv[2] - dr[2] * dv[1];
            j[1] = dr[2] * dv[0] - dr[0] * dv[2];
            j[2] = dr[0] * dv[1] - dr[1] * dv[0];
            double o[3] = {-dr[0] * dr[2], -dr[1] * dr[2], (dr[0] * dr[0] + dr[1] * dr[1])};
            for (int k = 0; k < 3; ++k) j[k] += cs.Omega * o[k];
            for (int k = 0; k < 3; ++k) c.J[k] += q.m * j[k];
            if (it.second < quarter)
                for (int k = 0; k < 3; ++k) c.Jq[k] += q.m * j[k];
        }
        for (int k = 0; k < 3; ++k) {
            c.J[k] /= hu;
            c.Jq[k] /= hu;
        }
    }
    /* PlanetesimalList::IsGravitationallyBound (src/tree.hpp:4867-4874) */
    bool grav_bound(const HClump &a, const HClump &b) const
    {
        double tm = a.total_mass + b.total_mass;
        double Pg = -cs.G / dist_pts(a.com, b.com);
        double dx = a.com[0] - b.com[0];
        double sh = cs.q * cs.Omega;
        double sv[3] = {0.0 * dx, sh * dx, 0.0 * dx};
        double n2 = 0;
        for (int k = 0; k < 3; ++k) {
            double t = (a.vel[k] - sv[k]) - b.vel[k];
            n2 += t * t;
        }
        double Ek = (0.5 / tm) * n2;
        return Pg + Ek < 0.;
    }
};

/* RemoveSmallMassAndLowPeak (src/tree.hpp:3524-3546); peak densities come
 * from the member table (the key is a member of its own clump or of an
 * absorbed one, both above threshold) */
void remove_small_mass_low_peak(std::map<uint32_t, HClump> &cl, const HostMembers &hm, const ClumpSettings &cs)
{
    std::vector<uint32_t> del;
    for (auto &it : cl) {
        uint32_t kr = hm.rec_of_tidx(it.first);
        double pk = kr == PLAN_NONE ? 0.0 : hm.rec[kr].rho;
        if (it.second.total_mass < cs.clump_mass_threshold || pk < cs.peak_density_threshold) {
            del.push_back(it.first);
            continue;
        }
        uint32_t lim = it.second.indices.size() < cs.K ? (uint32_t)it.second.indices.size() : cs.K;
        double ori = 0;
        for (uint32_t i = 0; i < lim; ++i) ori = std::max(ori, hm.rec[it.second.indices[i]].ath);
        if (ori < cs.peak_density_threshold) del.push_back(it.first);
    }
    for (auto k : del) cl.erase(k);
}

double rho_of_key(const HostMembers &hm, uint32_t tidx)
{
    uint32_t r = hm.rec_of_tidx(tidx);
    if (r == PLAN_NONE) throw std::runtime_error("clump peak is not a particle above the density threshold");
    return hm.rec[r].rho;
}

/* Keys, in key order, of the clumps (the masked ones only, if asked) for
 * which IsPositionDispersion2Large holds; the checks, a sequential sum over
 * each clump's members, run in parallel, the largest clumps first. */
static std::vector<uint32_t> dispersed_keys(std::map<uint32_t, HClump> &cl, double tol, bool masked_only)
{
    std::vector<std::pair<uint32_t, const HClump *>> v;
    for (auto &it : cl)
        if (!masked_only || it.second.mask) v.push_back(std::make_pair(it.first, &it.second));
    std::vector<size_t> order(v.size());
    for (size_t i = 0; i < v.size(); ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) {
        return v[x].second->particles.size() > v[y].second->particles.size();
    });
    std::vector<uint8_t> f(v.size(), 0);
    parallel_for((long)v.size(), [&](long k) { f[order[k]] = Ops::dispersion_too_large(*v[order[k]].second, tol); });
    std::vector<uint32_t> out;
    for (size_t i = 0; i < v.size(); ++i)
        if (f[i]) out.push_back(v[i].first);
    return out;
}

/* one merge loop (src/tree.hpp:3863-4006 with one10th radii, 4011-4152 with
 * Hill radii), serial semantics: predators in key order; a clump eaten in a
 * pass cannot eat in it; merges applied after the pass in pair order.
 * num is num_planetesimals, carried from one loop to the next: the
 * reference sizes its list of keys to delete to num, zero-filled
 * (3874 / 4019), and erases every entry after the loop (4000-4003 /
 * 4146-4149), so the clump keyed by tree index 0 goes too whenever fewer
 * than num clumps were deleted. */
void merge_loop(std::map<uint32_t, HClump> &cl, const HostMembers &hm, const ClumpSettings &cs, const Ops &ops,
                ClumpServices &svc, bool hill, uint32_t &num, ClumpStats &st)
{
    double max_radius = 0.;
    const uint32_t num_start = num;
    std::vector<uint32_t> deleted;
    bool happened = true;
    while (happened) {
        st.merge_passes++;
        /* BuildClumpTree (src/tree.hpp:5251-5284) over the masked clumps */
        std::vector<double4> pts, vel, queries;
        std::vector<std::map<uint32_t, HClump>::iterator> p1s;
        for (auto it = cl.begin(); it != cl.end(); ++it) {
            if (!it->second.mask) continue;
            pts.push_back(make_double4(it->second.com[0], it->second.com[1], it->second.com[2], (double)it->first));
            vel.push_back(make_double4(it->second.vel[0], it->second.vel[1], it->second.vel[2], 0.0));
            if (hill) max_radius = std::max(it->second.hill * cs.hill_frac, max_radius);
            else max_radius = std::max(it->second.one10th, max_radius);
        }
        for (auto it = cl.begin(); it != cl.end(); ++it) {
            if (!it->second.mask) continue;
            double r1 = hill ? it->second.hill * cs.hill_frac : it->second.one10th;
            queries.push_back(make_double4(it->second.com[0], it->second.com[1], it->second.com[2], r1 + max_radius));
            p1s.push_back(it);
        }
        std::vector<std::vector<uint32_t>> near;
        auto t0 = std::chrono::steady_clock::now();
        svc.clump_tree_balls(pts, vel, queries, near);
        auto t1 = std::chrono::steady_clock::now();
        st.mparts[0] += std::chrono::duration<double, std::milli>(t1 - t0).count();
        /* every saddle density this pass could ask for, in one device batch */
        struct Cand {
            uint32_t p1k, p2k;
            int kind; /* 1 contains, 2 needs saddle */
            size_t sq;
        };
        std::vector<std::vector<Cand>> cands(p1s.size());
        std::vector<double4> sq;
        for (size_t a = 0; a < p1s.size(); ++a) {
            const HClump &c1 = p1s[a]->second;
            double r1 = hill ? c1.hill * cs.hill_frac : c1.one10th;
            for (uint32_t id2 : near[a]) {
                if (id2 == p1s[a]->first) continue;
                auto p2 = cl.find(id2);
                if (p2 == cl.end()) {
                    std::fprintf(stderr, "Error: Cannot find a clump with peak_id=%u. Proceed for now.\n", id2);
                    continue;
                }
                const HClump &c2 = p2->second;
                if (!c2.mask) continue;
                double r2 = hill ? c2.hill * cs.hill_frac : c2.one10th;
                double cd = Ops::dist_pts(c1.com, c2.com);
                Cand cd_;
                cd_.p1k = p1s[a]->first;
                cd_.p2k = id2;
                cd_.sq = 0;
                if (cd < std::max(r1, r2)) {
                    cd_.kind = 1;
                } else if (cd < (r1 + r2) && ops.grav_bound(c1, c2)) {
                    /* IsSaddlePointDeepEnough / IsHillSaddlePointDeepEnough (src/tree.hpp:4880-4909) */
                    double ra = hill ? c1.hill : r1, rb = hill ? c2.hill : r2;
                    double est = hill ? std::min(c1.hill, c2.hill) / cs.hill_frac : std::min(r1, r2);
                    double f = ra / (ra + rb);
                    double sp[3];
                    for (int k = 0; k < 3; ++k) {
                        double r12 = c2.com[k] - c1.com[k];
                        r12 *= f;
                        sp[k] = c1.com[k] + r12;
                    }
                    if (!(est > 0)) throw std::runtime_error("saddle query with zero start radius");
                    cd_.kind = 2;
                    cd_.sq = sq.size();
                    sq.push_back(make_double4(sp[0], sp[1], sp[2], est));
                } else {
                    continue;
                }
                cands[a].push_back(cd_);
                st.pair_candidates++;
            }
        }
        std::vector<double> srho;
        if (!sq.empty()) svc.saddle_density(sq, srho);
        st.saddle_queries += (uint32_t)sq.size();
        auto t2 = std::chrono::steady_clock::now();
        st.mparts[1] += std::chrono::duration<double, std::milli>(t2 - t1).count();
        /* greedy replay */
        happened = false;
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        for (size_t a = 0; a < p1s.size(); ++a) {
            if (!p1s[a]->second.mask) continue;
            size_t before = pairs.size();
            for (const Cand &c : cands[a]) {
                auto p2 = cl.find(c.p2k);
                if (!p2->second.mask) continue;
                bool eat = c.kind == 1 || !(srho[c.sq] < 2.5 * cs.density_threshold);
                if (eat) {
                    p2->second.mask = false;
                    pairs.push_back(std::make_pair(c.p1k, c.p2k));
                    deleted.push_back(c.p2k);
                }
            }
            if (pairs.size() > before) happened = true;
        }
        /* The pairs in order, grouped into runs with one predator.  A run
         * reads its preys' final states, so it waits for the runs in which
         * they ate (earlier in the list); runs of one level are independent
         * and go in parallel. */
        {
            struct Run {
                HClump *a;
                std::vector<const HClump *> bs;
                uint32_t level;
            };
            std::vector<Run> runs;
            std::map<uint32_t, uint32_t> wlevel, rlevel; /* last level writing / reading a clump */
            uint32_t max_level = 0;
            for (size_t i = 0; i < pairs.size();) {
                size_t j = i;
                while (j < pairs.size() && pairs[j].first == pairs[i].first) ++j;
                Run r;
                r.a = &cl[pairs[i].first];
                uint32_t lv = 0;
                auto w = wlevel.find(pairs[i].first);
                if (w != wlevel.end()) lv = std::max(lv, w->second);
                auto rd = rlevel.find(pairs[i].first);
                if (rd != rlevel.end()) lv = std::max(lv, rd->second);
                for (size_t k = i; k < j; ++k) {
                    r.bs.push_back(&cl[pairs[k].second]);
                    auto wb = wlevel.find(pairs[k].second);
                    if (wb != wlevel.end()) lv = std::max(lv, wb->second);
                }
                r.level = lv + 1;
                max_level = std::max(max_level, r.level);
                wlevel[pairs[i].first] = r.level;
                for (size_t k = i; k < j; ++k) {
                    uint32_t &x = rlevel[pairs[k].second];
                    x = std::max(x, r.level);
                }
                runs.push_back(std::move(r));
                i = j;
            }
            for (uint32_t lv = 1; lv <= max_level; ++lv) {
                std::vector<Run *> now;
                for (auto &r : runs)
                    if (r.level == lv) now.push_back(&r);
                std::sort(now.begin(), now.end(),
                          [](const Run *x, const Run *y) { return x->a->particles.size() > y->a->particles.size(); });
                parallel_for((long)now.size(), [&](long k) { ops.merge_batch(*now[k]->a, now[k]->bs, hm, hill); });
            }
        }
        st.mparts[2] += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t2).count();
        for (uint32_t k : dispersed_keys(cl, 0.55, true)) {
            cl[k].mask = false;
            deleted.push_back(k);
        }
        num = (uint32_t)(cl.size() - deleted.size());
    }
    for (auto k : deleted) cl.erase(k);
    if (num_start > deleted.size()) st.key0_erased += (uint32_t)cl.erase(0u);
}

}  // namespace

/* State of one clump's rebinding: a superset of its Hill-sphere candidates
 * (records above threshold, ascending tree index, fetched once from the
 * device for a ball of radius R around c0) and what it claimed. */
struct Rebind {
    std::vector<uint32_t> sup;
    double c0[3] = {0, 0, 0};
    double R = 0;
    std::vector<uint32_t> claimed;
    uint32_t exact_searches = 0;
    uint32_t iters = 0;
    double ms = 0, ms_exact = 0, ms_sort = 0, ms_scan = 0, ms_energy = 0, ms_setup = 0;
    /* live records at the start; candidate energies evaluated */
    uint64_t live0 = 0, evals = 0;
    std::chrono::steady_clock::time_point t0, t1; /* start and end of the run */
};

static void to_records(const HostMembers &hm, const std::vector<uint32_t> &tidx, std::vector<uint32_t> &out)
{
    out.clear();
    out.reserve(tidx.size());
    for (uint32_t t : tidx) {
        uint32_t r = hm.rec_of_tidx(t);
        if (r == PLAN_NONE) throw std::runtime_error("rebinding candidate above threshold missing from member table");
        out.push_back(r);
    }
}

/* A member's terms of the kinematic sums (src/tree.hpp:4446-4460): m, x m,
 * y m, z m, vx m, vy m, vz m, computed once when it joins; a sum over the
 * members then reads one cache line per member. */
struct alignas(64) KinRow {
    double t[8];
};

/* A clump's members during rebinding, structure of arrays in the order they
 * joined (the members at the start, then each claim), with the current
 * distance order `ord` over those positions.  Only mass, centre and velocity
 * feed back into the loop, so member lists and radii are rebuilt once at the
 * end. */
struct RbLocal {
    std::vector<double> x, y, z, m, rho, d;
    std::vector<KinRow> kin;
    std::vector<uint32_t> h, ord, seq;
    std::vector<PD> v, tmp;
    std::vector<uint64_t> ka, kb;
    size_t n_sorted = 0;
    size_t size() const { return h.size(); }
    /* room for the members plus the superset they are claimed from */
    void reserve(size_t n)
    {
        for (auto *a : {&x, &y, &z, &m, &rho}) a->reserve(n);
        kin.reserve(n);
        h.reserve(n);
    }
    void push(const MemberRec &r, uint32_t hh)
    {
        x.push_back(r.x);
        y.push_back(r.y);
        z.push_back(r.z);
        m.push_back(r.m);
        kin.push_back(KinRow{{r.m, r.x * r.m, r.y * r.m, r.z * r.m, r.vx * r.m, r.vy * r.m, r.vz * r.m, 0.0}});
        rho.push_back(r.rho);
        h.push_back(hh);
    }
};

static bool cpu_has_avx512()
{
    static const bool v = __builtin_cpu_supports("avx512f") && __builtin_cpu_supports("avx512bw") &&
                          __builtin_cpu_supports("avx512vl");
    return v;
}

/* D[p] = |(X, Y, Z)[p] - o|, summed x, y, z as Ops::dist_pts; compiled for
 * each instruction set (IEEE operations give the same bits at any width) */
#define PLAN_DIST_BODY                                          \
    for (size_t p = 0; p < n; ++p) {                            \
        const double dx = X[p] - ox, dy = Y[p] - oy, dz = Z[p] - oz; \
        double q = 0;                                           \
        q += dx * dx;                                           \
        q += dy * dy;                                           \
        q += dz * dz;                                           \
        D[p] = std::sqrt(q);                                    \
    }
static void dist_base(double ox, double oy, double oz, const double *X, const double *Y, const double *Z, size_t n,
                      double *D)
{
    PLAN_DIST_BODY
}
__attribute__((target("avx2"))) static void dist_avx2(double ox, double oy, double oz, const double *X,
                                                       const double *Y, const double *Z, size_t n, double *D)
{
    PLAN_DIST_BODY
}
__attribute__((target("avx512f,avx512bw,avx512vl"))) static void dist_avx512(double ox, double oy, double oz,
                                                                              const double *X, const double *Y,
                                                                              const double *Z, size_t n, double *D)
{
    PLAN_DIST_BODY
}
#undef PLAN_DIST_BODY
typedef void (*DistFn)(double, double, double, const double *, const double *, const double *, size_t, double *);
static DistFn dist_fn()
{
    return cpu_has_avx512() ? dist_avx512 : __builtin_cpu_supports("avx2") ? dist_avx2 : dist_base;
}

/* out[k] = sum of the kinematic rows' term k over ord[0, ns) and then the
 * positions ns .. n - 1, each term accumulated in that order from 0 (the
 * seven sums of SortParticles); one vector add per member with AVX-512,
 * whose lanes round as the scalar adds do */
#define PLAN_KIN_SUM_ARGS const KinRow *__restrict K, const uint32_t *__restrict ord, size_t ns, size_t n, double *out
static void kin_sum_base(PLAN_KIN_SUM_ARGS)
{
    double s[7] = {0, 0, 0, 0, 0, 0, 0};
    for (size_t i = 0; i < n; ++i) {
        const double *t = K[i < ns ? ord[i] : i].t;
        for (int k = 0; k < 7; ++k) s[k] += t[k];
    }
    for (int k = 0; k < 7; ++k) out[k] = s[k];
}
__attribute__((target("avx512f"))) static void kin_sum_avx512(PLAN_KIN_SUM_ARGS)
{
    typedef double v8d __attribute__((vector_size(64)));
    v8d acc = {0, 0, 0, 0, 0, 0, 0, 0};
    for (size_t i = 0; i < ns; ++i) acc += *reinterpret_cast<const v8d *>(K[ord[i]].t);
    for (size_t p = ns; p < n; ++p) acc += *reinterpret_cast<const v8d *>(K[p].t);
    for (int k = 0; k < 7; ++k) out[k] = acc[k];
}
#undef PLAN_KIN_SUM_ARGS
typedef void (*KinSumFn)(const KinRow *, const uint32_t *, size_t, size_t, double *);
static KinSumFn kin_sum_fn()
{
    return cpu_has_avx512() ? kin_sum_avx512 : kin_sum_base;
}

/* SortParticles after claims (src/tree.hpp:4431-4512 via 4705-4709):
 * kinematics summed in member order (the last sorted order, then the claims
 * in claim order), distances to the new centre, sort. */
static void rb_resort(HClump &c, RbLocal &L)
{
    const size_t n = L.size(), ns = L.n_sorted;
    static const KinSumFn kin_sum = kin_sum_fn();
    double sum[7];
    kin_sum(L.kin.data(), L.ord.data(), ns, n, sum);
    const double M = sum[0], cx = sum[1], cy = sum[2], cz = sum[3], ux = sum[4], uy = sum[5], uz = sum[6];
    c.total_mass = M;
    c.com[0] = cx / M;
    c.com[1] = cy / M;
    c.com[2] = cz / M;
    c.vel[0] = ux / M;
    c.vel[1] = uy / M;
    c.vel[2] = uz / M;
    L.d.resize(n);
    static const DistFn dist_to_centre = dist_fn();
    double *D = L.d.data();
    dist_to_centre(c.com[0], c.com[1], c.com[2], L.x.data(), L.y.data(), L.z.data(), n, D);
    /* the sequence summed: the last order, then the claims in claim order */
    L.seq.resize(n);
    std::memcpy(L.seq.data(), L.ord.data(), ns * sizeof(uint32_t));
    for (size_t p = ns; p < n; ++p) L.seq[p] = (uint32_t)p;
    L.ord.resize(n);
    if (!radix_order(D, L.seq.data(), n, L.ka, L.kb, L.ord.data())) {
        L.v.resize(n);
        for (size_t i = 0; i < n; ++i) L.v[i] = PD(L.seq[i], D[L.seq[i]]);
        sort_by_distance_keys(L.v, L.tmp, L.ka, L.kb);
        for (size_t i = 0; i < n; ++i) L.ord[i] = L.v[i].first;
    }
    L.n_sorted = n;
}

/* member lists and radii of SortParticles (src/tree.hpp:4446-4512) from the
 * last sort */
static void rb_finish(HClump &c, const RbLocal &L)
{
    const size_t n = L.size();
    c.particles.resize(n);
    c.indices.resize(n);
    for (size_t i = 0; i < n; ++i) {
        c.particles[i] = PD(L.h[L.ord[i]], L.d[L.ord[i]]);
        c.indices[i] = c.particles[i].first;
    }
    const double tenth = L.rho[L.ord[n / 2]] / 10.;
    size_t i = 0;
    for (; i < n; ++i)
        if (L.rho[L.ord[i]] < tenth) break;
    c.inner = (i < n) ? c.particles[i].second : c.particles.back().second;
    size_t j = n;
    while (j > 0 && !(L.rho[L.ord[j - 1]] > tenth)) --j;
    c.outer = c.particles[j > 0 ? j - 1 : 0].second;
    c.one10th = (c.inner + c.outer) / 2.;
    double tm = 0., lim = c.total_mass * (1.0 - 0.95);
    size_t k2 = n;
    while (k2 > 0) {
        tm += L.m[L.ord[k2 - 1]];
        if (tm > lim) break;
        --k2;
    }
    c.two_sigma = c.particles[k2 > 0 ? k2 - 1 : 0].second;
    if (c.inner > 0 && c.outer / c.inner > 2.5) c.one10th = std::min(c.outer, c.two_sigma);
    else c.one10th = std::min(c.one10th, c.two_sigma);
}

/* Superset records that may still be claimed (neither flagged nor claimed by
 * this run), ascending tree index, structure of arrays, with their distance
 * from the superset's centre c0. */
struct LiveSet {
    std::vector<double> x, y, z, m, vx, vy, vz, lsd, s;
    std::vector<uint32_t> h;
    void clear()
    {
        for (auto *a : {&x, &y, &z, &m, &vx, &vy, &vz, &lsd}) a->clear();
        h.clear();
    }
    void reserve(size_t n)
    {
        for (auto *a : {&x, &y, &z, &m, &vx, &vy, &vz, &lsd}) a->reserve(n);
        h.reserve(n);
    }
    void push(const MemberRec &q, uint32_t hh, double l)
    {
        x.push_back(q.x);
        y.push_back(q.y);
        z.push_back(q.z);
        m.push_back(q.m);
        vx.push_back(q.vx);
        vy.push_back(q.vy);
        vz.push_back(q.vz);
        lsd.push_back(l);
        h.push_back(hh);
    }
    size_t size() const { return h.size(); }
    /* drop the claimed entries (lsd = +inf), keeping the order */
    void compact()
    {
        size_t w = 0;
        for (size_t k = 0; k < h.size(); ++k)
            if (!std::isinf(lsd[k])) {
                x[w] = x[k];
                y[w] = y[k];
                z[w] = z[k];
                m[w] = m[k];
                vx[w] = vx[k];
                vy[w] = vy[k];
                vz[w] = vz[k];
                lsd[w] = lsd[k];
                h[w] = h[k];
                ++w;
            }
        for (auto *a : {&x, &y, &z, &m, &vx, &vy, &vz, &lsd}) a->resize(w);
        h.resize(w);
    }
};

/* The energy of a candidate against the clump's current state
 * (src/tree.hpp:4693-4696), as one expression so batches of candidates can
 * be evaluated ahead of the state's updates. */
static inline double rebind_energy(const HClump &c, double sh, double G, const double *qp, const double *qv, double qm)
{
    const double tm = c.total_mass + qm;
    const double a = c.com[0] - qp[0];
    const double sv[3] = {0.0 * a, sh * a, 0.0 * a};
    double n2 = 0;
    for (int k = 0; k < 3; ++k) {
        const double dv = (c.vel[k] - sv[k]) - qv[k];
        n2 += dv * dv;
    }
    return (0.5 / tm) * n2 - G / Ops::dist_pts(c.com, qp);
}

/* The clump state a block of candidate energies is evaluated against. */
struct EnergyState {
    double M, cx, cy, cz, vx, vy, vz, sh, G;
};

/* rebind_energy for the candidates idx[0, n) of the live arrays: the same
 * expression element by element (written once, compiled for the baseline,
 * AVX2 and AVX-512; IEEE operations give the same bits at any vector width
 * and nothing is fused) */
#define PLAN_ENERGY_BLOCK_BODY                                                    \
    for (size_t j = 0; j < n; ++j) {                                              \
        const uint32_t k = idx[j];                                                \
        const double tm = S.M + m[k];                                             \
        const double a = S.cx - x[k];                                             \
        const double svx = 0.0 * a, svy = S.sh * a, svz = 0.0 * a;                \
        double n2 = 0;                                                            \
        const double d0 = (S.vx - svx) - vx[k];                                   \
        n2 += d0 * d0;                                                            \
        const double d1 = (S.vy - svy) - vy[k];                                   \
        n2 += d1 * d1;                                                            \
        const double d2 = (S.vz - svz) - vz[k];                                   \
        n2 += d2 * d2;                                                            \
        double q = 0;                                                             \
        const double t0 = S.cx - x[k];                                            \
        q += t0 * t0;                                                             \
        const double t1 = S.cy - y[k];                                            \
        q += t1 * t1;                                                             \
        const double t2 = S.cz - z[k];                                            \
        q += t2 * t2;                                                             \
        e[j] = (0.5 / tm) * n2 - S.G / std::sqrt(q);                              \
    }
#define PLAN_ENERGY_BLOCK_ARGS                                                                                      \
    const EnergyState &S, const double *__restrict x, const double *__restrict y, const double *__restrict z,    \
        const double *__restrict vx, const double *__restrict vy, const double *__restrict vz,                    \
        const double *__restrict m, const uint32_t *__restrict idx, size_t n, double *__restrict e
static void energy_block_base(PLAN_ENERGY_BLOCK_ARGS)
{
    PLAN_ENERGY_BLOCK_BODY
}
__attribute__((target("avx2"))) static void energy_block_avx2(PLAN_ENERGY_BLOCK_ARGS)
{
    PLAN_ENERGY_BLOCK_BODY
}
__attribute__((target("avx512f,avx512bw,avx512vl"))) static void energy_block_avx512(PLAN_ENERGY_BLOCK_ARGS)
{
    PLAN_ENERGY_BLOCK_BODY
}
#undef PLAN_ENERGY_BLOCK_BODY
#undef PLAN_ENERGY_BLOCK_ARGS
typedef void (*EnergyBlockFn)(const EnergyState &, const double *, const double *, const double *, const double *,
                              const double *, const double *, const double *, const uint32_t *, size_t, double *);
static EnergyBlockFn energy_block_fn()
{
    static const EnergyBlockFn f = cpu_has_avx512()                 ? energy_block_avx512
                                   : __builtin_cpu_supports("avx2") ? energy_block_avx2
                                                                    : energy_block_base;
    return f;
}

/* Screening of candidate energies without divisions or roots: with
 * h = |u|^2 / 2 (u as in rebind_energy) and q the squared distance, the
 * energy h / (M + m) - G / sqrt(q) is positive exactly when
 * h^2 q > (G (M + m))^2.  Both sides are formed in binary64 (relative error
 * below 1e-14); when the left exceeds the right by a relative 2e-9, the
 * energy's own binary64 evaluation (a few roundings of 2^-53 on each term)
 * is positive too: the record is unbound and would not be claimed.
 * need[j] = 1 for a record in the ball (F bit 0) without that proof (also
 * when a side leaves 1e-290 .. 1e290 or is not a number).  Records [0, n)
 * of the live arrays, contiguous. */
#define PLAN_ENERGY_SCREEN_BODY                                                   \
    for (size_t j = 0; j < n; ++j) {                                              \
        const double a = S.cx - x[j];                                             \
        const double d0 = (S.vx - 0.0 * a) - vx[j];                               \
        const double d1 = (S.vy - S.sh * a) - vy[j];                              \
        const double d2 = (S.vz - 0.0 * a) - vz[j];                               \
        const double t1 = S.cy - y[j], t2 = S.cz - z[j];                          \
        const double h = 0.5 * (d0 * d0 + d1 * d1 + d2 * d2);                     \
        const double q = a * a + t1 * t1 + t2 * t2;                               \
        const double gt = S.G * (S.M + m[j]);                                     \
        const double lhs = h * h * q, rhs = gt * gt;                              \
        const bool sure = (rhs > 1e-290) & (lhs < 1e290) & (lhs * (1.0 - 1e-9) > rhs * (1.0 + 1e-9)); \
        need[j] = (uint8_t)((F[j] & 1) & !sure);                                  \
    }
#define PLAN_ENERGY_SCREEN_ARGS                                                                                     \
    const EnergyState &S, const double *__restrict x, const double *__restrict y, const double *__restrict z,    \
        const double *__restrict vx, const double *__restrict vy, const double *__restrict vz,                    \
        const double *__restrict m, const uint8_t *__restrict F, size_t n, uint8_t *__restrict need
static void energy_screen_base(PLAN_ENERGY_SCREEN_ARGS)
{
    PLAN_ENERGY_SCREEN_BODY
}
__attribute__((target("avx2"))) static void energy_screen_avx2(PLAN_ENERGY_SCREEN_ARGS)
{
    PLAN_ENERGY_SCREEN_BODY
}
__attribute__((target("avx512f,avx512bw,avx512vl"))) static void energy_screen_avx512(PLAN_ENERGY_SCREEN_ARGS)
{
    PLAN_ENERGY_SCREEN_BODY
}
#undef PLAN_ENERGY_SCREEN_BODY
#undef PLAN_ENERGY_SCREEN_ARGS
typedef void (*EnergyScreenFn)(const EnergyState &, const double *, const double *, const double *, const double *,
                               const double *, const double *, const double *, const uint8_t *, size_t, uint8_t *);
static EnergyScreenFn energy_screen_fn()
{
    static const EnergyScreenFn f = cpu_has_avx512()                 ? energy_screen_avx512
                                    : __builtin_cpu_supports("avx2") ? energy_screen_avx2
                                                                     : energy_screen_base;
    return f;
}

/* One iteration's scan of the live records: squared distance s to the clump
 * centre, bit 0 = within the ball (s <= r2) and near c0, bit 1 = in the band
 * around the sphere; returns the number of band records.  The same
 * expression is compiled for each instruction set, as the energies are. */
struct LiveScan {
    double ox, oy, oz, bound, blo, bhi, r2;
};
#define PLAN_LIVE_SCAN_BODY                                       \
    size_t nband = 0;                                             \
    for (size_t k = 0; k < nl; ++k) {                             \
        const double t0 = q.ox - X[k], t1 = q.oy - Y[k], t2 = q.oz - Z[k]; \
        double s = 0;                                             \
        s += t0 * t0;                                             \
        s += t1 * t1;                                             \
        s += t2 * t2;                                             \
        S[k] = s;                                                 \
        const bool near = LS[k] <= q.bound;                       \
        const bool band = near & (s >= q.blo) & (s <= q.bhi);     \
        F[k] = (uint8_t)((near & (s <= q.r2)) | (band << 1));     \
        nband += band;                                            \
    }                                                             \
    return nband;
static size_t live_scan_base(const LiveScan &q, const double *X, const double *Y, const double *Z, const double *LS,
                             size_t nl, double *S, uint8_t *F)
{
    PLAN_LIVE_SCAN_BODY
}
__attribute__((target("avx2"))) static size_t live_scan_avx2(const LiveScan &q, const double *X, const double *Y,
                                                              const double *Z, const double *LS, size_t nl,
                                                              double *S, uint8_t *F)
{
    PLAN_LIVE_SCAN_BODY
}
__attribute__((target("avx512f,avx512bw,avx512vl"))) static size_t
live_scan_avx512(const LiveScan &q, const double *X, const double *Y, const double *Z, const double *LS, size_t nl,
                 double *S, uint8_t *F)
{
    PLAN_LIVE_SCAN_BODY
}
#undef PLAN_LIVE_SCAN_BODY
typedef size_t (*LiveScanFn)(const LiveScan &, const double *, const double *, const double *, const double *, size_t,
                             double *, uint8_t *);
static LiveScanFn live_scan_fn()
{
    static const LiveScanFn f = cpu_has_avx512()                 ? live_scan_avx512
                                : __builtin_cpu_supports("avx2") ? live_scan_avx2
                                                                 : live_scan_base;
    return f;
}

/* Open-addressing set of record handles (the particles a rebinding run
 * claimed). */
struct HandleSet {
    std::vector<uint32_t> t;
    size_t n = 0;
    HandleSet() : t(64, PLAN_NONE) {}
    static size_t hash(uint32_t h) { return (size_t)(h * 2654435761u); }
    bool count(uint32_t h) const
    {
        const size_t mask = t.size() - 1;
        for (size_t i = hash(h) & mask;; i = (i + 1) & mask) {
            if (t[i] == h) return true;
            if (t[i] == PLAN_NONE) return false;
        }
    }
    void insert(uint32_t h)
    {
        if (2 * (n + 1) > t.size()) {
            std::vector<uint32_t> old(t.size() * 2, PLAN_NONE);
            old.swap(t);
            n = 0;
            for (uint32_t v : old)
                if (v != PLAN_NONE) insert(v);
        }
        const size_t mask = t.size() - 1;
        for (size_t i = hash(h) & mask;; i = (i + 1) & mask) {
            if (t[i] == h) return;
            if (t[i] == PLAN_NONE) {
                t[i] = h;
                ++n;
                return;
            }
        }
    }
};

/* The do-while of SearchBoundParticlesWithinHillRadius + CalculateHillRadius
 * for one clump (src/tree.hpp:4201-4208, 4678-4711).  Each iteration's ball
 * (COM, R_Hill) is cut from the superset with the reference's own distance
 * expression; a candidate within `margin` of the sphere, where the tree's
 * box tests could disagree with the plain distance test, or a ball leaving
 * the superset, sends the iteration to the exact device ball search, which
 * also refreshes the superset.  The margin test needs the distance only
 * where the squared distance lies near (r +- margin)^2; elsewhere the
 * squared distance decides it (the band is widened far beyond rounding).
 * Candidates are tested in order; energies are evaluated in blocks against
 * the current state, and after a claim (which changes the state) the
 * evaluation restarts at the next candidate, so every test sees the state
 * the serial loop would. */
static void rebind_one(HClump &c, Rebind &rb, const HostMembers &hm, const std::vector<uint8_t> &flags,
                       const ClumpSettings &cs, const Ops &ops, ClumpServices &svc)
{
    const auto t_start = std::chrono::steady_clock::now();
    rb.t0 = t_start;
    const double sh = cs.q * cs.Omega;
    HandleSet mine;
    LiveSet live;
    auto rebuild = [&]() {
        live.clear();
        live.reserve(rb.sup.size());
        for (uint32_t h : rb.sup)
            if (!flags[h] && !mine.count(h)) live.push(hm.rec[h], h, Ops::dist_pts(rb.c0, &hm.rec[h].x));
    };
    rebuild();
    rb.live0 = live.size();
    RbLocal L;
    L.reserve(c.indices.size() + rb.sup.size());
    for (uint32_t h : c.indices) L.push(hm.rec[h], h);
    L.ord.resize(L.size());
    for (size_t i = 0; i < L.size(); ++i) L.ord[i] = (uint32_t)i;
    L.n_sorted = L.size();
    const size_t n0 = L.size();
    rb.ms_setup = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t_start).count();
    std::vector<uint32_t> cand;
    std::vector<uint8_t> flag;
    size_t ndead = 0;
    constexpr size_t SCREEN = 256;
    const double inf = std::numeric_limits<double>::infinity();
    const EnergyBlockFn energy_block = energy_block_fn();
    const EnergyScreenFn energy_screen = energy_screen_fn();
    const LiveScanFn live_scan = live_scan_fn();
    alignas(8) uint8_t need[SCREEN];
    size_t before;
    auto claim = [&](uint32_t h, const double *qp, const double *qv, double qm) {
        c.total_mass = c.total_mass + qm;
        for (int k = 0; k < 3; ++k) c.com[k] = (c.com[k] * c.total_mass + qm * qp[k]) / c.total_mass;
        for (int k = 0; k < 3; ++k) c.vel[k] = (c.total_mass * c.vel[k] + qm * qv[k]) / c.total_mass;
        mine.insert(h);
        rb.claimed.push_back(h);
        L.push(hm.rec[h], h);
    };
    do {
        rb.iters++;
        before = L.size();
        const double r = c.hill, r2 = r * r;
        const double margin = cs.ball_margin_abs + 1e-12 * r;
        const double dc = Ops::dist_pts(c.com, rb.c0);
        bool exact = dc + r + margin > rb.R;
        cand.clear();
        size_t klo = live.size(), khi = 0; /* records in the ball lie in [klo, khi) */
        const auto tsc = std::chrono::steady_clock::now();
        if (!exact) {
            /* a member farther than this from c0 is outside the ball by more
             * than the margin (triangle inequality, generous rounding slack) */
            const double bound = (dc + r + margin) * (1 + 1e-9) + 1e-300; /* finite */
            const double blo = r - margin > 0 ? (r - margin) * (r - margin) * (1 - 1e-9) : -1.0;
            const double bhi = (r + margin) * (r + margin) * (1 + 1e-9);
            const double ox = c.com[0], oy = c.com[1], oz = c.com[2];
            const size_t nl = live.size();
            live.s.resize(nl);
            const double *X = live.x.data(), *Y = live.y.data(), *Z = live.z.data();
            double *S = live.s.data();
            const double *LS = live.lsd.data();
            flag.resize(nl);
            uint8_t *F = flag.data();
            const size_t nband = live_scan(LiveScan{ox, oy, oz, bound, blo, bhi, r2}, X, Y, Z, LS, nl, S, F);
            /* a candidate within the margin of the sphere anywhere sends the
             * iteration to the exact search (its candidates are then the
             * device's), so the order of this test does not matter */
            if (nband)
                for (size_t k = 0; k < nl && !exact; ++k)
                    if ((F[k] & 2) && std::fabs(std::sqrt(S[k]) - r) <= margin) exact = true;
            if (!exact) {
                /* first and last record in the ball (flag bit 0), 8 flags
                 * at a time */
                constexpr uint64_t B0 = 0x0101010101010101ull;
                size_t k = 0;
                for (; k + 8 <= nl; k += 8) {
                    uint64_t w;
                    std::memcpy(&w, F + k, 8);
                    if (w & B0) {
                        klo = k + (size_t)(__builtin_ctzll(w & B0) >> 3);
                        break;
                    }
                }
                if (klo == nl)
                    for (; k < nl; ++k)
                        if (F[k] & 1) {
                            klo = k;
                            break;
                        }
                if (klo < nl) {
                    /* F[klo] bit 0 is set, so the walk back stops above klo */
                    size_t t = nl;
                    while (t > klo) {
                        if (t >= klo + 8) {
                            uint64_t w;
                            std::memcpy(&w, F + t - 8, 8);
                            if (!(w & B0)) {
                                t -= 8;
                                continue;
                            }
                            t = t - 8 + (size_t)((63 - __builtin_clzll(w & B0)) >> 3) + 1;
                            break;
                        }
                        if (F[t - 1] & 1) break;
                        --t;
                    }
                    khi = t;
                }
            }
        }
        const auto ten = std::chrono::steady_clock::now();
        rb.ms_scan += std::chrono::duration<double, std::milli>(ten - tsc).count();
        if (exact) {
            auto te = std::chrono::steady_clock::now();
            std::vector<double4> qs{make_double4(c.com[0], c.com[1], c.com[2], r),
                                    make_double4(c.com[0], c.com[1], c.com[2], 2.0 * r)};
            std::vector<std::vector<uint32_t>> l;
            svc.particle_balls(qs, cs.density_threshold, l);
            rb.exact_searches++;
            to_records(hm, l[0], cand);
            to_records(hm, l[1], rb.sup);
            rb.ms_exact += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - te).count();
            for (int k = 0; k < 3; ++k) rb.c0[k] = c.com[k];
            rb.R = 2.0 * r;
            for (uint32_t h : cand) {
                if (flags[h] || mine.count(h)) continue;
                const MemberRec &q = hm.rec[h];
                if (rebind_energy(c, sh, cs.G, &q.x, &q.vx, q.m) < 0) claim(h, &q.x, &q.vx, q.m);
            }
            rebuild();
            ndead = 0;
        } else if (klo < khi) {
            /* the live records in order, in chunks: those in the ball that
             * the screen cannot prove unbound get the exact energy against
             * the current state; after a claim (which changes the state) the
             * screen restarts at the next record with a chunk of 16, and the
             * chunk doubles (up to SCREEN) while no claim interrupts it, so
             * little screening is redone after a claim */
            const uint8_t *F = flag.data();
            size_t k = klo, chunk = 16;
            const size_t kend = khi;
            while (k < kend) {
                const size_t e = std::min(kend, k + chunk);
                const EnergyState es{c.total_mass, c.com[0], c.com[1], c.com[2], c.vel[0], c.vel[1], c.vel[2], sh, cs.G};
                energy_screen(es, live.x.data() + k, live.y.data() + k, live.z.data() + k, live.vx.data() + k,
                              live.vy.data() + k, live.vz.data() + k, live.m.data() + k, F + k, e - k, need);
                const size_t ne = e - k;
                size_t f = e;
                for (size_t t0 = 0; t0 < ne; t0 += 8) {
                    uint64_t w;
                    std::memcpy(&w, need + t0, 8);
                    if (t0 + 8 > ne) w &= ~0ull >> (8 * (t0 + 8 - ne));
                    while (w) {
                        const size_t t = t0 + (size_t)(__builtin_ctzll(w) >> 3);
                        w &= ~(0xffull << (8 * (t - t0)));
                        const uint32_t kk = (uint32_t)(k + t);
                        double ex;
                        energy_block(es, live.x.data(), live.y.data(), live.z.data(), live.vx.data(),
                                     live.vy.data(), live.vz.data(), live.m.data(), &kk, 1, &ex);
                        rb.evals++;
                        if (ex < 0) {
                            f = k + t;
                            break;
                        }
                    }
                    if (f != e) break;
                }
                if (f == e) {
                    k = e;
                    chunk = std::min(SCREEN, 2 * chunk);
                    continue;
                }
                chunk = 16;
                const double qp[3] = {live.x[f], live.y[f], live.z[f]};
                const double qv[3] = {live.vx[f], live.vy[f], live.vz[f]};
                claim(live.h[f], qp, qv, live.m[f]);
                live.lsd[f] = inf; /* beyond every bound: skipped from now on */
                ++ndead;
                k = f + 1;
            }
            if (ndead > live.size() / 4) {
                live.compact();
                ndead = 0;
            }
            rb.ms_energy += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ten).count();
        }
        if (L.size() > L.n_sorted) {
            auto ts = std::chrono::steady_clock::now();
            rb_resort(c, L);
            rb.ms_sort += std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - ts).count();
        }
        ops.hill_radius(c);
    } while (L.size() > before);
    if (L.size() > n0) rb_finish(c, L);
    rb.t1 = std::chrono::steady_clock::now();
    rb.ms = std::chrono::duration<double, std::milli>(rb.t1 - t_start).count();
}

/* Rebinding of all clumps in key order (src/tree.hpp:4196-4208).  Every
 * clump first runs in parallel against the flags set before rebinding (the
 * device ball searches an iteration may need do not depend on flags).  Then,
 * in key order, a run is kept when none of its claims was claimed by an
 * earlier clump; otherwise the clump is rerun from its saved state with the
 * flags as they now stand.  A particle a run only tested, without claiming
 * it, leaves the run's state unchanged whoever owns it, so every kept run is
 * exactly the serial one. */
static void rebind_all(std::map<uint32_t, HClump> &cl, const HostMembers &hm, const ClumpSettings &cs,
                       const Ops &ops, ClumpServices &svc, std::vector<uint8_t> &in_clump, ClumpStats &st)
{
    std::vector<HClump *> v;
    std::vector<double4> qs;
    for (auto &it : cl) {
        v.push_back(&it.second);
        qs.push_back(make_double4(it.second.com[0], it.second.com[1], it.second.com[2], 2.0 * it.second.hill));
    }
    if (v.empty()) return;
    const size_t nc = v.size();
    auto tt = std::chrono::steady_clock::now();
    auto part = [&](int k) {
        auto t = std::chrono::steady_clock::now();
        st.rparts[k] += std::chrono::duration<double, std::milli>(t - tt).count();
        tt = t;
    };
    std::vector<std::vector<uint32_t>> sup;
    svc.particle_balls(qs, cs.density_threshold, sup);
    st.rebind_searches += (uint32_t)qs.size();
    std::vector<Rebind> rb(nc);
    std::vector<HClump> saved(nc);
    std::vector<std::vector<uint32_t>> suprec(nc);
    auto start = [&](size_t i) {
        *v[i] = saved[i];
        rb[i] = Rebind();
        rb[i].sup = suprec[i];
        rb[i].c0[0] = qs[i].x;
        rb[i].c0[1] = qs[i].y;
        rb[i].c0[2] = qs[i].z;
        rb[i].R = qs[i].w;
    };
    for (size_t i = 0; i < nc; ++i) saved[i] = *v[i];
    part(0);
    /* longest first; the cost estimate is the number of superset records a
     * clump may still claim (above threshold, in no clump): a small clump in
     * a rich envelope grows over hundreds of iterations */
    std::vector<size_t> order(nc), cost(nc, 0);
    parallel_for((long)nc, [&](long i) {
        to_records(hm, sup[i], suprec[i]);
        for (uint32_t h : suprec[i]) cost[i] += !in_clump[h];
    });
    for (size_t i = 0; i < nc; ++i) order[i] = i;
    /* a clump that can claim many times its own size grows over many
     * iterations, each re-sorting all of its members */
    std::vector<double> est(nc);
    for (size_t i = 0; i < nc; ++i)
        est[i] = (double)cost[i] * (1.0 + (double)cost[i] / (double)(v[i]->indices.size() + 1));
    std::stable_sort(order.begin(), order.end(), [&](size_t x, size_t y) { return est[x] > est[y]; });
    parallel_for((long)nc, [&](long k) {
        const size_t i = order[k];
        start(i);
        rebind_one(*v[i], rb[i], hm, in_clump, cs, ops, svc);
    });
    part(1);
    auto first = rb[order[0]].t0;
    for (size_t k = 0; k < nc; ++k) first = std::min(first, rb[order[k]].t0);
    for (size_t k = 0; k < nc; ++k) {
        const Rebind &r = rb[order[k]];
        st.rebind_work_ms += r.ms;
        st.rebind_exact_total_ms += r.ms_exact;
        if (r.ms > rb[order[st.rebind_max_slot]].ms) st.rebind_max_slot = (uint32_t)k;
        const double end = std::chrono::duration<double, std::milli>(r.t1 - first).count();
        if (end > st.rebind_last_end_ms) {
            st.rebind_last_end_ms = end;
            st.rebind_last_slot = (uint32_t)k;
            st.rebind_last_ms = r.ms;
        }
    }
    for (size_t i = 0; i < nc; ++i) {
        bool ok = true;
        for (uint32_t h : rb[i].claimed)
            if (in_clump[h]) {
                ok = false;
                break;
            }
        if (!ok) {
            start(i);
            rebind_one(*v[i], rb[i], hm, in_clump, cs, ops, svc);
            st.rebind_reruns++;
        }
        for (uint32_t h : rb[i].claimed) in_clump[h] = 1;
        st.rebind_searches += rb[i].exact_searches;
        st.rebind_iters += rb[i].iters;
        if (rb[i].ms > st.rebind_max_ms) {
            st.rebind_max_ms = rb[i].ms;
            st.rebind_max_iters = rb[i].iters;
            st.rebind_max_n = (uint32_t)v[i]->indices.size();
            st.rebind_max_exact = rb[i].exact_searches;
            st.rebind_max_exact_ms = rb[i].ms_exact;
            st.rebind_max_sort_ms = rb[i].ms_sort;
            st.rebind_max_scan_ms = rb[i].ms_scan;
            st.rebind_max_energy_ms = rb[i].ms_energy;
            st.rebind_max_live = (double)rb[i].live0;
            st.rebind_max_evals = (double)rb[i].evals;
        }
        st.rebind_scan_total_ms += rb[i].ms_scan;
        st.rebind_energy_total_ms += rb[i].ms_energy;
        st.rebind_sort_total_ms += rb[i].ms_sort;
        st.rebind_setup_total_ms += rb[i].ms_setup;
    }
    part(2);
}

/* f on every clump (independent per clump), the largest first */
template <class F> static void for_each_clump_parallel(std::map<uint32_t, HClump> &cl, F f)
{
    std::vector<HClump *> v;
    v.reserve(cl.size());
    for (auto &it : cl) v.push_back(&it.second);
    std::stable_sort(v.begin(), v.end(),
                     [](const HClump *a, const HClump *b) { return a->indices.size() > b->indices.size(); });
    parallel_for((long)v.size(), [&](long i) { f(*v[i]); });
}

void run_clump_pipeline(HostMembers &hm, const ClumpSettings &cs, ClumpServices &svc,
                        std::map<uint32_t, HClump> &cl, double &Mp_max, double &Mp_tot, ClumpStats &st)
{
    Ops ops{hm.rec.data(), cs};
    cl.clear();
    Mp_max = Mp_tot = 0;
    auto tl = std::chrono::steady_clock::now();
    auto lap = [&](int k) {
        auto t = std::chrono::steady_clock::now();
        st.ms[k] += std::chrono::duration<double, std::milli>(t - tl).count();
        tl = t;
    };
    const uint32_t nc = (uint32_t)hm.cpeak.size();
    /* naive clumps in key order, members in the serial walk's discovery order
     * (src/tree.hpp:3766-3809) */
    {
        /* the map in key order, then each clump's member list and
         * SortParticles in parallel: in record order while the records
         * still arrive (each clump waits for its own), else the largest
         * first */
        std::vector<HClump *> hv(nc);
        for (uint32_t c = 0; c < nc; ++c) {
            auto it = cl.emplace_hint(cl.end(), hm.cpeak[c], HClump());
            it->second.peak_index = hm.cpeak[c];
            hv[c] = &it->second;
        }
        std::vector<uint32_t> order(nc);
        for (uint32_t c = 0; c < nc; ++c) order[c] = c;
        if (!hm.wait_records)
            std::stable_sort(order.begin(), order.end(), [&](uint32_t x, uint32_t y) {
                return hm.cstart[x + 1] - hm.cstart[x] > hm.cstart[y + 1] - hm.cstart[y];
            });
        parallel_for((long)nc, [&](long k) {
            const uint32_t c = order[k];
            HClump &h = *hv[c];
            const uint32_t s = hm.cstart[c], e = hm.cstart[c + 1];
            if (hm.wait_records) hm.wait_records(e);
            h.indices.resize(e - s);
            for (uint32_t i = s; i < e; ++i) h.indices[i - s] = i;
            ops.sort_particles(h);
        });
        if (hm.wait_records) hm.wait_records(hm.cstart[nc]);
    }
    std::vector<uint32_t> del;
    if (cs.diffuse < 0.525)
        for (uint32_t k : dispersed_keys(cl, cs.diffuse, false)) cl.erase(k);
    if (!cs.enough_resolution) remove_small_mass_low_peak(cl, hm, cs);
    st.naive = (uint32_t)cl.size();
    lap(0);

    /* 4: merging */
    uint32_t num = (uint32_t)cl.size();
    merge_loop(cl, hm, cs, ops, svc, false, num, st);
    lap(1);
    for (auto &it : cl) ops.hill_radius(it.second);
    merge_loop(cl, hm, cs, ops, svc, true, num, st);
    lap(2);
    st.after_merge = (uint32_t)cl.size();

    /* real peaks into keys (src/tree.hpp:4156-4182) */
    {
        std::vector<uint32_t> old_keys, new_keys;
        for (auto &it : cl) {
            uint32_t pk = it.first;
            for (uint32_t o : it.second.subpeaks)
                if (rho_of_key(hm, pk) < rho_of_key(hm, o)) pk = o;
            if (pk != it.first) {
                old_keys.push_back(it.first);
                new_keys.push_back(pk);
            }
        }
        for (size_t i = 0; i < old_keys.size(); ++i) {
            auto ins = cl.emplace(new_keys[i], HClump());
            if (ins.second) {
                auto orig = cl.find(old_keys[i]);
                std::swap(orig->second, ins.first->second);
                ins.first->second.subpeaks.push_back(orig->first);
                cl.erase(orig);
            } else {
                std::fprintf(stderr, "Real peak index replacing failed.\n");
            }
        }
    }

    lap(3);
    /* 5: unbinding and rebinding */
    for_each_clump_parallel(cl, [&](HClump &h) {
        ops.remove_unbound(h);
        ops.hill_radius(h);
    });
    lap(4);
    std::vector<uint8_t> in_clump(hm.rec.size(), 0);
    for (auto &it : cl)
        for (uint32_t r : it.second.indices) in_clump[r] = 1;
    rebind_all(cl, hm, cs, ops, svc, in_clump, st);

    lap(5);
    /* 6: filters */
    remove_small_mass_low_peak(cl, hm, cs);
    for (uint32_t k : dispersed_keys(cl, std::min(cs.diffuse + 0.1, 0.475), true)) cl.erase(k);

    lap(6);
    /* 7: envelope ripping (src/tree.hpp:4229-4249) */
    {
        std::vector<HClump *> v;
        std::vector<uint32_t> keys;
        for (auto &it : cl) {
            v.push_back(&it.second);
            keys.push_back(it.first);
        }
        std::vector<uint8_t> empty(v.size(), 0);
        std::vector<size_t> big(v.size());
        for (size_t i = 0; i < v.size(); ++i) big[i] = i;
        std::stable_sort(big.begin(), big.end(),
                         [&](size_t x, size_t y) { return v[x]->particles.size() > v[y]->particles.size(); });
        parallel_for((long)v.size(), [&](long k) {
            const size_t i = big[k];
            HClump &c = *v[i];
            while (c.hill < c.particles.back().second) {
                auto low = std::lower_bound(c.particles.begin(), c.particles.end(), PD(0, c.hill), pd_less);
                c.particles.resize(low - c.particles.begin());
                c.indices.resize(c.particles.size());
                if (c.indices.empty()) {
                    empty[i] = 1;
                    break;
                }
                ops.kinematics(c);
                ops.remove_unbound(c);
                ops.hill_radius(c);
            }
        });
        for (size_t i = 0; i < v.size(); ++i)
            if (empty[i]) cl.erase(keys[i]);
    }
    lap(7);
    /* 8 */
    remove_small_mass_low_peak(cl, hm, cs);
    /* 9: clumps centred in the ghost zones */
    for (auto &it : cl) {
        bool in = true;
        for (int k = 0; k < 3; ++k)
            in = in && (it.second.com[k] >= cs.box_min[k] && it.second.com[k] <= cs.box_max[k]);
        if (!in) del.push_back(it.first);
    }
    for (auto k : del) cl.erase(k);
    del.clear();
    /* 10: angular momentum */
    for_each_clump_parallel(cl, [&](HClump &h) { ops.angular_momentum(h); });
    lap(8);
    /* ranking and optional cap (src/tree.hpp:4291-4323) */
    if (!cl.empty()) {
        std::vector<std::pair<uint32_t, double>> pm;
        for (auto &it : cl) pm.push_back(std::make_pair(it.first, it.second.total_mass));
        std::sort(pm.begin(), pm.end(), [](const std::pair<uint32_t, double> &a, const std::pair<uint32_t, double> &b) {
            if (a.second == b.second) return a.first < b.first;
            return a.second < b.second;
        });
        if (cs.num_peaks > 0 && cl.size() > cs.num_peaks) {
            size_t nd = cl.size() - cs.num_peaks;
            for (size_t i = 0; i < nd; ++i) cl.erase(pm[i].first);
            pm.erase(pm.begin(), pm.begin() + nd);
        }
        Mp_tot = 0.;
        for (auto &q : pm) Mp_tot = Mp_tot + q.second;
        Mp_max = pm.back().second;
    }
    lap(9);
}

}  // namespace plan
