/*
 * plan_oracle.cpp — CPU restatement of PLAN's clump-finding hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the parity checker for the HIP
 * implementation under plan_amd/: only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg load it, and only as the checker or as the
 * timed CPU baseline.  Nothing in the product links or calls it.
 *
 * It restates, in its own code, the algorithm of the reference C++ sources
 * (astroboylrx/PLAN, src/tree.hpp, src/tree.cpp, src/global.cpp), keeping
 * every floating-point operation in the reference's order so that decisions
 * made on IEEE binary64 values come out identical.  Build it without FMA
 * contraction (-ffp-contract=off, baseline x86-64), as the reference's own
 * build is (CMakeLists.txt:47,76: -O3 -DNDEBUG, no -march reaches the target).
 * The reference cannot be compiled in this image (src/global.hpp:43-46 needs
 * Boost headers that are not installed), so this restatement is pinned by
 * closed forms, brute force and invariants in tests/ instead.
 *
 * libstdc++'s std::sort / std::push_heap / std::pop_heap / std::sort_heap are
 * used exactly where the reference uses them, on the same element types and
 * comparators, so the tie order of equal keys matches the reference too.
 *
 * Parity status per function (what pins it) is listed in the header comment
 * of each function and in tests/test_oracle_*.py.
 */
#include "../include/plan_params.h"

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>
#include <functional>
#include <map>
#include <numeric>
#include <unordered_map>
#include <utility>
#include <vector>
#ifdef _OPENMP
#include <omp.h>
#endif

namespace orc {

typedef __uint128_t u128;
static const double kPI = 3.14159265358979323846;       /* src/global.hpp:782 */
static const int kMaxLevel = 32;                          /* src/tree.hpp:2459 */
static const unsigned kMaxLeaf = 8;                       /* src/tree.hpp:2492 */

static int g_threads = 1;

/* ---------------------------------------------------------------- params */

/* ReadLisFile's parameter block (src/tree.hpp:1597-1612) after
 * CalculateNewParameters (src/global.cpp:28-47).  Pinned by
 * tests/test_oracle_tree.py::test_derived_parameters_closed_form. */
static void derive(const plan_params *p, uint32_t N, plan_derived *d)
{
    std::memset(d, 0, sizeof(*d));
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
    for (int k = 0; k < 3; ++k)
        d->cell_length[k] = d->box_length[k] / (double)p->box_resolution[k];
    d->cell_volume = ((1.0 * d->cell_length[0]) * d->cell_length[1]) * d->cell_length[2];
    const double four_PI = 4. * kPI;
    d->grav_constant = p->four_pi_G / four_PI;
    d->shear_speed = (d->q * d->Omega) * d->box_length[0];
    d->min_trusted_mass_code_unit =
        ((((3 * std::pow(d->cell_length[0], 3)) * d->Omega) * d->Omega) / 1.0) / d->grav_constant;
    d->mass_total_code_units = ((std::sqrt(2 * kPI) * d->box_length[0]) * d->box_length[1]) * p->solid_to_gas_ratio;
    uint32_t nt = p->num_types > PLAN_MAX_TYPES ? PLAN_MAX_TYPES : p->num_types;
    double frac[PLAN_MAX_TYPES];
    double tmp_sum = 0;
    for (uint32_t i = 0; i < nt; ++i) { frac[i] = p->type_info[i]; tmp_sum += p->type_info[i]; }
    for (uint32_t i = 0; i < nt; ++i) {
        frac[i] /= tmp_sum;
        d->mass_total_code_units = ((p->solid_to_gas_ratio * d->box_length[0]) * d->box_length[1]) * std::sqrt(2. * kPI);
        /* a later snapshot of a series keeps the first one's masses: the
         * reference derives them at loop_begin only (src/tree.hpp:1563-1581) */
        d->mass_per_particle[i] = (frac[i] * d->mass_total_code_units) / (double)(p->series_first_n ? p->series_first_n : N);
    }
}

/* ------------------------------------------------------------- particles */

struct Par {           /* Particle<3> (src/tree.hpp:1154-1198) */
    double pos[3], vel[3];
    double density;
    uint32_t id;
    int32_t property_index;
};

/* ParticleSet::MakeGhostParticles (src/tree.hpp:1626-1754): x copies with the
 * shearing offset and the int-truncation wrap in y, then y copies of the x
 * copies, then y copies of the originals; ids continue from N in that order.
 * Pinned by tests/test_oracle_tree.py::test_ghosts_are_shearing_periodic_images (independent periodic images). */
static void make_ghosts(const plan_params *pp, const plan_derived *d, std::vector<Par> &P)
{
    const uint32_t N = (uint32_t)P.size();
    uint32_t next_id = N;
    double ngw[3], ngmin[3], ngmax[3];
    for (int k = 0; k < 3; ++k) {
        ngw[k] = d->box_half_width[k] - d->ghost_zone_width[k];
        if (ngw[k] < 0) ngw[k] = 0;
        ngmin[k] = d->box_center[k] - ngw[k];
        ngmax[k] = d->box_center[k] + ngw[k];
    }
    const double Lx = d->box_length[0], Ly = d->box_length[1];
    const double t = pp->time;
    std::vector<Par> gx, gy;
    for (uint32_t i = 0; i < N; ++i) {
        if (P[i].pos[0] > ngmax[0]) {
            Par g = P[i];
            g.id = next_id++;
            g.pos[0] -= Lx;
            double ny = g.pos[1] + d->shear_speed * t;
            g.pos[1] = ny - (double)(int)((ny - pp->box_min[1]) / Ly) * Ly;
            gx.push_back(g);
        }
        if (P[i].pos[0] < ngmin[0]) {
            Par g = P[i];
            g.id = next_id++;
            g.pos[0] += Lx;
            double ny = g.pos[1] - d->shear_speed * t;
            g.pos[1] = ny + (double)(int)((pp->box_max[1] - ny) / Ly) * Ly;
            gx.push_back(g);
        }
    }
    for (size_t i = 0; i < gx.size(); ++i) {
        if (gx[i].pos[1] < ngmin[1]) {
            Par g = gx[i]; g.id = next_id++; g.pos[1] += Ly; gy.push_back(g);
        }
        if (gx[i].pos[1] > ngmax[1]) {
            Par g = gx[i]; g.id = next_id++; g.pos[1] -= Ly; gy.push_back(g);
        }
    }
    for (uint32_t i = 0; i < N; ++i) {
        if (P[i].pos[1] < ngmin[1]) {
            Par g = P[i]; g.id = next_id++; g.pos[1] += Ly; gy.push_back(g);
        }
        if (P[i].pos[1] > ngmax[1]) {
            Par g = P[i]; g.id = next_id++; g.pos[1] -= Ly; gy.push_back(g);
        }
    }
    P.insert(P.end(), gx.begin(), gx.end());
    P.insert(P.end(), gy.begin(), gy.end());
}

/* ------------------------------------------------------------ Morton key */

/* BaseMortonKey::Double2Int (src/tree.cpp:106-113): low 32 bits of the
 * binary64 d*(2^32-2) + 1.5*2^52, i.e. round-half-even of d*(2^32-2) for
 * d in [0,1).  Pinned by tests/test_oracle_tree.py::test_double_to_grid_edge_values and test_morton_keys_and_order. */
static inline uint32_t double2int(double d)
{
    double t = d * 4294967294.0;
    t = t + 6755399441055744.0;
    uint64_t b;
    std::memcpy(&b, &t, 8);
    return (uint32_t)b;
}

/* Dilate3_Int32 (src/tree.cpp:162-174): bit i of v moves to bit 3i.
 * Written here as the plain bit scatter; the reference's magic-multiply form
 * yields the same 96-bit value for every 32-bit input (tests pin it). */
static inline u128 dilate3(uint32_t v)
{
    u128 r = 0;
    for (int i = 0; i < 32; ++i)
        if (v >> i & 1u) r |= (u128)1 << (3 * i);
    return r;
}

static inline int key8level(u128 key, int level)         /* src/tree.hpp:2242-2245 */
{
    int shr = 93 - 3 * (level - 1);
    return (int)((key >> shr) & 7u);
}

/* -------------------------------------------------------------- the tree */

struct IPar {          /* BHtree::InternalParticle (src/tree.hpp:2367-2410) */
    double pos[3], vel[3];
    double mass;
    double ath_density;
    double new_density;
    uint32_t original_id;
    bool sink_flag;
    bool in_clump_flag;
};

struct Node {          /* BHtree::TreeNode (src/tree.hpp:2419-2455) */
    double center[3];
    double half_width;
    uint32_t begin, end, parent, first_daughter;
    uint16_t orthant;
    uint8_t num_daughters, level;
};

typedef std::pair<int, double> HeapItem;
struct LessSecond {    /* less_second (src/tree.hpp:3053-3058) */
    bool operator()(const HeapItem &a, const HeapItem &b) const { return a.second < b.second; }
};

static inline double norm2_diff(const double *a, const double *b)
{   /* SmallVec Norm2 of (a - b): ((0 + x^2) + y^2) + z^2 (src/global.hpp:416-424) */
    double s = 0;
    for (int k = 0; k < 3; ++k) { double t = a[k] - b[k]; s += t * t; }
    return s;
}

struct Tree {
    uint32_t n = 0;
    std::vector<IPar> p;
    std::vector<Node> nodes;
    std::vector<int> leaf_nodes, node2leaf;
    std::vector<u128> sorted_keys;       /* kept for tests (the reference frees them) */
    std::unordered_map<uint32_t, std::vector<IPar>> sinks;
    double root_center[3] = {0, 0, 0};
    double half_width = 0;
    double epsilon = 1. / 4294967296.;
    double boxmin[3], boxmax[3], scale[3];
    int level_count[kMaxLevel], level_ptr[kMaxLevel];
    uint32_t num_leaf = 0;
    double to_diagonal = std::sqrt(3.0);
    double cell_length[3], half_cell[3], cell_sq[3], three_half_cell[3];

    void init_morton()     /* MortonKey::InitMortonKey (src/tree.hpp:2304-2311) */
    {
        for (int k = 0; k < 3; ++k) {
            boxmin[k] = root_center[k] - half_width;
            boxmax[k] = root_center[k] + half_width;
            scale[k] = 1.0 / (boxmax[k] - boxmin[k]);
        }
    }
    u128 morton(const double *pos, uint32_t index) const    /* src/tree.hpp:2315-2336 */
    {
        u128 r = (u128)index << 96;
        for (int k = 0; k < 3; ++k) {
            double s = (pos[k] - boxmin[k]) * scale[k];
            r |= dilate3(double2int(s)) << k;
        }
        return r;
    }

    void count_nodes(const std::vector<u128> &key, int level, uint32_t b, uint32_t e, int &num_nodes)
    {   /* CountNodesLeaves (src/tree.hpp:2655-2684) */
        while (b < e) {
            int o = key8level(key[b], level);
            uint32_t s = b;
            while (b < e && key8level(key[b], level) == o) ++b;
            level_count[level]++;
            num_nodes++;
            if (b - s <= kMaxLeaf || level == kMaxLevel - 1) num_leaf++;
            else count_nodes(key, level + 1, s, b, num_nodes);
        }
    }
    void fill(const std::vector<u128> &key, int level, uint32_t b, uint32_t e, uint32_t parent,
              const double *c, double hw)
    {   /* FillTree (src/tree.hpp:2688-2755) */
        while (b < e) {
            int o = key8level(key[b], level);
            uint32_t s = b;
            while (b < e && key8level(key[b], level) == o) ++b;
            uint32_t dgt = (uint32_t)level_ptr[level]++;
            if (nodes[parent].first_daughter == 0) {
                nodes[parent].first_daughter = dgt;
                nodes[parent].num_daughters = 1;
            } else {
                nodes[parent].num_daughters++;
            }
            Node &q = nodes[dgt];
            q.level = (uint8_t)(level + 1);
            q.parent = parent;
            q.begin = s;
            q.end = b;
            q.half_width = hw;
            for (int k = 0; k < 3; ++k) {
                double sign = (o >> k & 1) ? 1.0 : -1.0;      /* Orthant<3> (src/tree.cpp:19-28) */
                q.center[k] = c[k] + hw * sign;
            }
            q.orthant = (uint16_t)o;
            q.first_daughter = 0;
            q.num_daughters = 0;
            if (b - s <= kMaxLeaf || level == kMaxLevel - 1) {
                leaf_nodes[num_leaf] = (int)dgt;
                node2leaf[dgt] = (int)num_leaf;
                num_leaf++;
            } else {
                double cc[3] = {q.center[0], q.center[1], q.center[2]};
                fill(key, level + 1, s, b, dgt, cc, 0.5 * hw);
            }
        }
    }

    /* BHtree::BuildTree (src/tree.hpp:2759-2908) with SortPoints
     * (2608-2651) and MakeSinkParticle (2591-2604).  Pinned by
     * tests/test_oracle_tree.py (invariants + brute-force key order). */
    void build(const plan_params *pp, const plan_derived *d, const std::vector<Par> &ps, bool check_pos)
    {
        (void)pp;
        if (check_pos) {
            double pmin[3], pmax[3];
            for (int k = 0; k < 3; ++k) pmin[k] = pmax[k] = ps[0].pos[k];
            for (size_t i = 1; i < ps.size(); ++i)
                for (int k = 0; k < 3; ++k) {
                    pmin[k] = std::min(ps[i].pos[k], pmin[k]);
                    pmax[k] = std::max(ps[i].pos[k], pmax[k]);
                }
            double extent = 0.0;
            for (int k = 0; k < 3; ++k) extent = std::max(extent, pmax[k] - pmin[k]);
            half_width = d->max_half_width + d->max_ghost_zone_width;
            for (int k = 0; k < 3; ++k) root_center[k] = d->box_center[k];
            if (extent < half_width) {
                half_width = (std::ceil(extent * 100.) + 1) / 100.;
                for (int k = 0; k < 3; ++k) {
                    if (pmax[k] - pmin[k] < 0.5 * d->box_length[k]) {
                        root_center[k] = (pmax[k] + pmin[k]) / 2.0;
                        root_center[k] = std::round(root_center[k] * 1000.) / 1000.;
                    } else {
                        root_center[k] = d->box_center[k];
                    }
                }
            }
        } else if (half_width < 1e-16) {
            half_width = d->max_half_width + d->max_ghost_zone_width;
            for (int k = 0; k < 3; ++k) root_center[k] = d->box_center[k];
        }
        n = (uint32_t)ps.size();
        std::vector<IPar> lst(n);
        for (uint32_t i = 0; i < n; ++i) {
            IPar &q = lst[i];
            for (int k = 0; k < 3; ++k) { q.pos[k] = ps[i].pos[k]; q.vel[k] = ps[i].vel[k]; }
            q.mass = d->mass_per_particle[ps[i].property_index];
            q.ath_density = ps[i].density;
            q.new_density = 0.0;
            q.original_id = ps[i].id;
            q.sink_flag = false;
            q.in_clump_flag = false;
        }
        init_morton();
        epsilon = 2 * half_width / std::pow(2, 31);
        std::vector<u128> key(n);
        for (uint32_t i = 0; i < n; ++i) key[i] = morton(lst[i].pos, i);
        std::sort(key.begin(), key.end(), [](u128 x, u128 y) { return (x << 32) < (y << 32); });
        /* SortPoints: runs of equal 96-bit keys become one sink particle whose
         * representative is the first of the run in sorted order.  The
         * reference reads one element past the end for the last particle
         * (src/tree.hpp:2616); that comparison is taken as "not equal". */
        std::vector<IPar> tmp;
        std::vector<u128> tkey;
        tmp.reserve(n);
        tkey.reserve(n);
        sinks.clear();
        for (uint32_t i = 0; i < n;) {
            uint32_t idx = (uint32_t)(key[i] >> 96);
            tkey.push_back(key[i]);
            if (i + 1 < n && (key[i] << 32) == (key[i + 1] << 32)) {
                IPar rep = lst[idx];
                rep.sink_flag = true;
                uint32_t nid = (uint32_t)tmp.size();
                std::vector<IPar> &mem = sinks[nid];
                mem.push_back(lst[idx]);
                mem.push_back(lst[(uint32_t)(key[i + 1] >> 96)]);
                uint32_t j = i + 2;
                while (j < n && (key[i] << 32) == (key[j] << 32)) {
                    mem.push_back(lst[(uint32_t)(key[j] >> 96)]);
                    ++j;
                }
                tmp.push_back(rep);
                i = j;
            } else {
                tmp.push_back(lst[idx]);
                ++i;
            }
        }
        p.swap(tmp);
        n = (uint32_t)p.size();
        for (auto &it : sinks) {        /* MakeSinkParticle */
            double mass = 0.0;
            double v[3] = {0, 0, 0};
            for (auto &m : it.second) {
                for (int k = 0; k < 3; ++k) v[k] += m.mass * m.vel[k];
                mass += m.mass;
            }
            for (int k = 0; k < 3; ++k) v[k] /= mass;
            p[it.first].mass = mass;
            for (int k = 0; k < 3; ++k) p[it.first].vel[k] = v[k];
        }
        /* level-major node array */
        num_leaf = 0;
        level_count[0] = 1;
        for (int l = 1; l < kMaxLevel; ++l) level_count[l] = 0;
        int num_nodes = 1;
        count_nodes(tkey, 1, 0, n, num_nodes);
        nodes.assign(num_nodes, Node());
        node2leaf.assign(num_nodes, -1);
        leaf_nodes.assign(num_leaf, 0);
        level_ptr[0] = 0;
        for (int l = 1; l < kMaxLevel; ++l) level_ptr[l] = level_ptr[l - 1] + level_count[l - 1];
        Node &r = nodes[0];
        r.first_daughter = 0;
        r.orthant = 0;
        r.num_daughters = 0;
        r.level = 1;
        for (int k = 0; k < 3; ++k) r.center[k] = root_center[k];
        r.half_width = half_width;
        r.begin = 0;
        r.end = n;
        r.parent = 0xffffffffu;
        num_leaf = 0;
        double rc[3] = {root_center[0], root_center[1], root_center[2]};
        fill(tkey, 1, 0, n, 0, rc, 0.5 * half_width);
        sorted_keys.swap(tkey);
        for (int k = 0; k < 3; ++k) {
            cell_length[k] = d->cell_length[k];
            half_cell[k] = cell_length[k] / 2.;
            cell_sq[k] = cell_length[k] * cell_length[k];
            three_half_cell[k] = half_cell[k] * 3.;
        }
    }

    bool within(const double *pos, const double *c, double hw) const     /* src/tree.hpp:2912-2919 */
    {
        for (int k = 0; k < 3; ++k)
            if (!(pos[k] >= c[k] - hw - epsilon && pos[k] <= c[k] + hw + epsilon)) return false;
        return true;
    }
    bool is_leaf(uint32_t nd) const { return nodes[nd].num_daughters == 0; }
    uint32_t node_size(uint32_t nd) const { return nodes[nd].end - nodes[nd].begin; }

    int64_t key2leaf(u128 key, uint32_t nd, int level) const          /* src/tree.hpp:2968-2989 */
    {
        while (!is_leaf(nd)) {
            int o = key8level(key, level);
            int64_t found = -1;
            for (uint32_t c = nodes[nd].first_daughter; c != nodes[nd].first_daughter + nodes[nd].num_daughters; ++c)
                if (nodes[c].orthant == o) { found = c; break; }
            if (found < 0) return -1;
            nd = (uint32_t)found;
            ++level;
        }
        return nd;
    }

    bool sphere_contain(const double *pc, double r, uint32_t nd) const   /* src/tree.hpp:3001-3015 */
    {
        const Node &q = nodes[nd];
        double t = q.half_width * to_diagonal;
        double s = 0;
        for (int k = 0; k < 3; ++k) {
            double m = std::fabs(q.center[k] - pc[k]) + t;
            s += m * m;
        }
        return s < r * r;
    }
    bool sphere_intersect(const double *pc, double r, uint32_t nd) const /* src/tree.hpp:3019-3049 */
    {
        const Node &q = nodes[nd];
        double c2c = norm2_diff(q.center, pc);
        double t = q.half_width * to_diagonal + r;
        if (c2c > t * t) return false;
        if (c2c < r || c2c < q.half_width) return true;     /* quirk: squared vs linear */
        double md = 0;
        for (int k = 0; k < 3; ++k) {
            double lo = q.center[k] - q.half_width;
            double hi = q.center[k] + q.half_width;
            if (pc[k] < lo) md += (pc[k] - lo) * (pc[k] - lo);
            else if (pc[k] > hi) md += (pc[k] - hi) * (pc[k] - hi);
        }
        return md <= r * r;
    }

    static void add2heap(std::vector<HeapItem> &h, unsigned k, int i, double d2)   /* src/tree.hpp:3161-3173 */
    {
        if (h.size() < k) {
            h.push_back(HeapItem(i, d2));
            std::push_heap(h.begin(), h.end(), LessSecond());
        } else if (d2 < h.front().second) {
            std::pop_heap(h.begin(), h.end(), LessSecond());
            h.pop_back();
            h.push_back(HeapItem(i, d2));
            std::push_heap(h.begin(), h.end(), LessSecond());
        }
    }
    void rec_knn(const double *q, uint32_t nd, double r, unsigned k, std::vector<HeapItem> &h) const
    {   /* RecursiveKNN_OpenMP (src/tree.hpp:3177-3193) */
        if (sphere_contain(q, r, nd)) {
            for (uint32_t i = nodes[nd].begin; i != nodes[nd].end; ++i) add2heap(h, k, (int)i, norm2_diff(q, p[i].pos));
        } else if (sphere_intersect(q, r, nd)) {
            if (is_leaf(nd)) {
                for (uint32_t i = nodes[nd].begin; i != nodes[nd].end; ++i) add2heap(h, k, (int)i, norm2_diff(q, p[i].pos));
            } else {
                for (uint32_t c = nodes[nd].first_daughter; c != nodes[nd].first_daughter + nodes[nd].num_daughters; ++c)
                    rec_knn(q, c, r, k, h);
            }
        }
    }
    /* KNN_Search_OpenMP (src/tree.hpp:3198-3251).  Returns false when the
     * start node cannot be located (the reference then reads out of
     * bounds; this restatement refuses instead).  Pinned against brute
     * force by tests/test_oracle_knn.py. */
    bool knn(const double *q, unsigned k, double &r_out, uint32_t *idx, std::vector<HeapItem> &h,
             double est = 0, double *r_final = nullptr) const
    {
        std::vector<HeapItem>().swap(h);
        h.reserve(k);
        double max_dr;
        if (est == 0) {
            double max_dr2 = 0;
            if (within(q, root_center, half_width)) {
                u128 key = morton(q, 0);
                int64_t nd = key2leaf(key, 0, 1);
                if (nd < 0) return false;
                uint32_t u = (uint32_t)nd;
                while (node_size(u) < k / 4) u = nodes[u].parent;
                max_dr2 = nodes[u].half_width * nodes[u].half_width;
            } else {
                for (int d = 0; d < 3; ++d) {
                    double a = root_center[d] - half_width - q[d];
                    double b = q[d] - root_center[d] - half_width;
                    double dx = std::max(std::max(a, 0.0), b);   /* MaxOf (src/global.hpp:1385-1411) */
                    max_dr2 += dx * dx;
                }
            }
            max_dr = std::sqrt(max_dr2);
        } else {
            max_dr = est;
        }
        double used;
        do {
            std::vector<HeapItem>().swap(h);
            h.reserve(k);
            used = max_dr;
            rec_knn(q, 0, max_dr, k, h);
            max_dr *= 2;
        } while (h.size() < k);
        if (r_final) *r_final = used;
        r_out = std::sqrt(h.front().second);
        std::sort_heap(h.begin(), h.end(), LessSecond());
        for (unsigned i = 0; i < h.size(); ++i) idx[i] = (uint32_t)h[i].first;
        return true;
    }
    /* number of candidates the traversal adds at radius r (|S(r)|) */
    void rec_count(const double *q, uint32_t nd, double r, uint64_t &c) const
    {
        if (sphere_contain(q, r, nd)) c += node_size(nd);
        else if (sphere_intersect(q, r, nd)) {
            if (is_leaf(nd)) c += node_size(nd);
            else for (uint32_t d = nodes[nd].first_daughter; d != nodes[nd].first_daughter + nodes[nd].num_daughters; ++d)
                rec_count(q, d, r, c);
        }
    }
    void rec_collect(const double *q, uint32_t nd, double r, std::vector<HeapItem> &all) const
    {
        if (sphere_contain(q, r, nd) || (sphere_intersect(q, r, nd) && is_leaf(nd))) {
            for (uint32_t i = nodes[nd].begin; i != nodes[nd].end; ++i) all.push_back(HeapItem((int)i, norm2_diff(q, p[i].pos)));
        } else if (sphere_intersect(q, r, nd)) {
            for (uint32_t d = nodes[nd].first_daughter; d != nodes[nd].first_daughter + nodes[nd].num_daughters; ++d)
                rec_collect(q, d, r, all);
        }
    }

    void rec_ball_count(const double *c, uint32_t nd, double r, uint32_t &cnt) const  /* src/tree.hpp:3255-3271 */
    {
        if (sphere_contain(c, r, nd)) cnt += nodes[nd].end - nodes[nd].begin;
        else if (sphere_intersect(c, r, nd)) {
            if (is_leaf(nd)) {
                for (uint32_t i = nodes[nd].begin; i != nodes[nd].end; ++i)
                    if (norm2_diff(c, p[i].pos) <= r * r) cnt++;
            } else {
                for (uint32_t d = nodes[nd].first_daughter; d != nodes[nd].first_daughter + nodes[nd].num_daughters; ++d)
                    rec_ball_count(c, d, r, cnt);
            }
        }
    }
    void rec_ball(const double *c, uint32_t nd, double r, std::vector<uint32_t> &out) const  /* src/tree.hpp:3275-3293 */
    {
        if (sphere_contain(c, r, nd)) {
            for (uint32_t i = nodes[nd].begin; i != nodes[nd].end; ++i) out.push_back(i);
        } else if (sphere_intersect(c, r, nd)) {
            if (is_leaf(nd)) {
                for (uint32_t i = nodes[nd].begin; i != nodes[nd].end; ++i)
                    if (norm2_diff(c, p[i].pos) <= r * r) out.push_back(i);
            } else {
                for (uint32_t d = nodes[nd].first_daughter; d != nodes[nd].first_daughter + nodes[nd].num_daughters; ++d)
                    rec_ball(c, d, r, out);
            }
        }
    }
    void ball(const double *c, double r, std::vector<uint32_t> &out) const        /* src/tree.hpp:3297-3304 */
    {
        out.clear();
        rec_ball(c, 0, r, out);
    }

    /* QuadraticSpline (src/tree.hpp:3406-3419) */
    double quadratic_spline(const double *dd) const
    {
        double w = 1;
        for (int k = 0; k < 3; ++k) {
            double x = std::fabs(dd[k]);
            if (x <= half_cell[k]) w *= (0.75 - x * x / cell_sq[k]);
            else if (x < three_half_cell[k]) w *= 0.5 * std::pow(1.5 - x / cell_length[k], 2.);
            else return 0.;
        }
        return w;
    }
};

/* Density kernels.  kind 0 = MedianSphericalDensityKernel (src/tree.hpp:3473-3491),
 * kind 1 = QseudoQuadraticSplinesKernel (src/tree.hpp:3427-3434).  `h` is the
 * sorted heap of (index, d^2) and idx[] the matching tree indices. */
static double density_kernel(const Tree &T, int kind, unsigned K, double cell_volume, uint32_t self_id,
                             const uint32_t *idx, const std::vector<HeapItem> &h)
{
    if (kind == 0) {
        int i = (int)K;
        double rho = 0;
        double tm = 4 * h[i / 2].second;
        for (i = i - 1; i >= 0; i--) {
            if (h[i].second > tm) continue;
            break;
        }
        tm = h[i].second;
        for (; i >= 0; i--) rho += T.p[idx[i]].mass;
        rho /= (kPI * 4. / 3.) * std::pow(tm, 1.5);
        return rho;
    }
    double rho = 0;
    for (unsigned i = 0; i < K; ++i) {
        double dd[3];
        for (int k = 0; k < 3; ++k) dd[k] = T.p[idx[i]].pos[k] - T.p[self_id].pos[k];
        rho += T.p[idx[i]].mass / cell_volume * T.quadratic_spline(dd);
    }
    return rho;
}

/* ------------------------------------------------------------ clumps */

struct Clump {          /* Planetesimal<3> (src/tree.hpp:4351-4809) */
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

typedef std::pair<uint32_t, double> PD;
static bool pd_less(const PD &a, const PD &b) { return a.second < b.second; }

static void kinematics(Clump &c, const std::vector<IPar> &P)    /* src/tree.hpp:4431-4442 */
{
    c.total_mass = 0;
    for (int k = 0; k < 3; ++k) c.com[k] = c.vel[k] = 0;
    for (uint32_t it : c.indices) {
        c.total_mass += P[it].mass;
        for (int k = 0; k < 3; ++k) c.com[k] += P[it].pos[k] * P[it].mass;
        for (int k = 0; k < 3; ++k) c.vel[k] += P[it].vel[k] * P[it].mass;
    }
    for (int k = 0; k < 3; ++k) c.com[k] /= c.total_mass;
    for (int k = 0; k < 3; ++k) c.vel[k] /= c.total_mass;
}

static inline double dist_to(const double *a, const double *c)
{
    return std::sqrt(norm2_diff(a, c));
}

/* radii block shared by SortParticles (src/tree.hpp:4461-4511) and
 * MergeAnotherPlanetesimal (src/tree.hpp:4559-4610) */
static void radii(Clump &c, const std::vector<IPar> &P)
{
    double tenth = P[c.indices[c.indices.size() / 2]].new_density / 10.;
    auto it = c.particles.begin();
    for (; it != c.particles.end(); ++it)
        if (P[it->first].new_density < tenth) break;
    c.inner = (it != c.particles.end()) ? it->second : c.particles.back().second;
    auto rit = c.particles.rbegin();
    for (; rit != c.particles.rend(); ++rit)
        if (P[rit->first].new_density > tenth) break;
    c.outer = rit->second;
    c.one10th = (c.inner + c.outer) / 2.;
    rit = c.particles.rbegin();
    double tm = 0., lim = c.total_mass * (1.0 - 0.95);
    for (; rit != c.particles.rend(); ++rit) {
        tm += P[rit->first].mass;
        if (tm > lim) break;
    }
    c.two_sigma = rit->second;
    if (c.inner > 0 && c.outer / c.inner > 2.5) c.one10th = std::min(c.outer, c.two_sigma);
    else c.one10th = std::min(c.one10th, c.two_sigma);
}

static void sort_particles(Clump &c, const std::vector<IPar> &P)  /* src/tree.hpp:4446-4512 */
{
    kinematics(c, P);
    for (uint32_t it : c.indices) c.particles.push_back(PD(it, dist_to(P[it].pos, c.com)));
    std::sort(c.particles.begin(), c.particles.end(), pd_less);
    for (size_t i = 0; i < c.particles.size(); ++i) c.indices[i] = c.particles[i].first;
    radii(c, P);
}

static bool dispersion_too_large(const Clump &c, double tol = 0.55)   /* src/tree.hpp:4516-4528 */
{
    size_t n = c.particles.size();
    double s = 0.0;
    for (size_t i = 0; i < n; ++i) s = s + c.particles[i].second * c.particles[i].second;
    return s / n > std::pow(tol * c.particles[n - 1].second, 2);
}

static void merge_into(Clump &a, const Clump &b, const std::vector<IPar> &P)   /* src/tree.hpp:4532-4611 */
{
    if (P[a.peak_index].new_density < P[b.peak_index].new_density) {
        a.subpeaks.push_back(a.peak_index);
        a.peak_index = b.peak_index;
    } else {
        a.subpeaks.push_back(b.peak_index);
    }
    double tot = a.total_mass + b.total_mass;
    for (int k = 0; k < 3; ++k) a.com[k] = (a.com[k] * a.total_mass + b.com[k] * b.total_mass) / tot;
    for (int k = 0; k < 3; ++k) a.vel[k] = (a.vel[k] * a.total_mass + b.vel[k] * b.total_mass) / tot;
    a.total_mass = tot;
    a.particles.insert(a.particles.end(), b.particles.begin(), b.particles.end());
    for (auto &it : a.particles) it.second = dist_to(P[it.first].pos, a.com);
    std::sort(a.particles.begin(), a.particles.end(), pd_less);
    a.indices.resize(a.particles.size());
    for (size_t i = 0; i < a.particles.size(); ++i) a.indices[i] = a.particles[i].first;
    radii(a, P);
}

struct Ctx {
    const plan_derived *d;
    double G;
    double shear;          /* shear_vector[1] = q * Omega */
};

static void remove_unbound(Clump &c, const std::vector<IPar> &P, const Ctx &X)   /* src/tree.hpp:4615-4674 */
{
    size_t i = c.particles.size() - 1;
    while (i != 0) {
        const IPar &q = P[c.particles[i].first];
        double tm = c.total_mass - q.mass;
        double tc[3], tv[3];
        for (int k = 0; k < 3; ++k) tc[k] = (c.com[k] * c.total_mass - q.pos[k] * q.mass) / tm;
        for (int k = 0; k < 3; ++k) tv[k] = (c.vel[k] * c.total_mass - q.vel[k] * q.mass) / tm;
        double a = tc[0] - q.pos[0];
        double sv[3] = {0.0 * a, X.shear * a, 0.0 * a};
        double dv[3];
        for (int k = 0; k < 3; ++k) dv[k] = (tv[k] - sv[k]) - q.vel[k];
        double n2 = 0;
        for (int k = 0; k < 3; ++k) n2 += dv[k] * dv[k];
        double e = (0.5 / c.total_mass) * n2 - X.G / dist_to(tc, q.pos);
        if (e > 0) {
            std::swap(c.indices[i], c.indices.back());
            c.indices.pop_back();
            std::swap(c.particles[i], c.particles.back());
            c.particles.pop_back();
            for (int k = 0; k < 3; ++k) { c.vel[k] = tv[k]; c.com[k] = tc[k]; }
            c.total_mass = tm;
        }
        i--;
    }
    std::vector<PD>().swap(c.particles);
    sort_particles(c, P);
}

static void hill_radius(Clump &c, const Ctx &X)     /* src/tree.hpp:4715-4717 */
{
    c.hill = std::pow((((c.total_mass * X.G) / 3.) / X.d->Omega) / X.d->Omega, 1. / 3.);
}

/* the candidate loop of SearchBoundParticlesWithinHillRadius
 * (src/tree.hpp:4685-4704): candidates in ball order; a bound one joins and
 * updates the state, with the reference's centre-of-mass update that uses
 * the new total mass on both sides (src/tree.hpp:4697-4700) */
static void rebind_candidates(Clump &c, std::vector<IPar> &P, const std::vector<uint32_t> &near, double thr,
                              const Ctx &X)
{
    for (uint32_t idx : near) {
        IPar &q = P[idx];
        if (q.in_clump_flag || q.new_density < thr) continue;
        double tm = c.total_mass + q.mass;
        double a = c.com[0] - q.pos[0];
        double sv[3] = {0.0 * a, X.shear * a, 0.0 * a};
        double dv[3];
        for (int k = 0; k < 3; ++k) dv[k] = (c.vel[k] - sv[k]) - q.vel[k];
        double n2 = 0;
        for (int k = 0; k < 3; ++k) n2 += dv[k] * dv[k];
        double e = (0.5 / tm) * n2 - X.G / dist_to(c.com, q.pos);
        if (e < 0) {
            c.indices.push_back(idx);
            c.total_mass = tm;
            for (int k = 0; k < 3; ++k) c.com[k] = (c.com[k] * c.total_mass + q.mass * q.pos[k]) / c.total_mass;
            for (int k = 0; k < 3; ++k) c.vel[k] = (c.total_mass * c.vel[k] + q.mass * q.vel[k]) / c.total_mass;
            q.in_clump_flag = true;
        }
    }
}

static void search_bound(Clump &c, Tree &T, double thr, const Ctx &X)   /* src/tree.hpp:4678-4711 */
{
    std::vector<uint32_t> near;
    T.ball(c.com, c.hill, near);
    rebind_candidates(c, T.p, near, thr, X);
    if (c.indices.size() > c.particles.size()) {
        std::vector<PD>().swap(c.particles);
        sort_particles(c, T.p);
    }
}

static void angular_momentum(Clump &c, const Tree &T, const Ctx &X)   /* src/tree.hpp:4721-4747 */
{
    double quarter = 0.25 * c.hill;
    double hu = ((c.total_mass * c.hill) * c.hill) * X.d->Omega;
    double s1 = X.d->q * X.d->Omega;
    for (const auto &it : c.particles) {
        const IPar &q = T.p[it.first];
        double dr[3], dv[3], j[3];
        for (int k = 0; k < 3; ++k) dr[k] = q.pos[k] - c.com[k];
        double sp[3] = {0. * q.pos[0], s1 * q.pos[0], 0. * q.pos[0]};
        double sc[3] = {0. * c.com[0], s1 * c.com[0], 0. * c.com[0]};
        for (int k = 0; k < 3; ++k) dv[k] = (q.vel[k] - sp[k]) - (c.vel[k] - sc[k]);
        j[0] = dr[1] * dv[2] - dr[2] * dv[1];
        j[1] = dr[2] * dv[0] - dr[0] * dv[2];
        j[2] = dr[0] * dv[1] - dr[1] * dv[0];
        double o[3] = {-dr[0] * dr[2], -dr[1] * dr[2], (dr[0] * dr[0] + dr[1] * dr[1])};
        for (int k = 0; k < 3; ++k) j[k] += X.d->Omega * o[k];
        for (int k = 0; k < 3; ++k) c.J[k] += q.mass * j[k];
        if (it.second < quarter)
            for (int k = 0; k < 3; ++k) c.Jq[k] += q.mass * j[k];
    }
    for (int k = 0; k < 3; ++k) { c.J[k] /= hu; c.Jq[k] /= hu; }
}

/* ranking and cap (src/tree.hpp:4291-4323): sort by (mass, peak index),
 * drop the lightest beyond num_peaks, total and maximum mass */
static void rank_and_cap(std::map<uint32_t, Clump> &cl, uint32_t num_peaks, struct Result &R);

/* ------------------------------------------------------- the whole path */

struct Stages {          /* per-particle intermediate results kept for tests */
    std::vector<uint32_t> densest;     /* densest_neighbor_id_list */
    std::vector<uint32_t> peak;        /* forest root per particle, T for below threshold */
    int kernel = 0;
    unsigned hop = 0;
    uint32_t naive_clumps = 0;
    uint32_t after_merge = 0;
};

struct Result {
    std::map<uint32_t, Clump> clumps;
    double density_threshold = 0, peak_density_threshold = 0, clump_mass_threshold = 0;
    double diffuse = 0.55, hill_frac = 0.25;
    double Mp_max = 0, Mp_tot = 0;
    uint32_t erased0 = 0;      /* clumps with key 0 erased after a merge loop */
};

static void rank_and_cap(std::map<uint32_t, Clump> &cl, uint32_t num_peaks, Result &R)
{
    if (cl.empty()) return;
    std::vector<std::pair<uint32_t, double>> pm;
    for (auto &it : cl) pm.push_back(std::make_pair(it.first, it.second.total_mass));
    std::sort(pm.begin(), pm.end(), [](const std::pair<uint32_t, double> &a, const std::pair<uint32_t, double> &b) {
        if (a.second == b.second) return a.first < b.first;
        return a.second < b.second;
    });
    if (num_peaks > 0 && cl.size() > num_peaks) {
        size_t nd = cl.size() - num_peaks;
        for (size_t i = 0; i < nd; ++i) cl.erase(pm[i].first);
        pm.erase(pm.begin(), pm.begin() + nd);
    }
    R.Mp_tot = 0.;
    for (auto &q : pm) R.Mp_tot = R.Mp_tot + q.second;
    R.Mp_max = pm.back().second;
}

static void remove_small_mass_low_peak(std::map<uint32_t, Clump> &cl, const Tree &T, const Result &R, unsigned K)
{   /* RemoveSmallMassAndLowPeak (src/tree.hpp:3524-3546) */
    std::vector<uint32_t> del;
    for (auto &it : cl) {
        if (it.second.total_mass < R.clump_mass_threshold || T.p[it.first].new_density < R.peak_density_threshold) {
            del.push_back(it.first);
            continue;
        }
        uint32_t lim = it.second.indices.size() < K ? (uint32_t)it.second.indices.size() : K;
        double ori = 0;
        for (uint32_t i = 0; i < lim; ++i) ori = std::max(ori, T.p[it.second.indices[i]].ath_density);
        if (ori < R.peak_density_threshold) del.push_back(it.first);
    }
    for (auto k : del) cl.erase(k);
}

/* BuildClumpTree (src/tree.hpp:5251-5284) */
static void build_clump_tree(Tree &CT, const Tree &T, std::map<uint32_t, Clump> &cl, uint32_t num,
                             double &max_radius, bool hill, double hill_frac,
                             const plan_params *pp, const plan_derived *d)
{
    std::vector<Par> cs(num);
    uint32_t t = 0;
    for (auto &it : cl) {
        if (!it.second.mask) continue;
        Par &q = cs[t];
        for (int k = 0; k < 3; ++k) { q.pos[k] = it.second.com[k]; q.vel[k] = it.second.vel[k]; }
        q.id = t;
        q.density = (double)it.first;
        q.property_index = 0;
        ++t;
        if (hill) max_radius = std::max(it.second.hill * hill_frac, max_radius);   /* MaxOf (src/global.hpp:1387-1390) */
        else max_radius = std::max(it.second.one10th, max_radius);
    }
    CT = Tree();
    for (int k = 0; k < 3; ++k) CT.root_center[k] = T.root_center[k];
    CT.half_width = T.half_width;
    CT.build(pp, d, cs, false);
}

static bool saddle_deep(const Result &R, const Clump &a, const Clump &b, double ra, double rb, double est,
                        const std::function<double(const double *, double)> &saddle)
{   /* IsSaddlePointDeepEnough / IsHillSaddlePointDeepEnough (src/tree.hpp:4880-4909): the
     * density at the point ra / (ra + rb) of the way from a's centre to b's */
    double r12[3], sp[3];
    double f = ra / (ra + rb);
    for (int k = 0; k < 3; ++k) r12[k] = b.com[k] - a.com[k];
    for (int k = 0; k < 3; ++k) r12[k] *= f;
    for (int k = 0; k < 3; ++k) sp[k] = a.com[k] + r12[k];
    return saddle(sp, est) < 2.5 * R.density_threshold;
}

/* the saddle-point density of the real path: KNN with start radius est and
 * the density kernel, with particle 0 as the QS kernel's "self" (4885-4890) */
static double saddle_density(const Tree &T, double *sp, double est, unsigned K, int kind, double cv)
{
    std::vector<HeapItem> h;
    std::vector<uint32_t> idx(K);
    double rk = 0;
    T.knn(sp, K, rk, idx.data(), h, est);
    return density_kernel(T, kind, K, cv, 0, idx.data(), h);
}

static bool grav_bound(const Clump &a, const Clump &b, const Ctx &X)   /* src/tree.hpp:4867-4874 */
{
    double tm = a.total_mass + b.total_mass;
    double P = -X.G / dist_to(a.com, b.com);
    double dx = a.com[0] - b.com[0];
    double sv[3] = {0.0 * dx, X.shear * dx, 0.0 * dx};
    double n2 = 0;
    for (int k = 0; k < 3; ++k) { double t = (a.vel[k] - sv[k]) - b.vel[k]; n2 += t * t; }
    double Ek = (0.5 / tm) * n2;
    return P + Ek < 0.;
}

/* The density at a saddle point (a point and its KNN start radius est),
 * IsSaddlePointDeepEnough's query (src/tree.hpp:4880-4892). */
typedef std::function<double(const double *pt, double est)> SaddleFn;

/* One of the two merge loops (src/tree.hpp:3863-4006 and 4011-4152) with
 * the serial semantics: predators in key order, a clump eaten in this pass
 * cannot eat, merges applied after the pass in order of first appearance;
 * the second loop recomputes the predator's Hill radius after each merge
 * (4119-4126).  num is PlanetesimalList::num_planetesimals: the clump-tree
 * size and the length of the zero-filled list of keys to delete
 * (peaks_to_be_deleted.resize(num), 3874 / 4019).  The list is erased key by
 * key after the loop (4000-4003 / 4146-4149), so its unused zero entries
 * erase the clump whose key is tree index 0 whenever fewer than num clumps
 * were deleted. */
static void merge_loop(Tree &T, Result &R, bool hill, uint32_t &num, const SaddleFn &saddle,
                       const plan_params *pp, const plan_derived *d, const Ctx &X, uint32_t &erased0)
{
    std::map<uint32_t, Clump> &cl = R.clumps;
    double max_radius = 0.;
    const uint32_t num_start = num;
    std::vector<uint32_t> deleted;
    bool happened = true;
    Tree CT;
    while (happened) {
        build_clump_tree(CT, T, cl, num, max_radius, hill, R.hill_frac, pp, d);
        happened = false;
        std::vector<std::pair<uint32_t, uint32_t>> pairs;
        std::vector<uint32_t> near;
        for (auto p1 = cl.begin(); p1 != cl.end(); ++p1) {
            double r1 = hill ? p1->second.hill * R.hill_frac : p1->second.one10th;
            if (!p1->second.mask) continue;
            CT.ball(p1->second.com, r1 + max_radius, near);
            size_t before = pairs.size();
            for (uint32_t ni : near) {
                uint32_t id2 = (uint32_t)CT.p[ni].ath_density;
                if (id2 == p1->first) continue;
                auto p2 = cl.find(id2);
                if (p2 == cl.end() || !p2->second.mask) continue;
                double r2 = hill ? p2->second.hill * R.hill_frac : p2->second.one10th;
                double cd = dist_to(p1->second.com, p2->second.com);
                bool eat = false;
                if (cd < std::max(r1, r2)) eat = true;
                else if (cd < (r1 + r2) && grav_bound(p1->second, p2->second, X)) {
                    double est = hill ? std::min(p1->second.hill, p2->second.hill) / R.hill_frac
                                      : std::min(r1, r2);
                    double ra = hill ? p1->second.hill : r1, rb = hill ? p2->second.hill : r2;
                    if (!saddle_deep(R, p1->second, p2->second, ra, rb, est, saddle)) eat = true;
                }
                if (eat) {
                    p2->second.mask = false;
                    pairs.push_back(std::make_pair(p1->first, p2->first));
                    deleted.push_back(p2->first);
                }
            }
            if (pairs.size() > before) happened = true;
        }
        for (auto &pr : pairs) {
            Clump &a = cl[pr.first];
            merge_into(a, cl[pr.second], T.p);
            if (hill) hill_radius(a, X);
        }
        for (auto &it : cl) {
            if (it.second.mask && dispersion_too_large(it.second)) {
                it.second.mask = false;
                deleted.push_back(it.first);
            }
        }
        num = (uint32_t)(cl.size() - deleted.size());
    }
    for (auto k : deleted) cl.erase(k);
    if (num_start > deleted.size()) erased0 += (uint32_t)cl.erase(0u);
}

/* Steps 4-10 of FindPlanetesimals (src/tree.hpp:3863-4323) on the naive
 * clumps in R, the stages selected by bits of `stages` run in this order:
 * merge loop A (3863-4006), merge loop B with the Hill radii recomputed
 * first (4011-4152), real-peak re-keying (4156-4182), unbinding
 * (4189-4192), rebinding in key order (4196-4208), the mass / peak /
 * grid-density and loose-stream filters (4213-4225), envelope ripping
 * (4229-4249), the second RemoveSmallMassAndLowPeak (4253), ghost-clump
 * removal (4258-4268), angular momenta (4273-4275) and the ranking
 * (4291-4323).  num is num_planetesimals carried between the merge loops;
 * after_merge (may be NULL) receives the clump count after merge loop B. */
enum {
    ST_MERGE_A = 1, ST_MERGE_B = 2, ST_REKEY = 4, ST_UNBIND = 8, ST_REBIND = 16, ST_FILTER6 = 32, ST_RIP = 64,
    ST_FILTER8 = 128, ST_GHOST = 256, ST_J = 512, ST_RANK = 1024, ST_ALL = 2047
};
static void clump_stages(Tree &T, Result &R, const plan_params *pp, const plan_derived *d, const Ctx &X,
                         unsigned K, unsigned stages, uint32_t &num, const SaddleFn &saddle, uint32_t *after_merge)
{
    std::map<uint32_t, Clump> &cl = R.clumps;
    std::vector<uint32_t> del;
    if (stages & ST_MERGE_A) merge_loop(T, R, false, num, saddle, pp, d, X, R.erased0);
    if (stages & ST_MERGE_B) {
        for (auto &it : cl) hill_radius(it.second, X);
        merge_loop(T, R, true, num, saddle, pp, d, X, R.erased0);
    }
    if (after_merge) *after_merge = (uint32_t)cl.size();

    /* real peaks into keys (src/tree.hpp:4156-4182): the densest of the key
     * and the subpeaks, replacing only on a strictly larger density */
    if (stages & ST_REKEY) {
        std::vector<uint32_t> old_keys, new_keys;
        for (auto &it : cl) {
            uint32_t pk = it.first;
            for (uint32_t o : it.second.subpeaks)
                if (T.p[pk].new_density < T.p[o].new_density) pk = o;
            if (pk != it.first) { old_keys.push_back(it.first); new_keys.push_back(pk); }
        }
        for (size_t i = 0; i < old_keys.size(); ++i) {
            auto ins = cl.emplace(new_keys[i], Clump());
            if (ins.second) {
                auto orig = cl.find(old_keys[i]);
                std::swap(orig->second, ins.first->second);
                ins.first->second.subpeaks.push_back(orig->first);
                cl.erase(orig);
            }
        }
    }

    /* 5: unbinding and rebinding */
    if (stages & ST_UNBIND)
        for (auto &it : cl) { remove_unbound(it.second, T.p, X); hill_radius(it.second, X); }
    if (stages & ST_REBIND) {
        for (auto &it : cl)
            for (uint32_t m : it.second.indices) T.p[m].in_clump_flag = true;
        for (auto &it : cl) {
            size_t before;
            do {
                before = it.second.indices.size();
                search_bound(it.second, T, R.density_threshold, X);
                hill_radius(it.second, X);
            } while (it.second.indices.size() > before);
        }
    }
    /* 6: filters */
    if (stages & ST_FILTER6) {
        remove_small_mass_low_peak(cl, T, R, K);
        for (auto &it : cl)
            if (it.second.mask && dispersion_too_large(it.second, std::min(R.diffuse + 0.1, 0.475)))
                del.push_back(it.first);
        for (auto k : del) cl.erase(k);
        del.clear();
    }
    /* 7: envelope ripping: cut at the Hill radius (lower_bound on the
     * distances), recompute, unbind, until the Hill sphere holds every member */
    if (stages & ST_RIP) {
        for (auto &it : cl) {
            Clump &c = it.second;
            while (c.hill < c.particles.back().second) {
                auto low = std::lower_bound(c.particles.begin(), c.particles.end(), PD(0, c.hill), pd_less);
                c.particles.resize(low - c.particles.begin());
                c.indices.resize(c.particles.size());
                if (c.indices.size() == 0) { del.push_back(it.first); break; }
                kinematics(c, T.p);
                remove_unbound(c, T.p, X);
                hill_radius(c, X);
            }
        }
        for (auto k : del) cl.erase(k);
        del.clear();
    }
    /* 8 */
    if (stages & ST_FILTER8) remove_small_mass_low_peak(cl, T, R, K);
    /* 9: clumps whose centre lies outside [box_min, box_max] (inclusive) */
    if (stages & ST_GHOST) {
        for (auto &it : cl) {
            bool in = true;
            for (int k = 0; k < 3; ++k)
                in = in && (it.second.com[k] >= pp->box_min[k] && it.second.com[k] <= pp->box_max[k]);
            if (!in) del.push_back(it.first);
        }
        for (auto k : del) cl.erase(k);
        del.clear();
    }
    /* 10 */
    if (stages & ST_J)
        for (auto &it : cl) angular_momentum(it.second, T, X);
    if (stages & ST_RANK) rank_and_cap(cl, pp->num_peaks, R);
}

/* BHtree::FindPlanetesimals (src/tree.hpp:3576-3612, 3619-4339), serial
 * semantics for the merge loops; the density and HOP loops are
 * embarrassingly parallel and run under OpenMP like the reference's. */
static void find_planetesimals(Tree &T, const plan_params *pp, const plan_derived *d, uint32_t N,
                               Result &R, Stages &S)
{
    Ctx X;
    X.d = d;
    X.G = d->grav_constant;
    X.shear = d->q * d->Omega;
    unsigned K = pp->num_knn, H = pp->num_hop;
    uint32_t hres = pp->box_resolution[0] * pp->box_resolution[1];
    if (T.half_width < d->max_half_width + d->max_ghost_zone_width - 1e-16)
        hres = (uint32_t)std::pow(T.half_width / d->cell_length[0] * ((T.half_width - d->max_ghost_zone_width) / T.half_width), 2.0);
    int kind;
    int enough = 0;
    if (N >= 4 * hres) {
        if (!pp->fixed_num_hop) {
            if (N >= 1.67e7) H = 64;
            else if (N >= 2.68e8) H = 128;
        }
        enough = 1;
        kind = 0;
    } else {
        kind = 1;
    }
    S.kernel = kind;
    S.hop = H;
    R.density_threshold = std::pow(d->Omega, 2.) / kPI / d->grav_constant * 2.;
    if (d->mass_total_code_units / N / d->cell_volume > R.density_threshold / 4) R.density_threshold *= 9. / 8.;
    double mn = 1e300;
    for (int k = 0; k < 3; ++k) mn = std::min(mn, (double)pp->box_resolution[k] / d->box_length[k]);
    unsigned hrh = (unsigned)mn;
    if (hrh <= 1024) { R.diffuse = 0.35; R.hill_frac = 0.75; }
    else if (hrh <= 1536) { R.diffuse = 0.4; R.hill_frac = 0.5; }
    else if (hrh <= 2048) { R.diffuse = 0.5; R.hill_frac = 0.35; }
    if (pp->clump_diffuse_threshold > 0) R.diffuse = pp->clump_diffuse_threshold;
    if (pp->Hill_fraction_for_merge > 0) R.hill_frac = pp->Hill_fraction_for_merge;
    R.clump_mass_threshold = d->min_trusted_mass_code_unit;
    R.peak_density_threshold = 3 * R.density_threshold;
    const uint32_t n = T.n;

    /* 1: densities */
#pragma omp parallel num_threads(g_threads)
    {
        std::vector<uint32_t> idx(K);
        std::vector<HeapItem> h;
        double rk;
#pragma omp for schedule(static)
        for (uint32_t i = 0; i < n; i++) {
            T.knn(T.p[i].pos, K, rk, idx.data(), h);
            T.p[i].new_density = density_kernel(T, kind, K, d->cell_volume, i, idx.data(), h);
        }
    }
    /* 2: densest neighbour */
    S.densest.assign(n, 0);
#pragma omp parallel num_threads(g_threads)
    {
        std::vector<uint32_t> idx(H);
        std::vector<HeapItem> h;
        double rk;
#pragma omp for schedule(dynamic, 256)
        for (uint32_t i = 0; i < n; i++) {
            S.densest[i] = i;
            if (T.p[i].new_density < R.density_threshold) continue;
            T.knn(T.p[i].pos, H, rk, idx.data(), h);
            double best = T.p[i].new_density;
            for (unsigned j = 0; j < H; ++j) {
                double dj = T.p[idx[j]].new_density;
                if (best < dj) { S.densest[i] = idx[j]; best = dj; }
                else if (best == dj && S.densest[i] > idx[j]) S.densest[i] = idx[j];
            }
        }
    }
    /* 3: hop chains (src/tree.hpp:3766-3809) */
    std::vector<uint8_t> mask(n, 1);
    S.peak.assign(n, 0);
    std::map<uint32_t, Clump> &cl = R.clumps;
    cl.clear();
    std::vector<uint32_t> chain_list;
    for (uint32_t i = 0; i < n; ++i) {
        if (!mask[i]) continue;
        if (T.p[i].new_density < R.density_threshold) {
            mask[i] = 0;
            S.peak[i] = n;
            continue;
        }
        chain_list.clear();
        S.peak[i] = i;
        uint32_t c = i;
        chain_list.push_back(c);
        mask[c] = 0;
        c = S.densest[i];
        while (chain_list.back() != c) {
            if (mask[c]) {
                chain_list.push_back(c);
                mask[c] = 0;
                c = S.densest[c];
            } else break;
        }
        uint32_t pk = (chain_list.back() == c) ? c : S.peak[c];
        for (uint32_t m : chain_list) S.peak[m] = pk;
        auto ins = cl.emplace(S.peak[i], Clump());
        if (ins.second) ins.first->second.peak_index = S.peak[i];
        ins.first->second.indices.insert(ins.first->second.indices.end(), chain_list.begin(), chain_list.end());
    }
    std::vector<uint32_t> del;
    for (auto &it : cl) {
        sort_particles(it.second, T.p);
        if (R.diffuse < 0.525 && dispersion_too_large(it.second, R.diffuse)) del.push_back(it.first);
    }
    if (R.diffuse < 0.525) for (auto k : del) cl.erase(k);
    del.clear();
    if (enough == 0) remove_small_mass_low_peak(cl, T, R, K);
    S.naive_clumps = (uint32_t)cl.size();

    /* 4-10 */
    const double cv = d->cell_volume;
    uint32_t num = (uint32_t)cl.size();
    clump_stages(T, R, pp, d, X, K, ST_ALL, num, [&](const double *pt, double est) {
        double sp[3] = {pt[0], pt[1], pt[2]};
        return saddle_density(T, sp, est, K, kind, cv);
    }, &S.after_merge);
}

} /* namespace orc */

/* ================================================================ C ABI */

using namespace orc;

struct orc_tree {
    plan_params pp;
    plan_derived d;
    uint32_t N;               /* real particles */
    Tree T;
    Result R;
    Stages S;
    bool found = false;
};

extern "C" {

void orc_set_threads(int n) { g_threads = n > 0 ? n : 1; }

void orc_derive(const plan_params *pp, uint32_t N, plan_derived *out) { derive(pp, N, out); }

/* Ghost particles for N particles given as AoS pos/vel (N*3), density,
 * property index and id.  Returns T = N + ghosts; call with out arrays NULL
 * to get T only. */
uint32_t orc_make_ghosts(const plan_params *pp, uint32_t N, const double *pos, const double *vel,
                         const double *dens, const int32_t *prop, const uint32_t *id,
                         double *opos, double *ovel, double *odens, int32_t *oprop, uint32_t *oid)
{
    plan_derived d;
    derive(pp, N, &d);
    std::vector<Par> P(N);
    for (uint32_t i = 0; i < N; ++i) {
        for (int k = 0; k < 3; ++k) { P[i].pos[k] = pos[3 * i + k]; P[i].vel[k] = vel[3 * i + k]; }
        P[i].density = dens[i];
        P[i].property_index = prop[i];
        P[i].id = id[i];
    }
    if (!pp->no_ghost) make_ghosts(pp, &d, P);
    uint32_t T = (uint32_t)P.size();
    if (opos) {
        for (uint32_t i = 0; i < T; ++i) {
            for (int k = 0; k < 3; ++k) { opos[3 * i + k] = P[i].pos[k]; ovel[3 * i + k] = P[i].vel[k]; }
            odens[i] = P[i].density;
            oprop[i] = P[i].property_index;
            oid[i] = P[i].id;
        }
    }
    return T;
}

/* Build the tree over T particles (after ghosts).  N = real particle count
 * (sets the particle masses and the kernel choice). */
orc_tree *orc_tree_build(const plan_params *pp, uint32_t N, uint32_t T, const double *pos, const double *vel,
                         const double *dens, const int32_t *prop, const uint32_t *id)
{
    orc_tree *h = new orc_tree();
    h->pp = *pp;
    h->N = N;
    derive(pp, N, &h->d);
    std::vector<Par> P(T);
    for (uint32_t i = 0; i < T; ++i) {
        for (int k = 0; k < 3; ++k) { P[i].pos[k] = pos[3 * i + k]; P[i].vel[k] = vel[3 * i + k]; }
        P[i].density = dens[i];
        P[i].property_index = prop[i];
        P[i].id = id[i];
    }
    h->T.build(pp, &h->d, P, true);
    return h;
}

void orc_tree_free(orc_tree *h) { delete h; }

uint32_t orc_tree_num_particles(const orc_tree *h) { return h->T.n; }
uint32_t orc_tree_num_nodes(const orc_tree *h) { return (uint32_t)h->T.nodes.size(); }
uint32_t orc_tree_num_leaves(const orc_tree *h) { return (uint32_t)h->T.leaf_nodes.size(); }
uint32_t orc_tree_num_sinks(const orc_tree *h) { return (uint32_t)h->T.sinks.size(); }
void orc_tree_root(const orc_tree *h, double *center, double *hw)
{
    for (int k = 0; k < 3; ++k) center[k] = h->T.root_center[k];
    *hw = h->T.half_width;
}

/* keys as three 32-bit words (bits 0-31, 32-63, 64-95) + particle index */
void orc_tree_keys(const orc_tree *h, uint32_t *w0, uint32_t *w1, uint32_t *w2, uint32_t *pidx)
{
    const Tree &T = h->T;
    for (size_t i = 0; i < T.sorted_keys.size(); ++i) {
        u128 k = T.sorted_keys[i];
        w0[i] = (uint32_t)k;
        w1[i] = (uint32_t)(k >> 32);
        w2[i] = (uint32_t)(k >> 64);
        pidx[i] = (uint32_t)(k >> 96);
    }
}

void orc_tree_particles(const orc_tree *h, double *pos, double *vel, double *mass, double *ath,
                        double *rho, uint32_t *oid, uint8_t *sink)
{
    const Tree &T = h->T;
    for (uint32_t i = 0; i < T.n; ++i) {
        const IPar &q = T.p[i];
        for (int k = 0; k < 3; ++k) { pos[3 * i + k] = q.pos[k]; vel[3 * i + k] = q.vel[k]; }
        mass[i] = q.mass;
        ath[i] = q.ath_density;
        rho[i] = q.new_density;
        oid[i] = q.original_id;
        sink[i] = q.sink_flag;
    }
}

/* ---- density against scale (BasicAnalysesWithTree, src/analyses.hpp:196-333) ---- */

struct ScaleQuery {
    const Tree &T;
    double boxmin[3], boxmax[3], boxlen[3];
    double maxd[3], shear;
};

/* ShearedCenter2Center (src/tree.hpp:3307-3330): the node centre's y moves
 * with the shear once the x offset wraps, and the moved y feeds the y offset */
static void sheared_c2c(const ScaleQuery &q, const double *c, const Node &nd, double *c2c)
{
    double nc[3] = {nd.center[0], nd.center[1], nd.center[2]};
    for (int i = 0; i < 3; ++i) {
        c2c[i] = c[i] - nc[i];
        if (std::fabs(c2c[i]) > q.maxd[i]) {
            if (c2c[i] > 0) c2c[i] -= q.boxlen[i];
            else c2c[i] += q.boxlen[i];
            if (i == 0) {
                if (c[0] < 0) {
                    nc[1] -= q.shear;
                    nc[1] += static_cast<int>((q.boxmax[1] - nc[1]) / q.boxlen[1]) * q.boxlen[1];
                } else {
                    nc[1] += q.shear;
                    nc[1] -= static_cast<int>((nc[1] - q.boxmin[1]) / q.boxlen[1]) * q.boxlen[1];
                }
            }
        }
    }
}

/* RecursiveBallSearchCountWithShear with SphereContainNodeWithShear and
 * SphereNodeIntersectWithShear (src/tree.hpp:3334-3402); leaf particles are
 * tested at their own (unshifted) positions */
static void count_sheared(const ScaleQuery &q, const double *c, uint32_t node, double r, uint32_t &count)
{
    const Node &nd = q.T.nodes[node];
    double c2c[3];
    sheared_c2c(q, c, nd, c2c);
    const double diag = nd.half_width * std::sqrt(3.0);
    double t[3];
    for (int i = 0; i < 3; ++i) t[i] = std::fabs(c2c[i]) + diag;
    double n2 = 0;
    for (int i = 0; i < 3; ++i) n2 += t[i] * t[i];
    if (n2 < r * r) {
        count += nd.end - nd.begin;
        return;
    }
    double c2 = 0;
    for (int i = 0; i < 3; ++i) c2 += c2c[i] * c2c[i];
    const double td = diag + r;
    if (c2 > td * td) return;
    bool hit = c2 < r || c2 < nd.half_width;
    if (!hit) {
        double md = 0;
        for (int i = 0; i < 3; ++i) {
            const double lo = (c[i] - c2c[i]) - nd.half_width, hi = (c[i] - c2c[i]) + nd.half_width;
            if (c[i] < lo) md += (c[i] - lo) * (c[i] - lo);
            else if (c[i] > hi) md += (c[i] - hi) * (c[i] - hi);
        }
        hit = md <= r * r;
    }
    if (!hit) return;
    if (q.T.is_leaf(node)) {
        for (uint32_t p = nd.begin; p != nd.end; ++p)
            if (norm2_diff(c, q.T.p[p].pos) <= r * r) ++count;
    } else {
        for (uint32_t dd = nd.first_daughter; dd != nd.first_daughter + nd.num_daughters; ++dd)
            count_sheared(q, c, dd, r, count);
    }
}

/* one sheared ball count from the root (for tests) */
uint32_t orc_sheared_count(const orc_tree *h, const double *c, double r, const double *maxd, double shear)
{
    ScaleQuery q{h->T, {}, {}, {}, {}, shear};
    for (int k = 0; k < 3; ++k) {
        q.boxmin[k] = h->pp.box_min[k];
        q.boxmax[k] = h->pp.box_max[k];
        q.boxlen[k] = h->d.box_length[k];
        q.maxd[k] = maxd[k];
    }
    uint32_t cnt = 0;
    count_sheared(q, c, 0, r, cnt);
    return cnt;
}

/* RecursiveFindExtreme along z (src/analyses.hpp:171-191) */
static void find_extreme(const Tree &T, uint32_t node, uint32_t &ext, bool smaller)
{
    auto comp = [smaller](double a, double b) { return smaller ? a < b : a > b; };
    const Node &nd = T.nodes[node];
    if (T.is_leaf(node)) {
        for (uint32_t p = nd.begin; p != nd.end; ++p)
            if (comp(T.p[p].pos[2], T.p[ext].pos[2])) ext = p;
        return;
    }
    for (uint32_t dd = nd.first_daughter; dd != nd.first_daughter + nd.num_daughters; ++dd) {
        if (comp(T.nodes[dd].center[2], nd.center[2])) find_extreme(T, dd, ext, smaller);
        else if (comp(nd.center[2], T.p[ext].pos[2])) find_extreme(T, dd, ext, smaller);
    }
}

/* the density-vs-scale pass: per level i, radius r_i = (cell/2) 2^i and
 * the largest sheared ball count over the cell-centre grid (levels 0-2 also
 * at the 8 half-cell offsets, level 0 at the quarter-cell offsets the
 * reference's loop visits), as max count x m_particle / (4/3 pi r^3).
 * out_sums[i] = sum of the counts of level i over every search made (the
 * reference repeats each level-0 quarter-cell search 8 times; counted once
 * here).  Returns the number of levels (at most cap written). */
uint32_t orc_density_vs_scale(const orc_tree *h, uint32_t cap, double *radii, double *rhop, uint64_t *sums,
                              double *zrange)
{
    const Tree &T = h->T;
    const plan_derived &d = h->d;
    const plan_params &pp = h->pp;
    uint32_t emin = 0, emax = 0;
    find_extreme(T, 0, emin, true);
    find_extreme(T, 0, emax, false);
    double z_min = d.box_center[2] - d.cell_length[2] / 2.0;
    const double z_max = T.p[emax].pos[2];
    while (z_min - d.cell_length[2] > T.p[emin].pos[2]) z_min -= d.cell_length[2];
    zrange[0] = z_min;
    zrange[1] = z_max;
    const int L = (int)std::round(std::log10(d.box_half_width[0] * 2 / d.cell_length[0]) / std::log10(2.0)) + 1;
    std::vector<double> rad(L, d.cell_length[0] / 2.0), best(L, 0.0);
    std::vector<uint64_t> sum(L, 0);
    ScaleQuery q{T, {}, {}, {}, {}, d.shear_speed * pp.time};
    for (int k = 0; k < 3; ++k) {
        q.boxmin[k] = pp.box_min[k];
        q.boxmax[k] = pp.box_max[k];
        q.boxlen[k] = d.box_length[k];
    }
    std::vector<std::array<double, 3>> maxd(L);
    for (int i = 0; i < L; ++i) {
        rad[i] *= std::pow(2.0, i);
        for (int k = 0; k < 3; ++k) maxd[i][k] = d.box_length[k] - rad[i];
        maxd[i][2] = 100 * d.max_half_width;
    }
    static const int orth[8][3] = {{-1, -1, -1}, {1, -1, -1}, {-1, 1, -1}, {1, 1, -1},
                                   {-1, -1, 1},  {1, -1, 1},  {-1, 1, 1},  {1, 1, 1}};
    double cmin[3], step[3], sub[3], subsub[3];
    for (int k = 0; k < 3; ++k) {
        cmin[k] = pp.box_min[k] + 0.5 * d.cell_length[k];
        step[k] = d.cell_length[k];
        sub[k] = step[k] / 2.0;
        subsub[k] = step[k] / 4.0;
    }
    auto search = [&](int i, const double *c) {
        for (int k = 0; k < 3; ++k) q.maxd[k] = maxd[i][k];
        uint32_t cnt = 0;
        count_sheared(q, c, 0, rad[i], cnt);
        if (cnt > best[i]) best[i] = cnt;
        sum[i] += cnt;
    };
    for (double z = z_min; z < z_max; z += step[2])
        for (double y = cmin[1]; y < pp.box_max[1]; y += step[1])
            for (double x = cmin[0]; x < pp.box_max[0]; x += step[0]) {
                const double cc[3] = {x, y, z};
                for (int i = 0; i < L; ++i) {
                    search(i, cc);
                    if (i < 3)
                        for (int o = 0; o < 8; ++o) {
                            double sc[3];
                            for (int k = 0; k < 3; ++k) sc[k] = cc[k] + sub[k] * orth[o][k];
                            search(i, sc);
                            if (i == 0) {
                                double ssc[3];
                                for (int k = 0; k < 3; ++k) ssc[k] = sc[k] + subsub[k] * orth[o][k];
                                search(i, ssc);
                            }
                        }
                }
            }
    for (int i = 0; i < L && i < (int)cap; ++i) {
        radii[i] = rad[i];
        rhop[i] = best[i] * (d.mass_per_particle[0] / (4.1887902047863910 * std::pow(rad[i], 3)));
        sums[i] = sum[i];
    }
    return (uint32_t)L;
}

/* positions of the particles merged into tree particle i (sink_particle_indices,
 * src/tree.hpp:2608-2651), in sorted order; a plain particle gives its own.
 * Writes at most cap positions; returns the count. */
uint32_t orc_tree_members(const orc_tree *h, uint32_t i, double *pos, uint32_t cap)
{
    const Tree &T = h->T;
    auto it = T.sinks.find(i);
    if (it == T.sinks.end()) {
        if (cap >= 1)
            for (int k = 0; k < 3; ++k) pos[k] = T.p[i].pos[k];
        return 1;
    }
    const uint32_t n = (uint32_t)it->second.size();
    for (uint32_t j = 0; j < n && j < cap; ++j)
        for (int k = 0; k < 3; ++k) pos[3 * j + k] = it->second[j].pos[k];
    return n;
}

void orc_tree_nodes(const orc_tree *h, double *center, double *hw, uint32_t *begin, uint32_t *end,
                    uint32_t *parent, uint32_t *first_daughter, uint8_t *num_daughters, uint8_t *level,
                    uint16_t *orthant, int32_t *node2leaf, int32_t *leaf_nodes)
{
    const Tree &T = h->T;
    for (size_t i = 0; i < T.nodes.size(); ++i) {
        const Node &q = T.nodes[i];
        for (int k = 0; k < 3; ++k) center[3 * i + k] = q.center[k];
        hw[i] = q.half_width;
        begin[i] = q.begin;
        end[i] = q.end;
        parent[i] = q.parent;
        first_daughter[i] = q.first_daughter;
        num_daughters[i] = q.num_daughters;
        level[i] = q.level;
        orthant[i] = q.orthant;
        node2leaf[i] = T.node2leaf[i];
    }
    for (size_t i = 0; i < T.leaf_nodes.size(); ++i) leaf_nodes[i] = T.leaf_nodes[i];
}

/* KNN queries.  q: nq points (AoS xyz); est[i] = 0 for the reference's
 * start-radius rule or the given start radius.  Outputs per query: K tree
 * indices and d^2 in ascending order, the K-th distance, the radius at
 * which the search stopped, |S(r*)| and the number of candidates in S(r*)
 * at exactly the K-th distance that the heap dropped (a tie at the
 * boundary, where the reference's heap order decides).  Returns 0 or the
 * index+1 of a query whose start node does not exist. */
uint32_t orc_knn(const orc_tree *h, const double *q, uint32_t nq, uint32_t K, const double *est,
                 uint32_t *idx, double *d2, double *rk, double *rstop, uint32_t *ncand, uint32_t *ndrop)
{
    const Tree &T = h->T;
    uint32_t bad = 0;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 64)
    for (uint32_t i = 0; i < nq; ++i) {
        std::vector<HeapItem> hp;
        double r = 0, rs = 0;
        if (!T.knn(q + 3 * i, K, r, idx + (size_t)K * i, hp, est ? est[i] : 0.0, &rs)) {
#pragma omp critical
            bad = i + 1;
            continue;
        }
        for (uint32_t j = 0; j < K; ++j) d2[(size_t)K * i + j] = hp[j].second;
        rk[i] = r;
        rstop[i] = rs;
        std::vector<HeapItem> all;
        T.rec_collect(q + 3 * i, 0, rs, all);
        ncand[i] = (uint32_t)all.size();
        double kth = hp[K - 1].second;
        uint32_t in_s = 0, in_h = 0;
        for (auto &a : all) if (a.second == kth) in_s++;
        for (uint32_t j = 0; j < K; ++j) if (hp[j].second == kth) in_h++;
        ndrop[i] = in_s - in_h;
    }
    return bad;
}

/* The reference's candidate set S(r*) of one KNN query (every particle the
 * traversal at the final radius offers to the heap, src/tree.hpp:3177-3193):
 * tree indices and squared distances, at most cap; returns |S(r*)|. */
uint32_t orc_knn_candidates(const orc_tree *h, const double *q, uint32_t K, double est, uint32_t *idx, double *d2,
                            uint32_t cap)
{
    const Tree &T = h->T;
    std::vector<HeapItem> hp;
    std::vector<uint32_t> tmp(K);
    double r = 0, rs = 0;
    if (!T.knn(q, K, r, tmp.data(), hp, est, &rs)) return 0;
    std::vector<HeapItem> all;
    T.rec_collect(q, 0, rs, all);
    for (size_t i = 0; i < all.size() && i < cap; ++i) {
        idx[i] = (uint32_t)all[i].first;
        d2[i] = all[i].second;
    }
    return (uint32_t)all.size();
}

/* ball search: returns count; fills out (cap entries) in traversal order */
uint32_t orc_ball(const orc_tree *h, const double *c, double r, uint32_t *out, uint32_t cap)
{
    std::vector<uint32_t> v;
    h->T.ball(c, r, v);
    for (size_t i = 0; i < v.size() && i < cap; ++i) out[i] = v[i];
    return (uint32_t)v.size();
}

/* BuildClumpTree's tree (src/tree.hpp:5251-5284: BuildTree over the given
 * points with the given root box, no position check) and ball searches
 * (3275-3299): for each query (x, y, z, r) the payloads (the points' density
 * field) of the particles found, in traversal order, at out[offsets[q] ..
 * offsets[q+1]).  Returns the total count (only cap entries are written). */
uint32_t orc_point_tree_balls(const plan_params *pp, uint32_t n, const double *pts4, const double *root_center,
                              double half_width, const double *q4, uint32_t nq, uint32_t *offsets, double *out,
                              uint32_t cap)
{
    plan_derived d;
    derive(pp, n, &d);
    std::vector<Par> P(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) { P[i].pos[k] = pts4[4 * i + k]; P[i].vel[k] = 0.0; }
        P[i].density = pts4[4 * i + 3];
        P[i].id = i;
        P[i].property_index = 0;
    }
    Tree CT;
    for (int k = 0; k < 3; ++k) CT.root_center[k] = root_center[k];
    CT.half_width = half_width;
    CT.build(pp, &d, P, false);
    uint32_t k = 0;
    std::vector<uint32_t> v;
    for (uint32_t q = 0; q < nq; ++q) {
        offsets[q] = k;
        CT.ball(q4 + 4 * q, q4[4 * q + 3], v);
        for (uint32_t j : v) {
            if (k < cap) out[k] = CT.p[j].ath_density;
            ++k;
        }
    }
    offsets[nq] = k;
    return k;
}

/* Run FindPlanetesimals on a built tree. */
void orc_find_planetesimals(orc_tree *h)
{
    find_planetesimals(h->T, &h->pp, &h->d, h->N, h->R, h->S);
    h->found = true;
}

/* ---- the clump stages on constructed clumps (pins of the multi-clump loops) ----
 * Replaces the tree particles' new_density (rho) and ath_density (ath) when
 * given (tree order). */
void orc_tree_set_densities(orc_tree *h, const double *rho, const double *ath)
{
    for (uint32_t i = 0; i < h->T.n; ++i) {
        if (rho) h->T.p[i].new_density = rho[i];
        if (ath) h->T.p[i].ath_density = ath[i];
    }
}

/* Steps 4-10 (clump_stages) on naive clumps of a built tree: clump c has
 * key keys[c] (a tree index, also its peak) and members members[off[c] ..
 * off[c+1]) (tree indices), set up by SortParticles as step 3 does
 * (src/tree.hpp:3816-3817).  settings = {density_threshold, hill_frac,
 * clump_mass_threshold, diffuse}; the peak threshold is 3x the density
 * threshold (3660).  saddle(pt, est, user) gives the saddle-point density,
 * NULL the tree's own KNN density.  The result is the handle's catalog
 * (orc_catalog) and subpeaks (orc_clump_subpeaks); returns the clump count. */
typedef double (*orc_saddle_fn)(const double *pt, double est, void *user);
uint32_t orc_clump_stages(orc_tree *h, uint32_t nc, const uint32_t *keys, const uint32_t *off,
                          const uint32_t *members, unsigned stages, const double *settings, orc_saddle_fn fn,
                          void *user)
{
    Result &R = h->R;
    R = Result();
    R.density_threshold = settings[0];
    R.hill_frac = settings[1];
    R.clump_mass_threshold = settings[2];
    R.diffuse = settings[3];
    R.peak_density_threshold = 3 * R.density_threshold;
    for (uint32_t i = 0; i < h->T.n; ++i) h->T.p[i].in_clump_flag = false;
    for (uint32_t c = 0; c < nc; ++c) {
        Clump &q = R.clumps[keys[c]];
        q.peak_index = keys[c];
        for (uint32_t i = off[c]; i < off[c + 1]; ++i) q.indices.push_back(members[i]);
        sort_particles(q, h->T.p);
    }
    Ctx X;
    X.d = &h->d;
    X.G = h->d.grav_constant;
    X.shear = h->d.q * h->d.Omega;
    const unsigned K = h->pp.num_knn;
    const double cv = h->d.cell_volume;
    const Tree &T = h->T;
    SaddleFn sf = [&](const double *pt, double est) {
        if (fn) return fn(pt, est, user);
        double sp[3] = {pt[0], pt[1], pt[2]};
        return saddle_density(T, sp, est, K, 0, cv);
    };
    uint32_t num = (uint32_t)R.clumps.size();
    clump_stages(h->T, R, &h->pp, &h->d, X, K, stages, num, sf, &h->S.after_merge);
    h->found = true;
    return (uint32_t)R.clumps.size();
}

/* clumps whose key (tree index 0) the merge loops' zero-filled delete list erased */
uint32_t orc_key0_erased(const orc_tree *h) { return h->R.erased0; }

/* subpeaks (potential_subpeak_indices) of the i-th clump in map order */
uint32_t orc_clump_subpeaks(const orc_tree *h, uint32_t i, uint32_t *out, uint32_t cap)
{
    auto it = h->R.clumps.begin();
    for (uint32_t k = 0; k < i && it != h->R.clumps.end(); ++k) ++it;
    if (it == h->R.clumps.end()) return 0;
    const std::vector<uint32_t> &s = it->second.subpeaks;
    for (size_t k = 0; k < s.size() && k < cap; ++k) out[k] = s[k];
    return (uint32_t)s.size();
}

void orc_densities(const orc_tree *h, double *rho)
{
    for (uint32_t i = 0; i < h->T.n; ++i) rho[i] = h->T.p[i].new_density;
}
void orc_hop(const orc_tree *h, uint32_t *densest, uint32_t *peak)
{
    for (uint32_t i = 0; i < h->T.n; ++i) { densest[i] = h->S.densest[i]; peak[i] = h->S.peak[i]; }
}
void orc_stage_info(const orc_tree *h, int *kernel, uint32_t *hop, uint32_t *naive, uint32_t *after_merge)
{
    *kernel = h->S.kernel;
    *hop = h->S.hop;
    *naive = h->S.naive_clumps;
    *after_merge = h->S.after_merge;
}

/* Catalog in map order. */
plan_catalog *orc_catalog(const orc_tree *h)
{
    const Result &R = h->R;
    const Tree &T = h->T;
    plan_catalog *c = (plan_catalog *)std::calloc(1, sizeof(plan_catalog));
    c->num_clumps = (uint32_t)R.clumps.size();
    uint32_t tot = 0;
    for (auto &it : R.clumps) tot += (uint32_t)it.second.particles.size();
    c->num_members_total = tot;
    c->clumps = (plan_clump *)std::calloc(c->num_clumps ? c->num_clumps : 1, sizeof(plan_clump));
    c->members = (uint32_t *)std::calloc(tot ? tot : 1, sizeof(uint32_t));
    c->member_original_ids = (uint32_t *)std::calloc(tot ? tot : 1, sizeof(uint32_t));
    c->density_threshold = R.density_threshold;
    c->peak_density_threshold = R.peak_density_threshold;
    c->clump_mass_threshold = R.clump_mass_threshold;
    c->Mp_max = R.Mp_max;
    c->Mp_tot = R.Mp_tot;
    uint32_t off = 0, i = 0;
    for (auto &it : R.clumps) {
        const Clump &q = it.second;
        plan_clump &o = c->clumps[i++];
        o.peak_index = it.first;
        o.num_members = (uint32_t)q.particles.size();
        uint32_t np = o.num_members;
        for (auto &m : q.indices) {
            auto s = T.sinks.find(m);
            if (s != T.sinks.end()) np += (uint32_t)s->second.size() - 1;
        }
        o.num_particles = np;
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
        o.peak_density = T.p[it.first].new_density;
        for (auto &m : q.particles) {
            c->members[off] = m.first;
            c->member_original_ids[off] = T.p[m.first].original_id;
            off++;
        }
    }
    return c;
}

void orc_catalog_free(plan_catalog *c)
{
    if (!c) return;
    std::free(c->clumps);
    std::free(c->members);
    std::free(c->member_original_ids);
    std::free(c);
}

/* Boundness of one clump by the reference's two-body rule, for tests of
 * the batched boundness entry points: members in distance order (tree
 * indices into the built tree are not needed; arrays are given directly).
 * Runs RemoveUnboundParticles once from the state computed by
 * SortParticles and returns the kept count; keep[] lists kept input
 * positions in final order. */
uint32_t orc_unbind_once(const plan_params *pp, uint32_t n, const double *pos, const double *vel,
                         const double *mass, uint32_t *keep, double *out_state)
{
    plan_derived d;
    derive(pp, n ? n : 1, &d);
    Ctx X;
    X.d = &d;
    X.G = d.grav_constant;
    X.shear = d.q * d.Omega;
    std::vector<IPar> P(n);
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) { P[i].pos[k] = pos[3 * i + k]; P[i].vel[k] = vel[3 * i + k]; }
        P[i].mass = mass[i];
        P[i].new_density = 1.0;
    }
    Clump c;
    for (uint32_t i = 0; i < n; ++i) c.indices.push_back(i);
    sort_particles(c, P);
    remove_unbound(c, P, X);
    for (size_t i = 0; i < c.indices.size(); ++i) keep[i] = c.indices[i];
    out_state[0] = c.total_mass;
    for (int k = 0; k < 3; ++k) { out_state[1 + k] = c.com[k]; out_state[4 + k] = c.vel[k]; }
    return (uint32_t)c.indices.size();
}

/* ---- single clump operations on plain arrays, for the closed-form pins ----
 * A clump is given as its members in order (positions, velocities, masses
 * n x 3 / n, new_density and ath_density per member); op:
 *   0 SortParticles                                    (src/tree.hpp:4446-4512)
 *   1 SortParticles, RemoveUnboundParticles             (4615-4674)
 *   2 SortParticles, CalculateHillRadius                (4715-4717)
 *   3 SortParticles, CalculateHillRadius,
 *     CalculateAngularMomentum                          (4721-4747)
 *   4 SortParticles, IsPositionDispersion2Large(0.55)   (4516-4528)
 * order[] receives the members (input positions) in final order; out[] =
 * {M, X[3], V[3], inner, outer, one10th, two_sigma, hill, J[3], Jq[3],
 * dispersion flag}.  Returns the member count. */
static void arrays_to_clump(uint32_t n, const double *pos, const double *vel, const double *mass, const double *rho,
                            const double *ath, std::vector<IPar> &P, Clump &c)
{
    P.assign(n, IPar());
    for (uint32_t i = 0; i < n; ++i) {
        for (int k = 0; k < 3; ++k) { P[i].pos[k] = pos[3 * i + k]; P[i].vel[k] = vel[3 * i + k]; }
        P[i].mass = mass[i];
        P[i].new_density = rho ? rho[i] : 1.0;
        P[i].ath_density = ath ? ath[i] : 0.0;
        P[i].in_clump_flag = false;
    }
    c = Clump();
    for (uint32_t i = 0; i < n; ++i) c.indices.push_back(i);
}

static void clump_out(const Clump &c, double *out)
{
    out[0] = c.total_mass;
    for (int k = 0; k < 3; ++k) { out[1 + k] = c.com[k]; out[4 + k] = c.vel[k]; }
    out[7] = c.inner; out[8] = c.outer; out[9] = c.one10th; out[10] = c.two_sigma; out[11] = c.hill;
    for (int k = 0; k < 3; ++k) { out[12 + k] = c.J[k]; out[15 + k] = c.Jq[k]; }
}

uint32_t orc_clump_op(const plan_params *pp, int op, uint32_t n, const double *pos, const double *vel,
                      const double *mass, const double *rho, const double *ath, uint32_t *order, double *out)
{
    plan_derived d;
    derive(pp, n ? n : 1, &d);
    Ctx X;
    X.d = &d;
    X.G = d.grav_constant;
    X.shear = d.q * d.Omega;
    std::vector<IPar> P;
    Clump c;
    arrays_to_clump(n, pos, vel, mass, rho, ath, P, c);
    sort_particles(c, P);
    out[18] = 0;
    if (op == 1) remove_unbound(c, P, X);
    if (op == 2 || op == 3) hill_radius(c, X);
    if (op == 3) {
        Tree T;
        T.p = P;
        angular_momentum(c, T, X);
    }
    if (op == 4) out[18] = dispersion_too_large(c) ? 1.0 : 0.0;
    clump_out(c, out);
    for (size_t i = 0; i < c.indices.size(); ++i) order[i] = c.indices[i];
    return (uint32_t)c.indices.size();
}

/* The candidate loop of SearchBoundParticlesWithinHillRadius on a clump
 * state {M, X[3], V[3]} (state[7]) and candidates in ball order (arrays
 * n x 3 / n; skip[i] marks in_clump_flag, rho[] against threshold thr).
 * accepted[i] = 1 for the candidates that joined; state is updated. */
uint32_t orc_rebind_candidates(const plan_params *pp, double *state, uint32_t n, const double *pos,
                               const double *vel, const double *mass, const double *rho, const uint8_t *skip,
                               double thr, uint8_t *accepted)
{
    plan_derived d;
    derive(pp, 1, &d);
    Ctx X;
    X.d = &d;
    X.G = d.grav_constant;
    X.shear = d.q * d.Omega;
    std::vector<IPar> P;
    Clump c;
    arrays_to_clump(n, pos, vel, mass, rho, nullptr, P, c);
    c.indices.clear();
    for (uint32_t i = 0; i < n; ++i) P[i].in_clump_flag = skip[i] != 0;
    c.total_mass = state[0];
    for (int k = 0; k < 3; ++k) { c.com[k] = state[1 + k]; c.vel[k] = state[4 + k]; }
    std::vector<uint32_t> near(n);
    for (uint32_t i = 0; i < n; ++i) near[i] = i;
    rebind_candidates(c, P, near, thr, X);
    for (uint32_t i = 0; i < n; ++i) accepted[i] = 0;
    for (uint32_t i : c.indices) accepted[i] = 1;
    state[0] = c.total_mass;
    for (int k = 0; k < 3; ++k) { state[1 + k] = c.com[k]; state[4 + k] = c.vel[k]; }
    return (uint32_t)c.indices.size();
}

/* MergeAnotherPlanetesimal (src/tree.hpp:4532-4611) of clump b (members
 * [na, na+nb) of the arrays) into clump a (members [0, na)); peak_a / peak_b
 * are member positions whose new_density decides the merged peak.  Each
 * clump is first set up by SortParticles.  out as orc_clump_op, plus
 * out[19] = merged peak (member position), out[20] = the subpeak recorded. */
uint32_t orc_merge_two(const plan_params *pp, uint32_t na, uint32_t nb, const double *pos, const double *vel,
                       const double *mass, const double *rho, uint32_t peak_a, uint32_t peak_b, uint32_t *order,
                       double *out)
{
    plan_derived d;
    derive(pp, na + nb, &d);
    std::vector<IPar> P;
    Clump all;
    arrays_to_clump(na + nb, pos, vel, mass, rho, nullptr, P, all);
    Clump a, b;
    for (uint32_t i = 0; i < na; ++i) a.indices.push_back(i);
    for (uint32_t i = na; i < na + nb; ++i) b.indices.push_back(i);
    a.peak_index = peak_a;
    b.peak_index = peak_b;
    sort_particles(a, P);
    sort_particles(b, P);
    merge_into(a, b, P);
    clump_out(a, out);
    out[18] = 0;
    out[19] = a.peak_index;
    out[20] = a.subpeaks.empty() ? -1.0 : (double)a.subpeaks.back();
    for (size_t i = 0; i < a.indices.size(); ++i) order[i] = a.indices[i];
    return (uint32_t)a.indices.size();
}

/* RemoveSmallMassAndLowPeak (src/tree.hpp:3524-3546) on clumps given as
 * member ranges [off[c], off[c+1]) of the member arrays (members in clump
 * order; key[c] = the peak member); keep[c] = 1 for the clumps kept. */
void orc_filter_clumps(uint32_t nc, const uint32_t *off, const uint32_t *key, const double *mass,
                       const double *rho, const double *ath, unsigned K, double mass_thr, double peak_thr,
                       uint8_t *keep)
{
    const uint32_t n = off[nc];
    Tree T;
    T.p.assign(n, IPar());
    for (uint32_t i = 0; i < n; ++i) {
        T.p[i].mass = mass[i];
        T.p[i].new_density = rho[i];
        T.p[i].ath_density = ath[i];
    }
    std::map<uint32_t, Clump> cl;
    for (uint32_t c = 0; c < nc; ++c) {
        Clump &q = cl[key[c]];
        q.peak_index = key[c];
        for (uint32_t i = off[c]; i < off[c + 1]; ++i) {
            q.indices.push_back(i);
            q.total_mass += mass[i];
        }
    }
    Result R;
    R.clump_mass_threshold = mass_thr;
    R.peak_density_threshold = peak_thr;
    remove_small_mass_low_peak(cl, T, R, K);
    for (uint32_t c = 0; c < nc; ++c) keep[c] = cl.count(key[c]) ? 1 : 0;
}

/* ranking and cap on clumps (key, total mass); keep[c] = 1 for the kept
 * ones; out = {Mp_tot, Mp_max} */
void orc_rank_clumps(uint32_t nc, const uint32_t *key, const double *mass, uint32_t num_peaks, uint8_t *keep,
                     double *out)
{
    std::map<uint32_t, Clump> cl;
    for (uint32_t c = 0; c < nc; ++c) cl[key[c]].total_mass = mass[c];
    Result R;
    rank_and_cap(cl, num_peaks, R);
    for (uint32_t c = 0; c < nc; ++c) keep[c] = cl.count(key[c]) ? 1 : 0;
    out[0] = R.Mp_tot;
    out[1] = R.Mp_max;
}

/* IsGravitationallyBound (src/tree.hpp:4867-4874) of two clump states
 * {M, X[3], V[3]} */
int orc_clumps_bound(const plan_params *pp, const double *sa, const double *sb)
{
    plan_derived d;
    derive(pp, 1, &d);
    Ctx X;
    X.d = &d;
    X.G = d.grav_constant;
    X.shear = d.q * d.Omega;
    Clump a, b;
    a.total_mass = sa[0];
    b.total_mass = sb[0];
    for (int k = 0; k < 3; ++k) {
        a.com[k] = sa[1 + k]; a.vel[k] = sa[4 + k];
        b.com[k] = sb[1 + k]; b.vel[k] = sb[4 + k];
    }
    return grav_bound(a, b, X) ? 1 : 0;
}

/* Pairwise-potential boundness: an extension of the path, which has no
 * counterpart in the reference (its boundness is the two-body test of
 * src/tree.hpp:4615-4674).  Brute force in binary64, potentials recomputed
 * from scratch every iteration, in member order:
 *   B = all members; repeat {
 *     M, X, V over B;  phi_i = -G sum_{j in B, x_j != x_i} m_j / |x_i - x_j|;
 *     dv_i = v_i - V + (0, q Omega (X_x - x_i,x), 0);  e_i = |dv_i|^2/2 + phi_i;
 *     E_kin = sum m |dv|^2 / 2, E_pot = sum m phi / 2;
 *     remove e_i > 0 } until nothing is removed (or max_iter iterations).
 * Pinned by tests/test_oracle_boundness.py (two- and three-body closed forms,
 * a numpy brute force, the Plummer potential energy -3 pi G M^2 / (32 a)).
 * out per clump: {n_bound, iterations, E_kin, E_pot, M, X, Y, Z}. */
void orc_boundness(uint32_t nc, const uint32_t *off, const double *pos, const double *vel, const double *mass,
                   double G, double Omega, double q, uint32_t max_iter, uint8_t *bound, double *out)
{
    const double qO = q * Omega;
#pragma omp parallel for num_threads(g_threads) schedule(dynamic, 1)
    for (uint32_t c = 0; c < nc; ++c) {
        std::vector<uint32_t> B;
        for (uint32_t i = off[c]; i < off[c + 1]; ++i) B.push_back(i);
        std::vector<double> phi;
        double M = 0, X[3] = {0, 0, 0}, V[3] = {0, 0, 0}, ek = 0, ep = 0;
        uint32_t it = 0;
        for (;;) {
            ++it;
            M = 0;
            double sx[3] = {0, 0, 0}, sv[3] = {0, 0, 0};
            for (uint32_t i : B) {
                M += mass[i];
                for (int k = 0; k < 3; ++k) {
                    sx[k] += mass[i] * pos[3 * (size_t)i + k];
                    sv[k] += mass[i] * vel[3 * (size_t)i + k];
                }
            }
            for (int k = 0; k < 3; ++k) {
                X[k] = M > 0 ? sx[k] / M : 0.0;
                V[k] = M > 0 ? sv[k] / M : 0.0;
            }
            phi.assign(B.size(), 0.0);
            for (size_t a = 0; a < B.size(); ++a) {
                const double *pa = pos + 3 * (size_t)B[a];
                double sum = 0;
                for (size_t b = 0; b < B.size(); ++b) {
                    const double *pb = pos + 3 * (size_t)B[b];
                    const double dx = pb[0] - pa[0], dy = pb[1] - pa[1], dz = pb[2] - pa[2];
                    const double r2 = (dx * dx + dy * dy) + dz * dz;
                    if (r2 > 0) sum += mass[B[b]] / std::sqrt(r2);
                }
                phi[a] = -G * sum;
            }
            ek = ep = 0;
            std::vector<uint32_t> keep;
            for (size_t a = 0; a < B.size(); ++a) {
                const uint32_t i = B[a];
                const double *p = pos + 3 * (size_t)i, *v = vel + 3 * (size_t)i;
                const double dvx = v[0] - V[0], dvy = (v[1] - V[1]) + qO * (X[0] - p[0]), dvz = v[2] - V[2];
                const double kin = 0.5 * ((dvx * dvx + dvy * dvy) + dvz * dvz);
                ek += mass[i] * kin;
                ep += 0.5 * mass[i] * phi[a];
                if (!(kin + phi[a] > 0)) keep.push_back(i);
            }
            const bool changed = keep.size() != B.size();
            if (!changed || it >= max_iter) break;
            B.swap(keep);
        }
        for (uint32_t i = off[c]; i < off[c + 1]; ++i) bound[i] = 0;
        for (uint32_t i : B) bound[i] = 1;
        double *o = out + 8 * (size_t)c;
        o[0] = (double)B.size();
        o[1] = it;
        o[2] = ek;
        o[3] = ep;
        o[4] = M;
        o[5] = X[0];
        o[6] = X[1];
        o[7] = X[2];
    }
}

} /* extern "C" */
