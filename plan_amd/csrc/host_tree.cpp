// host_tree.cpp — BuildTree and BallSearch on the host for small point sets
// (the clump trees of the merge passes), with the device tree's arithmetic.
#include "host_tree.h"

#include "morton.h"

#include <algorithm>
#include <cmath>
#include <numeric>

namespace plan {

namespace {

constexpr uint32_t LEAF_MAX = 8;  /* particles of a leaf (src/tree.hpp:2655-2755) */
constexpr int MAX_LEVEL = 32;

bool key_less96(const uint4 &a, const uint4 &b)
{
    if (a.z != b.z) return a.z < b.z;
    if (a.y != b.y) return a.y < b.y;
    return a.x < b.x;
}

/* SphereContainNode (src/tree.hpp:3001-3012): the sum over axes of
 * (|centre offset| + hw sqrt 3)^2 below r^2 */
bool sphere_contain_h(double gx, double gy, double gz, double hw, double qx, double qy, double qz, double r,
                      double to_diag)
{
    const double t = hw * to_diag;
    const double mx = std::fabs(gx - qx) + t;
    const double my = std::fabs(gy - qy) + t;
    const double mz = std::fabs(gz - qz) + t;
    const double s = (mx * mx + my * my) + mz * mz;
    return s < r * r;
}

/* SphereNodeIntersect (src/tree.hpp:3016-3049), with the squared-vs-linear
 * comparison of 3030 */
bool sphere_intersect_h(double gx, double gy, double gz, double hw, double qx, double qy, double qz, double r,
                        double to_diag)
{
    const double dx = gx - qx, dy = gy - qy, dz = gz - qz;
    const double c2c = (dx * dx + dy * dy) + dz * dz;
    const double t = hw * to_diag + r;
    if (c2c > t * t) return false;
    if (c2c < r || c2c < hw) return true;
    double md = 0;
    double lo = gx - hw, hi = gx + hw;
    if (qx < lo) md = md + (qx - lo) * (qx - lo);
    else if (qx > hi) md = md + (qx - hi) * (qx - hi);
    lo = gy - hw;
    hi = gy + hw;
    if (qy < lo) md = md + (qy - lo) * (qy - lo);
    else if (qy > hi) md = md + (qy - hi) * (qy - hi);
    lo = gz - hw;
    hi = gz + hw;
    if (qz < lo) md = md + (qz - lo) * (qz - lo);
    else if (qz > hi) md = md + (qz - hi) * (qz - hi);
    return md <= r * r;
}

}  // namespace

void HostTree::build(const std::vector<double4> &pts, const double *root_center, double half_width,
                     const plan_derived &d)
{
    nodes_.clear();
    pos_.clear();
    keys_.clear();
    to_diag_ = std::sqrt(3.0);
    const uint32_t T = (uint32_t)pts.size();
    if (T == 0) return;
    /* root box and key frame as the device build (BuildTree with a given
     * root, src/tree.hpp:2759-2808) */
    double rc[3], hw = half_width;
    for (int k = 0; k < 3; ++k) rc[k] = root_center[k];
    if (hw < 1e-16) {
        hw = d.max_half_width + d.max_ghost_zone_width;
        for (int k = 0; k < 3; ++k) rc[k] = d.box_center[k];
    }
    double boxmin[3], scale[3];
    for (int k = 0; k < 3; ++k) {
        volatile double lo = rc[k] - hw;
        volatile double hi = rc[k] + hw;
        boxmin[k] = lo;
        scale[k] = 1.0 / (hi - lo);
    }
    std::vector<uint4> keys(T);
    for (uint32_t i = 0; i < T; ++i) {
        const uint32_t ix = double_to_grid((pts[i].x - boxmin[0]) * scale[0]);
        const uint32_t iy = double_to_grid((pts[i].y - boxmin[1]) * scale[1]);
        const uint32_t iz = double_to_grid((pts[i].z - boxmin[2]) * scale[2]);
        keys[i] = morton_key(ix, iy, iz, i);
    }
    /* stable sort by the 96-bit key; a run of equal keys becomes one
     * particle, its first (smallest input index) */
    std::stable_sort(keys.begin(), keys.end(), key_less96);
    for (uint32_t i = 0; i < T; ++i) {
        if (i > 0 && key_eq(keys[i], keys[i - 1])) continue;
        keys_.push_back(keys[i]);
        pos_.push_back(pts[keys[i].w]);
    }
    nodes_.push_back(Node{rc[0], rc[1], rc[2], hw, 0u, (uint32_t)keys_.size(), 0u, 0u});
    split(0, 1);
}

/* children of node id (level `level`) by the octal digit of the next level,
 * in digit order; a node of more than LEAF_MAX particles below level 32
 * splits (the device node array follows the same rule) */
void HostTree::split(uint32_t id, int level)
{
    const uint32_t b = nodes_[id].begin, e = nodes_[id].end;
    if (e - b <= LEAF_MAX || level >= MAX_LEVEL) return;
    const double hwc = 0.5 * nodes_[id].hw;
    const uint32_t first = (uint32_t)nodes_.size();
    uint32_t count = 0;
    for (uint32_t s = b; s < e;) {
        const uint32_t o = key_digit(keys_[s], level);
        uint32_t t = s + 1;
        while (t < e && key_digit(keys_[t], level) == o) ++t;
        const Node &p = nodes_[id];
        const double cx = p.cx + ((o & 1u) ? hwc : -hwc);
        const double cy = p.cy + ((o & 2u) ? hwc : -hwc);
        const double cz = p.cz + ((o & 4u) ? hwc : -hwc);
        nodes_.push_back(Node{cx, cy, cz, hwc, s, t, 0u, 0u});
        ++count;
        s = t;
    }
    nodes_[id].first = first;
    nodes_[id].count = count;
    for (uint32_t c = 0; c < count; ++c) split(first + c, level + 1);
}

void HostTree::walk(uint32_t id, double cx, double cy, double cz, double r, std::vector<uint32_t> &out) const
{
    const Node &n = nodes_[id];
    if (sphere_contain_h(n.cx, n.cy, n.cz, n.hw, cx, cy, cz, r, to_diag_)) {
        for (uint32_t p = n.begin; p < n.end; ++p) out.push_back(p);
    } else if (sphere_intersect_h(n.cx, n.cy, n.cz, n.hw, cx, cy, cz, r, to_diag_)) {
        if (n.count == 0) {
            const double r2 = r * r;
            for (uint32_t p = n.begin; p < n.end; ++p) {
                const double dx = cx - pos_[p].x, dy = cy - pos_[p].y, dz = cz - pos_[p].z;
                if ((dx * dx + dy * dy) + dz * dz <= r2) out.push_back(p);
            }
        } else {
            for (uint32_t c = 0; c < n.count; ++c) walk(n.first + c, cx, cy, cz, r, out);
        }
    }
}

void HostTree::ball(double cx, double cy, double cz, double r, std::vector<uint32_t> &out) const
{
    out.clear();
    if (!nodes_.empty()) walk(0, cx, cy, cz, r, out);
}

}  // namespace plan
