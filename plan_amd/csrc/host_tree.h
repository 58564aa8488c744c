// host_tree.h — BuildTree and BallSearch on the host for small point sets.
//
// The merge passes build a tree over the clump centres (a few hundred to a
// few thousand points, src/tree.hpp:5251-5284) and ball-search it once per
// clump.  On the device each pass costs a chain of small launches and host
// round trips; on the host the same tree is a few tens of microseconds.  The
// keys, the sink rule, the node boxes (the chain of binary64 adds from the
// root), the leaf rule and the sphere predicates are those of the device
// tree (build.hip, query.hip), so the lists are identical.
#pragma once

#include "common.h"
#include "../../include/plan_params.h"

#include <vector>

namespace plan {

class HostTree {
public:
    /* tree over pts (x, y, z, payload) with the given root box (a half width
     * below 1e-16 takes the box of the derived parameters, as the device
     * build does) */
    void build(const std::vector<double4> &pts, const double *root_center, double half_width,
               const plan_derived &d);
    /* RecursiveBallSearch (src/tree.hpp:3275-3293): tree indices of the
     * particles within r of (cx, cy, cz), in the walk's (ascending) order */
    void ball(double cx, double cy, double cz, double r, std::vector<uint32_t> &out) const;
    /* the payload (input .w) of tree particle i: the first input point of its
     * run of equal keys */
    double payload(uint32_t i) const { return pos_[i].w; }
    uint32_t size() const { return (uint32_t)pos_.size(); }

private:
    struct Node {
        double cx, cy, cz, hw;
        uint32_t begin, end, first, count; /* children [first, first + count) */
    };
    std::vector<Node> nodes_;
    std::vector<double4> pos_;
    std::vector<uint4> keys_;
    double to_diag_ = 0;
    void split(uint32_t id, int level);
    void walk(uint32_t id, double cx, double cy, double cz, double r, std::vector<uint32_t> &out) const;
};

}  // namespace plan
