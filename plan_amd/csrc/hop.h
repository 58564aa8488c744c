// hop.h — HOP results shared between the device stage and the host clump logic.
#pragma once

#include "ctx.h"

namespace plan {

/* One particle above the density threshold, as the clump logic needs it
 * (80 bytes: the InternalParticle fields of src/tree.hpp:2367-2410 minus flags). */
struct MemberRec {
    double x, y, z, m;
    double vx, vy, vz, ath;
    double rho;
    uint32_t tidx;
    uint32_t oid;
};

struct HopResult {
    uint32_t na = 0;              /* particles with rho >= threshold ("A") */
    uint32_t *a_idx = nullptr;    /* A position -> tree index (ascending) */
    uint32_t *apos = nullptr;     /* tree index -> A position or NONE */
    uint32_t *densest = nullptr;  /* per A position: densest neighbour (tree index) */
    uint32_t *up = nullptr;       /* per A position: densest neighbour (A position) */
    uint32_t *root = nullptr;     /* per A position: peak (A position) */
    uint32_t *depth = nullptr;    /* hops to the peak */
    uint32_t *fmin = nullptr;     /* smallest A position in the subtree */
    uint4 *order = nullptr;       /* members grouped: {~depth, f, root, A position} */
    uint32_t nclumps = 0;
    uint32_t *cstart = nullptr;   /* first member of each naive clump in order[] */
    MemberRec *members = nullptr; /* member data in grouped order */
    uint32_t *msink = nullptr;    /* per member: particles merged into it (sinks) */
    uint32_t knn_overflow_waves = 0; /* densest-neighbour waves whose include list overflowed */
    uint32_t knn_res_queries = 0;    /* of their queries, those walked in several fills */
};

void dev_hop(const DevTree &t, Arena &a, const plan_derived &d, double thr, uint32_t H, HopResult &res,
             hipStream_t s);

}  // namespace plan
