// boundness.h — pairwise-potential boundness engine (device state of one batch).
#pragma once

#include "devmem.h"
#include "../../include/plan_hip.h"

#include <vector>

namespace plan {

struct PairShape;

class BoundnessEngine {
public:
    /* stage a batch in HBM: member positions, velocities, masses per clump */
    void upload(const plan_boundness_in &in, hipStream_t s);
    /* the same from members already on the device (tree particles by index) */
    void upload_tree(uint32_t nc, const uint32_t *offsets_h, const uint32_t *members_h, const double4 *pm,
                     const double4 *va, double G, double Omega, double q, hipStream_t s);
    /* form 0: binary64 per clump; form 1: binary32 compensated, batched */
    void run(int form, uint32_t max_iter, hipStream_t s, plan_boundness_stats *st);
    void download(uint8_t *bound, plan_boundness_clump *out, hipStream_t s);
    bool loaded() const { return pos_ != nullptr; }
    uint32_t num_clumps() const { return nc_; }
    uint32_t num_members() const { return nm_; }

private:
    template <typename R>
    void iteration(uint32_t c0, uint32_t c1, const PairShape &sh, bool near, uint32_t max_iter, hipStream_t s);
    void reset(uint32_t c0, uint32_t c1, hipStream_t s);
    void alloc_run_buffers(hipStream_t s);
    uint32_t read_active(hipStream_t s);
    void run_fp64_per_clump(const std::vector<uint32_t> &clumps, uint32_t max_iter, hipStream_t s);
    void run_fp32_batched(uint32_t max_iter, hipStream_t s);
    void autotune(hipStream_t s);

    Arena a_;
    uint32_t nc_ = 0, nm_ = 0;
    double G_ = 0, qO_ = 0;
    std::vector<uint32_t> off_h_;
    double4 *pos_ = nullptr, *vel_ = nullptr, *ref_ = nullptr, *pj_ = nullptr, *pr_ = nullptr;
    float4 *posf_ = nullptr;
    uint32_t *cid_ = nullptr, *off_ = nullptr, *flag_ = nullptr;
    double *phi_ = nullptr, *tot_ = nullptr, *eout_ = nullptr;
    uint32_t *inb_ = nullptr, *inr_ = nullptr, *sb_ = nullptr, *sr_ = nullptr, *blist_ = nullptr, *rlist_ = nullptr;
    uint4 *seg_ = nullptr;
    uint32_t *mode_ = nullptr, *tiles_ = nullptr, *toff_ = nullptr, *state_ = nullptr, *iters_ = nullptr;
    uint32_t *misc_ = nullptr, *near_ = nullptr, *nearz_ = nullptr;
    unsigned long long *pairs_ = nullptr;
    int persistent_blocks_ = 1024;
    int shape_idx_ = -1;
    uint32_t graph_replays_ = 0;
    std::vector<float> tune_ms_;
};

}  // namespace plan
