// vrgi_internal.h — host/device interfaces shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <stdexcept>
#include <string>

#include "../../include/vrgi.h"

namespace vrgi {

struct HipError : std::runtime_error {
    using std::runtime_error::runtime_error;
};

#define HIP_CHECK(expr)                                                                     \
    do {                                                                                    \
        hipError_t e_ = (expr);                                                             \
        if (e_ != hipSuccess)                                                               \
            throw ::vrgi::HipError(std::string(#expr) + ": " + hipGetErrorString(e_));      \
    } while (0)

inline void dev_alloc(void **p, size_t bytes) {
    if (bytes == 0) bytes = 16;
    HIP_CHECK(hipMalloc(p, bytes));
}

// largest subtree built by one wave with its primitives in LDS
#ifndef VRGI_SUB_MAX
#define VRGI_SUB_MAX 128
#endif
// largest range built by one wave entirely in LDS (bottom phase of the builder)
static constexpr int SUB_MAX_DEFAULT = VRGI_SUB_MAX;

// inputs of one batched build: n primitives, prim9 = 9 planes of n floats
// (box min xyz, box max xyz, centroid xyz), per-primitive power and id; tree t
// owns primitives [tree_beg[t], tree_beg[t] + tree_n[t]) and writes its
// 2*tree_n[t]-1 nodes at nodes + tree_base[t].
struct BuildArgs {
    const float *prim9;
    const float *power;
    const uint32_t *ids;
    uint32_t n;
    uint32_t n_trees;
    const uint32_t *tree_beg;
    const uint32_t *tree_n;
    const uint32_t *tree_base;
    vrgi_bvh_node *nodes;
    uint32_t *depth;
    // sum over split nodes of their primitive count (= sum over leaves of their depth)
    unsigned long long *split_count_sum;
    uint32_t *perm;  // working permutation -> final primitive order per leaf position
    float *pc;       // centroids permuted along with perm (3 planes)
    // per primitive p, one 32-byte record gathered by primitive index:
    // box8[2p] = (min x, min y, min z, max x), box8[2p+1] = (max y, max z, power, id bits)
    float4 *box8;
    int sub_max;
    uint32_t *powered;  // per tree: 1 when some primitive has non-zero power
};

struct BuildState {
    uint32_t *rank_tmp = nullptr;
    uint32_t *pair = nullptr;
    void *open = nullptr;
    void *open_next = nullptr;
    void *jobs = nullptr;
    void *items = nullptr;
    uint32_t *item_true = nullptr;
    uint32_t *item_beg = nullptr;
    uint32_t *item_end = nullptr;
    uint32_t *bins = nullptr;
    int *bounds = nullptr;
    void *split = nullptr;
    uint32_t *level_nodes = nullptr;
    uint32_t *n_open = nullptr;
    uint32_t *n_open_next = nullptr;
    uint32_t *n_jobs = nullptr;
    uint32_t *n_items = nullptr;
    uint32_t *n_level_nodes = nullptr;
    void *mid = nullptr;          // nodes for the middle phase (one workgroup each)
    uint32_t *n_mid = nullptr;
};

// top-phase levels a build may take (entries of its level list)
static constexpr uint32_t BUILDER_LEVEL_CAP = 65536;

// Owns the builder's workspace (grown on demand, reused across builds), or works in a
// region the caller binds (no device allocation by the builder then).
// the counts of a build, which a build of the same input repeats exactly
struct BuildPlan {
    uint32_t n = 0, trees = 0;
    uint32_t levels = 0;        // top-phase levels holding open nodes
    uint32_t power_levels = 0;  // of those, levels holding nodes of trees with non-zero power
    uint32_t n_mid = 0, n_jobs = 0;
    bool valid = false;
};

class Builder {
  public:
    ~Builder();
    // bytes of the workspace of builds of up to n primitives in up to n_trees trees
    static size_t workspace_bytes(uint32_t n, uint32_t n_trees);
    // works in [ws, ws + bytes) from now on: builds above n_max / trees_max then fail
    void bind(void *ws, size_t bytes, uint32_t n_max, uint32_t trees_max);
    // Queues a batched build on `stream`; synchronises the stream after each group of
    // top-phase levels (to see whether any node is still open): the first group has the
    // previous same-sized build's level count, later groups 1, 2, 4, 4, ... levels. prim_order_out (may be null) receives the
    // final primitive index per leaf position (global over all trees).
    // With `replay` set, the caller states that the input equals that of the last build
    // made without it (the plan()); the build then repeats that build's launch sequence
    // with no readback and no host synchronisation (so a stream capture can record it), and
    // a check on the device ORs 1 into *check when a count differs from the plan.
    int build(hipStream_t stream, const float *prim9, const float *power, const uint32_t *ids,
              uint32_t n, uint32_t n_trees, const uint32_t *d_tree_beg, const uint32_t *d_tree_n,
              const uint32_t *d_tree_base, const uint32_t *h_tree_base, vrgi_bvh_node *nodes,
              uint32_t *d_depth, uint32_t *prim_order_out, bool replay = false, uint32_t *check = nullptr);
    const BuildPlan &plan() const { return last_plan; }
    const uint32_t *prim_order() const { return perm; }
    // device address of the last build's sum over leaves of their depth
    const unsigned long long *leaf_depth_sum() const { return (const unsigned long long *)(counters + 8); }

  private:
    void reserve(uint32_t n, uint32_t n_trees);
    void free_all();
    uint32_t cap_n = 0, cap_t = 0, open_capacity = 0;
    uint32_t *perm = nullptr;
    float *pc = nullptr;
    float4 *box8 = nullptr;
    uint32_t *counters = nullptr;
    uint32_t *level_start = nullptr;  // first level-node entry of each top-phase level
    uint32_t *powered = nullptr;      // per tree: some primitive has non-zero power
    uint32_t level_cap = 0;
    uint32_t prev_n = 0, prev_trees = 0, prev_levels = 0;  // last build's size and non-empty top levels
    BuildPlan last_plan;  // counts of the last build made without replay
    bool external = false;  // the workspace is a bound region, not owned
    BuildState st;
};

}  // namespace vrgi
