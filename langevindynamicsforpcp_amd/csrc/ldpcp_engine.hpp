// ldpcp_engine.hpp -- host side of the first-passage engine (internal).
//
// The engine owns a pool of trajectory slots in HBM and a table of resident
// parameter sets.  A set is admitted into a contiguous slot range and joins
// the live list between two launches; each launch advances every live
// trajectory by K steps and retires the finished ones.  Both C-ABI front ends
// (ldpcp_sweep_* and ldpcp_server_*) drive it.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <vector>

#include "ldpcp_common.hpp"

namespace ldpcp {

struct SetDev;

// First-fit allocator of contiguous slot ranges.
class RangeAllocator {
public:
    void reset(uint64_t capacity);
    bool alloc(uint64_t n, uint64_t* first);
    void release(uint64_t first, uint64_t n);
    uint64_t free_total() const { return free_total_; }

private:
    std::map<uint64_t, uint64_t> free_;  // first -> length
    uint64_t free_total_ = 0;
};

// Device pointers of the slot pool (structure of arrays).
struct SlotPtrs {
    float *xp, *comp, *fv, *fR;  // FP32 engine
    int32_t* cell;
    double *x, *v, *R;           // FP64 engine
    uint32_t* flags;
    uint4* ts;                   // last-left step of the four barrier slots
    uint32_t *n_catch, *tpt0, *xc_first, *xc_last, *absorbed, *set_of;
};

// Device pointers of the per-set statistics.
struct StatsPtrs {
    unsigned long long* hist;   // [set][LDPCP_HIST_BINS]
    // [set][12]: for fpt, tpt, tptb0, tptb1 (in steps) the sum, and the sum of
    // squares as a 128-bit integer (low, high word): exact, so the same
    // whatever order the wavefronts add in
    unsigned long long* sums;
    unsigned long long* cnt;    // [set][4]: n_absorbed, n_finished, traj_steps, unused
    double* msd_sum;            // [set][LDPCP_MSD_LAGS]
    unsigned long long* msd_n;  // [set][LDPCP_MSD_LAGS]
};

// Launch control words, in device memory so that a captured launch needs no
// host arguments that change from launch to launch.
struct Ctrl {
    uint32_t n_live[2];       // live-list lengths (ping-pong)
    uint32_t work;            // next chunk to hand out in the running launch
    uint32_t pad;
    unsigned long long estep; // engine steps done (sum of K over launches)
    unsigned long long launches;
};

constexpr int kSumsPerSet = 12;
constexpr int kCntPerSet = 4;

class Engine {
public:
    ~Engine();
    // state bytes per trajectory slot, per resident set
    static uint64_t bytes_per_traj(int precision);
    static uint64_t bytes_per_set();
    int init(int precision, uint64_t capacity, int max_sets);
    int set_config(const ldpcp_engine_config& cfg);
    const ldpcp_engine_config& config() const { return cfg_; }
    // validates a set (0 = ok, else LDPCP_EINVAL with the error set)
    static int validate(const ldpcp_sweep_set& s);
    // admits a set now if slots and an id are free; *set_id = -1 otherwise
    int admit(const ldpcp_sweep_set& s, int* set_id);
    // enqueues one launch of K steps (no wait)
    int launch(uint32_t K);
    // enqueues an asynchronous copy of the control words and per-set counters
    // into pinned host memory; `slot` picks one of two buffers
    int enqueue_counters(int slot);
    int wait_counters(int slot);
    const Ctrl& counters_ctrl(int slot) const { return *h_ctrl_[slot]; }
    const unsigned long long* counters_cnt(int slot) const { return h_cnt_[slot]; }
    // blocking reads; stats() of a set whose trajectories have all finished
    // may use another stream, so that it does not wait for launches in flight
    int stats(int set_id, ldpcp_set_stats* out, hipStream_t s = nullptr);
    int per_traj(int set_id, uint64_t* absorbed, uint64_t* tpt0, uint64_t* xc_first, uint64_t* xc_last,
                 hipStream_t s = nullptr);
    int live_count(uint64_t* live);
    // frees a set's slots and id (its trajectories must all have finished)
    void release(int set_id);
    int sync();
    // milliseconds of launches that have completed, and their count
    int kernel_time(double* ms, uint64_t* launches);
    int tune(uint32_t steps_hint, ldpcp_engine_config* chosen);

    int precision() const { return precision_; }
    uint64_t capacity() const { return capacity_; }
    int max_sets() const { return max_sets_; }
    uint64_t estep() const { return estep_; }
    int parity() const { return parity_; }
    const ldpcp_sweep_set& params(int id) const { return h_params_[id]; }
    hipStream_t stream() const { return stream_; }

private:
    int launch_plain(uint32_t K);
    int graph_for(uint32_t K, hipGraphExec_t* exec);
    void drop_graphs();
    int collect_times();

    int precision_ = 32;
    uint64_t capacity_ = 0;
    int max_sets_ = 0;
    ldpcp_engine_config cfg_{};
    int n_cu_ = 0;

    void* d_block_ = nullptr;
    SlotPtrs P_{};
    StatsPtrs St_{};
    SetDev* d_sets_ = nullptr;
    Ctrl* d_ctrl_ = nullptr;
    uint32_t* d_live_[2] = {nullptr, nullptr};
    unsigned long long* d_mask_ = nullptr;
    uint32_t* d_off_ = nullptr;

    SetDev* h_sets_ = nullptr;  // pinned staging, one entry per set id
    Ctrl* h_ctrl_[2] = {nullptr, nullptr};
    unsigned long long* h_cnt_[2] = {nullptr, nullptr};
    hipEvent_t ev_cnt_[2] = {nullptr, nullptr};

    std::vector<ldpcp_sweep_set> h_params_;
    std::vector<uint64_t> first_;
    std::vector<int> free_ids_;
    RangeAllocator slots_;

    int parity_ = 0;
    uint64_t estep_ = 0;
    hipStream_t stream_ = nullptr;
    std::map<uint64_t, hipGraphExec_t> graphs_;  // (parity, K) -> executable graph
    std::vector<hipGraph_t> graph_defs_;

    // launch timing: event pairs not yet read
    std::vector<std::pair<hipEvent_t, hipEvent_t>> ev_pending_, ev_free_;
    double kernel_ms_ = 0.0;
    uint64_t timed_launches_ = 0;
};

}  // namespace ldpcp
