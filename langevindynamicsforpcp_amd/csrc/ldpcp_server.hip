// ldpcp_server.hip -- request server over the first-passage engine
// (continuous batching on one GPU).
//
// Requests wait in FIFO order until their HBM footprint (trajectory state +
// per-set statistics) fits the budget and the slot pool has a contiguous range
// for them; they are then admitted between two launches, so their trajectories
// run in the same persistent launches as everything already resident.  After
// every launch the per-set finished counters are copied to pinned host memory;
// the host reads the copy of launch i while launch i+1 runs, and a request
// whose last trajectory has finished has its statistics (and, when the
// server keeps them, its per-trajectory first-passage lists) read on a second
// stream, its slots freed and its id reported as completed.
#include <hip/hip_runtime.h>
#include <stdint.h>
#include <string.h>

#include <deque>
#include <unordered_map>
#include <vector>

#include "ldpcp_common.hpp"
#include "ldpcp_engine.hpp"

using namespace ldpcp;

namespace {

struct Request {
    ldpcp_sweep_set set;
    uint64_t bytes = 0;
    int state = LDPCP_REQ_QUEUED;
    int engine_id = -1;
    ldpcp_set_stats result;
    // per trajectory, when the server keeps them: absorption step (0 = capped),
    // TPT start, first and last xc crossing (double barrier)
    std::vector<uint64_t> fpt[4];
};

}  // namespace

struct ldpcp_server {
    Engine eng;
    uint64_t budget = 0;
    bool keep_fpt = false;
    uint64_t resident_bytes = 0;
    std::deque<uint64_t> queue;
    std::unordered_map<uint64_t, Request> reqs;
    std::vector<uint64_t> by_engine_id;  // engine set id -> request id (0 = none)
    std::vector<uint64_t> completed;
    uint64_t next_id = 1;
    uint64_t live = 0;
    uint64_t steps_retired = 0;  // trajectory-steps of completed requests
    int slot = 0;                // counter buffer for the next launch
    int pending_slot = -1;       // counter buffer written by a launch not yet read
    int pending_parity = 0;      // live-list parity after that launch
    std::vector<uint64_t> steps_of;  // engine set id -> trajectory-steps at the last read
    // A set id freed by a retirement can be given to a new request before the
    // counters of the launch in flight are read; those counters still hold the
    // previous owner's values.  admitted_at[id] is the index of the first launch
    // that runs the id's current request, counter_launch[k] the index of the
    // launch whose counters buffer k holds; process() skips ids admitted after it.
    std::vector<uint64_t> admitted_at;
    uint64_t n_launched = 0;
    uint64_t counter_launch[2] = {0, 0};
    hipStream_t copy_stream = nullptr;

    ~ldpcp_server()
    {
        if (copy_stream) (void)hipStreamDestroy(copy_stream);
    }

    uint64_t footprint(const ldpcp_sweep_set& s) const
    {
        return s.n_traj * Engine::bytes_per_traj(eng.precision()) + Engine::bytes_per_set();
    }

    int admit_queued()
    {
        while (!queue.empty()) {
            Request& r = reqs[queue.front()];
            if (resident_bytes + r.bytes > budget) break;
            int id = -1;
            const int rc = eng.admit(r.set, &id);
            if (rc) return rc;
            if (id < 0) break;  // no contiguous slot range or set id free yet
            r.state = LDPCP_REQ_RUNNING;
            r.engine_id = id;
            admitted_at[id] = n_launched;
            resident_bytes += r.bytes;
            by_engine_id[id] = queue.front();
            queue.pop_front();
        }
        return LDPCP_OK;
    }

    // Reads the counter copy in buffer k and retires completed requests.
    int process(int k, int parity)
    {
        int rc = eng.wait_counters(k);
        if (rc) return rc;
        live = eng.counters_ctrl(k).n_live[parity];
        const unsigned long long* cnt = eng.counters_cnt(k);
        for (int id = 0; id < eng.max_sets(); ++id) {
            const uint64_t rid = by_engine_id[id];
            if (!rid) continue;
            if (admitted_at[id] > counter_launch[k]) continue;  // counters predate this request
            Request& r = reqs[rid];
            steps_of[id] = cnt[(uint64_t)id * kCntPerSet + 2];
            if (cnt[(uint64_t)id * kCntPerSet + 1] < r.set.n_traj) continue;
            // every trajectory has finished in a launch that has completed:
            // the set's statistics are final
            rc = eng.stats(id, &r.result, copy_stream);
            if (rc) return rc;
            if (keep_fpt) {
                const size_t n = (size_t)r.set.n_traj;
                for (auto& v : r.fpt) v.assign(n, 0);
                const bool dbl = r.set.model == 1;
                rc = eng.per_traj(id, r.fpt[0].data(), r.fpt[1].data(), dbl ? r.fpt[2].data() : nullptr,
                                  dbl ? r.fpt[3].data() : nullptr, copy_stream);
                if (rc) return rc;
            }
            eng.release(id);
            by_engine_id[id] = 0;
            resident_bytes -= r.bytes;
            steps_retired += r.result.traj_steps;
            steps_of[id] = 0;
            r.state = LDPCP_REQ_DONE;
            r.engine_id = -1;
            completed.push_back(rid);
        }
        return LDPCP_OK;
    }
};

extern "C" int ldpcp_server_create(const ldpcp_server_config* cfg, ldpcp_server** out)
{
    if (!cfg || !out || (cfg->precision != 32 && cfg->precision != 64) || cfg->max_requests < 0 ||
        (cfg->keep_fpt != 0 && cfg->keep_fpt != 1)) {
        set_error("ldpcp_server_create: need a config with precision 32 or 64, max_requests >= 0 and keep_fpt 0 or 1");
        return LDPCP_EINVAL;
    }
    *out = nullptr;
    const int max_sets = cfg->max_requests ? cfg->max_requests : 4096;
    const uint64_t per_traj = Engine::bytes_per_traj(cfg->precision);
    const uint64_t sets_bytes = (uint64_t)max_sets * Engine::bytes_per_set();
    if (cfg->budget_bytes <= sets_bytes + per_traj * 64) {
        set_error("ldpcp_server_create: budget of %llu bytes does not hold the tables of %d requests",
                  (unsigned long long)cfg->budget_bytes, max_sets);
        return LDPCP_EINVAL;
    }
    uint64_t capacity = (cfg->budget_bytes - sets_bytes) / per_traj;
    if (capacity > (1ull << 32) - 128) capacity = (1ull << 32) - 128;
    ldpcp_server* s = new ldpcp_server();
    int rc = s->eng.init(cfg->precision, capacity, max_sets);
    if (!rc) rc = s->eng.set_config(cfg->engine);
    if (!rc) {
        const hipError_t e = hipStreamCreateWithFlags(&s->copy_stream, hipStreamNonBlocking);
        if (e != hipSuccess) {
            set_error("ldpcp_server_create: hipStreamCreateWithFlags failed: %s", hipGetErrorString(e));
            rc = LDPCP_EDEVICE;
        }
    }
    if (rc) {
        delete s;
        return rc;
    }
    s->budget = cfg->budget_bytes;
    s->keep_fpt = cfg->keep_fpt == 1;
    s->by_engine_id.assign((size_t)max_sets, 0);
    s->steps_of.assign((size_t)max_sets, 0);
    s->admitted_at.assign((size_t)max_sets, 0);
    *out = s;
    return LDPCP_OK;
}

extern "C" int ldpcp_server_footprint(const ldpcp_server* s, const ldpcp_sweep_set* set, uint64_t* bytes)
{
    if (!s || !set || !bytes) {
        set_error("ldpcp_server_footprint: null argument");
        return LDPCP_EINVAL;
    }
    *bytes = s->footprint(*set);
    return LDPCP_OK;
}

extern "C" int ldpcp_server_submit(ldpcp_server* s, const ldpcp_sweep_set* set, uint64_t* request_id)
{
    if (!s || !set || !request_id) {
        set_error("ldpcp_server_submit: null argument");
        return LDPCP_EINVAL;
    }
    int rc = Engine::validate(*set);
    if (rc) return rc;
    const uint64_t bytes = s->footprint(*set);
    if (bytes > s->budget || set->n_traj > s->eng.capacity()) {
        set_error("ldpcp_server_submit: request of %llu trajectories (%llu bytes) exceeds the budget",
                  (unsigned long long)set->n_traj, (unsigned long long)bytes);
        return LDPCP_ERANGE;
    }
    const uint64_t id = s->next_id++;
    Request r;
    r.set = *set;
    r.bytes = bytes;
    memset(&r.result, 0, sizeof(r.result));
    s->reqs.emplace(id, r);
    s->queue.push_back(id);
    *request_id = id;
    return LDPCP_OK;
}

extern "C" int ldpcp_server_pump(ldpcp_server* s, uint32_t n_launches, int32_t wait)
{
    if (!s) {
        set_error("ldpcp_server_pump: null server");
        return LDPCP_EINVAL;
    }
    int rc;
    for (uint32_t i = 0; i < n_launches; ++i) {
        rc = s->admit_queued();
        if (rc) return rc;
        bool resident = false;
        for (uint64_t rid : s->by_engine_id)
            if (rid) {
                resident = true;
                break;
            }
        if (!resident) break;
        rc = s->eng.launch(0);
        if (rc) return rc;
        const int k = s->slot;
        s->counter_launch[k] = s->n_launched++;
        rc = s->eng.enqueue_counters(k);
        if (rc) return rc;
        const int parity_now = s->eng.parity();
        // read the previous launch's counters while this one runs
        if (s->pending_slot >= 0) {
            rc = s->process(s->pending_slot, s->pending_parity);
            if (rc) return rc;
        }
        s->pending_slot = k;
        s->pending_parity = parity_now;
        s->slot ^= 1;
    }
    if (wait && s->pending_slot >= 0) {
        rc = s->process(s->pending_slot, s->pending_parity);
        if (rc) return rc;
        s->pending_slot = -1;
        rc = s->admit_queued();
        if (rc) return rc;
    }
    return LDPCP_OK;
}

extern "C" int ldpcp_server_status(ldpcp_server* s, uint64_t request_id, int32_t* state)
{
    if (!s || !state) {
        set_error("ldpcp_server_status: null argument");
        return LDPCP_EINVAL;
    }
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end()) {
        set_error("ldpcp_server_status: unknown request %llu", (unsigned long long)request_id);
        return LDPCP_EINVAL;
    }
    *state = it->second.state;
    return LDPCP_OK;
}

extern "C" int ldpcp_server_completed(ldpcp_server* s, uint64_t* ids, int32_t max, int32_t* n)
{
    if (!s || !n || (max > 0 && !ids)) {
        set_error("ldpcp_server_completed: null argument");
        return LDPCP_EINVAL;
    }
    const int32_t k = (int32_t)std::min<size_t>((size_t)std::max(max, 0), s->completed.size());
    for (int32_t i = 0; i < k; ++i) ids[i] = s->completed[(size_t)i];
    s->completed.erase(s->completed.begin(), s->completed.begin() + k);
    *n = k;
    return LDPCP_OK;
}

extern "C" int ldpcp_server_result(ldpcp_server* s, uint64_t request_id, ldpcp_set_stats* out)
{
    if (!s || !out) {
        set_error("ldpcp_server_result: null argument");
        return LDPCP_EINVAL;
    }
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end() || it->second.state != LDPCP_REQ_DONE) {
        set_error("ldpcp_server_result: request %llu is unknown or not done", (unsigned long long)request_id);
        return LDPCP_EINVAL;
    }
    *out = it->second.result;
    s->reqs.erase(it);
    return LDPCP_OK;
}

extern "C" int ldpcp_server_fpt(ldpcp_server* s, uint64_t request_id, uint64_t* absorbed_step, uint64_t* tpt_start,
                                uint64_t* xc_first, uint64_t* xc_last)
{
    if (!s || !absorbed_step) {
        set_error("ldpcp_server_fpt: null argument");
        return LDPCP_EINVAL;
    }
    if (!s->keep_fpt) {
        set_error("ldpcp_server_fpt: the server was created without keep_fpt");
        return LDPCP_EINVAL;
    }
    auto it = s->reqs.find(request_id);
    if (it == s->reqs.end() || it->second.state != LDPCP_REQ_DONE) {
        set_error("ldpcp_server_fpt: request %llu is unknown, not done or its result was read",
                  (unsigned long long)request_id);
        return LDPCP_EINVAL;
    }
    const Request& r = it->second;
    uint64_t* dst[4] = {absorbed_step, tpt_start, xc_first, xc_last};
    for (int k = 0; k < 4; ++k)
        if (dst[k]) {
            if (r.fpt[k].empty()) {
                memset(dst[k], 0, (size_t)r.set.n_traj * 8);
            } else {
                memcpy(dst[k], r.fpt[k].data(), r.fpt[k].size() * 8);
            }
        }
    return LDPCP_OK;
}

extern "C" int ldpcp_server_info(ldpcp_server* s, uint64_t* queued, uint64_t* resident, uint64_t* live_traj,
                                 uint64_t* resident_bytes, uint64_t* traj_steps)
{
    if (!s) {
        set_error("ldpcp_server_info: null server");
        return LDPCP_EINVAL;
    }
    uint64_t nres = 0, steps = s->steps_retired;
    for (int id = 0; id < s->eng.max_sets(); ++id)
        if (s->by_engine_id[id]) {
            nres++;
            steps += s->steps_of[id];
        }
    if (queued) *queued = s->queue.size();
    if (resident) *resident = nres;
    if (live_traj) *live_traj = s->live;
    if (resident_bytes) *resident_bytes = s->resident_bytes;
    if (traj_steps) *traj_steps = steps;
    return LDPCP_OK;
}

extern "C" int ldpcp_server_destroy(ldpcp_server* s)
{
    delete s;
    return LDPCP_OK;
}
