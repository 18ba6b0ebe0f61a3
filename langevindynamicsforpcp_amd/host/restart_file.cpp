// restart_file.cpp -- see restart_file.h.
#include "restart_file.h"

#include <math.h>
#include <stddef.h>
#include <stdio.h>
#include <string.h>

#include "ldpcp.h"

namespace ldpcp_host {

namespace {

constexpr int kBodyStart = 9998;  // TrajectoryData::WriteRestartFile's leading marker

class Writer {
public:
    explicit Writer(FILE* f) : f_(f) {}
    template <typename T>
    void put(T v) { ok_ = ok_ && fwrite(&v, sizeof(T), 1, f_) == 1; }
    void i32(int v) { put<int32_t>(v); }
    void u64(uint64_t v) { put<uint64_t>(v); }
    void f64(double v) { put<double>(v); }
    void bytes(const void* p, size_t n) { ok_ = ok_ && fwrite(p, 1, n, f_) == n; }
    bool ok() const { return ok_; }

private:
    FILE* f_;
    bool ok_ = true;
};

class Reader {
public:
    explicit Reader(FILE* f) : f_(f) {}
    template <typename T>
    T get()
    {
        T v{};
        if (fread(&v, sizeof(T), 1, f_) != 1) eof_ = true;
        return v;
    }
    int i32() { return get<int32_t>(); }
    uint64_t u64() { return get<uint64_t>(); }
    double f64() { return get<double>(); }
    bool bytes(void* p, size_t n) { return fread(p, 1, n, f_) == n; }
    bool eof() const { return eof_; }
    // a size_t count followed by that many doubles (StreamerWriter::WriteVector)
    bool vec(std::vector<double>& v)
    {
        const uint64_t n = u64();
        if (eof_ || n > (1ull << 36)) return eof_ = true, false;
        v.resize(n);
        if (n && !bytes(v.data(), n * sizeof(double))) return eof_ = true, false;
        return true;
    }

private:
    FILE* f_;
    bool eof_ = false;
};

// The saved-field flags the runner sets on the integrator (Single.h:3261-3271):
// x always; v, f, e, state and barrier with WRITEX, except v for the
// overdamped integrator, which clears it (:2822-2823); eta never (no input key).
struct SaveFlags {
    int x, v, f, e, eta, state, barrier;
};

SaveFlags save_flags(const RunnerConfig& c)
{
    const int w = c.write_x ? 1 : 0;
    return {1, c.overdamped ? 0 : w, w, w, 0, w, w};
}

// bRandomizeEta: the programs pass true, their input has no key for it
// (langevinDynamicsSingleBarrier.cpp:136)
constexpr int kRandomizeEta = 1;

// m_Xc0 of the double barrier (Double.h:1135-1146): the apex of the two
// inner slopes, written to the header and not checked on read.
double apex_x(const RunnerConfig& c)
{
    const double x1b = c.position[0], w1 = c.half_width[0], x2b = c.position[1], w2 = c.half_width[1];
    const double y1 = c.height[0][0], y2 = c.height[1][0];
    const double x1h = x1b + w1, x2l = x2b - w2;
    return fma(x1h * y1, x2b - x2l, x2l * y2 * (x1h - x1b)) / fma(y1, x2b - x2l, y2 * (x1h - x1b));
}

// The double barrier's per-state potential constants a lane carries in its
// restart record (IntegrateInitializeVariables, Double.h:1845-1899, selected by
// UpdateBarrierStateEtaParamter, :2737-2753): [barrier][state][0 falling side
// (m_SlopeNegative / m_InterceptNegative), 1 rising side (…Positive)].  The
// programs recompute them at every step, so they are written for the
// reference's reader and ignored on reading.
struct DoubleBarrierConstants {
    double slope[2][2][2], intercept[2][2][2];
};

DoubleBarrierConstants double_barrier_constants(const RunnerConfig& c)
{
    const double x1b = c.position[0], w1 = c.half_width[0], x2b = c.position[1], w2 = c.half_width[1];
    const double xc = apex_x(c);
    const double yc = (x1b + w1 - xc) / (x1b + w1 - x1b) * c.height[0][0];  // GetXc, Double.h:1730
    const double low0 = x1b - w1, high1 = x2b + w2;
    const double xb_xl0 = x1b - low0, xh_xb1 = high1 - x2b;
    DoubleBarrierConstants k;
    for (int s = 0; s < 2; ++s) {
        const double h0 = c.height[0][s], h1 = c.height[1][s];
        k.slope[0][s][0] = h0 / xb_xl0;
        k.slope[0][s][1] = (yc - h0) / (xc - x1b);
        k.slope[1][s][0] = (h1 - yc) / (x2b - xc);
        k.slope[1][s][1] = 0.0 - h1 / xh_xb1;
        k.intercept[0][s][0] = 0.0 - h0 * (low0 / xb_xl0);
        k.intercept[0][s][1] = (xc * h0 - yc * x1b) / (xc - x1b);
        k.intercept[1][s][0] = (yc * x2b - xc * h1) / (x2b - xc);
        k.intercept[1][s][1] = h1 * (high1 / xh_xb1);
    }
    return k;
}

constexpr int kDoubleTablesEnd = 9948;  // after the double barrier's two tables (Double.h:640)

// ReadRestartFile's error exit: message, close, false (Single.h:1143-1190)
bool fail(FILE* f, const char* what)
{
    fprintf(stderr, "Error: %s\n", what);
    fclose(f);
    return false;
}

// One lane of TrajectoryData as WriteRestartFile lays it out
// (langevinDynamicsSingleBarrier.h:436-572; langevinDynamicsDoubleBarrier.h:465-634,
// which adds xp, the per-barrier slopes, intercepts and heights and the local
// barrier indices).
struct RefLane {
    double x, v, f, e, eta, barrier, xp = 0.0;
    std::vector<double> traj[8];  // x, v, f, pe, ke, eta, state, barrier frames
    double state, prev_state, count0, count1, well, periodic, well_prev, periodic_prev;
    double local = 0.0;  // m_LocalBarrierIndex (double barrier)
    double v0_instance, eta0_instance, state0_instance;
    double tick_prev = 0.0;
    int passed = 0;
    std::vector<double> ticks;
    std::vector<double> fpt, fpt2, tpt, tpt2, inst, init;
    double fpt_init_state[2] = {0.0, 0.0};
    int fpt_init_index[2] = {0, 0};
    double fpt_last = 0.0, fpt2_last = 0.0, tpt_last = 0.0, tpt2_last = 0.0;
    int fpt_up = 0, fpt2_up = 0, tpt_up = 0, tpt2_up = 0, tpt2_well = 0, fpt2_idx[2] = {0, 0};
    std::vector<int> cg_size;
    std::vector<std::vector<double>> cg;
    std::vector<double> cg_last;
    std::vector<int> cg_up, cg_well;
};

bool read_ref_lane(Reader& r, const RunnerConfig& c, bool wt, bool fpt, RefLane& L)
{
    const bool dbl = c.model != 0;
    L.x = r.f64();
    L.v = r.f64();
    L.f = r.f64();
    L.e = r.f64();
    L.eta = r.f64();
    L.barrier = r.f64();
    if (dbl) L.xp = r.f64();
    for (auto& t : L.traj)
        if (!r.vec(t)) return false;
    L.state = r.f64();
    L.prev_state = r.f64();
    // slopes (and intercepts) and the current height(s): recomputed by the next step
    for (int k = 0; k < (dbl ? 10 : 3); ++k) r.f64();
    L.count0 = r.f64();
    L.count1 = r.f64();
    L.well = r.f64();
    L.periodic = r.f64();
    L.well_prev = r.f64();
    L.periodic_prev = r.f64();
    r.i32();  // entering flags: recomputed by the next step
    r.i32();
    if (dbl) {
        L.local = r.f64();
        r.f64();  // previous local barrier index, its entering flag and the *Stored copies:
        r.i32();  // recomputed by the next step
        r.f64();
        r.f64();
        r.i32();
    }
    L.v0_instance = r.f64();
    L.eta0_instance = r.f64();
    L.state0_instance = r.f64();
    for (int k = 0; k < 4; ++k) r.f64();  // the *Stored indices of the SIMD bookkeeping
    r.i32();
    r.i32();
    if (wt) {
        L.tick_prev = r.f64();
        L.passed = r.i32();
        if (!r.vec(L.ticks)) return false;
    }
    if (fpt) {
        if (!r.vec(L.fpt) || !r.vec(L.inst) || !r.vec(L.init)) return false;
        L.fpt_last = r.f64();
        L.fpt_up = r.i32();
        L.fpt_init_state[0] = r.f64();
        L.fpt_init_state[1] = r.f64();
        L.fpt_init_index[0] = r.i32();
        L.fpt_init_index[1] = r.i32();
        if (!r.vec(L.fpt2)) return false;
        L.fpt2_last = r.f64();
        L.fpt2_up = r.i32();
        L.fpt2_idx[0] = r.i32();
        L.fpt2_idx[1] = r.i32();
    }
    if (c.tpt) {
        if (!r.vec(L.tpt)) return false;
        L.tpt_last = r.f64();
        L.tpt_up = r.i32();
        if (!r.vec(L.tpt2)) return false;
        L.tpt2_last = r.f64();
        L.tpt2_up = r.i32();
        L.tpt2_well = r.i32();
    }
    if (c.fptcg) {
        const uint64_t n = r.u64();
        if (r.eof() || n > 1024) return false;
        L.cg_size.resize(n);
        for (auto& v : L.cg_size) v = r.i32();
        L.cg.resize(n);
        L.cg_last.resize(n);
        L.cg_up.resize(n);
        L.cg_well.resize(n);
        for (uint64_t j = 0; j < n; ++j) {
            if (!r.vec(L.cg[j])) return false;
            L.cg_last[j] = r.f64();
            L.cg_up[j] = r.i32();
            L.cg_well[j] = r.i32();
        }
    }
    return !r.eof();
}

// STATISTICS::AVERAGE sums a list front to back; the device keeps the same sum as
// it goes, so the running sum of an imported list is this loop.
double list_sum(const std::vector<double>& v)
{
    double s = 0.0;
    for (double x : v) s += x;
    return s;
}

// A barrier table as the reference's file holds it ([id][lane]).
struct RefTable {
    uint64_t n = 0;
    int min_id = 0;
    std::vector<int32_t> state;
    std::vector<uint64_t> step;
};

bool read_ref_table(Reader& r, int W, RefTable& t)
{
    t.n = r.u64();
    if (r.eof() || t.n > (1u << 24)) return false;
    t.state.resize(t.n * W);
    t.step.resize(t.n * W);
    for (uint64_t b = 0; b < t.n; ++b)
        for (int i = 0; i < W; ++i) {
            t.state[b * W + i] = r.i32();
            t.step[b * W + i] = r.u64();
        }
    t.min_id = r.i32();
    return !r.eof();
}

// The reference program's body (either model) -> checkpoint image.
bool import_reference_body(Reader& r, FILE* f, const RunnerConfig& c, int W, const ImageSpec& spec,
                           uint64_t nsteps_total, std::vector<char>& image)
{
    if (r.i32() != kBodyStart) return fail(f, "restart file corrupted");
    const bool wt = r.i32() == 1;
    const bool fpt = r.i32() == 1;
    if (wt != c.waiting) return fail(f, "inconsistent bWaitingTimeDistribution");
    if (fpt != c.fpt) return fail(f, "inconsistent bFirstPassageTime");
    std::vector<RefLane> lanes(W);
    for (int i = 0; i < W; ++i)
        if (!read_ref_lane(r, c, wt, fpt, lanes[i])) return fail(f, "restart file corrupted");
    const int n_arrays = c.model != 0 ? 2 : 1;
    RefTable tables[2];
    for (int a = 0; a < n_arrays; ++a)
        if (!read_ref_table(r, W, tables[a])) return fail(f, "restart file corrupted");
    // the double barrier's tables end with their own marker, which the
    // reference's reader only reports (Double.h:890-894)
    if (c.model != 0 && r.i32() != kDoubleTablesEnd) return fail(f, "wrong signiture");
    std::vector<double> r_prev(W, 0.0);
    if (c.overdamped)
        for (auto& v : r_prev) v = r.f64();
    if (r.i32() != kRestartEnd || r.eof()) return fail(f, "restart file corrupted");
    // a file this program wrote carries its exact state after the end marker
    const int next = fgetc(f);
    if (next != EOF) {
        ungetc(next, f);
        const uint64_t magic = r.u64(), size = r.u64();
        if (r.eof() || magic != LDPCP_CHECKPOINT_MAGIC || size > (1ull << 40))
            return fail(f, "restart file corrupted");
        image.resize(size);
        if (!r.bytes(image.data(), size) || fgetc(f) != EOF) return fail(f, "restart file corrupted");
        fclose(f);
        return true;
    }

    // checks the resumed run relies on: at least the step-0 frame, at most one per
    // FREQSAVE steps of the file's clock (the resumed run records its frames every
    // FREQSAVE steps from that clock, as the reference's does)
    const uint64_t n_frames = lanes[0].traj[0].size();
    if (n_frames < 1 || n_frames > nsteps_total / c.save_freq + 1)
        return fail(f, "the restart file's frames do not match its step count");
    const int nfields = spec.n_fields;
    const int src_field[7] = {0, 1, 2, 3, 4, 6, 7};  // x v f pe ke state barrier (eta is not kept)
    for (int i = 0; i < W; ++i) {
        // the overdamped integrator saves no velocity frames (Single.h:2822-2823); its
        // frames of v are zeros here, as the device records them
        if (c.overdamped && nfields > 1 && lanes[i].traj[1].empty()) lanes[i].traj[1].assign(n_frames, 0.0);
        for (int k = 0; k < nfields; ++k)
            if (lanes[i].traj[src_field[k]].size() != n_frames)
                return fail(f, "the restart file lacks frames this run saves (WRITEX)");
        if (c.fptcg) {
            bool same = lanes[i].cg_size.size() == c.cg_size.size();
            for (size_t j = 0; same && j < c.cg_size.size(); ++j) same = lanes[i].cg_size[j] == c.cg_size[j];
            if (!same) return fail(f, "inconsistent FPTCG sizes");
        }
    }

    const uint64_t n = (uint64_t)W, wb = (uint64_t)spec.barrier_window, cap = spec.tick_capacity;
    const uint64_t o_lanes = sizeof(ldpcp_checkpoint_header);
    const uint64_t o_wbase = o_lanes + n * sizeof(ldpcp_lane);
    const uint64_t o_win = o_wbase + ((n * sizeof(int32_t) + 7) & ~7ull);
    const uint64_t o_frames = o_win + n * 2 * wb * sizeof(uint64_t);
    const uint64_t o_ticks = o_frames + (uint64_t)nfields * n_frames * n * sizeof(double);
    const uint64_t o_events = o_ticks + n * cap * sizeof(double);
    // the event lists in the library's order (LDPCP_LIST_*), at most event_capacity entries each
    const uint32_t n_lists = spec.n_lists;
    const uint64_t ecap = spec.event_capacity;
    auto list_of = [&](const RefLane& R, uint32_t l) -> const std::vector<double>& {
        switch (l) {
        case LDPCP_LIST_FPT: return R.fpt;
        case LDPCP_LIST_FPT2: return R.fpt2;
        case LDPCP_LIST_TPT: return R.tpt;
        case LDPCP_LIST_TPT2: return R.tpt2;
        case LDPCP_LIST_INSTANT: return R.inst;
        case LDPCP_LIST_INITIAL: return R.init;
        default: return R.cg[l - LDPCP_LIST_CG0];
        }
    };
    uint64_t n_events = 0;
    bool complete = n_lists > 0;
    for (int i = 0; i < W; ++i)
        for (uint32_t l = 0; l < n_lists; ++l) {
            const uint64_t len = list_of(lanes[i], l).size();
            n_events += len < ecap ? len : ecap;
            complete = complete && len <= ecap;
        }
    const uint64_t total = o_events + n_events * sizeof(double);
    image.assign(total, 0);
    char* b = image.data();

    ldpcp_checkpoint_header hd;
    memset(&hd, 0, sizeof(hd));
    hd.magic = LDPCP_CHECKPOINT_MAGIC;
    hd.header_bytes = sizeof(ldpcp_checkpoint_header);
    hd.lane_bytes = sizeof(ldpcp_lane);
    hd.n_traj = (uint32_t)n;
    hd.traj_offset = spec.traj_offset;
    hd.model = c.model;
    hd.overdamped = c.overdamped ? 1 : 0;
    hd.barrier_window = spec.barrier_window;
    hd.n_fields = nfields;
    hd.seed = spec.seed;
    hd.steps_done = nsteps_total;
    hd.n_frames = n_frames;
    hd.save_freq = c.save_freq;
    hd.tick_capacity = cap;
    hd.timestep = c.timestep;
    hd.bytes = total;
    hd.event_capacity = ecap;
    hd.n_lists = n_lists;
    hd.events_complete = complete ? 1u : 0u;
    memcpy(b, &hd, sizeof(hd));

    constexpr uint64_t kValid = 1ull << 63, kState = 1ull << 62, kStepMask = kState - 1;
    for (int i = 0; i < W; ++i) {
        const RefLane& R = lanes[i];
        ldpcp_lane L;
        memset(&L, 0, sizeof(L));
        L.x = R.x;
        L.v = R.v;
        L.f = R.f;
        L.e = R.e;
        L.r_prev = r_prev[i];
        L.state = R.state;
        L.barrier = R.barrier;
        L.xp = R.xp;
        L.local_barrier = R.local;
        L.state_count[0] = R.count0;
        L.state_count[1] = R.count1;
        L.well = R.well;
        L.periodic = R.periodic;
        L.v0_instance = R.v0_instance;
        L.state0_instance = R.state0_instance;
        L.fpt_sum = list_sum(R.fpt);
        L.fpt_n = R.fpt.size();
        L.fpt_last_tick = R.fpt_last;
        L.fpt_to_update = R.fpt_up;
        L.fpt2_sum = list_sum(R.fpt2);
        L.fpt2_n = R.fpt2.size();
        L.fpt2_last_tick = R.fpt2_last;
        L.fpt2_to_update = R.fpt2_up;
        L.fpt2_idx[0] = R.fpt2_idx[0];
        L.fpt2_idx[1] = R.fpt2_idx[1];
        L.tpt_sum = list_sum(R.tpt);
        L.tpt_n = R.tpt.size();
        L.tpt_last_tick = R.tpt_last;
        L.tpt_to_update = R.tpt_up;
        L.tpt2_sum = list_sum(R.tpt2);
        L.tpt2_n = R.tpt2.size();
        L.tpt2_last_tick = R.tpt2_last;
        L.tpt2_to_update = R.tpt2_up;
        L.tpt2_initial_well = R.tpt2_well;
        for (size_t j = 0; j < R.cg.size() && j < LDPCP_MAX_CG; ++j) {
            L.cg_sum[j] = list_sum(R.cg[j]);
            L.cg_n[j] = R.cg[j].size();
            L.cg_last_tick[j] = R.cg_last[j];
            L.cg_to_update[j] = R.cg_up[j];
            L.cg_initial_well[j] = R.cg_well[j];
        }
        // waiting time: the barrier ticks and the moments of their spacings, in
        // the order the device accumulates them
        L.wt_tick_prev = R.tick_prev;
        L.wt_passed = R.passed;
        L.wt_ticks = R.ticks.size();
        for (size_t k = 0; k < R.ticks.size(); ++k) {
            if (k > 0) {
                const double w = (R.ticks[k] - L.wt_last) * c.timestep;
                L.wt_sum += w;
                L.wt_sum2 = fma(w, w, L.wt_sum2);
                if (L.wt_n == 0 || w > L.wt_max) L.wt_max = w;
                L.wt_n++;
            }
            L.wt_last = R.ticks[k];
        }
        L.steps_done = nsteps_total;
        L.fpt_init_state[0] = R.fpt_init_state[0];
        L.fpt_init_state[1] = R.fpt_init_state[1];
        L.fpt_init_index[0] = R.fpt_init_index[0];
        L.fpt_init_index[1] = R.fpt_init_index[1];
        memcpy(b + o_lanes + i * sizeof(ldpcp_lane), &L, sizeof(L));
        // barrier tables -> the lane's window, centred on its current barrier
        const int32_t base = (int32_t)R.periodic - spec.barrier_window / 2;
        memcpy(b + o_wbase + i * sizeof(int32_t), &base, sizeof(base));
        for (int a = 0; a < n_arrays; ++a) {
            const RefTable& T = tables[a];
            for (uint64_t k = 0; k < T.n; ++k) {
                const int64_t slot = (int64_t)T.min_id + (int64_t)k - base;
                if (slot < 0 || slot >= (int64_t)wb) continue;  // beyond reach of the window
                const uint64_t e = kValid | (T.state[k * W + i] ? kState : 0ull) | (T.step[k * W + i] & kStepMask);
                memcpy(b + o_win + (((uint64_t)i * 2 + a) * wb + (uint64_t)slot) * sizeof(uint64_t), &e, sizeof(e));
            }
        }
        for (int fld = 0; fld < nfields; ++fld)
            for (uint64_t fr = 0; fr < n_frames; ++fr)
                memcpy(b + o_frames + (((uint64_t)fld * n_frames + fr) * n + i) * sizeof(double),
                       &R.traj[src_field[fld]][fr], sizeof(double));
        const uint64_t keep = R.ticks.size() < cap ? R.ticks.size() : cap;
        if (keep) memcpy(b + o_ticks + (uint64_t)i * cap * sizeof(double), R.ticks.data(), keep * sizeof(double));
    }
    char* ev = b + o_events;
    for (int i = 0; i < W; ++i)
        for (uint32_t l = 0; l < n_lists; ++l) {
            const std::vector<double>& v = list_of(lanes[i], l);
            const uint64_t keep = v.size() < ecap ? v.size() : ecap;
            if (keep) memcpy(ev, v.data(), keep * sizeof(double));
            ev += keep * sizeof(double);
        }
    fclose(f);
    return true;
}

}  // namespace

std::string restart_file_name(const std::string& base, int rank, int trial)
{
    char buf[64];
    snprintf(buf, sizeof(buf), "_%05d_%05d.restart", rank, trial);
    return base + buf;
}

namespace {

// The settings header and the three noise positions (Single.h:1054-1101;
// Double.h:1208-1255).  The noise positions are those of this program: the
// Philox counter of the step noise is the step itself.
void write_header(Writer& w, const RunnerConfig& c, int signature, int simd_width, int rank, uint64_t nsteps_total)
{
    const SaveFlags sf = save_flags(c);
    w.i32(signature);
    w.i32(simd_width);
    w.f64(c.x0);
    w.f64(c.v0);
    w.f64(c.gamma);
    w.f64(c.kb);
    w.f64(c.temperature);
    w.f64(c.mass);
    w.f64(c.timestep);
    w.u64(nsteps_total);
    w.i32(c.waiting ? 1 : 0);
    w.u64(c.stat_freq);
    w.u64(c.save_freq);
    w.i32(sf.x);
    w.i32(sf.v);
    w.i32(sf.f);
    w.i32(sf.e);
    w.i32(c.randomize_v ? 1 : 0);
    w.i32(kRandomizeEta);
    w.f64(c.periodic_length);
    if (c.model == 0) {
        w.f64(c.height[0][0]);
        w.f64(c.height[0][1]);
        w.f64(c.initial_state);
    } else {
        w.f64(c.height[0][0]);
        w.f64(c.height[0][1]);
        w.f64(c.height[1][0]);
        w.f64(c.height[1][1]);
        w.f64(c.initial_state);
        w.f64(c.position[0]);
        w.f64(c.position[1]);
        w.f64(apex_x(c));
    }
    w.i32(c.randomize_state ? 1 : 0);
    w.i32(c.same_initial_barrier ? 1 : 0);
    w.f64(c.q);
    w.f64(c.tau);
    w.f64(c.eta0);
    w.i32(sf.eta);
    w.i32(sf.state);
    w.i32(sf.barrier);
    w.i32(rank);
    w.put<int64_t>((int64_t)nsteps_total);
    w.u64(0);
    w.put<int64_t>((int64_t)nsteps_total);
    w.u64(0);
    w.put<int64_t>(0);
    w.u64(0);
}

bool finish(FILE* f, const Writer& w, const std::string& name)
{
    const bool ok = w.ok();
    if (fclose(f) != 0 || !ok) {
        fprintf(stderr, "Error: can't open restart file (%s)\n", name.c_str());
        return false;
    }
    return true;
}

}  // namespace

bool write_restart_file(const RunnerConfig& c, int simd_width, int rank, int trial, uint64_t nsteps_total,
                        const std::vector<char>& image)
{
    const std::string name = restart_file_name(c.output_restart, rank, trial);
    FILE* f = fopen(name.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "Error: can't open restart file (%s)\n", name.c_str());
        return false;
    }
    Writer w(f);
    write_header(w, c, kRestartSignature, simd_width, rank, nsteps_total);
    w.i32(kBodyStart);
    w.u64(image.size());
    w.bytes(image.data(), image.size());
    w.i32(kRestartEnd);
    return finish(f, w, name);
}

bool reference_layout_possible(const RunnerConfig& c, const std::vector<char>& image)
{
    if (image.size() < sizeof(ldpcp_checkpoint_header)) return false;
    ldpcp_checkpoint_header hd;
    memcpy(&hd, image.data(), sizeof(hd));
    if (!hd.events_complete) return false;
    const ldpcp_lane* lanes = (const ldpcp_lane*)(image.data() + sizeof(hd));
    for (uint32_t t = 0; t < hd.n_traj; ++t)
        if (c.waiting && lanes[t].wt_ticks > hd.tick_capacity) return false;
    return true;
}

int step0_barrier_reach(const RunnerConfig& c) { return c.model != 0 ? 10000 : 100; }

void barrier_span(const std::vector<char>& image, int array, int* lo, int* hi)
{
    ldpcp_checkpoint_header hd;
    memcpy(&hd, image.data(), sizeof(hd));
    const uint64_t n = hd.n_traj, wb = (uint64_t)hd.barrier_window;
    const char* b = image.data() + sizeof(hd) + n * sizeof(ldpcp_lane);
    const char* win = b + ((n * sizeof(int32_t) + 7) & ~7ull);
    for (uint64_t t = 0; t < n; ++t) {
        int32_t base;
        memcpy(&base, b + t * sizeof(int32_t), sizeof(base));
        for (uint64_t k = 0; k < wb; ++k) {
            uint64_t e;
            memcpy(&e, win + ((t * 2 + (uint64_t)array) * wb + k) * sizeof(uint64_t), sizeof(e));
            if (!(e >> 63)) continue;  // never created
            const int id = base + (int)k;
            if (id < *lo) *lo = id;
            if (id > *hi) *hi = id;
        }
    }
}

bool write_reference_restart_file(const RunnerConfig& c, int simd_width, int rank, int trial, uint64_t nsteps_total,
                                  const std::vector<char>& image, const std::vector<BarrierTable>& tables)
{
    const bool dbl = c.model != 0;
    if (tables.size() != (dbl ? 2u : 1u)) {
        fprintf(stderr, "Error: restart file: %zu barrier tables for a model with %d\n", tables.size(), dbl ? 2 : 1);
        return false;
    }
    ldpcp_checkpoint_header hd;
    memcpy(&hd, image.data(), sizeof(hd));
    const uint64_t n = hd.n_traj, wb = (uint64_t)hd.barrier_window, nf = hd.n_frames, cap = hd.tick_capacity;
    const char* b = image.data();
    const ldpcp_lane* lanes = (const ldpcp_lane*)(b + sizeof(hd));
    const uint64_t o_wbase = sizeof(hd) + n * sizeof(ldpcp_lane);
    const uint64_t o_frames = o_wbase + ((n * sizeof(int32_t) + 7) & ~7ull) + n * 2 * wb * sizeof(uint64_t);
    const uint64_t o_ticks = o_frames + (uint64_t)hd.n_fields * nf * n * sizeof(double);
    const uint64_t o_events = o_ticks + n * cap * sizeof(double);
    auto frame = [&](int field, uint64_t fr, uint64_t t) {
        double v;
        memcpy(&v, b + o_frames + (((uint64_t)field * nf + fr) * n + t) * sizeof(double), sizeof(v));
        return v;
    };
    // packed event lists: trajectory after trajectory, list after list
    std::vector<uint64_t> list_at(n * hd.n_lists);
    {
        uint64_t k = 0;
        for (uint64_t t = 0; t < n; ++t)
            for (uint32_t l = 0; l < hd.n_lists; ++l) {
                list_at[t * hd.n_lists + l] = k;
                const uint64_t cnt = ldpcp_list_count(lanes + t, (int)l);
                k += cnt < hd.event_capacity ? cnt : hd.event_capacity;
            }
    }
    const std::string name = restart_file_name(c.output_restart, rank, trial);
    FILE* f = fopen(name.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "Error: can't open restart file (%s)\n", name.c_str());
        return false;
    }
    Writer w(f);
    const SaveFlags sf = save_flags(c);
    write_header(w, c, kReferenceSignature, simd_width, rank, nsteps_total);
    // TrajectoryData::WriteRestartFile (Single.h:433-572; Double.h:470-640)
    const DoubleBarrierConstants dk = dbl ? double_barrier_constants(c) : DoubleBarrierConstants{};
    w.i32(kBodyStart);
    w.i32(c.waiting ? 1 : 0);
    w.i32(c.fpt ? 1 : 0);
    auto list = [&](uint64_t t, int l) {
        const uint64_t cnt = ldpcp_list_count(lanes + t, l);
        w.u64(cnt);
        w.bytes(b + o_events + list_at[t * hd.n_lists + l] * sizeof(double), cnt * sizeof(double));
    };
    // saved-frame lists x, v, f, pe, ke, eta, state, barrier -> this image's fields
    const int field_of[8] = {0, sf.v ? 1 : -1, sf.f ? 2 : -1, sf.e ? 3 : -1, sf.e ? 4 : -1, -1, sf.state ? 5 : -1,
                             sf.barrier ? 6 : -1};
    for (uint64_t t = 0; t < n; ++t) {
        const ldpcp_lane& L = lanes[t];
        const int st = L.state == 1.0 ? 1 : 0;
        w.f64(L.x);
        w.f64(L.v);
        w.f64(L.f);
        w.f64(L.e);
        w.f64(0.0);  // eta: the programs have no eta noise on
        w.f64(L.barrier);
        if (dbl) w.f64(L.xp);
        for (int k = 0; k < 8; ++k) {
            const int fld = field_of[k];
            if (fld < 0 || fld >= hd.n_fields) {
                w.u64(0);
                continue;
            }
            w.u64(nf);
            for (uint64_t fr = 0; fr < nf; ++fr) w.f64(frame(fld, fr, t));
        }
        w.f64(L.state);
        w.f64(L.state);  // previous state: overwritten by the next step before any use
        if (!dbl) {
            // UpdateBarrierStateEtaParamter (Single.h:2154-2165)
            const double slope = c.height[0][st] / c.half_width[0];
            w.f64(slope);
            w.f64(0.0 - slope);
            w.f64(c.height[0][st]);
        } else {
            for (int side = 1; side >= 0; --side)  // m_SlopePositive[0..1], then m_SlopeNegative[0..1]
                for (int k = 0; k < 2; ++k) w.f64(dk.slope[k][st][side]);
            for (int side = 1; side >= 0; --side)  // m_InterceptPositive, m_InterceptNegative
                for (int k = 0; k < 2; ++k) w.f64(dk.intercept[k][st][side]);
            w.f64(c.height[0][st]);
            w.f64(c.height[1][st]);
        }
        w.f64(L.state_count[0]);
        w.f64(L.state_count[1]);
        w.f64(L.well);
        w.f64(L.periodic);
        w.f64(L.well);  // previous indices: overwritten by the next step before any use
        w.f64(L.periodic);
        w.i32(0);
        w.i32(0);
        if (dbl) {
            // local barrier index, previous (overwritten by the next step), entering
            // flag, and the same three as the *Stored copies (Double.h:558-563)
            for (int k = 0; k < 2; ++k) {
                w.f64(L.local_barrier);
                w.f64(L.local_barrier);
                w.i32(0);
            }
        }
        w.f64(L.v0_instance);
        w.f64(0.0);
        w.f64(L.state0_instance);
        w.f64(L.well);
        w.f64(L.periodic);
        w.f64(L.well);
        w.f64(L.periodic);
        w.i32(0);
        w.i32(0);
        if (c.waiting) {
            w.f64(L.wt_tick_prev);
            w.i32(L.wt_passed);
            w.u64(L.wt_ticks);
            w.bytes(b + o_ticks + t * cap * sizeof(double), L.wt_ticks * sizeof(double));
        }
        if (c.fpt) {
            list(t, LDPCP_LIST_FPT);
            list(t, LDPCP_LIST_INSTANT);
            list(t, LDPCP_LIST_INITIAL);
            w.f64(L.fpt_last_tick);
            w.i32(L.fpt_to_update);
            w.f64(L.fpt_init_state[0]);
            w.f64(L.fpt_init_state[1]);
            w.i32(L.fpt_init_index[0]);
            w.i32(L.fpt_init_index[1]);
            list(t, LDPCP_LIST_FPT2);
            w.f64(L.fpt2_last_tick);
            w.i32(L.fpt2_to_update);
            w.i32(L.fpt2_idx[0]);
            w.i32(L.fpt2_idx[1]);
        }
        if (c.tpt) {
            list(t, LDPCP_LIST_TPT);
            w.f64(L.tpt_last_tick);
            w.i32(L.tpt_to_update);
            list(t, LDPCP_LIST_TPT2);
            w.f64(L.tpt2_last_tick);
            w.i32(L.tpt2_to_update);
            w.i32(L.tpt2_initial_well);
        }
        if (c.fptcg) {
            w.u64(c.cg_size.size());
            for (int v : c.cg_size) w.i32(v);
            for (size_t j = 0; j < c.cg_size.size(); ++j) {
                list(t, LDPCP_LIST_CG0 + (int)j);
                w.f64(L.cg_last_tick[j]);
                w.i32(L.cg_to_update[j]);
                w.i32(L.cg_initial_well[j]);
            }
        }
    }
    // the barrier table(s) (DataLocator.h:163-169; the double barrier's two and
    // their marker, Double.h:637-640), then R_prev (Single.h:2776-2785, Double.h:3391-3400)
    for (const BarrierTable& T : tables) {
        w.u64((uint64_t)T.n);
        for (int k = 0; k < T.n; ++k)
            for (uint64_t t = 0; t < n; ++t) {
                w.i32(T.states[t * T.n + k]);
                w.u64(T.steps[t * T.n + k]);
            }
        w.i32(T.lo);
    }
    if (dbl) w.i32(kDoubleTablesEnd);
    if (c.overdamped)
        for (uint64_t t = 0; t < n; ++t) w.f64(lanes[t].r_prev);
    w.i32(kRestartEnd);
    // this program's exact state after the reference's end marker, which the
    // reference's reader stops at
    w.u64(LDPCP_CHECKPOINT_MAGIC);
    w.u64(image.size());
    w.bytes(image.data(), image.size());
    return finish(f, w, name);
}

bool read_restart_file(const RunnerConfig& c, int simd_width, int rank, int trial, const ImageSpec& spec,
                       std::vector<char>& image, uint64_t* nsteps_total)
{
    const std::string name = restart_file_name(c.input_restart, rank, trial);
    FILE* f = fopen(name.c_str(), "rb");
    if (!f) {
        fprintf(stderr, "Error: can't open restart file (%s)\n", name.c_str());
        return false;
    }
    Reader r(f);
    const SaveFlags sf = save_flags(c);
    // the reference's checks, in its order and words (Single.h:1143-1190; Double.h:1302-1350)
    const int signature = r.i32();
    const bool reference_file = (signature == kReferenceSignature);
    if (signature != kRestartSignature && !reference_file) return fail(f, "inconsistent signature");
    if (r.i32() != simd_width) return fail(f, "inconsistent simd witdth");
    if (r.f64() != c.x0) return fail(f, "inconsistent x0");
    if (r.f64() != c.v0) return fail(f, "inconsistent v0");
    if (r.f64() != c.gamma) return fail(f, "inconsistent gamma0");
    if (r.f64() != c.kb) return fail(f, "inconsistent kb0");
    if (r.f64() != c.temperature) return fail(f, "inconsistent T0");
    if (r.f64() != c.mass) return fail(f, "inconsistent m0");
    if (r.f64() != c.timestep) return fail(f, "inconsistent time step0");
    *nsteps_total = r.u64();
    if (r.i32() != (c.waiting ? 1 : 0)) return fail(f, "inconsistent bWaitingTimeDistribution");
    if (r.u64() != c.stat_freq) return fail(f, "inconsistent statisticsFreq");
    if (r.u64() != c.save_freq) return fail(f, "inconsistent save freq");
    if (r.i32() != sf.x) return fail(f, "inconsistent bSaveX");
    r.i32();  // bSaveV is not checked (Single.h:1159)
    if (r.i32() != sf.f) return fail(f, "inconsistent bSaveF");
    if (r.i32() != sf.e) return fail(f, "inconsistent bSaveE");
    if (r.i32() != (c.randomize_v ? 1 : 0)) return fail(f, "inconsistent randomize velocity");
    if (r.i32() != kRandomizeEta) return fail(f, "inconsistent randomize eta");
    if (r.f64() != c.periodic_length) return fail(f, "inconsistent periodic length0");
    if (c.model == 0) {
        if (r.f64() != c.height[0][0]) return fail(f, "inconsistent barrier height0 [0]");
        if (r.f64() != c.height[0][1]) return fail(f, "inconsistent barrier height0 [1]");
        if (r.f64() != c.initial_state) return fail(f, "inconsistent initialState0");
    } else {
        if (r.f64() != c.height[0][0]) return fail(f, "inconsistent barrier height0 [0][0]");
        if (r.f64() != c.height[0][1]) return fail(f, "inconsistent barrier height0 [0][1]");
        if (r.f64() != c.height[1][0]) return fail(f, "inconsistent barrier height0 [1][0]");
        if (r.f64() != c.height[1][1]) return fail(f, "inconsistent barrier height0 [1][1]");
        if (r.f64() != c.initial_state) return fail(f, "inconsistent initialState0");
        if (r.f64() != c.position[0]) return fail(f, "inconsistent barrierPosition0[0]");
        if (r.f64() != c.position[1]) return fail(f, "inconsistent barrierPosition0[1]");
        r.f64();  // m_Xc0 is computed later (Double.h:1335)
    }
    if (r.i32() != (c.randomize_state ? 1 : 0)) return fail(f, "inconsistent randomizeState0");
    if (r.i32() != (c.same_initial_barrier ? 1 : 0)) return fail(f, "inconsistent bSameInitialBarrier");
    if (r.f64() != c.q) return fail(f, "inconsistent q0");
    if (r.f64() != c.tau) return fail(f, "inconsistent tau0");
    if (r.f64() != c.eta0) return fail(f, "inconsistent eta0");
    if (r.i32() != sf.eta) return fail(f, "inconsistent bSaveEta");
    if (r.i32() != sf.state) return fail(f, "inconsistent bSaveState");
    if (r.i32() != sf.barrier) return fail(f, "inconsistent save barrier");
    if (r.i32() != rank) return fail(f, "inconsistent rank");
    for (int k = 0; k < 3; ++k) {  // noise positions: the step count locates this program's noise
        r.get<int64_t>();
        r.u64();
    }
    if (reference_file) return import_reference_body(r, f, c, simd_width, spec, *nsteps_total, image);
    if (r.i32() != kBodyStart) return fail(f, "restart file corrupted");
    const uint64_t n = r.u64();
    if (r.eof() || n > (1ull << 40)) return fail(f, "restart file corrupted");
    image.resize(n);
    if (!r.bytes(image.data(), n)) return fail(f, "restart file corrupted");
    if (r.i32() != kRestartEnd || r.eof()) return fail(f, "restart file corrupted");
    fclose(f);
    return true;
}

}  // namespace ldpcp_host
