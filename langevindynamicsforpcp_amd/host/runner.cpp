// runner.cpp -- see runner.h.  Output formats follow
// langevinDynamicsSingleBarrier.h:3040-3124 (headers) and :3594-4036 (files).
#include "runner.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <time.h>

#include <algorithm>
#include <array>
#include <string.h>
#include <chrono>
#include <string>
#include <thread>
#include <vector>

#include "ldpcp.h"
#include "restart_file.h"
#include "text_util.h"

namespace ldpcp_host {

namespace {

// Rank handles integrating at once (one stream each).
constexpr int kMaxHandlesInFlight = 30;
// Upper bound on the hardware queues the programs ask the HIP runtime for.
constexpr int kMaxHwQueues = 32;

// Each rank's handle runs its kernels on a stream of its own.  The runtime
// gives a process GPU_MAX_HW_QUEUES hardware queues (4 by default); streams
// beyond that share queues and their kernels run by turns, so a run of more
// ranks than queues takes proportionally longer.  Before the first HIP call,
// raise the variable to the handles in flight plus two for the runtime's own
// copies (at most kMaxHwQueues), unless it is already that high or
// LDPCP_KEEP_HW_QUEUES is set.
void reserve_hw_queues(int handles)
{
    if (getenv("LDPCP_KEEP_HW_QUEUES")) return;
    const int want = std::min(kMaxHwQueues, handles + 2);
    const char* cur = getenv("GPU_MAX_HW_QUEUES");
    if (cur && atoi(cur) >= want) return;
    char buf[16];
    snprintf(buf, sizeof(buf), "%d", want);
    setenv("GPU_MAX_HW_QUEUES", buf, 1);
}

// HISTOGRAM::HISTOGRAM (clsHistogram.cpp:96-263): AddValue truncates
// (v - min) / step to int; HistogramVector never counts (its loop bounds are
// taken before `last = -1` is clamped, clsHistogram.cpp:242-254), so the
// barrier-height and waiting-time histograms the runner fills through it stay 0.
struct Histogram {
    double min = 0, max = 0, step = 1, step_inv = 1;
    int size = 0;
    std::vector<double> count, percent;
    void allocate(double lo, double hi, int n)
    {
        min = lo;
        max = hi;
        size = n;
        step = (hi - lo) / n;
        step_inv = 1.0 / step;
        count.assign(n, 0.0);
        percent.assign(n, 0.0);
    }
    void add(double v)
    {
        const int index = (int)((v - min) * step_inv);
        if (index >= 0 && index < size) count[index]++;
    }
    void normalize()
    {
        double sum = 0.0;
        for (int i = 0; i < size; ++i) sum += count[i];
        if (sum != 0.0)
            for (int i = 0; i < size; ++i) percent[i] = count[i] / sum;
    }
    void print(FILE* f) const
    {
        fprintf(f, "# HMin %23.16E\n", min);
        fprintf(f, "# HMax %23.16E\n", max);
        fprintf(f, "# HBin %23.16E\n", step);
        fprintf(f, "# NBin %23d\n", size);
        for (int i = 0; i < size; ++i)
            fprintf(f, "%23.16E %23.16E %23.16E %23.0f %23.16E\n", fma(step, (double)i, min),
                    fma(step, i + 0.5, min), fma(step, i + 1.0, min), count[i], percent[i]);
    }
};

// STATISTICS::AVERAGE / RMSD (STATISTICS.cpp:100-134); the squared terms are
// one statement each, contracted by the reference's compiler.
double average(const std::vector<double>& v)
{
    if (v.empty()) return 0.0;
    double s = 0.0;
    for (double x : v) s += x;
    return s / (double)v.size();
}

double rmsd(const std::vector<double>& v)
{
    if (v.empty()) return 0.0;
    const double a = average(v);
    double r = 0.0;
    for (double x : v) r = fma(x - a, x - a, r);
    return sqrt(r / (double)v.size());
}

double xc_of(const RunnerConfig& c, double* yc)
{
    // LangevinEquationSIMDMultiple::GetXc (langevinDynamicsDoubleBarrier.h:1722-1732)
    const double x1b = c.position[0], w1 = c.half_width[0], x2b = c.position[1], w2 = c.half_width[1];
    const double y1 = c.height[0][0], y2 = c.height[1][0];
    const double x1h = x1b + w1, x2l = x2b - w2;
    const double xc = fma(x1h * y1, x2b - x2l, x2l * y2 * (x1h - x1b)) / fma(y1, x2b - x2l, y2 * (x1h - x1b));
    *yc = (x1h - xc) / (x1h - x1b) * y1;
    return xc;
}

struct Header {
    const RunnerConfig* c;
    int nrun = 0;
    int seed_used = -1;
    double traj_sets = 0, memory = 0, save_dt = 0;

    // PrintParameter (Single.h:3040-3124; Double.h lines for [b][s] heights, Xc/Yc, positions)
    void print(FILE* f) const
    {
        const RunnerConfig& r = *c;
        const char* bar = "# ==============================================\n";
        fprintf(f, "%s", bar);
        fprintf(f, "# Overdamped           = %13s\n", bool_text(r.overdamped));
        fprintf(f, "%s", bar);
        fprintf(f, "# Number of node       = %13d\n", r.ranks);
        fprintf(f, "%s", bar);
        fprintf(f, "# Time step            = %13.7f\n", r.timestep);
        fprintf(f, "# Save frequency       = %13zd\n", (size_t)r.save_freq);
        fprintf(f, "# Save time step       = %13.7f\n", save_dt);
        fprintf(f, "# Statistics freq      = %13zd\n", (size_t)r.stat_freq);
        fprintf(f, "# Number of step (pre) = %13zd\n", (size_t)r.nsteps_previous);
        fprintf(f, "# Number of step       = %13zd\n", (size_t)r.nsteps);
        fprintf(f, "# Simulation time      = %13.6E\n", r.timestep * (double)r.nsteps);
        fprintf(f, "# Number of trial      = %13d\n", r.ntrial);
        fprintf(f, "# Number of run/trial  = %13d\n", r.simd_width);
        fprintf(f, "# Number of run        = %13d\n", nrun);
        fprintf(f, "%s", bar);
        if (r.model == 0) {
            fprintf(f, "# Barrier height [0]   = %13.7f\n", r.height[0][0]);
            fprintf(f, "# Barrier height [1]   = %13.7f\n", r.height[0][1]);
            fprintf(f, "# Barrier periodicity  = %13.7f\n", r.periodic_length);
            fprintf(f, "# slope half length    = %13.7f\n", r.half_width[0]);
        } else {
            double yc = 0.0;
            const double xc = xc_of(r, &yc);
            fprintf(f, "# Barrier height [0][0]= %13.7f\n", r.height[0][0]);
            fprintf(f, "# Barrier height [0][1]= %13.7f\n", r.height[0][1]);
            fprintf(f, "# Barrier height [1][0]= %13.7f\n", r.height[1][0]);
            fprintf(f, "# Barrier height [1][1]= %13.7f\n", r.height[1][1]);
            fprintf(f, "# Xc                   = %13.7f\n", xc);
            fprintf(f, "# Yc                   = %13.7f\n", yc);
            fprintf(f, "# Barrier periodicity  = %13.7f\n", r.periodic_length);
            fprintf(f, "# Slope half length[0] = %13.7f\n", r.half_width[0]);
            fprintf(f, "# Slope half length[1] = %13.7f\n", r.half_width[1]);
            fprintf(f, "# Barrier position [0] = %13.7f\n", r.position[0]);
            fprintf(f, "# Barrier position [1] = %13.7f\n", r.position[1]);
        }
        fprintf(f, "# Equal probability    = %13s\n", bool_text(r.equal_probability));
        fprintf(f, "# Equilibrium constant = %13.7E\n", r.eq_constant);
        fprintf(f, "%s", bar);
        fprintf(f, "# Mass                 = %13.7f\n", r.mass);
        fprintf(f, "# Gamma                = %13.7f\n", r.gamma);
        fprintf(f, "# Temperature          = %13.7f\n", r.temperature);
        fprintf(f, "# kB                   = %13.7f\n", r.kb);
        fprintf(f, "# Eta variance (Q)     = %13.7f\n", r.q);
        fprintf(f, "# Eta tau              = %13.7f\n", r.tau);
        fprintf(f, "# Sample per order     = %13d\n", r.sample_per_order);
        fprintf(f, "%s", bar);
        fprintf(f, "# X0                   = %13.7f\n", r.x0);
        if (r.randomize_v) fprintf(f, "# V0                   = %13s\n", "Random");
        else fprintf(f, "# V0                   = %13.7f\n", r.v0);
        fprintf(f, "# Eta0                 = %13s\n", "Random");
        if (r.randomize_state) fprintf(f, "# State0               = %13s\n", "Random");
        else fprintf(f, "# State0               = %13d\n", (int)r.initial_state);
        fprintf(f, "# Same initial barrier = %13s\n", bool_text(r.same_initial_barrier));
        fprintf(f, "%s", bar);
        fprintf(f, "# Random number seed0  = %13d\n", r.seed);
        fprintf(f, "# Random number seed   = %13d\n", seed_used);
        fprintf(f, "%s", bar);
        fprintf(f, "# Number of traj(each) = %13.0f\n", traj_sets);
        fprintf(f, "# Memory usage (each)  = %13s\n", bytes_readable(memory).c_str());
        fprintf(f, "# Memory usage (total) = %13s\n", bytes_readable(memory * r.ranks).c_str());
        fprintf(f, "%s", bar);
        fprintf(f, "# MSD                  = %13s\n", "Mandatory");
        fprintf(f, "# Waiting time         = %13s\n", bool_text(r.waiting));
        fprintf(f, "# Write trajectories   = %13s\n", bool_text(r.write_x));
        fprintf(f, "# V0 Distribution      = %13s\n", bool_text(r.v0_distribution));
        fprintf(f, "# Eta0 Distribution    = %13s\n", bool_text(false));
        fprintf(f, "# State0 Distribution  = %13s\n", bool_text(r.state0_distribution));
        fprintf(f, "# Eta result           = %13s\n", bool_text(false));
        fprintf(f, "# First passage time   = %13s\n", bool_text(r.fpt));
        if (r.fpt) {
            fprintf(f, "# Barrier hist hmin    = %13.7f\n", 0.0);
            fprintf(f, "# Barrier hist hmax    = %13.7f\n", 10.0);
            fprintf(f, "# Barrier hist nbin    = %13d\n", 100);
        }
        fprintf(f, "# Transtion path time  = %13s\n", bool_text(r.tpt));
        fprintf(f, "# First passage time CG= %13s\n", bool_text(r.fptcg));
        if (r.fptcg)
            for (size_t j = 0; j < r.cg_size.size(); ++j)
                fprintf(f, "# FFT CG size %4d     = %13d\n", (int)j, r.cg_size[j]);
        fprintf(f, "# X correlation        = %13s\n", bool_text(false));
        fprintf(f, "%s", bar);
        fprintf(f, "# Read restart         = %13s\n", bool_text(r.read_restart));
        fprintf(f, "# Write restart        = %13s\n", bool_text(r.write_restart));
        fprintf(f, "# Input restart (base) = %13s\n", r.input_restart.c_str());
        fprintf(f, "# Output restart (base)= %13s\n", r.output_restart.c_str());
        fprintf(f, "%s", bar);
    }
};

FILE* open_output(const std::string& name)
{
    backup_existing(name);
    FILE* f = fopen(name.c_str(), "w+");
    if (!f) fprintf(stderr, "Error: can't open output file (%s)\n", name.c_str());
    return f;
}

void write_dat(const std::string& name, const std::vector<double>& v)
{
    backup_existing(name);
    FILE* f = fopen(name.c_str(), "wb");
    if (!f) {
        fprintf(stderr, "Error: can't open output file (%s)\n", name.c_str());
        return;
    }
    if (!v.empty()) fwrite(v.data(), sizeof(double), v.size(), f);
    fclose(f);
}

void write_mean_log(const std::string& name, const Header& h, const std::vector<double>& v)
{
    FILE* f = open_output(name);
    if (!f) return;
    h.print(f);
    fprintf(f, "# average %23.16E\n", average(v));
    fprintf(f, "# sd      %23.16E\n", rmsd(v));
    fprintf(f, "# count   %23zd\n", v.size());
    fclose(f);
}

[[noreturn]] void die(const char* what)
{
    fprintf(stderr, "Error: %s: %s\n", what, ldpcp_last_error());
    exit(1);
}

}  // namespace

// Phase times of the integrate and restart path, on stderr when LDPCP_TIMING is set.
void phase_time(const char* what, int rank, std::chrono::steady_clock::time_point t0,
                std::chrono::steady_clock::time_point t1)
{
    static const bool on = getenv("LDPCP_TIMING") != nullptr;
    if (!on) return;
    fprintf(stderr, "ldpcp timing: rank %d %-26s %10.3f s\n", rank, what,
            std::chrono::duration<double>(t1 - t0).count());
}

std::vector<int> rank_seeds(int master_seed, int ranks, int* used_master)
{
    const int value = master_seed == -1 ? (int)time(nullptr) : master_seed;
    srand(value);
    if (used_master) *used_master = value;
    std::vector<int> out;
    while ((int)out.size() < ranks) {
        const int v = rand() % RAND_MAX;
        if (std::find(out.begin(), out.end(), v) == out.end()) out.push_back(v);
    }
    return out;
}

std::vector<unsigned long> log10_scale(double timestep, unsigned long n_frames, unsigned long minimal,
                                       unsigned long per_order, bool include_zero)
{
    const double step = 1.0 / per_order;
    const double max_time = timestep * (double)(n_frames - minimal);
    const double t0 = log10(timestep), t1 = log10(max_time);
    std::vector<unsigned long> out;
    if (include_zero) out.push_back(0);
    for (double t = t0; t <= t1; t += step) {
        const unsigned long iv = (unsigned long)round(pow(10.0, t) / timestep);
        if (out.empty() || out.back() != iv) out.push_back(iv);
    }
    return out;
}

int run(const RunnerConfig& cfg)
{
    using clock = std::chrono::steady_clock;
    const auto t_start = clock::now();
    fprintf(stderr, "LangevinEquationMultipleRunner\n");
    const int W = cfg.simd_width, R = cfg.ranks;
    const unsigned long n_frames = (cfg.nsteps + cfg.nsteps_previous) / cfg.save_freq + 1;
    if (n_frames < 1000) {
        fprintf(stderr, "Error: numberOfFrame < 1000\n");
        return 0;
    }
    Header hd;
    hd.c = &cfg;
    hd.nrun = cfg.ntrial * W * R;
    const std::vector<int> seeds = rank_seeds(cfg.seed, R, nullptr);
    hd.seed_used = seeds[0];
    hd.traj_sets = cfg.write_x ? 7 : 1;
    hd.memory = hd.traj_sets * (double)(cfg.nsteps / cfg.save_freq + 1) * sizeof(double) * W;
    hd.save_dt = cfg.timestep * (double)cfg.save_freq;
    hd.print(stderr);

    const std::vector<unsigned long> interval = log10_scale(hd.save_dt, n_frames, 1000, cfg.sample_per_order, true);
    if (interval.size() < 20) {
        fprintf(stderr, "Error: interval is less than 20 (%zd)\n", interval.size());
        exit(0);
    } else if (interval.size() < 100) {
        fprintf(stderr, "Warning: interval is less than 100 (%zd)\n", interval.size());
    } else {
        fprintf(stderr, "Interval = %10zd\n", interval.size());
    }

    const size_t nl = interval.size();
    const size_t ncg = cfg.fptcg ? cfg.cg_size.size() : 0;
    // per emulated rank, as each MPI rank accumulates them (Single.h:3339-3414)
    std::vector<std::vector<double>> msd(R, std::vector<double>(nl, 0.0));
    std::vector<std::array<double, 2>> state_count(R, {0.0, 0.0});
    std::vector<std::vector<double>> fpt(R), fpt2(R), tpt(R), tpt2(R);
    std::vector<std::vector<std::vector<double>>> cg(ncg, std::vector<std::vector<double>>(R));
    std::vector<double> wt_sum(R, 0.0), wt_sum2(R, 0.0), wt_max(R, 0.0), wt_n(R, 0.0);
    std::vector<Histogram> h_v0(R), h_s0(R);
    const double vsigma = sqrt(cfg.kb * cfg.temperature / cfg.mass);
    for (int r = 0; r < R; ++r) {
        if (cfg.v0_distribution) h_v0[r].allocate(-vsigma * 5.0, vsigma * 5.0, 200);
        if (cfg.state0_distribution) h_s0[r].allocate(0.0, 2.0, 2);
    }
    Histogram h_init, h_inst;
    if (cfg.fpt) {
        h_init.allocate(0.0, 10.0, 100);
        h_inst.allocate(0.0, 10.0, 100);
    }
    const uint64_t tick_cap = cfg.waiting ? std::min<uint64_t>(1u << 16, (1ull << 27) / (uint64_t)W + 1) : 0;
    bool wt_exact = true;
    std::vector<std::vector<double>> wt_list(R);
    double t_integrate = 0.0, t_analysis = 0.0;

    ldpcp_run_params p;
    memset(&p, 0, sizeof(p));
    p.model = cfg.model;
    p.overdamped = cfg.overdamped;
    p.x0 = cfg.x0;
    p.v0 = cfg.v0;
    p.gamma = cfg.gamma;
    p.kb = cfg.kb;
    p.temperature = cfg.temperature;
    p.mass = cfg.mass;
    p.randomize_v = cfg.randomize_v;
    p.randomize_state = cfg.randomize_state;
    p.same_initial_barrier = cfg.same_initial_barrier;
    p.periodic_length = cfg.periodic_length;
    for (int b = 0; b < 2; ++b) {
        for (int s = 0; s < 2; ++s) p.height[b][s] = cfg.height[b][s];
        p.half_width[b] = cfg.half_width[b];
        p.position[b] = cfg.position[b];
    }
    p.timestep = cfg.timestep;
    p.nsteps = cfg.nsteps;
    p.save_freq = cfg.save_freq;
    p.tau = cfg.tau;
    p.initial_state = cfg.initial_state;
    p.equal_probability = cfg.equal_probability;
    p.eq_constant = cfg.eq_constant;
    p.b_waiting = cfg.waiting;
    p.b_fpt = cfg.fpt;
    p.b_tpt = cfg.tpt;
    p.b_fptcg = cfg.fptcg;
    p.n_cg = (int)ncg;
    for (size_t j = 0; j < ncg && j < LDPCP_MAX_CG; ++j) p.cg_size[j] = cfg.cg_size[j];
    p.save_all = cfg.write_x;
    p.barrier_window = 4096;
    p.tick_capacity = tick_cap;
    // restart files in the reference's layout need the raw event lists: a
    // trajectory keeps up to 256 MB / (W x lists) doubles per list (1M at most)
    const uint32_t n_lists = LDPCP_LIST_CG0 + (cfg.fptcg ? (uint32_t)ncg : 0u);
    p.event_capacity =
        cfg.write_restart ? std::min<uint64_t>(1u << 20, (256ull << 20) / ((uint64_t)W * n_lists * sizeof(double))) : 0;

    std::vector<ldpcp_lane> lanes(W);
    std::vector<double> msd0((size_t)W * nl);
    std::vector<double> ticks(cfg.waiting ? (size_t)W * tick_cap : 0);
    std::vector<double> frames;

    // each rank's lanes' initial barrier states, carried into the next trial as
    // UserData::m_bInitialValue (Single.h:1378-1380)
    std::vector<std::vector<double>> state0_prev(R, std::vector<double>(W, 0.0));
    std::vector<char> image;

    // Trials are independent unless UNIFORMINITALSTATES carries each lane's
    // initial state into the next one, so up to 30 handles (ranks x trials) are
    // launched before any is waited for; the post analysis then runs in the
    // reference's order, trial by trial and rank by rank.
    const int batch = cfg.same_initial_barrier ? 1 : std::max(1, kMaxHandlesInFlight / R);
    reserve_hw_queues(std::min(cfg.ntrial, batch) * R);
    for (int i0 = 0; i0 < cfg.ntrial; i0 += batch) {
        const int i1 = std::min(cfg.ntrial, i0 + batch);
        const auto t0 = clock::now();
        std::vector<ldpcp_run*> handles((size_t)(i1 - i0) * R, nullptr);
        std::vector<uint64_t> before(handles.size(), 0);
        for (int i = i0; i < i1; ++i)
            for (int r = 0; r < R; ++r) {
                p.seed = (uint64_t)(uint32_t)seeds[r];
                ldpcp_run* h = nullptr;
                if (ldpcp_run_create(&p, (uint32_t)W, (uint32_t)((uint64_t)i * W), &h)) die("ldpcp_run_create");
                if (cfg.same_initial_barrier && ldpcp_run_set_barrier_default(h, state0_prev[r].data()))
                    die("ldpcp_run_set_barrier_default");
                uint64_t nsteps_before = 0;  // m_NumberOfStepDonePreviously
                if (cfg.read_restart) {
                    const auto tr0 = clock::now();
                    // PreIntegrate -> ReadRestartFile (Single.h:1346-1351); a failure ends the
                    // run as the runner's integrator-failure check does (:3436-3446)
                    ImageSpec spec;  // what a reference program's file must become for this handle
                    spec.seed = p.seed;
                    spec.traj_offset = (uint32_t)((uint64_t)i * W);
                    spec.barrier_window = p.barrier_window;
                    spec.n_fields = cfg.write_x ? 7 : 1;
                    spec.tick_capacity = p.tick_capacity;
                    spec.event_capacity = p.event_capacity;
                    spec.n_lists = p.event_capacity ? n_lists : 0;
                    bool ok = read_restart_file(cfg, W, r, i, spec, image, &nsteps_before);
                    const auto tr1 = clock::now();
                    if (ok && ldpcp_run_resume(h, image.data(), image.size())) {
                        fprintf(stderr, "Error: %s\n", ldpcp_last_error());
                        ok = false;
                    }
                    phase_time("read restart file", r, tr0, tr1);
                    phase_time("resume (image to device)", r, tr1, clock::now());
                    if (!ok) {
                        fprintf(stderr, "Error: can't read restart file on rank %10d\n", r);
                        fprintf(stderr, "Error: Integrator failure\n");
                        ldpcp_run_destroy(h);
                        for (ldpcp_run* q : handles) ldpcp_run_destroy(q);  // waits for their runs
                        exit(0);
                    }
                }
                handles[(size_t)(i - i0) * R + r] = h;
                before[(size_t)(i - i0) * R + r] = nsteps_before;
            }
        // Each handle is launched from a thread of its own: a launch enqueues its
        // run's kernels on the handle's stream and waits while the stream's hardware
        // queue is full, so launches made one after another from one thread would
        // run a long trial's ranks one after another.
        std::vector<std::string> launch_error(handles.size());
        {
            std::vector<std::thread> launchers;
            for (size_t k = 0; k < handles.size(); ++k)
                launchers.emplace_back([&, k] {
                    const auto tl0 = clock::now();
                    if (ldpcp_run_launch(handles[k], cfg.steps_per_launch)) launch_error[k] = ldpcp_last_error();
                    phase_time("launch (enqueue)", (int)(k % R), tl0, clock::now());
                });
            for (std::thread& th : launchers) th.join();
        }
        for (const std::string& msg : launch_error)
            if (!msg.empty()) {
                fprintf(stderr, "Error: ldpcp_run_launch: %s\n", msg.c_str());
                exit(1);
            }
        for (size_t k = 0; k < handles.size(); ++k) {
            const auto tw0 = clock::now();
            if (ldpcp_run_wait(handles[k])) die("ldpcp_run_integrate");
            phase_time("wait", (int)(k % R), tw0, clock::now());
        }
        const auto t1 = clock::now();
        t_integrate += std::chrono::duration<double>(t1 - t0).count();
        for (int i = i0; i < i1; ++i) {
            fprintf(stderr, "##################################################################\n");
            fprintf(stderr, "Trial %10d/%10d\n", i, cfg.ntrial);
            fprintf(stderr, "##################################################################\n");
            fprintf(stderr, "   Integrate\n");
            for (int r = 0; r < R; ++r) {
                ldpcp_run* h = handles[(size_t)(i - i0) * R + r];
                const uint64_t nsteps_before = before[(size_t)(i - i0) * R + r];
                if (cfg.write_restart) {
                    const auto tc0 = clock::now();
                    uint64_t bytes = 0;
                    if (ldpcp_run_checkpoint_size(h, &bytes)) die("ldpcp_run_checkpoint_size");
                    image.resize(bytes);
                    if (ldpcp_run_checkpoint(h, image.data(), bytes)) die("ldpcp_run_checkpoint");
                    bool ok;
                    if (reference_layout_possible(cfg, image)) {
                        // each barrier table spans every barrier a trajectory touched and the
                        // ids the reference creates at step 0
                        std::vector<BarrierTable> tables(cfg.model != 0 ? 2 : 1);
                        for (size_t a = 0; a < tables.size(); ++a) {
                            BarrierTable& T = tables[a];
                            int lo = -step0_barrier_reach(cfg), hi = step0_barrier_reach(cfg);
                            barrier_span(image, (int)a, &lo, &hi);
                            T.lo = lo;
                            T.n = hi - lo + 1;
                            T.states.resize((size_t)W * T.n);
                            T.steps.resize((size_t)W * T.n);
                            if (ldpcp_run_barrier_table(h, (int32_t)a, lo, hi, T.states.data(), T.steps.data()))
                                die("ldpcp_run_barrier_table");
                        }
                        ok = write_reference_restart_file(cfg, W, r, i, nsteps_before + cfg.nsteps, image, tables);
                    } else {
                        ok = write_restart_file(cfg, W, r, i, nsteps_before + cfg.nsteps, image);
                    }
                    if (!ok) fprintf(stderr, "Error: can't open restart file on rank %10d\n", r);
                    phase_time("checkpoint + restart file", r, tc0, clock::now());
                }
                const auto t2 = clock::now();
                if (ldpcp_run_lanes(h, 0, (uint32_t)W, lanes.data())) die("ldpcp_run_lanes");
                if (ldpcp_run_msd(h, interval.data(), (uint32_t)nl, msd0.data())) die("ldpcp_run_msd");
                if (cfg.waiting && ldpcp_run_ticks(h, 0, (uint32_t)W, ticks.data())) die("ldpcp_run_ticks");
                // frames of this run, after a restart's frames when it continues one
                const uint64_t nf = ldpcp_run_n_frames(h);
                if (cfg.write_x) frames.resize((size_t)7 * W * nf);
                if (cfg.write_x)
                    for (int field = 0; field < 7; ++field)
                        if (ldpcp_run_frames(h, field, 0, (uint32_t)W, frames.data() + (size_t)field * W * nf))
                            die("ldpcp_run_frames");
                ldpcp_run_destroy(h);
                if (r == 0) fprintf(stderr, "   Post Analysis\n");
                // per-lane post analysis (Single.h:3452-3580)
                for (int ch = 0; ch < W; ++ch) {
                    const ldpcp_lane& L = lanes[ch];
                    for (size_t t = 0; t < nl; ++t) {
                        const double v = msd0[(size_t)ch * nl + t];
                        msd[r][t] += v;
                        if (isnan(v) || isinf(v)) fprintf(stderr, "Error: on node %10d, ill msd0\n", r);
                    }
                    if (cfg.waiting) {
                        if (L.wt_ticks <= tick_cap) {
                            const double* tk = ticks.data() + (size_t)ch * tick_cap;
                            for (uint64_t t = 1; t < L.wt_ticks; ++t)
                                wt_list[r].push_back((tk[t] - tk[t - 1]) * cfg.timestep);
                        } else {
                            wt_exact = false;
                        }
                        wt_sum[r] += L.wt_sum;
                        wt_sum2[r] += L.wt_sum2;
                        wt_n[r] += (double)L.wt_n;
                        if (L.wt_n > 0) wt_max[r] = std::max(wt_max[r], L.wt_max);
                    }
                    state0_prev[r][ch] = L.state0_instance;
                    if (cfg.v0_distribution) h_v0[r].add(L.v0_instance);
                    if (cfg.state0_distribution) h_s0[r].add(L.state0_instance);
                    if (cfg.write_x) {
                        char buf[512];
                        snprintf(buf, sizeof(buf), "%s_trajectory_%010d_%010d_%010d.log", cfg.output_base.c_str(), i,
                                 ch, r);
                        FILE* f = open_output(buf);
                        if (f) {
                            hd.print(f);
                            auto fr = [&](int field, uint64_t t) { return frames[((size_t)field * W + ch) * nf + t]; };
                            for (uint64_t t = 0; t < nf; ++t)
                                fprintf(f, " %23.16f %13.7f %13.7f %13.7f %13.7f %13.7f %13.7f %13.7f\n",
                                        hd.save_dt * (double)t, fr(0, t), cfg.overdamped ? 0.0 : fr(1, t), fr(2, t),
                                        fr(6, t), fr(3, t), cfg.overdamped ? 0.0 : fr(4, t), fr(5, t));
                            if (cfg.waiting) {
                                fprintf(f, "# event tick\n");
                                const uint64_t n = std::min<uint64_t>(L.wt_ticks, tick_cap);
                                for (uint64_t t = 0; t < n; ++t)
                                    fprintf(f, "# %10zd %23.16E\n", (size_t)t,
                                            ticks[(size_t)ch * tick_cap + t] * cfg.timestep);
                            }
                            fclose(f);
                        }
                    }
                    if (cfg.fpt) {
                        fpt[r].push_back(L.fpt_n ? L.fpt_sum / (double)L.fpt_n : 0.0);
                        fpt2[r].push_back(L.fpt2_n ? L.fpt2_sum / (double)L.fpt2_n : 0.0);
                    }
                    if (cfg.tpt) {
                        tpt[r].push_back(L.tpt_n ? L.tpt_sum / (double)L.tpt_n : 0.0);
                        tpt2[r].push_back(L.tpt2_n ? L.tpt2_sum / (double)L.tpt2_n : 0.0);
                    }
                    for (size_t j = 0; j < ncg; ++j)
                        cg[j][r].push_back(L.cg_n[j] ? L.cg_sum[j] / (double)L.cg_n[j] : 0.0);
                    state_count[r][0] += L.state_count[0];
                    state_count[r][1] += L.state_count[1];
                }
                t_analysis += std::chrono::duration<double>(clock::now() - t2).count();
            }
            fprintf(stderr, "   Done %10d/%10d\n", i + 1, cfg.ntrial);
        }
    }
    const auto t_post0 = clock::now();
    const std::string& ob = cfg.output_base;
    auto combined = [&](const std::vector<std::vector<double>>& per_rank) {
        std::vector<double> out;  // MPIKernel::CombineVectorDoubleToMaster: rank-major
        for (const auto& v : per_rank) out.insert(out.end(), v.begin(), v.end());
        return out;
    };
    if (cfg.fpt) {
        const auto a = combined(fpt), b = combined(fpt2);
        write_dat(ob + "_firstPassageTime.dat", a);
        write_dat(ob + "_firstPassageTime2.dat", b);
        h_init.normalize();
        h_inst.normalize();
        write_mean_log(ob + "_firstPassageTime.log", hd, a);
        write_mean_log(ob + "_firstPassageTime2.log", hd, b);
        for (int k = 0; k < 2; ++k) {
            FILE* f = open_output(ob + (k ? "_instantBarrierHeightHistogram.log" : "_initialBarrierHeightHistogram.log"));
            if (f) {
                hd.print(f);
                (k ? h_inst : h_init).print(f);
                fclose(f);
            }
        }
    }
    {
        double s0 = 0.0, s1 = 0.0;
        for (int r = 0; r < R; ++r) {
            s0 += state_count[r][0];
            s1 += state_count[r][1];
        }
        FILE* f = open_output(ob + "_stateCount.log");
        if (f) {
            hd.print(f);
            fprintf(f, "# state 0 %23.0f %13.7f\n", s0, s0 / (s0 + s1));
            fprintf(f, "# state 1 %23.0f %13.7f\n", s1, s1 / (s0 + s1));
            fclose(f);
        }
    }
    if (cfg.tpt) {
        const auto a = combined(tpt), b = combined(tpt2);
        write_dat(ob + "_transitionPathTime.dat", a);
        write_dat(ob + "_transitionPathTime2.dat", b);
        write_mean_log(ob + "_transitionPathTime.log", hd, a);
        write_mean_log(ob + "_transitionPathTime2.log", hd, b);
    }
    if (cfg.fptcg) {
        std::vector<std::vector<double>> all(ncg);
        for (size_t j = 0; j < ncg; ++j) {
            all[j] = combined(cg[j]);
            char buf[512];
            snprintf(buf, sizeof(buf), "%s_firstPassageTimeCG2_%04d_%04d.dat", ob.c_str(), (int)j, cfg.cg_size[j]);
            write_dat(buf, all[j]);
        }
        FILE* f = open_output(ob + "_firstPassageTimeCG2.log");
        if (f) {
            hd.print(f);
            for (size_t j = 0; j < ncg; ++j) {
                fprintf(f, "# %4d %4d\n", (int)j, cfg.cg_size[j]);
                fprintf(f, "# average %23.16E\n", average(all[j]));
                fprintf(f, "# sd      %23.16E\n", rmsd(all[j]));
                fprintf(f, "# count   %23zd\n", all[j].size());
            }
            fclose(f);
        }
    }
    auto sum_hist = [&](std::vector<Histogram>& hs) {
        Histogram out = hs[0];  // MPIKernel::SumVectorDouble into the master, rank by rank
        for (int r = 1; r < R; ++r)
            for (int k = 0; k < out.size; ++k) out.count[k] += hs[r].count[k];
        out.normalize();
        return out;
    };
    if (cfg.v0_distribution) {
        Histogram h = sum_hist(h_v0);
        FILE* f = open_output(ob + "_V0Distribution.log");
        if (f) {
            hd.print(f);
            h.print(f);
            fclose(f);
        }
    }
    if (cfg.state0_distribution) {
        Histogram h = sum_hist(h_s0);
        FILE* f = open_output(ob + "_State0Distribution.log");
        if (f) {
            hd.print(f);
            h.print(f);
            fclose(f);
        }
    }
    {
        std::vector<double> total = msd[0];
        for (int r = 1; r < R; ++r)
            for (size_t t = 0; t < nl; ++t) total[t] += msd[r][t];
        FILE* f = open_output(ob + "_msd.log");
        if (f) {
            hd.print(f);
            for (size_t t = 0; t < nl; ++t) {
                const double v = total[t] / (double)hd.nrun;
                if (cfg.model == 0) fprintf(f, "%14.7E %14.7E\n", hd.save_dt * (double)interval[t], v);
                else fprintf(f, "%13.7f %13.7f\n", hd.save_dt * (double)interval[t], v);
            }
            fclose(f);
        }
    }
    if (cfg.waiting) {
        double sum = 0.0, sum2 = 0.0, n = 0.0, mx = 0.0;
        for (int r = 0; r < R; ++r) {
            if (wt_exact) {
                // STATISTICS::SUM / SUM2 / MAX over the rank's list, then the all-reduce
                double s = 0.0, s2 = 0.0, m = 0.0;
                for (double w : wt_list[r]) {
                    s += w;
                    s2 = fma(w, w, s2);
                }
                if (!wt_list[r].empty()) {
                    m = -1.7976931348623157e308;
                    for (double w : wt_list[r]) m = std::max(w, m);
                }
                sum += s;
                sum2 += s2;
                n += (double)wt_list[r].size();
                mx = std::max(mx, m);
            } else {
                sum += wt_sum[r];
                sum2 += wt_sum2[r];
                n += wt_n[r];
                mx = std::max(mx, wt_max[r]);
            }
        }
        const double avg = sum / n, avg2 = sum2 / n;
        const double Rv = fma(-avg, avg, avg2) / avg / avg;
        int nbin = 100;
        if (n < 1000) nbin = std::max<int>(10, (int)n / 100);
        Histogram h;
        h.allocate(0.0, mx + 0.1, nbin);
        h.normalize();
        FILE* f = open_output(ob + "_waitingTime.log");
        if (f) {
            hd.print(f);
            fprintf(f, "# R      = %13.7f\n", Rv);
            fprintf(f, "# Nwt    = %13zd\n", (size_t)n);
            h.print(f);
            fclose(f);
        }
    }
    const double t_post = std::chrono::duration<double>(clock::now() - t_post0).count();
    const double t_total = std::chrono::duration<double>(clock::now() - t_start).count();
    {
        FILE* f = open_output(ob + "_time.log");
        if (f) {
            for (int r = 0; r < R; ++r) fprintf(f, "Rank %10d Seed %10d\n", r, seeds[r]);
            fprintf(f, "Time consumed %s %s\n", time_readable(t_total).c_str(), time_readable(0.0).c_str());
            fprintf(f, "Integrate     %s %s\n", time_readable(t_integrate).c_str(), time_readable(0.0).c_str());
            fprintf(f, "Analysis      %s %s\n", time_readable(t_analysis).c_str(), time_readable(0.0).c_str());
            fprintf(f, "Post          %s %s\n", time_readable(t_post).c_str(), time_readable(0.0).c_str());
            fclose(f);
        }
    }
    return 0;
}

}  // namespace ldpcp_host
