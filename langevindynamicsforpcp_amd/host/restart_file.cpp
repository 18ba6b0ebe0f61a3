// restart_file.cpp -- see restart_file.h.
#include "restart_file.h"

#include <math.h>
#include <stdio.h>
#include <string.h>

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

private:
    FILE* f_;
    bool eof_ = false;
};

// The saved-field flags the runner sets on the integrator (Single.h:3261-3271):
// x always; v, f, e, state and barrier with WRITEX; eta never (no input key).
struct SaveFlags {
    int x, v, f, e, eta, state, barrier;
};

SaveFlags save_flags(const RunnerConfig& c)
{
    const int w = c.write_x ? 1 : 0;
    return {1, w, w, w, 0, w, w};
}

// m_Xc0 of the double barrier (Double.h:1135-1146): the apex of the two
// inner slopes, written to the header and not checked on read.
double apex_x(const RunnerConfig& c)
{
    const double x1b = c.position[0], w1 = c.half_width[0], x2b = c.position[1], w2 = c.half_width[1];
    const double y1 = c.height[0][0], y2 = c.height[1][0];
    const double x1h = x1b + w1, x2l = x2b - w2;
    return fma(x1h * y1, x2b - x2l, x2l * y2 * (x1h - x1b)) / fma(y1, x2b - x2l, y2 * (x1h - x1b));
}

// ReadRestartFile's error exit: message, close, false (Single.h:1143-1190)
bool fail(FILE* f, const char* what)
{
    fprintf(stderr, "Error: %s\n", what);
    fclose(f);
    return false;
}

}  // namespace

std::string restart_file_name(const std::string& base, int rank, int trial)
{
    char buf[64];
    snprintf(buf, sizeof(buf), "_%05d_%05d.restart", rank, trial);
    return base + buf;
}

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
    const SaveFlags sf = save_flags(c);
    w.i32(kRestartSignature);
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
    w.i32(0);  // RandomizeEta: no input key, always false
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
    // noise positions: the Philox counter of the step noise is the step itself;
    // the scalar catch-up positions live per trajectory in the image
    w.put<int64_t>((int64_t)nsteps_total);
    w.u64(0);
    w.put<int64_t>((int64_t)nsteps_total);
    w.u64(0);
    w.put<int64_t>(0);
    w.u64(0);
    w.i32(kBodyStart);
    w.u64(image.size());
    w.bytes(image.data(), image.size());
    w.i32(kRestartEnd);
    const bool ok = w.ok();
    if (fclose(f) != 0 || !ok) {
        fprintf(stderr, "Error: can't open restart file (%s)\n", name.c_str());
        return false;
    }
    return true;
}

bool read_restart_file(const RunnerConfig& c, int simd_width, int rank, int trial, std::vector<char>& image,
                       uint64_t* nsteps_total)
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
    if (r.i32() != kRestartSignature) return fail(f, "inconsistent signature");
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
    if (r.i32() != 0) return fail(f, "inconsistent randomize eta");
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
    for (int k = 0; k < 3; ++k) {
        r.get<int64_t>();
        r.u64();
    }
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
