// ldpcp_engine.hip -- batched first-passage engine (the throughput path).
//
// Every trajectory of every parameter set starts at its set's X0 and is
// integrated until its first passage event or its step cap:
//   single barrier: first barrier-top crossing (the reference's first FPT1
//                   event, langevinDynamicsSingleBarrier.h:2394-2411),
//   double barrier: first arrival at a non-adjacent well centre (the
//                   reference's first FPT2 event, :2445-2489, the statistic
//                   the double-barrier code keeps, langevinDynamicsDoubleBarrier.h:2974-3006).
// Before that event a trajectory can only sit on two periodic cells, so its
// lazily created barrier states (the reference's BarrierStatus table,
// langevinDynamicsSingleBarrier.h:66-118) fit in four register slots.
//
// Work layout.  Trajectory state is SoA in a slot pool in HBM; a resident set
// owns a contiguous slot range.  The live list holds slot indices, each set's
// entries contiguous (sets are appended on admission, compaction keeps the
// order).  One persistent kernel per launch: every wavefront takes 64-entry
// chunks of the live list from an atomic counter and integrates each lane's
// trajectory for K steps with its state in VGPRs.  A wavefront covering one
// set (all but those at set boundaries) moves the set's constants and the
// Philox key to scalar registers; a boundary wavefront runs once per set it
// covers.  Each set has its own step origin (the engine step at admission), so
// requests join between any two launches and their noise -- keyed by
// (seed, trajectory, step) -- does not depend on when they joined or what ran
// beside them.  A scan and a scatter kernel then compact the live list
// (order preserving).  All three kernels read their sizes from device control
// words, so a launch is one replay of a captured hipGraph.
//
// Two arithmetics share the same Philox noise (ldpcp_device.h):
//   FP64: the reference's operations on the absolute position (plain FP64);
//   FP32: position as (integer cell, FP32 offset from the cell's barrier)
//         with a Kahan-compensated offset update, FP32 force, noise and
//         velocity, integer clock; only the statistics sums are FP64.
#include <hip/hip_runtime.h>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <vector>

#include "ldpcp_common.hpp"
#include "ldpcp_device.h"
#include "ldpcp_engine.hpp"

namespace ldpcp {

constexpr int kMaxBlock = 256;
constexpr uint32_t kNoLag = 0xffffffffu;

// per-set constants (FP64 and FP32 copies of IntegrateInitializeVariables,
// langevinDynamicsSingleBarrier.h:1520-1618, langevinDynamicsDoubleBarrier.h:1733-1937)
struct SetDev {
    int32_t model, overdamped, target, equal_prob, randomize_state, randomize_v, tau_zero, kind;
    double dt, L, Linv, halfL, w, S[2], H[2], dtgm, var, minv, halfdt, expmgdt, c_lap, K, p0, p_create, vsigma, v0;
    double xb[2], lo[2], hi[2], xc, slope[2][2][2], icept[2][2][2];
    double p_flip, p_asym[2];
    double x0;
    float fL, fhalfL, fw, fS[2], fdtgm, fvar, fminv_halfdt, fhalfdt, fexpmgdt, fc_lap, fK;
    float fxb[2], flo[2], fhi[2], fxc, fslope[2][2][2];
    float fxp0, fbm;                                // fbm = -2 ln2 var^2 (scaled FP32 Box-Muller)
    uint64_t T_flip, T_asym[2], T_create, T_p0le;  // integer thresholds on Philox words
    uint32_t Tl[2];                                 // per-step flip thresholds leaving state 0 / 1, below 2^32
    int32_t Tall[2];                                // 1: that threshold is 2^32 (always flip)
    int32_t fixed_state, cell0;                     // initial state when not randomised; cell of X0
    uint32_t key0, key1;
    int32_t slot_base, well0;                       // first of the two cells (slot 0)
    uint64_t n_traj, max_steps, first, start;       // first slot; engine step at admission
    double hist_lo, hist_inv;                       // log10 histogram edges
    uint32_t hist_edge[LDPCP_HIST_BINS];            // first absorption step of each bin (0xffffffff: past the cap)
    uint32_t lag[LDPCP_MSD_LAGS];                   // MSD lag steps, ascending, kNoLag past the cap
};

namespace {

// flags layout
constexpr uint32_t F_STATE = 1u << 0;
constexpr uint32_t F_ALIVE = 1u << 1;
constexpr uint32_t F_SLOT_VALID = 1u << 4;  // + slot (4 slots)
constexpr uint32_t F_SLOT_STATE = 1u << 8;  // + slot

__device__ __forceinline__ int slot_index(int model, int cell, int slot_base, int local)
{
    const int k = cell - slot_base;
    if (k < 0 || k > 1) return -1;
    return model == 0 ? k : k * 2 + local;
}

__device__ __forceinline__ float unit_f32(uint32_t w)
{
    return __uint2float_rn(w) * 2.3283064365386963e-10f + 1.16415321826934814e-10f;
}

// FP32 Box-Muller with the hardware transcendental units: v_log_f32 (log2),
// v_sqrt_f32 and v_sin_f32 / v_cos_f32 (arguments in revolutions).
__device__ __forceinline__ void box_muller_f32(uint32_t a, uint32_t b, float& g0, float& g1)
{
    const float u1 = unit_f32(a);
    const float u2 = unit_f32(b);
    const float r = __builtin_amdgcn_sqrtf(-1.3862943611198906f * __builtin_amdgcn_logf(u1));
    g0 = r * __builtin_amdgcn_cosf(u2);
    g1 = r * __builtin_amdgcn_sinf(u2);
}

// The step loop's pair of Gaussians already multiplied by the noise amplitude
// s: r = sqrt(c log2 u1) with c = -2 ln2 s^2 (one multiply by a scalar), and
// u = w 2^-32 + 2^-33 in one fused multiply-add.
__device__ __forceinline__ void box_muller_f32_scaled(uint32_t a, uint32_t b, float c, float& g0, float& g1)
{
    const float u1 = __builtin_fmaf(__uint2float_rn(a), 2.3283064365386963e-10f, 1.16415321826934814e-10f);
    const float u2 = __builtin_fmaf(__uint2float_rn(b), 2.3283064365386963e-10f, 1.16415321826934814e-10f);
    const float r = __builtin_amdgcn_sqrtf(c * __builtin_amdgcn_logf(u1));
    g0 = r * __builtin_amdgcn_cosf(u2);
    g1 = r * __builtin_amdgcn_sinf(u2);
}

// A wave-uniform value copied to a vector register: a per-lane select between
// it and a scalar is then one v_cndmask (a select reads at most one scalar
// operand; two scalars cost two extra moves on every use).
__device__ __forceinline__ float in_vgpr(float x)
{
    float v;
    asm volatile("v_mov_b32 %0, %1" : "=v"(v) : "s"(x));
    return v;
}

__device__ __forceinline__ bool below(uint32_t w, uint64_t T) { return (uint64_t)w < T; }

// Wave-uniform copy of a set: the fields the step loop touches every step are
// moved to scalar registers, so Philox's round keys and the integrator's
// constants cost no VGPRs and their updates run on the scalar unit.
__device__ __forceinline__ uint32_t uni(uint32_t v) { return (uint32_t)__builtin_amdgcn_readfirstlane((int)v); }
__device__ __forceinline__ int32_t uni(int32_t v) { return __builtin_amdgcn_readfirstlane(v); }
__device__ __forceinline__ float uni(float v)
{
    return __int_as_float(__builtin_amdgcn_readfirstlane(__float_as_int(v)));
}
__device__ __forceinline__ uint64_t uni(uint64_t v)
{
    return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v);
}
__device__ __forceinline__ double uni(double v)
{
    return __longlong_as_double((long long)uni((uint64_t)__double_as_longlong(v)));
}

// The constants the step loop reads every step, loaded field by field into
// scalar registers (the set is wave-uniform wherever a body runs), so Philox's
// round keys and the integrator's constants cost no VGPRs and their updates run
// on the scalar unit.  Rarely used fields are read through the set pointer.
struct Hot {
    int32_t target, equal_prob, tau_zero, slot_base;
    uint32_t key0, key1, first, max_steps;
    uint32_t Tl0, Tl1;    // per-step flip thresholds for leaving state 0 / 1 ...
    int32_t all0, all1;   // ... or 1 when the threshold is 2^32 (always flip)
    // FP32 engine
    float fL, fhalfL, fdtgm, fminv_halfdt, fhalfdt, fexpmgdt, fvar, fbm, fw, fS0, fS1, fxc;
    float fxb[2], flo[2], fhi[2], fsl[2][2][2];  // fsl[barrier][state][side]
    // FP64 engine
    double L, Linv, dtgm, minv, halfdt, expmgdt, var, w, S0, S1, xc;
    double xb[2], lo[2], hi[2], sl[2][2][2];
};

template <bool F32, int MODEL, bool OVER>
__device__ __forceinline__ Hot load_hot(const SetDev* __restrict__ g)
{
    Hot h;
    h.target = uni(g->target);
    h.equal_prob = uni(g->equal_prob);
    h.tau_zero = uni(g->tau_zero);
    h.slot_base = uni(g->slot_base);
    h.key0 = uni(g->key0);
    h.key1 = uni(g->key1);
    h.first = uni((uint32_t)g->first);
    h.max_steps = uni((uint32_t)g->max_steps);
    h.Tl0 = uni(g->Tl[0]);
    h.Tl1 = uni(g->Tl[1]);
    h.all0 = uni(g->Tall[0]);
    h.all1 = uni(g->Tall[1]);
    if (F32) {
        h.fL = uni(g->fL);
        h.fhalfL = uni(g->fhalfL);
        if (OVER) {
            h.fdtgm = uni(g->fdtgm);
        } else {
            h.fminv_halfdt = uni(g->fminv_halfdt);
            h.fhalfdt = uni(g->fhalfdt);
            h.fexpmgdt = uni(g->fexpmgdt);
        }
        h.fvar = uni(g->fvar);
        h.fbm = uni(g->fbm);
        if (MODEL == 0) {
            h.fw = uni(g->fw);
            h.fS0 = uni(g->fS[0]);
            h.fS1 = uni(g->fS[1]);
        } else {
            h.fxc = uni(g->fxc);
            for (int b = 0; b < 2; ++b) {
                h.fxb[b] = uni(g->fxb[b]);
                h.flo[b] = uni(g->flo[b]);
                h.fhi[b] = uni(g->fhi[b]);
                for (int k = 0; k < 2; ++k)
                    for (int side = 0; side < 2; ++side) h.fsl[b][k][side] = uni(g->fslope[b][k][side]);
            }
        }
    } else {
        h.L = uni(g->L);
        h.Linv = uni(g->Linv);
        if (OVER) {
            h.dtgm = uni(g->dtgm);
        } else {
            h.minv = uni(g->minv);
            h.halfdt = uni(g->halfdt);
            h.expmgdt = uni(g->expmgdt);
        }
        h.var = uni(g->var);
        if (MODEL == 0) {
            h.w = uni(g->w);
            h.S0 = uni(g->S[0]);
            h.S1 = uni(g->S[1]);
        } else {
            h.xc = uni(g->xc);
            for (int b = 0; b < 2; ++b) {
                h.xb[b] = uni(g->xb[b]);
                h.lo[b] = uni(g->lo[b]);
                h.hi[b] = uni(g->hi[b]);
                for (int k = 0; k < 2; ++k)
                    for (int side = 0; side < 2; ++side) h.sl[b][k][side] = uni(g->slope[b][k][side]);
            }
        }
    }
    return h;
}

// number of lag steps <= s (the index of the first lag still to come)
__device__ __forceinline__ int lag_index(const SetDev* Sg, uint32_t s)
{
    int lo = 0, hi = LDPCP_MSD_LAGS;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (Sg->lag[mid] <= s) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// A lane's first passage in this launch window (step 0: none).
struct Absorb {
    uint32_t step, tpt0, xf, xl;
};

// Bytes of LDS per wavefront in engine_step: MSD partial sums and counts (one
// lag per lane) and the wavefront's private FPT histogram.
constexpr size_t kLdsPerWave = LDPCP_MSD_LAGS * 8 + LDPCP_MSD_LAGS * 4 + LDPCP_HIST_BINS * 4;

// The first passages of one wavefront's lanes of one set (lanes of other sets
// and lanes without one pass step 0): the FPT/TPT step sums are reduced
// across the wavefront and added once by the set's leading lane; each passage
// is counted in the wavefront's LDS histogram, whose non-zero bins are then
// added to the set's histogram.  The bin of absorption step s is the last k
// with hist_edge[k] <= s (integer edges, so no logarithm on the device).
template <int MODEL>
__device__ __forceinline__ void flush_absorptions(const SetDev* __restrict__ Sg, const StatsPtrs& St, int set,
                                                  const Absorb a, uint64_t absorbed_mask, int lane, int leader,
                                                  uint32_t* __restrict__ hist)
{
    const bool on = a.step != 0u;
    // The durations are whole numbers of steps: their sums and sums of squares
    // are added as integers (the squares as 128-bit words), so the totals do
    // not depend on the order in which wavefronts add them.  One sum at a time,
    // each total moved to scalar registers.
    unsigned long long* sm = St.sums + (uint64_t)set * kSumsPerSet;
    auto add = [&](int k, uint32_t n) __attribute__((always_inline)) {
        uint64_t s1 = n;
        uint64_t lo = (uint64_t)n * n, hi = 0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) {
            s1 += (uint64_t)__shfl_xor((unsigned long long)s1, off);
            const uint64_t olo = (uint64_t)__shfl_xor((unsigned long long)lo, off);
            const uint64_t ohi = (uint64_t)__shfl_xor((unsigned long long)hi, off);
            const uint64_t sum = lo + olo;
            hi += ohi + (sum < lo ? 1u : 0u);
            lo = sum;
        }
        s1 = uni(s1);
        lo = uni(lo);
        hi = uni(hi);
        if (lane == leader) {
            atomicAdd(sm + 3 * k, (unsigned long long)s1);
            // the high word takes the carry out of this addition to the low word
            const uint64_t old = atomicAdd(sm + 3 * k + 1, (unsigned long long)lo);
            atomicAdd(sm + 3 * k + 2, (unsigned long long)(hi + (old + lo < old ? 1u : 0u)));
        }
    };
    add(0, on ? a.step : 0u);
    add(1, on ? a.step - a.tpt0 : 0u);
    if (MODEL == 1) {
        add(2, (on && a.xf) ? a.xf - a.tpt0 : 0u);
        add(3, on ? a.step - a.xl : 0u);
    }
    if (lane == leader)
        atomicAdd(St.cnt + (uint64_t)set * kCntPerSet + 0, (unsigned long long)__popcll(absorbed_mask));
    if (on) {
        int b = 0;
#pragma unroll
        for (int h = LDPCP_HIST_BINS / 2; h > 0; h >>= 1)
            if (Sg->hist_edge[b + h] <= a.step) b += h;
        atomicAdd(hist + b, 1u);
    }
    __threadfence_block();
#pragma unroll
    for (int k = 0; k < LDPCP_HIST_BINS / 64; ++k) {
        const int b = lane + 64 * k;
        const uint32_t n = hist[b];
        if (n) {
            atomicAdd(St.hist + (uint64_t)set * LDPCP_HIST_BINS + b, (unsigned long long)n);
            hist[b] = 0u;
        }
    }
    __threadfence_block();
}

// The word of the trajectory's n-th catch-up uniform: word n mod 4 of the
// role-2 call at index n / 4.  cw holds words 1-3 of the call at index
// n_catch / 4 whenever n_catch mod 4 != 0, so one call serves four hand-overs
// (a wavefront runs the call only when one of its lanes starts a new four).
__device__ __forceinline__ uint32_t catch_word(const Hot& H, uint32_t tid, uint32_t& n_catch, uint32_t* cw)
{
    const uint32_t k = n_catch & 3u;
    uint32_t word;
    if (k != 0u) {
        word = k == 1u ? cw[0] : (k == 2u ? cw[1] : cw[2]);
    } else {
        const Philox4 w = philox4x32_10(n_catch >> 2, 0u, tid, 2u, H.key0, H.key1);
        cw[0] = w.w[1];
        cw[1] = w.w[2];
        cw[2] = w.w[3];
        word = w.w[0];
    }
    ++n_catch;
    return word;
}

// the catch-up cache at the start of a launch (catch_word)
__device__ __forceinline__ void catch_cache_init(const Hot& H, uint32_t tid, uint32_t n_catch, uint32_t* cw)
{
    cw[0] = cw[1] = cw[2] = 0u;
    if (!H.tau_zero && (n_catch & 3u)) {
        const Philox4 w = philox4x32_10(n_catch >> 2, 0u, tid, 2u, H.key0, H.key1);
        cw[0] = w.w[1];
        cw[1] = w.w[2];
        cw[2] = w.w[3];
    }
}

// catch-up flip on entering a barrier (langevinDynamicsSingleBarrier.h:2039-2051):
// state of the barrier entered, aged by the steps since it was last left.
// cw: the lane's catch-up word cache (catch_word).
template <typename Real, int MODEL>
__device__ __forceinline__ int handover(const Hot& H, const SetDev* __restrict__ Sg, uint32_t& flags, uint32_t* ts,
                                        int kp, int kc, int state, uint32_t step, uint32_t tid, uint32_t& n_catch,
                                        uint32_t* cw)
{
    // the single barrier uses slots 0 and 1 only (one per cell): two-way
    // selects instead of four-way ones for the lane's slot index
    auto ts_set = [&](int k, uint32_t v) __attribute__((always_inline)) {
        if (MODEL == 0) {
            ts[0] = k == 0 ? v : ts[0];
            ts[1] = k == 1 ? v : ts[1];
        } else {
            ts[k] = v;
        }
    };
    if (kp >= 0) {
        flags = (flags & ~(F_SLOT_STATE << kp)) | (F_SLOT_VALID << kp) | (state ? (F_SLOT_STATE << kp) : 0u);
        ts_set(kp, step);
    }
    if (kc < 0) return state;  // leaving the two-cell window: absorbed this step
    int prev;
    uint32_t ts_c;
    if (flags & (F_SLOT_VALID << kc)) {
        prev = (flags >> (8 + kc)) & 1;
        ts_c = MODEL == 0 ? (kc == 0 ? ts[0] : ts[1]) : ts[kc];
    } else {
        // first touch: the state the barrier was created with, at step 0
        const int array = MODEL == 0 ? 0 : (kc & 1);
        const int cell = H.slot_base + (MODEL == 0 ? kc : (kc >> 1));
        const Philox4 w = philox4x32_10((uint32_t)cell, 0u, tid, 3u + 256u * (uint32_t)array, H.key0, H.key1);
        prev = below(w.w[0], Sg->T_create) ? 0 : 1;
        ts_c = 0;
        flags |= (F_SLOT_VALID << kc) | (prev ? (F_SLOT_STATE << kc) : 0u);
        ts_set(kc, 0u);
    }
    if (H.tau_zero) return prev;
    const uint32_t word = catch_word(H, tid, n_catch, cw);
    const Real lap = (Real)(step - ts_c);
    Real pc;
    if (sizeof(Real) == 4) {
        const float fK = Sg->fK;
        const float e = __expf(Sg->fc_lap * (float)lap);
        const float scale = H.equal_prob ? 0.5f : (prev == 0 ? fK / (1.0f + fK) : 1.0f / (1.0f + fK));
        pc = (Real)(scale * (1.0f - e));
    } else {
        const double K = Sg->K;
        const double e = exp(Sg->c_lap * (double)lap);
        const double scale = H.equal_prob ? 0.5 : (prev == 0 ? K / (1.0 + K) : 1.0 / (1.0 + K));
        pc = (Real)(scale * (1.0 - e));
    }
    // the catch-up uniform in the stepping precision
    const bool flip = (sizeof(Real) == 4 ? (Real)unit_f32(word) : (Real)unit_f64(word)) < pc;
    return flip ? (prev ^ 1) : prev;
}

// per-step flip (langevinDynamicsSingleBarrier.h:2010-2015), u < p as the
// integer test w < T.  EQ: equally probable states share one threshold.
template <bool EQ>
__device__ __forceinline__ int flip_step(const Hot& H, int state, uint32_t word)
{
    if (H.tau_zero) return state;
    bool f;
    if (EQ) {
        f = H.all0 || word < H.Tl0;
    } else {
        const uint32_t T = state ? H.Tl1 : H.Tl0;
        f = (state ? H.all1 : H.all0) || word < T;
    }
    return f ? (state ^ 1) : state;
}

// Launch window of one set: local steps s0+1 .. s_end, MSD lags [j0, j1).
struct Window {
    uint32_t s0, s_end;
    int j0;
};

// ---------------------------------------------------------------------------
// FP32 stepping.  MODEL 0 single, 1 double barrier.
// ---------------------------------------------------------------------------
// state-0 slopes held in vector registers (in_vgpr), so that choosing the
// slope of the lane's state is one select
struct Slope0 {
    float s0;        // single barrier
    float a[2][2];   // double barrier [barrier][side]
};

template <int MODEL>
__device__ __forceinline__ Slope0 slope0_f32(const Hot& H)
{
    Slope0 z;
    if (MODEL == 0) {
        z.s0 = in_vgpr(H.fS0);
    } else {
        for (int b = 0; b < 2; ++b)
            for (int side = 0; side < 2; ++side) z.a[b][side] = in_vgpr(H.fsl[b][0][side]);
    }
    return z;
}

template <int MODEL>
__device__ __forceinline__ float force_f32(const Hot& H, const Slope0& z, float xp, int state)
{
    if (MODEL == 0) {
        // -slope on (-w, 0), +slope on (0, w), 0 elsewhere (strict compares,
        // langevinDynamicsSingleBarrier.h:1292-1324): the slope with xp's sign bit
        // xor-ed in, where 0 < |xp| < w
        const float sp = state ? H.fS1 : z.s0;
        const float signed_sp = __uint_as_float(__float_as_uint(sp) ^ (__float_as_uint(xp) & 0x80000000u));
        return (fabsf(xp) < H.fw && xp != 0.0f) ? signed_sp : 0.0f;
    } else {
        const float a0 = state ? H.fsl[0][1][0] : z.a[0][0], a1 = state ? H.fsl[0][1][1] : z.a[0][1];
        const float b0 = state ? H.fsl[1][1][0] : z.a[1][0], b1 = state ? H.fsl[1][1][1] : z.a[1][1];
        float s = 0.0f;
        s = (xp < H.fxb[0] && xp > H.flo[0]) ? a0 : s;
        s = (xp < H.fxc && xp > H.fxb[0]) ? a1 : s;
        s = (xp < H.fxb[1] && xp > H.fxc) ? b0 : s;
        s = (xp < H.fhi[1] && xp > H.fxb[1]) ? b1 : s;
        return 0.0f - s;
    }
}

// One trajectory for one launch window.  H holds the set's hot constants in
// scalar registers; EQ: the two barrier states are equally probable.
template <int MODEL, bool OVER, bool EQ>
__device__ __forceinline__ void body_f32(const SlotPtrs& P, const StatsPtrs& St, const Hot H,
                                         const SetDev* __restrict__ Sg, int set, uint32_t slot, const Window W,
                                         double* lds_sum, uint32_t* lds_n, bool& fin, uint32_t& nsteps,
                                         Absorb& ab)
{
    const uint32_t tid = slot - H.first;
    float xp = P.xp[slot], comp = P.comp[slot], v = P.fv[slot], R = P.fR[slot];
    int cell = P.cell[slot];
    uint32_t flags = P.flags[slot];
    const uint4 t4 = P.ts[slot];
    uint32_t ts[4] = {t4.x, t4.y, t4.z, t4.w};
    uint32_t n_catch = P.n_catch[slot];
    uint32_t cw[3];
    catch_cache_init(H, tid, n_catch, cw);
    uint32_t tpt0 = P.tpt0[slot];
    uint32_t xf = 0, xl = 0;
    if (MODEL == 1) {
        xf = P.xc_first[slot];
        xl = P.xc_last[slot];
    }
    int state = flags & F_STATE;
    int local = (MODEL == 1) ? (xp < H.fxc ? 0 : 1) : 0;
    const Slope0 z = slope0_f32<MODEL>(H);
    float F = force_f32<MODEL>(H, z, xp, state);
    int j = W.j0;
    uint32_t next_lag = j < LDPCP_MSD_LAGS ? Sg->lag[j] : kNoLag;
    // the step counter is wave-uniform; the exit step of a lane is kept apart
    // so that the counter is not live after the loop (it stays scalar)
    uint32_t last = W.s_end;
    bool done = false;
    // one step with Gaussian g and flip word wflip; true when the lane is absorbed
    auto one = [&](const uint32_t step, const float g, const uint32_t wflip) __attribute__((always_inline)) {
        // position increment (overdamped MA(1) scheme, langevinDynamicsSingleBarrier.h:2860-2876;
        // underdamped B-A-O-A-B, :1680-1706), compensated offset update
        // (g carries the noise amplitude: box_muller_f32_scaled)
        float dx;
        if (OVER) {
            dx = __builtin_fmaf(F, H.fdtgm, g + R);
            R = g;
        } else {
            v = __builtin_fmaf(F, H.fminv_halfdt, v);
            const float d1 = v * H.fhalfdt;
            v = __builtin_fmaf(H.fexpmgdt, v, g);
            dx = __builtin_fmaf(v, H.fhalfdt, d1);
        }
        const float y = dx - comp;
        const float t = xp + y;
        comp = (t - xp) - y;
        // the offset's sign changes within a cell: a barrier-top crossing
        const bool sign_change = (int32_t)(__float_as_uint(t) ^ __float_as_uint(xp)) < 0;
        xp = t;
        // periodic (well-centre) crossing: offset kept in [-L/2, L/2)
        bool ent_period = false;
        const int cell_prev = cell;
        if (__builtin_expect(xp >= H.fhalfL || xp < -H.fhalfL, 0)) {
            const bool up = xp >= H.fhalfL;
            xp = up ? xp - H.fL : xp + H.fL;
            cell = up ? cell + 1 : cell - 1;
            ent_period = true;
        }
        // well (barrier-top) crossing
        const bool ent_well = !ent_period && sign_change;
        state = flip_step<EQ>(H, state, wflip);
        if (MODEL == 0) {
            if (ent_period) {
                const int kp = slot_index(0, cell_prev, H.slot_base, 0);
                const int kc = slot_index(0, cell, H.slot_base, 0);
                state = handover<float, MODEL>(H, Sg, flags, ts, kp, kc, state, step, tid, n_catch, cw);
            }
        } else {
            const int local_prev = local;
            local = xp < H.fxc ? 0 : 1;
            if (local != local_prev) {
                const int kp = slot_index(1, cell_prev, H.slot_base, local_prev);
                const int kc = slot_index(1, cell, H.slot_base, local);
                state = handover<float, MODEL>(H, Sg, flags, ts, kp, kc, state, step, tid, n_catch, cw);
                if (!ent_period) {
                    xl = step;
                    if (!xf) xf = step;
                }
            }
        }
        F = force_f32<MODEL>(H, z, xp, state);
        if (!OVER) v = __builtin_fmaf(F, H.fminv_halfdt, v);
        if (step == next_lag) {
            const double x = (double)cell * (double)H.fL + (double)xp - (double)comp;
            const double d = x - Sg->x0;
            atomicAdd(lds_sum + j, d * d);
            atomicAdd(lds_n + j, 1u);
            // every lane still stepping has passed the same lags: j and the next lag
            // stay wave-uniform (scalar) inside the loop
            j = uni(j + 1);
            next_lag = uni(j < LDPCP_MSD_LAGS ? Sg->lag[j] : kNoLag);
        }
        // single barrier: first barrier-top crossing; double: first well centre
        // outside the two cells of the start
        const bool hit =
            (MODEL == 0) ? ent_well : (ent_period && (unsigned)(cell - H.slot_base) > 1u);
        if (!hit && ent_period) {
            tpt0 = step;
            if (MODEL == 1) xf = 0;
        }
        return hit;
    };
    // Steps 2q+1 and 2q+2 share one Philox call (role 0, counter q): the loop
    // runs over whole pairs, with half a pair before it when the window starts
    // after an odd step and after it when the window ends on one.
    uint32_t begin = W.s0;
    if ((begin & 1u) && begin < W.s_end) {
        const Philox4 w = philox4x32_10_lane_ctr(begin >> 1, 0u, tid, 0u, H.key0, H.key1);
        float g0, g1;
        box_muller_f32_scaled(w.w[0], w.w[1], H.fbm, g0, g1);
        ++begin;
        if (one(begin, g1, w.w[3])) {
            done = true;
            last = begin;
        }
    }
    const uint32_t pair_end = begin + ((W.s_end - begin) & ~1u);
    if (!done) {
        for (uint32_t step = begin; step < pair_end; step += 2) {
            const Philox4 w = philox4x32_10_lane_ctr(step >> 1, 0u, tid, 0u, H.key0, H.key1);
            float g0, g1;
            box_muller_f32_scaled(w.w[0], w.w[1], H.fbm, g0, g1);
            if (one(step + 1, g0, w.w[2])) {
                done = true;
                last = step + 1;
                break;
            }
            if (one(step + 2, g1, w.w[3])) {
                done = true;
                last = step + 2;
                break;
            }
        }
    }
    if (!done && pair_end < W.s_end) {
        const Philox4 w = philox4x32_10_lane_ctr(pair_end >> 1, 0u, tid, 0u, H.key0, H.key1);
        float g0, g1;
        box_muller_f32_scaled(w.w[0], w.w[1], H.fbm, g0, g1);
        if (one(pair_end + 1, g0, w.w[2])) {
            done = true;
            last = pair_end + 1;
        }
    }
    if (done) P.absorbed[slot] = last;
    ab.step = done ? last : 0u;
    if (!done && W.s_end >= H.max_steps) done = true;  // capped
    if (done) flags &= ~F_ALIVE;
    flags = (flags & ~F_STATE) | (uint32_t)state;
    P.xp[slot] = xp;
    P.comp[slot] = comp;
    P.fv[slot] = v;
    P.fR[slot] = R;
    P.cell[slot] = cell;
    P.flags[slot] = flags;
    P.ts[slot] = make_uint4(ts[0], ts[1], ts[2], ts[3]);
    P.n_catch[slot] = n_catch;
    P.tpt0[slot] = tpt0;
    if (MODEL == 1) {
        P.xc_first[slot] = xf;
        P.xc_last[slot] = xl;
    }
    fin = done;
    nsteps = last - W.s0;
    ab.tpt0 = tpt0;
    ab.xf = xf;
    ab.xl = xl;
}

// ---------------------------------------------------------------------------
// FP64 stepping: the reference arithmetic on the absolute position
// ---------------------------------------------------------------------------
template <int MODEL>
__device__ __forceinline__ double force_f64(const Hot& H, double x, double periodic, int state)
{
    const double xp = x - periodic * H.L;
    if (MODEL == 0) {
        const double sp = state ? H.S1 : H.S0, sn = 0.0 - sp;
        double F = (xp < 0.0 && xp > -H.w) ? sn : 0.0;
        F = (xp < H.w && xp > 0.0) ? sp : F;
        return F;
    } else {
        const double a0 = state ? H.sl[0][1][0] : H.sl[0][0][0], a1 = state ? H.sl[0][1][1] : H.sl[0][0][1];
        const double b0 = state ? H.sl[1][1][0] : H.sl[1][0][0], b1 = state ? H.sl[1][1][1] : H.sl[1][0][1];
        double s = 0.0;
        s = (xp < H.xb[0] && xp > H.lo[0]) ? a0 : s;
        s = (xp < H.xc && xp > H.xb[0]) ? a1 : s;
        s = (xp < H.xb[1] && xp > H.xc) ? b0 : s;
        s = (xp < H.hi[1] && xp > H.xb[1]) ? b1 : s;
        return 0.0 - s;
    }
}

template <int MODEL, bool OVER, bool EQ>
__device__ __forceinline__ void body_f64(const SlotPtrs& P, const StatsPtrs& St, const Hot H,
                                         const SetDev* __restrict__ Sg, int set, uint32_t slot, const Window W,
                                         double* lds_sum, uint32_t* lds_n, bool& fin, uint32_t& nsteps,
                                         Absorb& ab)
{
    const uint32_t tid = slot - H.first;
    double x = P.x[slot], v = P.v[slot], R = P.R[slot];
    uint32_t flags = P.flags[slot];
    const uint4 t4 = P.ts[slot];
    uint32_t ts[4] = {t4.x, t4.y, t4.z, t4.w};
    uint32_t n_catch = P.n_catch[slot];
    uint32_t cw[3];
    catch_cache_init(H, tid, n_catch, cw);
    uint32_t tpt0 = P.tpt0[slot];
    uint32_t xf = 0, xl = 0;
    if (MODEL == 1) {
        xf = P.xc_first[slot];
        xl = P.xc_last[slot];
    }
    int state = flags & F_STATE;
    double r = x * H.Linv;
    double well = floor(r), periodic = rint(r);
    int local = (MODEL == 1) ? ((x - periodic * H.L) < H.xc ? 0 : 1) : 0;
    double F = force_f64<MODEL>(H, x, periodic, state);
    int j = W.j0;
    uint32_t next_lag = j < LDPCP_MSD_LAGS ? Sg->lag[j] : kNoLag;
    uint32_t last = W.s_end;
    bool done = false;
    // one step with Gaussian g and flip word wflip; true when the lane is absorbed
    auto one = [&](const uint32_t step, const double g, const uint32_t wflip) __attribute__((always_inline)) {
        if (OVER) {
            const double rf = H.var * g;
            x = fma(F, H.dtgm, x);
            x = x + rf;
            x = x + R;
            R = rf;
        } else {
            v = v + (H.minv * F) * H.halfdt;
            x = x + v * H.halfdt;
            v = H.expmgdt * v + H.var * g;
            x = x + v * H.halfdt;
        }
        r = x * H.Linv;
        const double well_prev = well, periodic_prev = periodic;
        well = floor(r);
        periodic = rint(r);
        const bool ent_well = well != well_prev;
        const bool ent_period = periodic != periodic_prev;
        state = flip_step<EQ>(H, state, wflip);
        if (MODEL == 0) {
            if (ent_period) {
                const int kp = slot_index(0, (int)periodic_prev, H.slot_base, 0);
                const int kc = slot_index(0, (int)periodic, H.slot_base, 0);
                state = handover<double, MODEL>(H, Sg, flags, ts, kp, kc, state, step, tid, n_catch, cw);
            }
        } else {
            const int local_prev = local;
            local = (x - periodic * H.L) < H.xc ? 0 : 1;
            if (local != local_prev) {
                const int kp = slot_index(1, (int)periodic_prev, H.slot_base, local_prev);
                const int kc = slot_index(1, (int)periodic, H.slot_base, local);
                state = handover<double, MODEL>(H, Sg, flags, ts, kp, kc, state, step, tid, n_catch, cw);
                if (!ent_period) {
                    xl = step;
                    if (!xf) xf = step;
                }
            }
        }
        F = force_f64<MODEL>(H, x, periodic, state);
        if (!OVER) v = v + (H.minv * F) * H.halfdt;
        if (step == next_lag) {
            const double d = x - Sg->x0;
            atomicAdd(lds_sum + j, d * d);
            atomicAdd(lds_n + j, 1u);
            // every lane still stepping has passed the same lags: j and the next lag
            // stay wave-uniform (scalar) inside the loop
            j = uni(j + 1);
            next_lag = uni(j < LDPCP_MSD_LAGS ? Sg->lag[j] : kNoLag);
        }
        const int kcell = (int)periodic - H.slot_base;
        const bool hit = (MODEL == 0) ? ent_well : (ent_period && (kcell < 0 || kcell > 1));
        if (!hit && ent_period) {
            tpt0 = step;
            if (MODEL == 1) xf = 0;
        }
        return hit;
    };
    // step pairs sharing one Philox call, as in body_f32
    uint32_t begin = W.s0;
    if ((begin & 1u) && begin < W.s_end) {
        const Philox4 w = philox4x32_10(begin >> 1, 0u, tid, 0u, H.key0, H.key1);
        double g0, g1;
        box_muller_f64(w.w[0], w.w[1], g0, g1);
        ++begin;
        if (one(begin, g1, w.w[3])) {
            done = true;
            last = begin;
        }
    }
    const uint32_t pair_end = begin + ((W.s_end - begin) & ~1u);
    if (!done) {
        for (uint32_t step = begin; step < pair_end; step += 2) {
            const Philox4 w = philox4x32_10(step >> 1, 0u, tid, 0u, H.key0, H.key1);
            double g0, g1;
            box_muller_f64(w.w[0], w.w[1], g0, g1);
            if (one(step + 1, g0, w.w[2])) {
                done = true;
                last = step + 1;
                break;
            }
            if (one(step + 2, g1, w.w[3])) {
                done = true;
                last = step + 2;
                break;
            }
        }
    }
    if (!done && pair_end < W.s_end) {
        const Philox4 w = philox4x32_10(pair_end >> 1, 0u, tid, 0u, H.key0, H.key1);
        double g0, g1;
        box_muller_f64(w.w[0], w.w[1], g0, g1);
        if (one(pair_end + 1, g0, w.w[2])) {
            done = true;
            last = pair_end + 1;
        }
    }
    if (done) P.absorbed[slot] = last;
    ab.step = done ? last : 0u;
    if (!done && W.s_end >= H.max_steps) done = true;
    if (done) flags &= ~F_ALIVE;
    flags = (flags & ~F_STATE) | (uint32_t)state;
    P.x[slot] = x;
    P.v[slot] = v;
    P.R[slot] = R;
    P.flags[slot] = flags;
    P.ts[slot] = make_uint4(ts[0], ts[1], ts[2], ts[3]);
    P.n_catch[slot] = n_catch;
    P.tpt0[slot] = tpt0;
    if (MODEL == 1) {
        P.xc_first[slot] = xf;
        P.xc_last[slot] = xl;
    }
    fin = done;
    nsteps = last - W.s0;
    ab.tpt0 = tpt0;
    ab.xf = xf;
    ab.xl = xl;
}

template <bool F32, int MODEL, bool OVER>
__device__ __forceinline__ void body(const SlotPtrs& P, const StatsPtrs& St, const SetDev* __restrict__ Sg, int set,
                                     uint32_t slot, const Window W, double* lds_sum, uint32_t* lds_n, bool& fin,
                                     uint32_t& nsteps, Absorb& ab)
{
    const bool eq = __builtin_amdgcn_readfirstlane(Sg->equal_prob) != 0;
    if (F32) {
        const Hot H = load_hot<true, MODEL, OVER>(Sg);
        if (eq) body_f32<MODEL, OVER, true>(P, St, H, Sg, set, slot, W, lds_sum, lds_n, fin, nsteps, ab);
        else body_f32<MODEL, OVER, false>(P, St, H, Sg, set, slot, W, lds_sum, lds_n, fin, nsteps, ab);
    } else {
        const Hot H = load_hot<false, MODEL, OVER>(Sg);
        if (eq) body_f64<MODEL, OVER, true>(P, St, H, Sg, set, slot, W, lds_sum, lds_n, fin, nsteps, ab);
        else body_f64<MODEL, OVER, false>(P, St, H, Sg, set, slot, W, lds_sum, lds_n, fin, nsteps, ab);
    }
}

// ---------------------------------------------------------------------------
// Kernels
// ---------------------------------------------------------------------------

// One launch: persistent workgroups; each wavefront takes 64-entry chunks of
// the live list until none is left, and records the chunk's survivors as a
// 64-bit mask for the compaction.  The FP32 kernel is held to 96 VGPRs (five
// wavefronts per SIMD); the step loop needs 93, the absorption flush after it
// would otherwise take a few more.
template <bool F32>
__global__ void __launch_bounds__(kMaxBlock) __attribute__((amdgpu_waves_per_eu(F32 ? 5 : 1))) engine_step(SlotPtrs P, StatsPtrs St, const SetDev* __restrict__ sets,
                                                         Ctrl* ctrl, const uint32_t* __restrict__ live,
                                                         unsigned long long* __restrict__ mask, int parity, uint32_t K)
{
    // dynamic LDS, kLdsPerWave bytes per wavefront: [waves][64] MSD sums,
    // [waves][64] MSD counts, [waves][256] FPT histogram bins
    extern __shared__ __align__(16) unsigned char lds_raw[];
    const int wv = threadIdx.x >> 6;
    const int lane = threadIdx.x & 63;
    const int n_waves = blockDim.x >> 6;
    double* const s_sum = (double*)lds_raw + wv * LDPCP_MSD_LAGS;
    uint32_t* const s_n = (uint32_t*)(lds_raw + (size_t)n_waves * LDPCP_MSD_LAGS * 8) + wv * LDPCP_MSD_LAGS;
    uint32_t* const s_hist =
        (uint32_t*)(lds_raw + (size_t)n_waves * LDPCP_MSD_LAGS * 12) + wv * LDPCP_HIST_BINS;
    s_sum[lane] = 0.0;
    s_n[lane] = 0u;
#pragma unroll
    for (int k = 0; k < LDPCP_HIST_BINS / 64; ++k) s_hist[lane + 64 * k] = 0u;
    __syncthreads();
    const uint32_t n_live = ctrl->n_live[parity];
    const unsigned long long estep = ctrl->estep;
    const uint32_t n_chunks = (n_live + 63u) >> 6;
    for (;;) {
        // Chunk index: lane 0 adds 1 and the other lanes add 0, so the value
        // broadcast from lane 0 is computed with every lane active.  (Writing
        // it under `if (lane == 0)` lets the optimizer route lanes 1-63 around
        // the atomic to the broadcast, which then reads one of them.)
        const uint32_t got = atomicAdd(&ctrl->work, lane == 0 ? 1u : 0u);
        const uint32_t c = (uint32_t)__builtin_amdgcn_readfirstlane((int)got);
        if (c >= n_chunks) break;
        const uint32_t li = c * 64u + (uint32_t)lane;
        const bool in = li < n_live;
        const uint32_t slot = in ? live[li] : 0u;
        const int set = in ? (int)P.set_of[slot] : -1;
        uint64_t pending = __ballot(in);
        uint64_t alive = 0;
        while (pending) {
            const int leader = (int)__builtin_ctzll(pending);
            const int s = __builtin_amdgcn_readlane(set, leader);
            const bool mine = in && set == s;
            const uint64_t group = __ballot(mine);
            const SetDev* Sg = sets + s;
            const uint32_t maxs = (uint32_t)Sg->max_steps;
            Window W;
            W.s0 = (uint32_t)(estep - Sg->start);
            const uint64_t e64 = (uint64_t)W.s0 + K;
            W.s_end = e64 < (uint64_t)maxs ? (uint32_t)e64 : maxs;
            W.j0 = lag_index(Sg, W.s0);
            bool fin = false;
            uint32_t nst = 0;
            Absorb ab = {0u, 0u, 0u, 0u};
            const int kind = __builtin_amdgcn_readfirstlane(Sg->kind);
            if (mine) {
                switch (kind) {
                case 0: body<F32, 0, false>(P, St, Sg, s, slot, W, s_sum, s_n, fin, nst, ab); break;
                case 1: body<F32, 0, true>(P, St, Sg, s, slot, W, s_sum, s_n, fin, nst, ab); break;
                case 2: body<F32, 1, false>(P, St, Sg, s, slot, W, s_sum, s_n, fin, nst, ab); break;
                default: body<F32, 1, true>(P, St, Sg, s, slot, W, s_sum, s_n, fin, nst, ab); break;
                }
            }
            const uint64_t abm = __ballot(ab.step != 0u);
            if (abm) {
                if (kind < 2) flush_absorptions<0>(Sg, St, s, ab, abm, lane, leader, s_hist);
                else flush_absorptions<1>(Sg, St, s, ab, abm, lane, leader, s_hist);
            }
            const uint64_t finm = __ballot(fin);
            alive |= group & ~finm;
            uint32_t sum = nst;
#pragma unroll
            for (int off = 32; off > 0; off >>= 1) sum += __shfl_xor(sum, off);
            if (lane == leader) {
                unsigned long long* cn = St.cnt + (uint64_t)s * kCntPerSet;
                if (finm) atomicAdd(cn + 1, (unsigned long long)__popcll(finm));
                atomicAdd(cn + 2, (unsigned long long)sum);
            }
            // flush the MSD samples of the lags this window crossed
            const int j1 = lag_index(Sg, W.s_end);
            __threadfence_block();
            const int jl = W.j0 + lane;
            if (jl < j1) {
                const uint32_t n = s_n[jl];
                if (n) {
                    atomicAdd(St.msd_sum + (uint64_t)s * LDPCP_MSD_LAGS + jl, s_sum[jl]);
                    atomicAdd(St.msd_n + (uint64_t)s * LDPCP_MSD_LAGS + jl, (unsigned long long)n);
                    s_sum[jl] = 0.0;
                    s_n[jl] = 0u;
                }
            }
            __threadfence_block();
            pending &= ~group;
        }
        mask[c] = alive;  // the same value from every lane: no lane-0 branch in the loop
    }
}

// Exclusive scan of the survivors per chunk (one workgroup); advances the
// control words for the next launch.
__global__ void __launch_bounds__(1024) engine_scan(Ctrl* ctrl, const unsigned long long* __restrict__ mask,
                                                    uint32_t* __restrict__ off, int parity, uint32_t K)
{
    __shared__ uint32_t sh[1024];
    const uint32_t n_live = ctrl->n_live[parity];
    const uint32_t nc = (n_live + 63u) >> 6;
    const uint32_t per = (nc + 1023u) / 1024u;
    const uint32_t b = min(threadIdx.x * per, nc), e = min(b + per, nc);
    uint32_t sum = 0;
    for (uint32_t i = b; i < e; ++i) sum += (uint32_t)__popcll(mask[i]);
    sh[threadIdx.x] = sum;
    __syncthreads();
    for (uint32_t d = 1; d < 1024; d <<= 1) {
        const uint32_t v = threadIdx.x >= d ? sh[threadIdx.x - d] : 0u;
        __syncthreads();
        sh[threadIdx.x] += v;
        __syncthreads();
    }
    uint32_t run = sh[threadIdx.x] - sum;
    for (uint32_t i = b; i < e; ++i) {
        off[i] = run;
        run += (uint32_t)__popcll(mask[i]);
    }
    if (threadIdx.x == 1023) {
        ctrl->n_live[parity ^ 1] = sh[1023];
        ctrl->work = 0;
        ctrl->estep += K;
        ctrl->launches += 1;
    }
}

// Order-preserving scatter of the survivors into the other live list.
__global__ void __launch_bounds__(256) engine_scatter(const Ctrl* ctrl, const unsigned long long* __restrict__ mask,
                                                      const uint32_t* __restrict__ off,
                                                      const uint32_t* __restrict__ live_in,
                                                      uint32_t* __restrict__ live_out, int parity)
{
    const uint32_t n_live = ctrl->n_live[parity];
    const uint32_t nc = (n_live + 63u) >> 6;
    const int lane = threadIdx.x & 63;
    const uint32_t nw = gridDim.x * (blockDim.x >> 6);
    for (uint32_t c = blockIdx.x * (blockDim.x >> 6) + (threadIdx.x >> 6); c < nc; c += nw) {
        const unsigned long long m = mask[c];
        if ((m >> lane) & 1ull) {
            const uint32_t pos = off[c] + (uint32_t)__popcll(m & ((1ull << lane) - 1ull));
            live_out[pos] = live_in[c * 64u + (uint32_t)lane];
        }
    }
}

// Admission: step-0 state of a set's trajectories (IntegrateStep0,
// langevinDynamicsSingleBarrier.h:1619-1652, with the Philox role-1 draws),
// appended to the live list after its current entries.
__global__ void __launch_bounds__(256) engine_admit(SlotPtrs P, const SetDev* __restrict__ sets, int set,
                                                    const Ctrl* ctrl, uint32_t* __restrict__ live, int parity,
                                                    uint32_t n, int fp32)
{
    const uint32_t j = blockIdx.x * 256u + threadIdx.x;
    if (j >= n) return;
    const SetDev& S = sets[set];
    const uint32_t slot = (uint32_t)S.first + j;
    const Philox4 w = philox4x32_10(0u, 0u, j, 1u, S.key0, S.key1);
    uint32_t flags = F_ALIVE;
    // u <= p0 -> state 0 (langevinDynamicsSingleBarrier.h:1888-1895)
    const int state = S.randomize_state ? ((uint64_t)w.w[2] < S.T_p0le ? 0 : 1) : S.fixed_state;
    if (state) flags |= F_STATE;
    // the barrier under X0 takes the initial state (:1906-1911)
    {
        const int local = (S.model == 1 && (S.x0 - (double)S.cell0 * S.L) >= S.xc) ? 1 : 0;
        const int k = slot_index(S.model, S.cell0, S.slot_base, local);
        if (k >= 0) flags |= (F_SLOT_VALID << k) | (state ? (F_SLOT_STATE << k) : 0u);
    }
    if (fp32) {
        P.xp[slot] = S.fxp0;
        P.comp[slot] = 0.0f;
        P.cell[slot] = S.cell0;
        // FP32 noise amplitude applied to the FP32 Gaussian, as in the step loop
        float g0, g1;
        box_muller_f32(w.w[0], w.w[1], g0, g1);
        P.fv[slot] = S.randomize_v ? (float)S.vsigma * g0 : (float)S.v0;
        P.fR[slot] = S.overdamped ? S.fvar * g1 : 0.0f;
    } else {
        double gv, gr;
        box_muller_f64(w.w[0], w.w[1], gv, gr);
        P.x[slot] = S.x0;
        P.v[slot] = S.randomize_v ? S.vsigma * gv : S.v0;
        P.R[slot] = S.overdamped ? S.var * gr : 0.0;
    }
    P.flags[slot] = flags;
    P.ts[slot] = make_uint4(0u, 0u, 0u, 0u);
    P.n_catch[slot] = 0;
    P.tpt0[slot] = 0;
    P.xc_first[slot] = 0;
    P.xc_last[slot] = 0;
    P.absorbed[slot] = 0;
    P.set_of[slot] = (uint32_t)set;
    live[ctrl->n_live[parity] + j] = slot;
}

__global__ void engine_bump(Ctrl* ctrl, int parity, uint32_t n) { ctrl->n_live[parity] += n; }

__global__ void __launch_bounds__(256) engine_zero_set(StatsPtrs St, int set)
{
    const int t = threadIdx.x;
    St.hist[(uint64_t)set * LDPCP_HIST_BINS + t] = 0ull;
    if (t < LDPCP_MSD_LAGS) {
        St.msd_sum[(uint64_t)set * LDPCP_MSD_LAGS + t] = 0.0;
        St.msd_n[(uint64_t)set * LDPCP_MSD_LAGS + t] = 0ull;
    }
    if (t < kSumsPerSet) St.sums[(uint64_t)set * kSumsPerSet + t] = 0ull;
    if (t < kCntPerSet) St.cnt[(uint64_t)set * kCntPerSet + t] = 0ull;
}

__global__ void engine_zero_ctrl(Ctrl* ctrl)
{
    ctrl->n_live[0] = ctrl->n_live[1] = 0;
    ctrl->work = 0;
    ctrl->pad = 0;
    ctrl->estep = 0;
    ctrl->launches = 0;
}

}  // namespace

// ---------------------------------------------------------------------------
// Host side
// ---------------------------------------------------------------------------
static_assert(LDPCP_HIST_BINS == 256, "engine_zero_set assumes 256 histogram bins");
static_assert(LDPCP_MSD_LAGS == 64, "one MSD lag per lane");

void RangeAllocator::reset(uint64_t capacity)
{
    free_.clear();
    if (capacity) free_[0] = capacity;
    free_total_ = capacity;
}

bool RangeAllocator::alloc(uint64_t n, uint64_t* first)
{
    for (auto it = free_.begin(); it != free_.end(); ++it) {
        if (it->second >= n) {
            *first = it->first;
            const uint64_t rest = it->second - n, at = it->first + n;
            free_.erase(it);
            if (rest) free_[at] = rest;
            free_total_ -= n;
            return true;
        }
    }
    return false;
}

void RangeAllocator::release(uint64_t first, uint64_t n)
{
    if (!n) return;
    auto it = free_.emplace(first, n).first;
    free_total_ += n;
    // merge with the next and previous ranges
    auto nx = std::next(it);
    if (nx != free_.end() && it->first + it->second == nx->first) {
        it->second += nx->second;
        free_.erase(nx);
    }
    if (it != free_.begin()) {
        auto pv = std::prev(it);
        if (pv->first + pv->second == it->first) {
            pv->second += it->second;
            free_.erase(it);
        }
    }
}

// slot state (per precision) + two live-list entries + per-chunk mask and offset
uint64_t Engine::bytes_per_traj(int precision)
{
    const uint64_t state = precision == 32 ? 5 * 4 : 3 * 8;
    const uint64_t common = 4 /*flags*/ + 16 /*ts*/ + 6 * 4 /*n_catch tpt0 xc_first xc_last absorbed set_of*/;
    return state + common + 2 * 4 + 1;
}

uint64_t Engine::bytes_per_set()
{
    return sizeof(SetDev) + 8ull * (LDPCP_HIST_BINS + kSumsPerSet + kCntPerSet + 2 * LDPCP_MSD_LAGS);
}

int Engine::validate(const ldpcp_sweep_set& s)
{
    if (s.n_traj == 0 || s.n_traj >= (1ull << 31) || s.max_steps == 0 || s.max_steps >= (1ull << 32) - (1ull << 21) ||
        !(s.timestep > 0.0) || !(s.periodic_length > 0.0) || (s.model != 0 && s.model != 1) || !(s.mass > 0.0) ||
        (s.overdamped && !(s.gamma > 0.0)) || !(s.tau >= 0.0)) {
        set_error("invalid parameter set (need 0 < n_traj < 2^31, 0 < max_steps < 2^32 - 2^21, timestep > 0, "
                  "periodic_length > 0, model 0 or 1, mass > 0, gamma > 0 when overdamped, tau >= 0)");
        return LDPCP_EINVAL;
    }
    if (s.model == 1) {
        const double xp0 = s.x0 - rint(s.x0 * (1.0 / s.periodic_length)) * s.periodic_length;
        if (!(xp0 <= s.position[0] - s.half_width[0] || xp0 >= s.position[1] + s.half_width[1])) {
            set_error("invalid parameter set: double-barrier X0 must lie in the outer well");
            return LDPCP_EINVAL;
        }
    }
    return LDPCP_OK;
}

static void fill_set_dev(const ldpcp_sweep_set& s, uint64_t first, uint64_t start, SetDev& d)
{
    memset(&d, 0, sizeof(d));
    d.model = s.model;
    d.overdamped = s.overdamped;
    d.kind = s.model * 2 + (s.overdamped ? 1 : 0);
    d.target = s.model == 0 ? 0 : 1;
    d.equal_prob = s.equal_probability;
    d.randomize_state = s.randomize_state;
    d.randomize_v = s.randomize_v;
    d.tau_zero = s.tau == 0.0;
    d.dt = s.timestep;
    d.L = s.periodic_length;
    d.Linv = 1.0 / s.periodic_length;
    d.halfL = s.periodic_length * 0.5;
    d.w = s.half_width[0];
    for (int k = 0; k < 2; ++k) {
        d.H[k] = s.height[0][k];
        d.S[k] = s.height[0][k] / s.half_width[0];
    }
    d.minv = 1.0 / s.mass;
    d.halfdt = s.timestep * 0.5;
    d.expmgdt = exp(-s.gamma * s.timestep);
    d.var = sqrt(s.kb * s.temperature / s.mass * (1.0 - exp(-2.0 * s.gamma * s.timestep)));
    if (s.overdamped) {
        const double mgi = 1.0 / (s.mass * s.gamma);
        d.dtgm = s.timestep * mgi;
        d.var = sqrt(s.kb * s.temperature * mgi * s.timestep * 0.5);
    }
    d.K = s.eq_constant;
    if (s.equal_probability) {
        d.c_lap = d.tau_zero ? 0.0 : -2.0 * s.timestep / s.tau;
        d.p_flip = d.tau_zero ? 0.0 : 0.5 * (1.0 - exp(-2.0 * s.timestep / s.tau));
        d.p0 = 0.5;
    } else {
        d.c_lap = d.tau_zero ? 0.0 : -(s.eq_constant + 1.0) * s.timestep / s.tau;
        if (!d.tau_zero) {
            d.p_asym[0] = s.eq_constant / (1.0 + s.eq_constant) * (1.0 - exp(d.c_lap));
            d.p_asym[1] = 1.0 / (1.0 + s.eq_constant) * (1.0 - exp(d.c_lap));
        }
        d.p0 = 1.0 / (1.0 + s.eq_constant);
    }
    d.p_create = 1.0 / (1.0 + s.eq_constant);
    d.vsigma = sqrt(s.kb * s.temperature / s.mass);
    d.v0 = s.v0;
    d.x0 = s.x0;
    if (s.model == 1) {
        // apex of the two inner segments (GetXc, langevinDynamicsDoubleBarrier.h:1722-1732)
        for (int b = 0; b < 2; ++b) d.xb[b] = s.position[b];
        const double x1b = s.position[0], w1 = s.half_width[0], x2b = s.position[1], w2 = s.half_width[1];
        const double y1 = s.height[0][0], y2 = s.height[1][0];
        const double x1h = x1b + w1, x2l = x2b - w2;
        const double xc = fma(x1h * y1, x2b - x2l, x2l * y2 * (x1h - x1b)) / fma(y1, x2b - x2l, y2 * (x1h - x1b));
        const double yc = (x1h - xc) / (x1h - x1b) * y1;
        d.xc = xc;
        d.lo[0] = x1b - w1;
        d.hi[0] = x1b + w1;
        d.lo[1] = x2b - w2;
        d.hi[1] = x2b + w2;
        const double xb_xl[2] = {d.xb[0] - d.lo[0], d.xb[1] - d.lo[1]};
        const double xh_xb[2] = {d.hi[0] - d.xb[0], d.hi[1] - d.xb[1]};
        for (int st = 0; st < 2; ++st) {
            const double H0 = s.height[0][st], H1 = s.height[1][st];
            d.slope[0][st][0] = H0 / xb_xl[0];
            d.slope[0][st][1] = (yc - H0) / (xc - d.xb[0]);
            d.slope[1][st][0] = (H1 - yc) / (d.xb[1] - xc);
            d.slope[1][st][1] = 0.0 - H1 / xh_xb[1];
            d.icept[0][st][0] = 0.0 - H0 * (d.lo[0] / xb_xl[0]);
            d.icept[0][st][1] = (xc * H0 - yc * d.xb[0]) / (xc - d.xb[0]);
            d.icept[1][st][0] = (yc * d.xb[1] - xc * H1) / (d.xb[1] - xc);
            d.icept[1][st][1] = H1 * (d.hi[1] / xh_xb[1]);
        }
    }
    // FP32 copies
    d.fL = (float)d.L;
    d.fhalfL = (float)d.halfL;
    d.fw = (float)d.w;
    d.fS[0] = (float)d.S[0];
    d.fS[1] = (float)d.S[1];
    d.fdtgm = (float)d.dtgm;
    d.fvar = (float)d.var;
    d.fbm = (float)(-1.3862943611198906 * d.var * d.var);
    d.fminv_halfdt = (float)(d.minv * d.halfdt);
    d.fhalfdt = (float)d.halfdt;
    d.fexpmgdt = (float)d.expmgdt;
    d.fc_lap = (float)d.c_lap;
    d.fK = (float)d.K;
    for (int b = 0; b < 2; ++b) {
        d.fxb[b] = (float)d.xb[b];
        d.flo[b] = (float)d.lo[b];
        d.fhi[b] = (float)d.hi[b];
        for (int st = 0; st < 2; ++st)
            for (int side = 0; side < 2; ++side) d.fslope[b][st][side] = (float)d.slope[b][st][side];
    }
    d.fxc = (float)d.xc;
    d.T_flip = flip_threshold(d.p_flip);
    d.T_asym[0] = flip_threshold(d.p_asym[0]);
    d.T_asym[1] = flip_threshold(d.p_asym[1]);
    d.T_create = flip_threshold(d.p_create);
    for (int k = 0; k < 2; ++k) {
        const uint64_t T = s.equal_probability ? d.T_flip : d.T_asym[k];
        d.Tall[k] = T >= (1ull << 32) ? 1 : 0;
        d.Tl[k] = d.Tall[k] ? 0u : (uint32_t)T;
    }
    if (s.randomize_state) {
        // u <= p0  <=>  w <= floor(p0 2^32 - 1/2)
        const double P = floor(d.p0 * 4294967296.0 - 0.5) + 1.0;
        d.T_p0le = P < 0.0 ? 0 : (P > 4294967296.0 ? 4294967296ull : (uint64_t)P);
    }
    d.fixed_state = (s.initial_state == 1.0) ? 1 : 0;
    d.key0 = (uint32_t)s.seed;
    d.key1 = (uint32_t)(s.seed >> 32);
    const double r_ref = s.x0 * d.Linv;
    d.cell0 = (int)rint(r_ref);
    d.well0 = (int)floor(r_ref);
    d.fxp0 = (float)(s.x0 - (double)d.cell0 * s.periodic_length);
    if (d.target == 0) d.slot_base = d.well0;
    else d.slot_base = (d.well0 == d.cell0) ? d.cell0 : d.cell0 - 1;
    d.n_traj = s.n_traj;
    d.max_steps = s.max_steps;
    d.first = first;
    d.start = start;
    d.hist_lo = log10(s.timestep);
    const double hi = log10(s.timestep * (double)s.max_steps) + 1e-9;
    d.hist_inv = LDPCP_HIST_BINS / (hi - d.hist_lo);
    // bin of absorption step k: floor((log10(dt k) - lo) * inv), clamped to
    // [0, 255]; it is non-decreasing in k, so each bin is a step range and the
    // device compares the integer step against the range starts
    const double lo = d.hist_lo, inv = d.hist_inv;
    auto bin_of = [&](uint64_t k) {
        const double b = floor((log10(s.timestep * (double)k) - lo) * inv);
        return b < 0.0 ? 0 : (b > LDPCP_HIST_BINS - 1 ? LDPCP_HIST_BINS - 1 : (int)b);
    };
    d.hist_edge[0] = 0;
    const int top = bin_of(s.max_steps);
    for (int b = 1; b < LDPCP_HIST_BINS; ++b) {
        if (b > top) {
            d.hist_edge[b] = 0xffffffffu;
            continue;
        }
        uint64_t a = 1, z = s.max_steps;  // smallest k in [1, max_steps] with bin_of(k) >= b
        while (a < z) {
            const uint64_t m = a + (z - a) / 2;
            if (bin_of(m) >= b) z = m;
            else a = m + 1;
        }
        d.hist_edge[b] = (uint32_t)a;
    }
    // 64 log-spaced lag steps in [1, max_steps], strictly increasing
    uint32_t prev = 0;
    for (int k = 0; k < LDPCP_MSD_LAGS; ++k) {
        const double f = (double)k / (LDPCP_MSD_LAGS - 1);
        uint64_t st = (uint64_t)llround(pow(10.0, f * log10((double)s.max_steps)));
        if (st <= prev) st = (uint64_t)prev + 1;
        d.lag[k] = st > s.max_steps ? kNoLag : (uint32_t)st;
        if (d.lag[k] != kNoLag) prev = d.lag[k];
        else prev = kNoLag - 1;
    }
}

Engine::~Engine()
{
    if (stream_) (void)hipStreamSynchronize(stream_);
    drop_graphs();
    for (auto& e : ev_pending_) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (auto& e : ev_free_) {
        (void)hipEventDestroy(e.first);
        (void)hipEventDestroy(e.second);
    }
    for (int i = 0; i < 2; ++i) {
        if (ev_cnt_[i]) (void)hipEventDestroy(ev_cnt_[i]);
        if (h_ctrl_[i]) (void)hipHostFree(h_ctrl_[i]);
        if (h_cnt_[i]) (void)hipHostFree(h_cnt_[i]);
    }
    if (h_sets_) (void)hipHostFree(h_sets_);
    if (d_block_) (void)hipFree(d_block_);
    if (stream_) (void)hipStreamDestroy(stream_);
}

int Engine::init(int precision, uint64_t capacity, int max_sets)
{
    if ((precision != 32 && precision != 64) || capacity == 0 || capacity >= (1ull << 32) - 64 || max_sets <= 0) {
        set_error("engine: precision 32 or 64, 0 < capacity < 2^32 - 64 and max_sets > 0 required");
        return LDPCP_EINVAL;
    }
    precision_ = precision;
    capacity_ = capacity;
    max_sets_ = max_sets;
    cfg_.block = 256;
    cfg_.blocks_per_cu = 8;
    cfg_.steps_per_launch = 4096;
    cfg_.use_graph = 1;
    int dev = 0;
    LDPCP_HIP_TRY(hipGetDevice(&dev));
    LDPCP_HIP_TRY(hipDeviceGetAttribute(&n_cu_, hipDeviceAttributeMultiprocessorCount, dev));
    const uint64_t n = capacity, nc = (capacity + 63) / 64 + 1, S = (uint64_t)max_sets;
    size_t bytes = 0;
    auto take = [&](size_t b) {
        const size_t o = bytes;
        bytes += (b + 255) & ~(size_t)255;
        return o;
    };
    const bool f32 = precision == 32;
    const size_t o_xp = f32 ? take(n * 4) : 0, o_comp = f32 ? take(n * 4) : 0, o_fv = f32 ? take(n * 4) : 0,
                 o_fR = f32 ? take(n * 4) : 0, o_cell = f32 ? take(n * 4) : 0;
    const size_t o_x = f32 ? 0 : take(n * 8), o_v = f32 ? 0 : take(n * 8), o_R = f32 ? 0 : take(n * 8);
    const size_t o_flags = take(n * 4), o_ts = take(n * 16), o_nc = take(n * 4), o_tpt = take(n * 4),
                 o_xf = take(n * 4), o_xl = take(n * 4), o_abs = take(n * 4), o_set = take(n * 4);
    const size_t o_live0 = take((n + 64) * 4), o_live1 = take((n + 64) * 4), o_mask = take(nc * 8),
                 o_off = take(nc * 4);
    const size_t o_sets = take(sizeof(SetDev) * S), o_hist = take(8 * S * LDPCP_HIST_BINS),
                 o_sums = take(8 * S * kSumsPerSet), o_cnt = take(8 * S * kCntPerSet),
                 o_msd = take(8 * S * LDPCP_MSD_LAGS), o_msdn = take(8 * S * LDPCP_MSD_LAGS), o_ctrl = take(sizeof(Ctrl));
    hipError_t e = hipMalloc(&d_block_, bytes);
    if (e != hipSuccess) {
        d_block_ = nullptr;
        set_error("engine: hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        return LDPCP_ENOMEM;
    }
    char* b = (char*)d_block_;
    memset(&P_, 0, sizeof(P_));
    if (f32) {
        P_.xp = (float*)(b + o_xp);
        P_.comp = (float*)(b + o_comp);
        P_.fv = (float*)(b + o_fv);
        P_.fR = (float*)(b + o_fR);
        P_.cell = (int32_t*)(b + o_cell);
    } else {
        P_.x = (double*)(b + o_x);
        P_.v = (double*)(b + o_v);
        P_.R = (double*)(b + o_R);
    }
    P_.flags = (uint32_t*)(b + o_flags);
    P_.ts = (uint4*)(b + o_ts);
    P_.n_catch = (uint32_t*)(b + o_nc);
    P_.tpt0 = (uint32_t*)(b + o_tpt);
    P_.xc_first = (uint32_t*)(b + o_xf);
    P_.xc_last = (uint32_t*)(b + o_xl);
    P_.absorbed = (uint32_t*)(b + o_abs);
    P_.set_of = (uint32_t*)(b + o_set);
    d_live_[0] = (uint32_t*)(b + o_live0);
    d_live_[1] = (uint32_t*)(b + o_live1);
    d_mask_ = (unsigned long long*)(b + o_mask);
    d_off_ = (uint32_t*)(b + o_off);
    d_sets_ = (SetDev*)(b + o_sets);
    St_.hist = (unsigned long long*)(b + o_hist);
    St_.sums = (unsigned long long*)(b + o_sums);
    St_.cnt = (unsigned long long*)(b + o_cnt);
    St_.msd_sum = (double*)(b + o_msd);
    St_.msd_n = (unsigned long long*)(b + o_msdn);
    d_ctrl_ = (Ctrl*)(b + o_ctrl);
    LDPCP_HIP_TRY(hipStreamCreateWithFlags(&stream_, hipStreamNonBlocking));
    LDPCP_HIP_TRY(hipHostMalloc((void**)&h_sets_, sizeof(SetDev) * S, hipHostMallocDefault));
    for (int i = 0; i < 2; ++i) {
        LDPCP_HIP_TRY(hipHostMalloc((void**)&h_ctrl_[i], sizeof(Ctrl), hipHostMallocDefault));
        LDPCP_HIP_TRY(hipHostMalloc((void**)&h_cnt_[i], 8 * S * kCntPerSet, hipHostMallocDefault));
        memset(h_ctrl_[i], 0, sizeof(Ctrl));
        memset(h_cnt_[i], 0, 8 * S * kCntPerSet);
        LDPCP_HIP_TRY(hipEventCreateWithFlags(&ev_cnt_[i], hipEventDisableTiming));
    }
    hipLaunchKernelGGL(engine_zero_ctrl, dim3(1), dim3(1), 0, stream_, d_ctrl_);
    LDPCP_HIP_TRY(hipGetLastError());
    LDPCP_HIP_TRY(hipStreamSynchronize(stream_));
    h_params_.assign(S, ldpcp_sweep_set{});
    first_.assign(S, 0);
    free_ids_.clear();
    for (int i = max_sets - 1; i >= 0; --i) free_ids_.push_back(i);
    slots_.reset(capacity);
    parity_ = 0;
    estep_ = 0;
    return LDPCP_OK;
}

int Engine::set_config(const ldpcp_engine_config& c)
{
    ldpcp_engine_config n = c;
    if (n.block == 0) n.block = 256;
    if (n.blocks_per_cu == 0) n.blocks_per_cu = 8;
    if (n.steps_per_launch == 0) n.steps_per_launch = 4096;
    if (n.use_graph == 0) n.use_graph = 1;
    if ((n.block != 64 && n.block != 128 && n.block != 256) || n.blocks_per_cu < 1 || n.blocks_per_cu > 64 ||
        n.steps_per_launch > (1u << 20) || (n.use_graph != 1 && n.use_graph != -1)) {
        set_error("engine config: block 64/128/256, 1 <= blocks_per_cu <= 64, steps_per_launch <= 2^20, "
                  "use_graph 1 or -1");
        return LDPCP_EINVAL;
    }
    if (n.block != cfg_.block || n.blocks_per_cu != cfg_.blocks_per_cu) drop_graphs();
    cfg_ = n;
    return LDPCP_OK;
}

int Engine::admit(const ldpcp_sweep_set& s, int* set_id)
{
    *set_id = -1;
    const int rc = validate(s);
    if (rc) return rc;
    if (free_ids_.empty()) return LDPCP_OK;
    uint64_t first = 0;
    if (!slots_.alloc(s.n_traj, &first)) return LDPCP_OK;
    const int id = free_ids_.back();
    free_ids_.pop_back();
    fill_set_dev(s, first, estep_, h_sets_[id]);
    h_params_[id] = s;
    first_[id] = first;
    LDPCP_HIP_TRY(hipMemcpyAsync(d_sets_ + id, h_sets_ + id, sizeof(SetDev), hipMemcpyHostToDevice, stream_));
    hipLaunchKernelGGL(engine_zero_set, dim3(1), dim3(256), 0, stream_, St_, id);
    LDPCP_HIP_TRY(hipGetLastError());
    const uint32_t n = (uint32_t)s.n_traj;
    hipLaunchKernelGGL(engine_admit, dim3((n + 255) / 256), dim3(256), 0, stream_, P_, (const SetDev*)d_sets_, id,
                       (const Ctrl*)d_ctrl_, d_live_[parity_], parity_, n, precision_ == 32 ? 1 : 0);
    LDPCP_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(engine_bump, dim3(1), dim3(1), 0, stream_, d_ctrl_, parity_, n);
    LDPCP_HIP_TRY(hipGetLastError());
    *set_id = id;
    return LDPCP_OK;
}

void Engine::release(int id)
{
    if (id < 0 || id >= max_sets_) return;
    slots_.release(first_[id], h_params_[id].n_traj);
    h_params_[id].n_traj = 0;
    free_ids_.push_back(id);
}

static unsigned scatter_grid(int n_cu) { return (unsigned)std::max(1, n_cu * 4); }

int Engine::launch_plain(uint32_t K)
{
    const dim3 grid((unsigned)(n_cu_ * cfg_.blocks_per_cu)), block((unsigned)cfg_.block);
    const size_t lds = (size_t)(cfg_.block / 64) * kLdsPerWave;
    if (precision_ == 32)
        hipLaunchKernelGGL(engine_step<true>, grid, block, lds, stream_, P_, St_, (const SetDev*)d_sets_, d_ctrl_,
                           (const uint32_t*)d_live_[parity_], d_mask_, parity_, K);
    else
        hipLaunchKernelGGL(engine_step<false>, grid, block, lds, stream_, P_, St_, (const SetDev*)d_sets_, d_ctrl_,
                           (const uint32_t*)d_live_[parity_], d_mask_, parity_, K);
    LDPCP_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(engine_scan, dim3(1), dim3(1024), 0, stream_, d_ctrl_, (const unsigned long long*)d_mask_,
                       d_off_, parity_, K);
    LDPCP_HIP_TRY(hipGetLastError());
    hipLaunchKernelGGL(engine_scatter, dim3(scatter_grid(n_cu_)), dim3(256), 0, stream_, (const Ctrl*)d_ctrl_,
                       (const unsigned long long*)d_mask_, (const uint32_t*)d_off_, (const uint32_t*)d_live_[parity_],
                       d_live_[parity_ ^ 1], parity_);
    LDPCP_HIP_TRY(hipGetLastError());
    return LDPCP_OK;
}

void Engine::drop_graphs()
{
    for (auto& g : graphs_) (void)hipGraphExecDestroy(g.second);
    graphs_.clear();
    for (auto g : graph_defs_) (void)hipGraphDestroy(g);
    graph_defs_.clear();
}

int Engine::graph_for(uint32_t K, hipGraphExec_t* exec)
{
    const uint64_t key = ((uint64_t)K << 1) | (uint64_t)parity_;
    auto it = graphs_.find(key);
    if (it != graphs_.end()) {
        *exec = it->second;
        return LDPCP_OK;
    }
    hipGraph_t g = nullptr;
    LDPCP_HIP_TRY(hipStreamBeginCapture(stream_, hipStreamCaptureModeThreadLocal));
    const int rc = launch_plain(K);
    const hipError_t ec = hipStreamEndCapture(stream_, &g);
    if (rc) return rc;
    LDPCP_HIP_TRY(ec);
    hipGraphExec_t x = nullptr;
    LDPCP_HIP_TRY(hipGraphInstantiate(&x, g, nullptr, nullptr, 0));
    graph_defs_.push_back(g);
    graphs_[key] = x;
    *exec = x;
    return LDPCP_OK;
}

int Engine::launch(uint32_t K)
{
    if (K == 0) K = cfg_.steps_per_launch;
    std::pair<hipEvent_t, hipEvent_t> ev;
    if (!ev_free_.empty()) {
        ev = ev_free_.back();
        ev_free_.pop_back();
    } else {
        LDPCP_HIP_TRY(hipEventCreate(&ev.first));
        LDPCP_HIP_TRY(hipEventCreate(&ev.second));
    }
    LDPCP_HIP_TRY(hipEventRecord(ev.first, stream_));
    if (cfg_.use_graph == 1) {
        hipGraphExec_t x = nullptr;
        const int rc = graph_for(K, &x);
        if (rc) return rc;
        LDPCP_HIP_TRY(hipGraphLaunch(x, stream_));
    } else {
        const int rc = launch_plain(K);
        if (rc) return rc;
    }
    LDPCP_HIP_TRY(hipEventRecord(ev.second, stream_));
    ev_pending_.push_back(ev);
    parity_ ^= 1;
    estep_ += K;
    return collect_times();
}

// reads the durations of completed launches without waiting
int Engine::collect_times()
{
    size_t k = 0;
    for (; k < ev_pending_.size(); ++k) {
        const hipError_t q = hipEventQuery(ev_pending_[k].second);
        if (q == hipErrorNotReady) break;
        LDPCP_HIP_TRY(q);
        float ms = 0.0f;
        LDPCP_HIP_TRY(hipEventElapsedTime(&ms, ev_pending_[k].first, ev_pending_[k].second));
        kernel_ms_ += ms;
        timed_launches_++;
        ev_free_.push_back(ev_pending_[k]);
    }
    ev_pending_.erase(ev_pending_.begin(), ev_pending_.begin() + (long)k);
    return LDPCP_OK;
}

int Engine::sync()
{
    LDPCP_HIP_TRY(hipStreamSynchronize(stream_));
    return collect_times();
}

int Engine::kernel_time(double* ms, uint64_t* launches)
{
    const int rc = sync();
    if (rc) return rc;
    if (ms) *ms = kernel_ms_;
    if (launches) *launches = timed_launches_;
    return LDPCP_OK;
}

int Engine::enqueue_counters(int slot)
{
    LDPCP_HIP_TRY(hipMemcpyAsync(h_ctrl_[slot], d_ctrl_, sizeof(Ctrl), hipMemcpyDeviceToHost, stream_));
    LDPCP_HIP_TRY(hipMemcpyAsync(h_cnt_[slot], St_.cnt, 8ull * max_sets_ * kCntPerSet, hipMemcpyDeviceToHost, stream_));
    LDPCP_HIP_TRY(hipEventRecord(ev_cnt_[slot], stream_));
    return LDPCP_OK;
}

int Engine::wait_counters(int slot)
{
    LDPCP_HIP_TRY(hipEventSynchronize(ev_cnt_[slot]));
    return collect_times();
}

int Engine::live_count(uint64_t* live)
{
    int rc = enqueue_counters(0);
    if (rc) return rc;
    rc = wait_counters(0);
    if (rc) return rc;
    *live = h_ctrl_[0]->n_live[parity_];
    return LDPCP_OK;
}

int Engine::stats(int id, ldpcp_set_stats* out, hipStream_t s)
{
    if (id < 0 || id >= max_sets_ || h_params_[id].n_traj == 0) {
        set_error("engine: no resident set %d", id);
        return LDPCP_EINVAL;
    }
    const hipStream_t st = s ? s : stream_;
    memset(out, 0, sizeof(*out));
    unsigned long long sums[kSumsPerSet];
    unsigned long long cnt[kCntPerSet];
    LDPCP_HIP_TRY(hipMemcpyAsync(sums, St_.sums + (uint64_t)id * kSumsPerSet, sizeof(sums), hipMemcpyDeviceToHost, st));
    LDPCP_HIP_TRY(hipMemcpyAsync(cnt, St_.cnt + (uint64_t)id * kCntPerSet, sizeof(cnt), hipMemcpyDeviceToHost, st));
    LDPCP_HIP_TRY(hipMemcpyAsync(out->hist, St_.hist + (uint64_t)id * LDPCP_HIST_BINS, 8ull * LDPCP_HIST_BINS,
                                 hipMemcpyDeviceToHost, st));
    LDPCP_HIP_TRY(hipMemcpyAsync(out->msd_sum, St_.msd_sum + (uint64_t)id * LDPCP_MSD_LAGS, 8ull * LDPCP_MSD_LAGS,
                                 hipMemcpyDeviceToHost, st));
    LDPCP_HIP_TRY(hipMemcpyAsync(out->msd_n, St_.msd_n + (uint64_t)id * LDPCP_MSD_LAGS, 8ull * LDPCP_MSD_LAGS,
                                 hipMemcpyDeviceToHost, st));
    if (s) {
        LDPCP_HIP_TRY(hipStreamSynchronize(s));
    } else {
        const int rc = sync();
        if (rc) return rc;
    }
    const SetDev& d = h_sets_[id];
    out->n_traj = h_params_[id].n_traj;
    // the integer step sums in time units: dt * sum and dt^2 * sum of squares
    const double dt = d.dt;
    auto sum1 = [&](int k) { return dt * (double)sums[3 * k]; };
    auto sum2 = [&](int k) {
        return dt * dt * ((double)sums[3 * k + 2] * 18446744073709551616.0 + (double)sums[3 * k + 1]);
    };
    out->fpt_sum = sum1(0);
    out->fpt_sum2 = sum2(0);
    out->tpt_sum = sum1(1);
    out->tpt_sum2 = sum2(1);
    out->tptb_sum[0] = sum1(2);
    out->tptb_sum2[0] = sum2(2);
    out->tptb_sum[1] = sum1(3);
    out->tptb_sum2[1] = sum2(3);
    out->n_absorbed = cnt[0];
    out->n_finished = cnt[1];
    out->n_capped = cnt[1] - cnt[0];
    out->traj_steps = cnt[2];
    out->hist_lo = d.hist_lo;
    out->hist_hi = d.hist_lo + LDPCP_HIST_BINS / d.hist_inv;
    for (int k = 0; k < LDPCP_MSD_LAGS; ++k) out->msd_step[k] = d.lag[k] == kNoLag ? 0 : d.lag[k];
    out->hist_step[0] = 1;
    for (int k = 1; k < LDPCP_HIST_BINS; ++k) out->hist_step[k] = d.hist_edge[k] == 0xffffffffu ? 0 : d.hist_edge[k];
    return LDPCP_OK;
}

int Engine::per_traj(int id, uint64_t* absorbed, uint64_t* tpt0, uint64_t* xc_first, uint64_t* xc_last,
                     hipStream_t s)
{
    if (id < 0 || id >= max_sets_ || h_params_[id].n_traj == 0 || !absorbed) {
        set_error("engine: no resident set %d or no output", id);
        return LDPCP_EINVAL;
    }
    const uint64_t n = h_params_[id].n_traj, first = first_[id];
    std::vector<uint32_t> buf(n);
    uint32_t* src[4] = {P_.absorbed, P_.tpt0, P_.xc_first, P_.xc_last};
    uint64_t* dst[4] = {absorbed, tpt0, xc_first, xc_last};
    for (int k = 0; k < 4; ++k) {
        if (!dst[k]) continue;
        // as in stats(): another stream reads a set whose trajectories have
        // all finished without waiting for the launches in flight
        LDPCP_HIP_TRY(hipMemcpyAsync(buf.data(), src[k] + first, n * 4, hipMemcpyDeviceToHost, s ? s : stream_));
        if (s) {
            LDPCP_HIP_TRY(hipStreamSynchronize(s));
        } else {
            const int rc = sync();
            if (rc) return rc;
        }
        for (uint64_t i = 0; i < n; ++i) dst[k][i] = buf[i];
    }
    return LDPCP_OK;
}

// Autotuner.  Launch shapes do not change any result (each trajectory's
// noise and arithmetic are fixed by its set, index and step), so candidates
// are timed on the engine's own resident trajectories: first the workgroup
// shapes at one launch length, then the launch lengths with the best shape.
// Each candidate gets one launch to build its graph, then one timed launch;
// the rate is trajectory-steps per second of wall time.
int Engine::tune(uint32_t steps_hint, ldpcp_engine_config* chosen)
{
    const ldpcp_engine_config saved = cfg_;
    auto steps_now = [&](uint64_t* total, uint64_t* live) -> int {
        int rc = enqueue_counters(0);
        if (rc) return rc;
        rc = wait_counters(0);
        if (rc) return rc;
        uint64_t t = 0;
        for (int i = 0; i < max_sets_; ++i)
            if (h_params_[i].n_traj) t += h_cnt_[0][(uint64_t)i * kCntPerSet + 2];
        *total = t;
        *live = h_ctrl_[0]->n_live[parity_];
        return LDPCP_OK;
    };
    // rate of config c; *live = 0 when nothing is left to time
    auto rate_of = [&](const ldpcp_engine_config& c, double* rate, uint64_t* live) -> int {
        int rc = set_config(c);
        if (rc) return rc;
        rc = launch(c.steps_per_launch);
        if (rc) return rc;
        uint64_t t0 = 0, t1 = 0;
        rc = steps_now(&t0, live);
        if (rc || *live == 0) return rc;
        const auto w0 = std::chrono::steady_clock::now();
        rc = launch(c.steps_per_launch);
        if (rc) return rc;
        rc = steps_now(&t1, live);
        if (rc) return rc;
        const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - w0).count();
        *rate = (double)(t1 - t0) / std::max(sec, 1e-9);
        return LDPCP_OK;
    };
    // workgroup size x persistent workgroups per CU
    const int shapes[][2] = {{256, 2}, {256, 4}, {256, 8}, {128, 8}, {128, 16}, {64, 16}, {64, 32}};
    double best = -1.0;
    ldpcp_engine_config best_cfg = saved;
    uint64_t live = 1;
    const uint32_t K0 = steps_hint ? steps_hint : 2048;
    for (const auto& sh : shapes) {
        ldpcp_engine_config c = saved;
        c.block = sh[0];
        c.blocks_per_cu = sh[1];
        c.steps_per_launch = K0;
        double r = -1.0;
        const int rc = rate_of(c, &r, &live);
        if (rc) return rc;
        if (live == 0) break;
        if (r > best) {
            best = r;
            best_cfg = c;
        }
    }
    if (!steps_hint && live) {
        best = -1.0;
        const ldpcp_engine_config shape = best_cfg;
        for (uint32_t K : {2048u, 4096u, 8192u}) {
            ldpcp_engine_config c = shape;
            c.steps_per_launch = K;
            double r = -1.0;
            const int rc = rate_of(c, &r, &live);
            if (rc) return rc;
            if (live == 0) break;
            if (r > best) {
                best = r;
                best_cfg = c;
            }
        }
    }
    const int rc = set_config(best_cfg);
    if (rc) return rc;
    if (chosen) *chosen = cfg_;
    return LDPCP_OK;
}

}  // namespace ldpcp

// ---------------------------------------------------------------------------
// C ABI: static sweep
// ---------------------------------------------------------------------------
using namespace ldpcp;

struct ldpcp_sweep {
    Engine eng;
    std::vector<int> ids;  // set index -> engine set id
    uint64_t live = 0;
};

extern "C" int ldpcp_sweep_create(const ldpcp_sweep_set* sets, int32_t nsets, int32_t precision, ldpcp_sweep** out)
{
    if (!out || !sets || nsets <= 0 || (precision != 32 && precision != 64)) {
        set_error("ldpcp_sweep_create: need sets, nsets > 0 and precision 32 or 64");
        return LDPCP_EINVAL;
    }
    *out = nullptr;
    uint64_t total = 0;
    for (int i = 0; i < nsets; ++i) {
        if (Engine::validate(sets[i])) {
            char msg[512];
            snprintf(msg, sizeof(msg), "%s", ldpcp_last_error());
            set_error("ldpcp_sweep_create: set %d: %s", i, msg);
            return LDPCP_EINVAL;
        }
        total += sets[i].n_traj;
    }
    if (total >= (1ull << 32) - 64) {
        set_error("ldpcp_sweep_create: at most 2^32-65 trajectories per sweep");
        return LDPCP_EINVAL;
    }
    ldpcp_sweep* h = new ldpcp_sweep();
    int rc = h->eng.init(precision, total, nsets);
    if (rc) {
        delete h;
        return rc;
    }
    for (int i = 0; i < nsets; ++i) {
        int id = -1;
        rc = h->eng.admit(sets[i], &id);
        if (rc || id < 0) {
            if (!rc) set_error("ldpcp_sweep_create: could not admit set %d", i);
            delete h;
            return rc ? rc : LDPCP_ENOMEM;
        }
        h->ids.push_back(id);
    }
    rc = h->eng.sync();
    if (rc) {
        delete h;
        return rc;
    }
    h->live = total;
    *out = h;
    return LDPCP_OK;
}

extern "C" int ldpcp_sweep_set_config(ldpcp_sweep* h, const ldpcp_engine_config* cfg)
{
    if (!h || !cfg) {
        set_error("ldpcp_sweep_set_config: null handle or config");
        return LDPCP_EINVAL;
    }
    return h->eng.set_config(*cfg);
}

extern "C" int ldpcp_sweep_get_config(ldpcp_sweep* h, ldpcp_engine_config* cfg)
{
    if (!h || !cfg) {
        set_error("ldpcp_sweep_get_config: null handle or config");
        return LDPCP_EINVAL;
    }
    *cfg = h->eng.config();
    return LDPCP_OK;
}

extern "C" int ldpcp_sweep_tune(ldpcp_sweep* h, ldpcp_engine_config* chosen)
{
    if (!h) {
        set_error("ldpcp_sweep_tune: null handle");
        return LDPCP_EINVAL;
    }
    int rc = h->eng.tune(0, chosen);
    if (rc) return rc;
    return h->eng.live_count(&h->live);
}

extern "C" int ldpcp_sweep_launch(ldpcp_sweep* h, uint32_t steps_per_launch, uint64_t* live)
{
    if (!h) {
        set_error("ldpcp_sweep_launch: null handle");
        return LDPCP_EINVAL;
    }
    if (h->live == 0) {
        if (live) *live = 0;
        return LDPCP_OK;
    }
    int rc = h->eng.launch(steps_per_launch);
    if (rc) return rc;
    rc = h->eng.live_count(&h->live);
    if (rc) return rc;
    if (live) *live = h->live;
    return LDPCP_OK;
}

extern "C" int ldpcp_sweep_run(ldpcp_sweep* h, uint32_t steps_per_launch)
{
    uint64_t live = 1;
    while (live) {
        const int rc = ldpcp_sweep_launch(h, steps_per_launch, &live);
        if (rc) return rc;
    }
    return LDPCP_OK;
}

extern "C" int ldpcp_sweep_stats(ldpcp_sweep* h, int32_t set, ldpcp_set_stats* out)
{
    if (!h || !out || set < 0 || set >= (int)h->ids.size()) {
        set_error("ldpcp_sweep_stats: bad handle, set or output");
        return LDPCP_EINVAL;
    }
    return h->eng.stats(h->ids[set], out);
}

extern "C" int ldpcp_sweep_fpt(ldpcp_sweep* h, int32_t set, uint64_t* absorbed_step, uint64_t* tpt_start,
                               uint64_t* xc_first, uint64_t* xc_last)
{
    if (!h || set < 0 || set >= (int)h->ids.size() || !absorbed_step) {
        set_error("ldpcp_sweep_fpt: bad handle, set or output");
        return LDPCP_EINVAL;
    }
    return h->eng.per_traj(h->ids[set], absorbed_step, tpt_start, xc_first, xc_last);
}

extern "C" int ldpcp_sweep_kernel_time(ldpcp_sweep* h, double* ms_total, uint64_t* launches)
{
    if (!h) {
        set_error("ldpcp_sweep_kernel_time: null handle");
        return LDPCP_EINVAL;
    }
    return h->eng.kernel_time(ms_total, launches);
}

extern "C" int ldpcp_sweep_destroy(ldpcp_sweep* h)
{
    delete h;
    return LDPCP_OK;
}
