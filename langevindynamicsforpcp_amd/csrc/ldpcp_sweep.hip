// ldpcp_sweep.hip -- batched first-passage engine (the throughput path).
//
// Every trajectory of every parameter set starts at its set's X0 and is
// integrated until the first passage event or its step cap:
//   target 0: first barrier-top crossing (the reference's first FPT1 event,
//             langevinDynamicsSingleBarrier.h:2394-2411),
//   target 1: first arrival at a non-adjacent well centre (the reference's
//             first FPT2 event, :2445-2489; the double-barrier statistic,
//             langevinDynamicsDoubleBarrier.h:2974-3006).
// Before that event a trajectory can only sit on two periodic cells, so its
// lazily created barrier states (the reference's BarrierStatus table,
// langevinDynamicsSingleBarrier.h:66-118) fit in four register slots.
//
// Work layout: one thread per live trajectory.  Trajectory state is SoA in
// HBM and stays in VGPRs for a launch of K steps; all live trajectories share
// the same step count, so after each launch an order-preserving compaction
// (hipcub::DeviceSelect::Flagged) drops absorbed ones and keeps the live list
// sorted by parameter set.  Launch boundaries also fall on the 64 log-spaced
// MSD lag steps, where the launch epilogue samples (x - x0)^2 and reduces it
// across the wavefront before one atomic per wave and set.
//
// Two arithmetics share the same Philox noise (ldpcp_device.h):
//   FP64: the reference's operations on the absolute position (plain FP64);
//   FP32: position as (integer cell, FP32 offset from the cell's barrier)
//         with a Kahan-compensated offset update, FP32 force, noise and
//         velocity, integer clock; only the statistics sums are FP64.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>
#include <math.h>
#include <stdint.h>
#include <string.h>

#include <algorithm>
#include <vector>

#include "ldpcp_common.hpp"
#include "ldpcp_device.h"

namespace ldpcp {
namespace {

constexpr int kBlock = 256;

// per-set constants (FP64 and FP32 copies of IntegrateInitializeVariables)
struct SetDev {
    int32_t model, overdamped, target, equal_prob, randomize_state, randomize_v, tau_zero, pad0;
    double dt, L, Linv, halfL, w, S[2], H[2], dtgm, var, minv, halfdt, expmgdt, c_lap, K, p0, p_create, vsigma, v0;
    double xb[2], lo[2], hi[2], xc, slope[2][2][2], icept[2][2][2];
    double p_flip, p_asym[2];
    double x0;
    float fL, fhalfL, fw, fS[2], fdtgm, fvar, fminv_halfdt, fhalfdt, fexpmgdt, fc_lap, fK;
    float fxb[2], flo[2], fhi[2], fxc, fslope[2][2][2];
    float fxp0;
    uint64_t T_flip, T_asym[2], T_create, T_p0le;  // integer thresholds on Philox words
    int32_t fixed_state, pad2;                      // initial state when not randomised
    uint32_t key0, key1;
    int32_t cell0, slot_base;                       // cell of X0; first of the two cells (slot 0)
    int32_t well0, pad1;
    uint64_t n_traj, max_steps, first;              // first: global index of the set's trajectory 0
    double hist_lo, hist_inv;                       // log10 histogram edges
};

struct SweepPtrs {
    // FP32 state
    float *xp, *comp, *fv, *fR;
    int32_t* cell;
    // FP64 state
    double *x, *v, *R;
    // shared
    uint32_t *flags, *ts, *n_catch, *tpt_start, *absorbed;
    int32_t* set_of;
    const SetDev* sets;
    // statistics (per set)
    unsigned long long* hist;   // [set][256]
    double *fpt_sum, *fpt_sum2, *tpt_sum, *tpt_sum2;
    unsigned long long *n_absorbed, *traj_steps;
    double* msd_sum;            // [set][64]
    unsigned long long* msd_n;  // [set][64]
    uint8_t* alive_flag;        // per live-list entry, for compaction
};

// flags layout
constexpr uint32_t F_STATE = 1u << 0;
constexpr uint32_t F_ALIVE = 1u << 1;
constexpr uint32_t F_SLOT_VALID = 1u << 4;   // + slot (4 slots)
constexpr uint32_t F_SLOT_STATE = 1u << 8;   // + slot

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
__device__ __forceinline__ uint64_t uni(uint64_t v) { return ((uint64_t)uni((uint32_t)(v >> 32)) << 32) | uni((uint32_t)v); }
__device__ __forceinline__ double uni(double v) { return __longlong_as_double((long long)uni((uint64_t)__double_as_longlong(v))); }

template <bool F32, int MODEL, bool OVER>
__device__ __forceinline__ SetDev uniform_copy(const SetDev& src)
{
    SetDev s = src;
    s.target = uni(s.target);
    s.equal_prob = uni(s.equal_prob);
    s.tau_zero = uni(s.tau_zero);
    s.key0 = uni(s.key0);
    s.key1 = uni(s.key1);
    s.slot_base = uni(s.slot_base);
    s.max_steps = uni(s.max_steps);
    s.T_flip = uni(s.T_flip);
    if (F32) {
        s.fL = uni(s.fL);
        s.fhalfL = uni(s.fhalfL);
        if (OVER) {
            s.fdtgm = uni(s.fdtgm);
        } else {
            s.fminv_halfdt = uni(s.fminv_halfdt);
            s.fhalfdt = uni(s.fhalfdt);
            s.fexpmgdt = uni(s.fexpmgdt);
        }
        s.fvar = uni(s.fvar);
        if (MODEL == 0) {
            s.fw = uni(s.fw);
            s.fS[0] = uni(s.fS[0]);
            s.fS[1] = uni(s.fS[1]);
        } else {
            s.fxc = uni(s.fxc);
            for (int b = 0; b < 2; ++b) {
                s.fxb[b] = uni(s.fxb[b]);
                s.flo[b] = uni(s.flo[b]);
                s.fhi[b] = uni(s.fhi[b]);
                for (int k = 0; k < 2; ++k)
                    for (int side = 0; side < 2; ++side) s.fslope[b][k][side] = uni(s.fslope[b][k][side]);
            }
        }
    } else {
        s.L = uni(s.L);
        s.Linv = uni(s.Linv);
        if (OVER) {
            s.dtgm = uni(s.dtgm);
        } else {
            s.minv = uni(s.minv);
            s.halfdt = uni(s.halfdt);
            s.expmgdt = uni(s.expmgdt);
        }
        s.var = uni(s.var);
        if (MODEL == 0) {
            s.w = uni(s.w);
            s.S[0] = uni(s.S[0]);
            s.S[1] = uni(s.S[1]);
        } else {
            s.xc = uni(s.xc);
            for (int b = 0; b < 2; ++b) {
                s.xb[b] = uni(s.xb[b]);
                s.lo[b] = uni(s.lo[b]);
                s.hi[b] = uni(s.hi[b]);
            }
        }
    }
    return s;
}

__device__ __forceinline__ void record_absorption(const SetDev& S, const SweepPtrs& P, int set, uint32_t gi,
                                                  uint32_t step, uint32_t tpt0)
{
    P.absorbed[gi] = step;
    P.tpt_start[gi] = tpt0;
    const double fpt = S.dt * (double)step;
    const double tpt = S.dt * (double)(step - tpt0);
    atomicAdd(P.fpt_sum + set, fpt);
    atomicAdd(P.fpt_sum2 + set, fpt * fpt);
    atomicAdd(P.tpt_sum + set, tpt);
    atomicAdd(P.tpt_sum2 + set, tpt * tpt);
    atomicAdd(P.n_absorbed + set, 1ull);
    // bin on log10(fpt) = log2(fpt) * log10(2), in FP32 (only the bin index is kept)
    int bin = (int)((__builtin_amdgcn_logf((float)fpt) * 0.30102999566f - (float)S.hist_lo) * (float)S.hist_inv);
    bin = bin < 0 ? 0 : (bin >= LDPCP_HIST_BINS ? LDPCP_HIST_BINS - 1 : bin);
    atomicAdd(P.hist + (uint64_t)set * LDPCP_HIST_BINS + bin, 1ull);
}

// Reduce v over the lanes of a wavefront that share `set`, one atomic per set
// present in the wave (the live list is set-sorted, so normally one).
__device__ __forceinline__ void wave_add_by_set(int set, bool active, double v, double* base, int stride_off,
                                                unsigned long long* nbase)
{
    const int lane = __lane_id();
    uint64_t pending = __ballot(active);
    while (pending) {
        const int leader = __ffsll((long long)pending) - 1;
        const int lset = __shfl(set, leader);
        const bool mine = active && (set == lset);
        const uint64_t group = __ballot(mine);
        double s = mine ? v : 0.0;
#pragma unroll
        for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off);
        if (lane == leader) {
            atomicAdd(base + (uint64_t)lset * LDPCP_MSD_LAGS + stride_off, s);
            atomicAdd(nbase + (uint64_t)lset * LDPCP_MSD_LAGS + stride_off, (unsigned long long)__popcll(group));
        }
        pending &= ~group;
    }
}

// ---------------------------------------------------------------------------
// Initialisation (IntegrateStep0, langevinDynamicsSingleBarrier.h:1619-1652,
// with the Philox role-1 draws)
// ---------------------------------------------------------------------------
__global__ void __launch_bounds__(kBlock) sweep_init_kernel(SweepPtrs P, uint64_t n_total, int fp32)
{
    const uint64_t gi = (uint64_t)blockIdx.x * kBlock + threadIdx.x;
    if (gi >= n_total) return;
    const int set = P.set_of[gi];
    const SetDev& S = P.sets[set];
    const uint32_t tid = (uint32_t)(gi - S.first);
    const Philox4 w = philox4x32_10(0u, 0u, tid, 1u, S.key0, S.key1);
    uint32_t flags = F_ALIVE;
    // u <= p0 -> state 0 (Single.h:1888-1895)
    const int state = S.randomize_state ? ((uint64_t)w.w[2] < S.T_p0le ? 0 : 1) : S.fixed_state;
    if (state) flags |= F_STATE;
    // the barrier under X0 takes the initial state (Single.h:1906-1911)
    {
        const int local = (S.model == 1 && (S.x0 - (double)S.cell0 * S.L) >= S.xc) ? 1 : 0;
        const int k = slot_index(S.model, S.cell0, S.slot_base, local);
        if (k >= 0) flags |= (F_SLOT_VALID << k) | (state ? (F_SLOT_STATE << k) : 0u);
    }
    double gv, gr;
    box_muller_f64(w.w[0], w.w[1], gv, gr);
    const double v0 = S.randomize_v ? S.vsigma * gv : S.v0;
    const double R0 = S.overdamped ? S.var * gr : 0.0;
    if (fp32) {
        P.xp[gi] = S.fxp0;
        P.comp[gi] = 0.0f;
        P.cell[gi] = S.cell0;
        P.fv[gi] = (float)v0;
        // FP32 noise amplitude applied to the FP32 Gaussian, as in the step loop
        float g0, g1;
        box_muller_f32(w.w[0], w.w[1], g0, g1);
        P.fR[gi] = S.overdamped ? S.fvar * g1 : 0.0f;
        if (S.randomize_v) P.fv[gi] = (float)S.vsigma * g0;
    } else {
        P.x[gi] = S.x0;
        P.v[gi] = v0;
        P.R[gi] = R0;
    }
    P.flags[gi] = flags;
    P.n_catch[gi] = 0;
    P.tpt_start[gi] = 0;
    P.absorbed[gi] = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) P.ts[gi * 4 + k] = 0;
}

// catch-up flip on entering a barrier (Single.h:2039-2051): state of the
// barrier entered, aged by the steps since it was last left.
template <typename Real>
__device__ __forceinline__ int handover(const SetDev& S, uint32_t& flags, uint32_t* ts, int kp, int kc, int state,
                                        uint32_t step, uint32_t tid, uint32_t& n_catch)
{
    if (kp >= 0) {
        flags = (flags & ~(F_SLOT_STATE << kp)) | (F_SLOT_VALID << kp) | (state ? (F_SLOT_STATE << kp) : 0u);
        ts[kp] = step;
    }
    if (kc < 0) return state;  // leaving the two-cell window: absorbed this step
    int prev;
    uint32_t ts_c;
    if (flags & (F_SLOT_VALID << kc)) {
        prev = (flags >> (8 + kc)) & 1;
        ts_c = ts[kc];
    } else {
        // first touch: the state the barrier was created with, at step 0
        const int array = S.model == 0 ? 0 : (kc & 1);
        const int cell = S.slot_base + (S.model == 0 ? kc : (kc >> 1));
        const Philox4 w = philox4x32_10((uint32_t)cell, 0u, tid, 3u + 256u * (uint32_t)array, S.key0, S.key1);
        prev = below(w.w[0], S.T_create) ? 0 : 1;
        ts_c = 0;
        flags |= (F_SLOT_VALID << kc) | (prev ? (F_SLOT_STATE << kc) : 0u);
        ts[kc] = 0;
    }
    if (S.tau_zero) return prev;
    const Philox4 w = philox4x32_10(n_catch, 0u, tid, 2u, S.key0, S.key1);
    n_catch++;
    const Real lap = (Real)(step - ts_c);
    Real pc;
    if (sizeof(Real) == 4) {
        const float e = __expf((float)S.fc_lap * (float)lap);
        const float scale = S.equal_prob ? 0.5f : (prev == 0 ? S.fK / (1.0f + S.fK) : 1.0f / (1.0f + S.fK));
        pc = (Real)(scale * (1.0f - e));
    } else {
        const double e = exp(S.c_lap * (double)lap);
        const double scale = S.equal_prob ? 0.5 : (prev == 0 ? S.K / (1.0 + S.K) : 1.0 / (1.0 + S.K));
        pc = (Real)(scale * (1.0 - e));
    }
    const bool flip = (Real)unit_f64(w.w[0]) < pc;
    return flip ? (prev ^ 1) : prev;
}

__device__ __forceinline__ int flip_step(const SetDev& S, int state, uint32_t word)
{
    if (S.tau_zero) return state;
    const uint64_t T = S.equal_prob ? S.T_flip : S.T_asym[state];
    return below(word, T) ? (state ^ 1) : state;
}

// ---------------------------------------------------------------------------
// FP32 stepping.  MODEL 0 single, 1 double barrier.
// ---------------------------------------------------------------------------
template <int MODEL>
__device__ __forceinline__ float force_f32(const SetDev& S, float xp, int state)
{
    if (MODEL == 0) {
        const float sp = S.fS[state];
        float F = 0.0f;
        F = (xp < 0.0f && xp > -S.fw) ? -sp : F;
        F = (xp < S.fw && xp > 0.0f) ? sp : F;
        return F;
    } else {
        float s = 0.0f;
        s = (xp < S.fxb[0] && xp > S.flo[0]) ? S.fslope[0][state][0] : s;
        s = (xp < S.fxc && xp > S.fxb[0]) ? S.fslope[0][state][1] : s;
        s = (xp < S.fxb[1] && xp > S.fxc) ? S.fslope[1][state][0] : s;
        s = (xp < S.fhi[1] && xp > S.fxb[1]) ? S.fslope[1][state][1] : s;
        return 0.0f - s;
    }
}

// One trajectory for K steps.  On set-uniform waves S is a copy whose hot
// fields were moved to scalar registers (uniform_copy); mixed waves read the
// set through a reference.
template <int MODEL, bool OVER, typename SetT>
__device__ __forceinline__ void step_f32_body(const SweepPtrs& P, SetT S, int set, uint32_t gi, uint32_t li,
                                              uint8_t* __restrict__ alive_out, uint32_t s0, uint32_t K,
                                              int sample_lag, bool& active, double& msd_v)
{
    {
        const uint32_t tid = (uint32_t)(gi - S.first);
        float xp = P.xp[gi], comp = P.comp[gi], v = P.fv[gi], R = P.fR[gi];
        int cell = P.cell[gi];
        uint32_t flags = P.flags[gi];
        uint32_t ts[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) ts[k] = P.ts[gi * 4 + k];
        uint32_t n_catch = P.n_catch[gi];
        uint32_t tpt0 = P.tpt_start[gi];
        int state = flags & F_STATE;
        int local = (MODEL == 1) ? (xp < S.fxc ? 0 : 1) : 0;
        float F = force_f32<MODEL>(S, xp, state);
        const uint32_t s_end = (s0 + K < (uint32_t)S.max_steps) ? s0 + K : (uint32_t)S.max_steps;
        float g_odd = 0.0f;
        uint32_t w_odd = 0;
        bool done = false;
        uint32_t step = s0;
        for (step = s0 + 1; step <= s_end; ++step) {
            float g;
            uint32_t wflip;
            if ((step & 1u) == 1u) {
                const uint32_t q = (step - 1) >> 1;
                const Philox4 w = philox4x32_10(q, 0u, tid, 0u, S.key0, S.key1);
                box_muller_f32(w.w[0], w.w[1], g, g_odd);
                wflip = w.w[2];
                w_odd = w.w[3];
            } else {
                if (step == s0 + 1) {
                    const uint32_t q = (step - 1) >> 1;
                    const Philox4 w = philox4x32_10(q, 0u, tid, 0u, S.key0, S.key1);
                    float g0;
                    box_muller_f32(w.w[0], w.w[1], g0, g_odd);
                    w_odd = w.w[3];
                }
                g = g_odd;
                wflip = w_odd;
            }
            // position increment (overdamped MA(1) scheme, Single.h:2860-2876;
            // underdamped B-A-O-A-B, :1680-1706), compensated offset update
            float dx;
            if (OVER) {
                const float rf = S.fvar * g;
                dx = __builtin_fmaf(F, S.fdtgm, rf + R);
                R = rf;
            } else {
                v = __builtin_fmaf(F, S.fminv_halfdt, v);
                const float d1 = v * S.fhalfdt;
                v = __builtin_fmaf(S.fexpmgdt, v, S.fvar * g);
                dx = __builtin_fmaf(v, S.fhalfdt, d1);
            }
            const float y = dx - comp;
            const float t = xp + y;
            comp = (t - xp) - y;
            const bool neg_before = xp < 0.0f;
            xp = t;
            // periodic (well-centre) and well (barrier-top) crossings
            bool ent_period = false;
            const int cell_prev = cell;
            if (xp >= S.fhalfL) {
                xp -= S.fL;
                cell += 1;
                ent_period = true;
            } else if (xp < -S.fhalfL) {
                xp += S.fL;
                cell -= 1;
                ent_period = true;
            }
            const bool ent_well = !ent_period && (neg_before != (xp < 0.0f));
            state = flip_step(S, state, wflip);
            if (MODEL == 0) {
                if (ent_period) {
                    const int kp = slot_index(0, cell_prev, S.slot_base, 0);
                    const int kc = slot_index(0, cell, S.slot_base, 0);
                    state = handover<float>(S, flags, ts, kp, kc, state, step, tid, n_catch);
                }
            } else {
                const int local_prev = local;
                local = xp < S.fxc ? 0 : 1;
                if (local != local_prev) {
                    const int kp = slot_index(1, cell_prev, S.slot_base, local_prev);
                    const int kc = slot_index(1, cell, S.slot_base, local);
                    state = handover<float>(S, flags, ts, kp, kc, state, step, tid, n_catch);
                }
            }
            F = force_f32<MODEL>(S, xp, state);
            if (!OVER) v = __builtin_fmaf(F, S.fminv_halfdt, v);
            const bool hit = (S.target == 0) ? ent_well : (ent_period && (cell - S.slot_base < 0 || cell - S.slot_base > 1));
            if (hit) {
                record_absorption(S, P, set, gi, step, tpt0);
                done = true;
                break;
            }
            if (ent_period) tpt0 = step;
        }
        const uint32_t last = done ? step : s_end;
        if (!done && s_end >= (uint32_t)S.max_steps) done = true;  // capped
        if (done) flags &= ~F_ALIVE;
        flags = (flags & ~F_STATE) | (uint32_t)state;
        P.xp[gi] = xp;
        P.comp[gi] = comp;
        P.fv[gi] = v;
        P.fR[gi] = R;
        P.cell[gi] = cell;
        P.flags[gi] = flags;
#pragma unroll
        for (int k = 0; k < 4; ++k) P.ts[gi * 4 + k] = ts[k];
        P.n_catch[gi] = n_catch;
        P.tpt_start[gi] = tpt0;
        alive_out[li] = (flags & F_ALIVE) ? 1 : 0;
        active = (sample_lag >= 0) && !(done && last != s_end) && (s_end == s0 + K);
        if (active) {
            const double x = (double)cell * (double)S.fL + (double)xp - (double)comp;
            msd_v = (x - S.x0) * (x - S.x0);
        }
    }
}

template <int MODEL, bool OVER>
__global__ void __launch_bounds__(kBlock) sweep_step_f32(SweepPtrs P, const SetDev* __restrict__ sets,
                                                         const uint32_t* __restrict__ live,
                                                         uint8_t* __restrict__ alive_out, uint32_t n_live,
                                                         uint32_t s0, uint32_t K, int sample_lag)
{
    const uint32_t li = blockIdx.x * kBlock + threadIdx.x;
    const bool in = li < n_live;
    const uint32_t gi = in ? live[li] : 0u;
    const int set = in ? P.set_of[gi] : -1;
    // lane 0 of a wave is live whenever any lane is (the live list is dense)
    // lane 0 of a wave is live whenever any lane is (the live list is dense)
    const int set0 = __builtin_amdgcn_readfirstlane(set);
    const bool uniform = __ballot(in && set != set0) == 0;
    bool active = false;
    double msd_v = 0.0;
    if (in) {
        // set-uniform waves (all but the few at set boundaries) keep the set's
        // constants in scalar registers
        if (uniform)
            step_f32_body<MODEL, OVER, const SetDev>(P, uniform_copy<true, MODEL, OVER>(sets[set0]), set, gi, li,
                                                      alive_out, s0, K, sample_lag, active, msd_v);
        else
            step_f32_body<MODEL, OVER, const SetDev&>(P, sets[set], set, gi, li, alive_out, s0, K, sample_lag, active,
                                                       msd_v);
    }
    if (sample_lag >= 0) wave_add_by_set(set, active, msd_v, P.msd_sum, sample_lag, P.msd_n);
}

// ---------------------------------------------------------------------------
// FP64 stepping: the reference arithmetic on the absolute position
// ---------------------------------------------------------------------------
template <int MODEL>
__device__ __forceinline__ double force_f64(const SetDev& S, double x, double periodic, int state)
{
    const double xp = x - periodic * S.L;
    if (MODEL == 0) {
        const double sp = S.S[state], sn = 0.0 - sp;
        double F = (xp < 0.0 && xp > -S.w) ? sn : 0.0;
        F = (xp < S.w && xp > 0.0) ? sp : F;
        return F;
    } else {
        double s = 0.0;
        s = (xp < S.xb[0] && xp > S.lo[0]) ? S.slope[0][state][0] : s;
        s = (xp < S.xc && xp > S.xb[0]) ? S.slope[0][state][1] : s;
        s = (xp < S.xb[1] && xp > S.xc) ? S.slope[1][state][0] : s;
        s = (xp < S.hi[1] && xp > S.xb[1]) ? S.slope[1][state][1] : s;
        return 0.0 - s;
    }
}

template <int MODEL, bool OVER, typename SetT>
__device__ __forceinline__ void step_f64_body(const SweepPtrs& P, SetT S, int set, uint32_t gi, uint32_t li,
                                              uint8_t* __restrict__ alive_out, uint32_t s0, uint32_t K,
                                              int sample_lag, bool& active, double& msd_v)
{
    {
        const uint32_t tid = (uint32_t)(gi - S.first);
        double x = P.x[gi], v = P.v[gi], R = P.R[gi];
        uint32_t flags = P.flags[gi];
        uint32_t ts[4];
#pragma unroll
        for (int k = 0; k < 4; ++k) ts[k] = P.ts[gi * 4 + k];
        uint32_t n_catch = P.n_catch[gi];
        uint32_t tpt0 = P.tpt_start[gi];
        int state = flags & F_STATE;
        double r = x * S.Linv;
        double well = floor(r), periodic = rint(r);
        int local = (MODEL == 1) ? ((x - periodic * S.L) < S.xc ? 0 : 1) : 0;
        double F = force_f64<MODEL>(S, x, periodic, state);
        const uint32_t s_end = (s0 + K < (uint32_t)S.max_steps) ? s0 + K : (uint32_t)S.max_steps;
        double g_odd = 0.0;
        uint32_t w_odd = 0;
        bool done = false;
        uint32_t step = s0;
        for (step = s0 + 1; step <= s_end; ++step) {
            double g;
            uint32_t wflip;
            if ((step & 1u) == 1u) {
                const uint32_t q = (step - 1) >> 1;
                const Philox4 w = philox4x32_10(q, 0u, tid, 0u, S.key0, S.key1);
                box_muller_f64(w.w[0], w.w[1], g, g_odd);
                wflip = w.w[2];
                w_odd = w.w[3];
            } else {
                if (step == s0 + 1) {
                    const uint32_t q = (step - 1) >> 1;
                    const Philox4 w = philox4x32_10(q, 0u, tid, 0u, S.key0, S.key1);
                    double g0;
                    box_muller_f64(w.w[0], w.w[1], g0, g_odd);
                    w_odd = w.w[3];
                }
                g = g_odd;
                wflip = w_odd;
            }
            if (OVER) {
                const double rf = S.var * g;
                x = fma(F, S.dtgm, x);
                x = x + rf;
                x = x + R;
                R = rf;
            } else {
                v = v + (S.minv * F) * S.halfdt;
                x = x + v * S.halfdt;
                v = S.expmgdt * v + S.var * g;
                x = x + v * S.halfdt;
            }
            r = x * S.Linv;
            const double well_prev = well, periodic_prev = periodic;
            well = floor(r);
            periodic = rint(r);
            const bool ent_well = well != well_prev;
            const bool ent_period = periodic != periodic_prev;
            state = flip_step(S, state, wflip);
            if (MODEL == 0) {
                if (ent_period) {
                    const int kp = slot_index(0, (int)periodic_prev, S.slot_base, 0);
                    const int kc = slot_index(0, (int)periodic, S.slot_base, 0);
                    state = handover<double>(S, flags, ts, kp, kc, state, step, tid, n_catch);
                }
            } else {
                const int local_prev = local;
                local = (x - periodic * S.L) < S.xc ? 0 : 1;
                if (local != local_prev) {
                    const int kp = slot_index(1, (int)periodic_prev, S.slot_base, local_prev);
                    const int kc = slot_index(1, (int)periodic, S.slot_base, local);
                    state = handover<double>(S, flags, ts, kp, kc, state, step, tid, n_catch);
                }
            }
            F = force_f64<MODEL>(S, x, periodic, state);
            if (!OVER) v = v + (S.minv * F) * S.halfdt;
            const int kcell = (int)periodic - S.slot_base;
            const bool hit = (S.target == 0) ? ent_well : (ent_period && (kcell < 0 || kcell > 1));
            if (hit) {
                record_absorption(S, P, set, gi, step, tpt0);
                done = true;
                break;
            }
            if (ent_period) tpt0 = step;
        }
        const uint32_t last = done ? step : s_end;
        if (!done && s_end >= (uint32_t)S.max_steps) done = true;
        if (done) flags &= ~F_ALIVE;
        flags = (flags & ~F_STATE) | (uint32_t)state;
        P.x[gi] = x;
        P.v[gi] = v;
        P.R[gi] = R;
        P.flags[gi] = flags;
#pragma unroll
        for (int k = 0; k < 4; ++k) P.ts[gi * 4 + k] = ts[k];
        P.n_catch[gi] = n_catch;
        P.tpt_start[gi] = tpt0;
        alive_out[li] = (flags & F_ALIVE) ? 1 : 0;
        active = (sample_lag >= 0) && !(done && last != s_end) && (s_end == s0 + K);
        if (active) msd_v = (x - S.x0) * (x - S.x0);
    }
}

template <int MODEL, bool OVER>
__global__ void __launch_bounds__(kBlock) sweep_step_f64(SweepPtrs P, const SetDev* __restrict__ sets,
                                                         const uint32_t* __restrict__ live,
                                                         uint8_t* __restrict__ alive_out, uint32_t n_live,
                                                         uint32_t s0, uint32_t K, int sample_lag)
{
    const uint32_t li = blockIdx.x * kBlock + threadIdx.x;
    const bool in = li < n_live;
    const uint32_t gi = in ? live[li] : 0u;
    const int set = in ? P.set_of[gi] : -1;
    // lane 0 of a wave is live whenever any lane is (the live list is dense)
    const int set0 = __builtin_amdgcn_readfirstlane(set);
    const bool uniform = __ballot(in && set != set0) == 0;
    bool active = false;
    double msd_v = 0.0;
    if (in) {
        // set-uniform waves (all but the few at set boundaries) keep the set's
        // constants in scalar registers
        if (uniform)
            step_f64_body<MODEL, OVER, const SetDev>(P, uniform_copy<false, MODEL, OVER>(sets[set0]), set, gi, li,
                                                      alive_out, s0, K, sample_lag, active, msd_v);
        else
            step_f64_body<MODEL, OVER, const SetDev&>(P, sets[set], set, gi, li, alive_out, s0, K, sample_lag, active,
                                                       msd_v);
    }
    if (sample_lag >= 0) wave_add_by_set(set, active, msd_v, P.msd_sum, sample_lag, P.msd_n);
}

__global__ void iota_kernel(uint32_t* out, uint32_t n, uint32_t base)
{
    const uint32_t i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) out[i] = base + i;
}

}  // namespace
}  // namespace ldpcp

using namespace ldpcp;

struct ldpcp_sweep {
    int precision;
    int nsets;
    uint64_t n_total;
    std::vector<ldpcp_sweep_set> host_sets;
    std::vector<SetDev> dev_sets;
    SweepPtrs P;
    void* block;
    // one live list per kind (model, overdamped)
    uint32_t* live[4];
    uint32_t* live_tmp[4];
    uint32_t n_live[4];
    uint64_t kind_first[4], kind_count[4];
    uint32_t* d_nsel;
    void* cub_tmp;
    size_t cub_bytes;
    uint32_t step;            // steps done by every live trajectory
    uint32_t max_cap;         // largest step cap
    std::vector<uint32_t> lag_steps;
    uint32_t next_lag;
    hipStream_t stream;
    hipEvent_t ev0, ev1;
    double kernel_ms;
    uint64_t launches;
};

static void fill_set_dev(const ldpcp_sweep_set& s, uint64_t first, SetDev& d)
{
    memset(&d, 0, sizeof(d));
    d.model = s.model;
    d.overdamped = s.overdamped;
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
        d.c_lap = -2.0 * s.timestep / s.tau;
        d.p_flip = d.tau_zero ? 0.0 : 0.5 * (1.0 - exp(-2.0 * s.timestep / s.tau));
        d.p0 = 0.5;
    } else {
        d.c_lap = -(s.eq_constant + 1.0) * s.timestep / s.tau;
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
    if (s.randomize_state) {
        // u <= p0  <=>  w <= floor(p0 2^32 - 1/2)
        double P = floor(d.p0 * 4294967296.0 - 0.5) + 1.0;
        d.T_p0le = P < 0.0 ? 0 : (P > 4294967296.0 ? 4294967296ull : (uint64_t)P);
    }
    d.fixed_state = (s.initial_state == 1.0) ? 1 : 0;
    d.key0 = (uint32_t)s.seed;
    d.key1 = (uint32_t)(s.seed >> 32);
    const double r = s.x0 / s.periodic_length;
    const double r_ref = s.x0 * d.Linv;
    (void)r;
    d.cell0 = (int)rint(r_ref);
    d.well0 = (int)floor(r_ref);
    d.fxp0 = (float)(s.x0 - (double)d.cell0 * s.periodic_length);
    if (d.target == 0) d.slot_base = d.well0;
    else d.slot_base = (d.well0 == d.cell0) ? d.cell0 : d.cell0 - 1;
    d.n_traj = s.n_traj;
    d.max_steps = s.max_steps;
    d.first = first;
    d.hist_lo = log10(s.timestep);
    const double hi = log10(s.timestep * (double)s.max_steps) + 1e-9;
    d.hist_inv = LDPCP_HIST_BINS / (hi - d.hist_lo);
}

static int sweep_kind(const ldpcp_sweep_set& s) { return s.model * 2 + (s.overdamped ? 1 : 0); }

extern "C" int ldpcp_sweep_create(const ldpcp_sweep_set* sets, int32_t nsets, int32_t precision, ldpcp_sweep** out)
{
    if (!out || !sets || nsets <= 0 || (precision != 32 && precision != 64)) {
        set_error("ldpcp_sweep_create: need sets, nsets > 0 and precision 32 or 64");
        return LDPCP_EINVAL;
    }
    *out = nullptr;
    uint64_t total = 0;
    uint32_t max_cap = 0;
    for (int i = 0; i < nsets; ++i) {
        const ldpcp_sweep_set& s = sets[i];
        if (s.n_traj == 0 || s.max_steps == 0 || s.max_steps >= (1ull << 32) || s.timestep <= 0.0 ||
            s.periodic_length <= 0.0 || (s.model != 0 && s.model != 1) || s.mass <= 0.0 ||
            (s.overdamped && s.gamma <= 0.0) || s.n_traj >= (1ull << 32)) {
            set_error("ldpcp_sweep_create: set %d is invalid (n_traj, max_steps < 2^32, timestep, length, model)", i);
            return LDPCP_EINVAL;
        }
        const double r = s.x0 / s.periodic_length;
        if (s.model == 1) {
            const double xp0 = s.x0 - rint(s.x0 * (1.0 / s.periodic_length)) * s.periodic_length;
            if (!(xp0 <= s.position[0] - s.half_width[0] || xp0 >= s.position[1] + s.half_width[1])) {
                set_error("ldpcp_sweep_create: set %d: double-barrier X0 must lie in the outer well", i);
                return LDPCP_EINVAL;
            }
        }
        (void)r;
        total += s.n_traj;
        max_cap = std::max<uint32_t>(max_cap, (uint32_t)s.max_steps);
    }
    if (total >= (1ull << 32)) {
        set_error("ldpcp_sweep_create: at most 2^32-1 trajectories per sweep");
        return LDPCP_EINVAL;
    }
    ldpcp_sweep* h = new ldpcp_sweep();
    h->precision = precision;
    h->nsets = nsets;
    h->n_total = total;
    h->host_sets.assign(sets, sets + nsets);
    // trajectories are laid out by kind, then by set, so each kind's live list is contiguous
    std::vector<int> order(nsets);
    for (int i = 0; i < nsets; ++i) order[i] = i;
    std::stable_sort(order.begin(), order.end(), [&](int a, int b) { return sweep_kind(sets[a]) < sweep_kind(sets[b]); });
    h->dev_sets.resize(nsets);
    std::vector<int32_t> set_of(total);
    uint64_t first = 0;
    for (int k = 0; k < 4; ++k) h->kind_first[k] = h->kind_count[k] = 0;
    for (int oi = 0; oi < nsets; ++oi) {
        const int i = order[oi];
        const int kind = sweep_kind(sets[i]);
        if (h->kind_count[kind] == 0) h->kind_first[kind] = first;
        h->kind_count[kind] += sets[i].n_traj;
        fill_set_dev(sets[i], first, h->dev_sets[i]);
        for (uint64_t t = 0; t < sets[i].n_traj; ++t) set_of[first + t] = i;
        first += sets[i].n_traj;
    }
    const uint64_t n = total;
    size_t bytes = 0;
    auto take = [&](size_t b) {
        size_t o = bytes;
        bytes += (b + 255) & ~(size_t)255;
        return o;
    };
    const bool f32 = precision == 32;
    const size_t o_xp = f32 ? take(n * 4) : 0, o_comp = f32 ? take(n * 4) : 0, o_fv = f32 ? take(n * 4) : 0,
                 o_fR = f32 ? take(n * 4) : 0, o_cell = f32 ? take(n * 4) : 0;
    const size_t o_x = f32 ? 0 : take(n * 8), o_v = f32 ? 0 : take(n * 8), o_R = f32 ? 0 : take(n * 8);
    const size_t o_flags = take(n * 4), o_ts = take(n * 16), o_nc = take(n * 4), o_tpt = take(n * 4),
                 o_abs = take(n * 4), o_set = take(n * 4), o_sets = take(sizeof(SetDev) * nsets),
                 o_hist = take(8ull * nsets * LDPCP_HIST_BINS), o_sums = take(8ull * nsets * 6),
                 o_msd = take(16ull * nsets * LDPCP_MSD_LAGS), o_alive = take(n), o_live = take(n * 4 * 2),
                 o_nsel = take(64);
    hipError_t e = hipMalloc(&h->block, bytes);
    if (e != hipSuccess) {
        set_error("ldpcp_sweep_create: hipMalloc(%zu) failed: %s", bytes, hipGetErrorString(e));
        delete h;
        return LDPCP_ENOMEM;
    }
    char* b = (char*)h->block;
    SweepPtrs& P = h->P;
    memset(&P, 0, sizeof(P));
    if (f32) {
        P.xp = (float*)(b + o_xp);
        P.comp = (float*)(b + o_comp);
        P.fv = (float*)(b + o_fv);
        P.fR = (float*)(b + o_fR);
        P.cell = (int32_t*)(b + o_cell);
    } else {
        P.x = (double*)(b + o_x);
        P.v = (double*)(b + o_v);
        P.R = (double*)(b + o_R);
    }
    P.flags = (uint32_t*)(b + o_flags);
    P.ts = (uint32_t*)(b + o_ts);
    P.n_catch = (uint32_t*)(b + o_nc);
    P.tpt_start = (uint32_t*)(b + o_tpt);
    P.absorbed = (uint32_t*)(b + o_abs);
    P.set_of = (int32_t*)(b + o_set);
    P.sets = (const SetDev*)(b + o_sets);
    P.hist = (unsigned long long*)(b + o_hist);
    double* sums = (double*)(b + o_sums);
    P.fpt_sum = sums;
    P.fpt_sum2 = sums + nsets;
    P.tpt_sum = sums + 2 * nsets;
    P.tpt_sum2 = sums + 3 * nsets;
    P.n_absorbed = (unsigned long long*)(sums + 4 * nsets);
    P.traj_steps = (unsigned long long*)(sums + 5 * nsets);
    P.msd_sum = (double*)(b + o_msd);
    P.msd_n = (unsigned long long*)(b + o_msd + 8ull * nsets * LDPCP_MSD_LAGS);
    P.alive_flag = (uint8_t*)(b + o_alive);
    uint32_t* lists = (uint32_t*)(b + o_live);
    h->d_nsel = (uint32_t*)(b + o_nsel);
    for (int k = 0; k < 4; ++k) {
        h->live[k] = lists + h->kind_first[k];
        h->live_tmp[k] = lists + n + h->kind_first[k];
        h->n_live[k] = (uint32_t)h->kind_count[k];
    }
    h->max_cap = max_cap;
    // 64 log-spaced lag steps in [1, max_cap]
    for (int k = 0; k < LDPCP_MSD_LAGS; ++k) {
        const double f = (LDPCP_MSD_LAGS == 1) ? 0.0 : (double)k / (LDPCP_MSD_LAGS - 1);
        uint32_t st = (uint32_t)llround(pow(10.0, f * log10((double)max_cap)));
        if (st < 1) st = 1;
        if (!h->lag_steps.empty() && st <= h->lag_steps.back()) st = h->lag_steps.back() + 1;
        h->lag_steps.push_back(std::min(st, max_cap));
    }
    h->next_lag = 0;
    h->step = 0;
    h->kernel_ms = 0.0;
    h->launches = 0;
    h->cub_tmp = nullptr;
    h->cub_bytes = 0;
#define SWEEP_TRY(expr)                                                                        \
    do {                                                                                        \
        hipError_t e_ = (expr);                                                                 \
        if (e_ != hipSuccess) {                                                                 \
            set_error("%s failed: %s", #expr, hipGetErrorString(e_));                           \
            ldpcp_sweep_destroy(h);                                                             \
            return LDPCP_EDEVICE;                                                               \
        }                                                                                       \
    } while (0)
    SWEEP_TRY(hipStreamCreate(&h->stream));
    SWEEP_TRY(hipEventCreate(&h->ev0));
    SWEEP_TRY(hipEventCreate(&h->ev1));
    SWEEP_TRY(hipMemcpyAsync(P.set_of, set_of.data(), n * 4, hipMemcpyHostToDevice, h->stream));
    SWEEP_TRY(hipMemcpyAsync((void*)P.sets, h->dev_sets.data(), sizeof(SetDev) * nsets, hipMemcpyHostToDevice,
                             h->stream));
    SWEEP_TRY(hipMemsetAsync(P.hist, 0, 8ull * nsets * LDPCP_HIST_BINS, h->stream));
    SWEEP_TRY(hipMemsetAsync(sums, 0, 8ull * nsets * 6, h->stream));
    SWEEP_TRY(hipMemsetAsync(P.msd_sum, 0, 16ull * nsets * LDPCP_MSD_LAGS, h->stream));
    hipLaunchKernelGGL(sweep_init_kernel, dim3((unsigned)((n + kBlock - 1) / kBlock)), dim3(kBlock), 0, h->stream, P,
                       n, f32 ? 1 : 0);
    SWEEP_TRY(hipGetLastError());
    for (int k = 0; k < 4; ++k)
        if (h->n_live[k]) {
            hipLaunchKernelGGL(iota_kernel, dim3((h->n_live[k] + 255) / 256), dim3(256), 0, h->stream, h->live[k],
                               h->n_live[k], (uint32_t)h->kind_first[k]);
            SWEEP_TRY(hipGetLastError());
        }
    size_t need = 0;
    SWEEP_TRY(hipcub::DeviceSelect::Flagged(nullptr, need, h->live[0], P.alive_flag, h->live_tmp[0], h->d_nsel,
                                            (int)std::max<uint64_t>(n, 1), h->stream));
    h->cub_bytes = need;
    SWEEP_TRY(hipMalloc(&h->cub_tmp, need));
    SWEEP_TRY(hipStreamSynchronize(h->stream));
#undef SWEEP_TRY
    *out = h;
    return LDPCP_OK;
}

template <int MODEL, bool OVER>
static void launch_kind(ldpcp_sweep* h, int kind, uint32_t K, int lag)
{
    const uint32_t n = h->n_live[kind];
    const dim3 grid((n + kBlock - 1) / kBlock), block(kBlock);
    if (h->precision == 32)
        hipLaunchKernelGGL((sweep_step_f32<MODEL, OVER>), grid, block, 0, h->stream, h->P, h->P.sets, h->live[kind],
                           h->P.alive_flag + h->kind_first[kind], n, h->step, K, lag);
    else
        hipLaunchKernelGGL((sweep_step_f64<MODEL, OVER>), grid, block, 0, h->stream, h->P, h->P.sets, h->live[kind],
                           h->P.alive_flag + h->kind_first[kind], n, h->step, K, lag);
}

extern "C" int ldpcp_sweep_launch(ldpcp_sweep* h, uint32_t steps_per_launch, uint64_t* live)
{
    if (!h || steps_per_launch == 0) {
        set_error("ldpcp_sweep_launch: null handle or zero steps");
        return LDPCP_EINVAL;
    }
    uint64_t total_live = 0;
    for (int k = 0; k < 4; ++k) total_live += h->n_live[k];
    if (total_live == 0 || h->step >= h->max_cap) {
        if (live) *live = 0;
        return LDPCP_OK;
    }
    // end this launch at the next MSD lag step if it comes first
    uint32_t K = steps_per_launch;
    int lag = -1;
    while (h->next_lag < h->lag_steps.size() && h->lag_steps[h->next_lag] <= h->step) h->next_lag++;
    if (h->next_lag < h->lag_steps.size()) {
        const uint32_t ls = h->lag_steps[h->next_lag];
        if (ls <= h->step + K) {
            K = ls - h->step;
            lag = (int)h->next_lag;
        }
    }
    if (h->step + K > h->max_cap) K = h->max_cap - h->step;
    LDPCP_HIP_TRY(hipEventRecord(h->ev0, h->stream));
    for (int kind = 0; kind < 4; ++kind) {
        if (!h->n_live[kind]) continue;
        switch (kind) {
        case 0: launch_kind<0, false>(h, kind, K, lag); break;
        case 1: launch_kind<0, true>(h, kind, K, lag); break;
        case 2: launch_kind<1, false>(h, kind, K, lag); break;
        default: launch_kind<1, true>(h, kind, K, lag); break;
        }
        LDPCP_HIP_TRY(hipGetLastError());
    }
    LDPCP_HIP_TRY(hipEventRecord(h->ev1, h->stream));
    // retire absorbed and capped trajectories (order-preserving)
    uint32_t nsel[4] = {0, 0, 0, 0};
    for (int kind = 0; kind < 4; ++kind) {
        if (!h->n_live[kind]) continue;
        // alive_flag is indexed by position in the kind's live list
        size_t bytes = h->cub_bytes;
        LDPCP_HIP_TRY(hipcub::DeviceSelect::Flagged(h->cub_tmp, bytes, h->live[kind], h->P.alive_flag + h->kind_first[kind],
                                                    h->live_tmp[kind], h->d_nsel + kind, (int)h->n_live[kind],
                                                    h->stream));
    }
    LDPCP_HIP_TRY(hipMemcpyAsync(nsel, h->d_nsel, sizeof(nsel), hipMemcpyDeviceToHost, h->stream));
    LDPCP_HIP_TRY(hipStreamSynchronize(h->stream));
    float ms = 0.0f;
    LDPCP_HIP_TRY(hipEventElapsedTime(&ms, h->ev0, h->ev1));
    h->kernel_ms += ms;
    h->launches++;
    total_live = 0;
    for (int kind = 0; kind < 4; ++kind) {
        if (!h->n_live[kind]) continue;
        std::swap(h->live[kind], h->live_tmp[kind]);
        h->n_live[kind] = nsel[kind];
        total_live += nsel[kind];
    }
    h->step += K;
    if (live) *live = total_live;
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
    if (!h || !out || set < 0 || set >= h->nsets) {
        set_error("ldpcp_sweep_stats: bad handle, set or output");
        return LDPCP_EINVAL;
    }
    memset(out, 0, sizeof(*out));
    const int ns = h->nsets;
    std::vector<double> sums(6ull * ns);
    LDPCP_HIP_TRY(hipMemcpy(sums.data(), h->P.fpt_sum, 8ull * ns * 6, hipMemcpyDeviceToHost));
    out->fpt_sum = sums[set];
    out->fpt_sum2 = sums[ns + set];
    out->tpt_sum = sums[2 * ns + set];
    out->tpt_sum2 = sums[3 * ns + set];
    uint64_t na;
    memcpy(&na, &sums[4 * ns + set], 8);
    out->n_absorbed = na;
    {
        // steps integrated: the absorption step, or every step done so far for live / capped ones
        const uint64_t n = h->host_sets[set].n_traj, first = h->dev_sets[set].first;
        const uint64_t cap = std::min<uint64_t>(h->host_sets[set].max_steps, h->step);
        std::vector<uint32_t> ab(n);
        LDPCP_HIP_TRY(hipMemcpy(ab.data(), h->P.absorbed + first, n * 4, hipMemcpyDeviceToHost));
        uint64_t steps = 0;
        for (uint64_t i = 0; i < n; ++i) steps += ab[i] ? ab[i] : cap;
        out->traj_steps = steps;
    }
    out->n_traj = h->host_sets[set].n_traj;
    out->n_capped = out->n_traj - na;
    out->hist_lo = h->dev_sets[set].hist_lo;
    out->hist_hi = h->dev_sets[set].hist_lo + LDPCP_HIST_BINS / h->dev_sets[set].hist_inv;
    LDPCP_HIP_TRY(hipMemcpy(out->hist, h->P.hist + (uint64_t)set * LDPCP_HIST_BINS, 8ull * LDPCP_HIST_BINS,
                            hipMemcpyDeviceToHost));
    LDPCP_HIP_TRY(hipMemcpy(out->msd_sum, h->P.msd_sum + (uint64_t)set * LDPCP_MSD_LAGS, 8ull * LDPCP_MSD_LAGS,
                            hipMemcpyDeviceToHost));
    LDPCP_HIP_TRY(hipMemcpy(out->msd_n, h->P.msd_n + (uint64_t)set * LDPCP_MSD_LAGS, 8ull * LDPCP_MSD_LAGS,
                            hipMemcpyDeviceToHost));
    for (size_t k = 0; k < h->lag_steps.size() && k < LDPCP_MSD_LAGS; ++k) out->msd_step[k] = h->lag_steps[k];
    return LDPCP_OK;
}

extern "C" int ldpcp_sweep_fpt(ldpcp_sweep* h, int32_t set, uint64_t* absorbed_step, uint64_t* tpt_start)
{
    if (!h || set < 0 || set >= h->nsets || !absorbed_step) {
        set_error("ldpcp_sweep_fpt: bad handle, set or output");
        return LDPCP_EINVAL;
    }
    const uint64_t n = h->host_sets[set].n_traj, first = h->dev_sets[set].first;
    std::vector<uint32_t> buf(n);
    LDPCP_HIP_TRY(hipMemcpy(buf.data(), h->P.absorbed + first, n * 4, hipMemcpyDeviceToHost));
    for (uint64_t i = 0; i < n; ++i) absorbed_step[i] = buf[i];
    if (tpt_start) {
        LDPCP_HIP_TRY(hipMemcpy(buf.data(), h->P.tpt_start + first, n * 4, hipMemcpyDeviceToHost));
        for (uint64_t i = 0; i < n; ++i) tpt_start[i] = buf[i];
    }
    return LDPCP_OK;
}

extern "C" int ldpcp_sweep_kernel_time(ldpcp_sweep* h, double* ms_total, uint64_t* launches)
{
    if (!h) {
        set_error("ldpcp_sweep_kernel_time: null handle");
        return LDPCP_EINVAL;
    }
    if (ms_total) *ms_total = h->kernel_ms;
    if (launches) *launches = h->launches;
    return LDPCP_OK;
}

extern "C" int ldpcp_sweep_destroy(ldpcp_sweep* h)
{
    if (!h) return LDPCP_OK;
    if (h->cub_tmp) (void)hipFree(h->cub_tmp);
    if (h->block) (void)hipFree(h->block);
    if (h->stream) (void)hipStreamDestroy(h->stream);
    if (h->ev0) (void)hipEventDestroy(h->ev0);
    if (h->ev1) (void)hipEventDestroy(h->ev1);
    delete h;
    return LDPCP_OK;
}
