// ldpcp_run.hip -- reference-semantics integrator on the GPU (the drop-in path).
//
// One thread owns one trajectory ("lane" of the reference's SIMD vector).
// The step arithmetic follows the reference operation by operation in FP64
// (this file is compiled with -ffp-contract=off; the single fused
// multiply-add the reference requests, _mmSIMD_fmadd_pd in the overdamped
// position update, langevinDynamicsSingleBarrier.h:2868, is an explicit
// fma()).  The hot state (position, velocity, force, barrier state, indices)
// lives in VGPRs for a whole launch; statistics and the per-trajectory barrier
// table live in HBM and are touched only on the rare crossing events.
#include <hip/hip_runtime.h>
#include <limits.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <chrono>
#include <mutex>
#include <type_traits>
#include <vector>

#include "ldpcp_common.hpp"
#include "ldpcp_device.h"
#include "ldpcp_schedule.h"

namespace ldpcp {

namespace {

constexpr int kFields = 7;  // x, v, f, pe, ke, state, barrier
constexpr uint64_t kValid = 1ull << 63;
constexpr uint64_t kState = 1ull << 62;
constexpr uint64_t kStepMask = kState - 1;

struct RunConst {
    int32_t model, overdamped, tau_zero, equal_prob, same_init, centre;
    int32_t b_wait, b_fpt, b_tpt, b_cg, n_cg, save_all, randomize_v, randomize_state, replay, wb;
    int32_t cg_size[LDPCP_MAX_CG];
    double cg_inv[LDPCP_MAX_CG];  // 1 / cg_size
    uint64_t t_flip, t_asym0, t_asym1;  // flip_threshold of p_flip, p_asym0, p_asym1
    double dt, minv, mhalf, halfdt, expmgdt, var, dtgm, L, Linv, c_lap, p_flip, p_asym0, p_asym1, K, p0,
        p_create;
    double H[2], S[2], w;
    double Hd[2][2], xb[2], lo[2], hi[2], xc;
    double slope[2][2][2], icept[2][2][2];
    double vsigma, x0, v0, initial_state;
    uint64_t save_freq, nsteps_eff, frames_cap;
    // this run's step clock (ldpcp_schedule.h): steps run_begin+1 .. run_end are
    // integrated, frames fall at run_begin + k save_freq with indices from frame0,
    // and the clock ends at clock_end (the dropped trailing steps skipped)
    uint64_t run_begin, run_end, clock_end, frame0;
    uint32_t key0, key1, n_traj, traj_offset;
    uint64_t n_gauss, n_flip, catch_cap, init_gauss, init_flip, tick_cap;
    uint64_t ev_cap;   // entries kept per event list per trajectory
    int32_t n_lists, pad_;
    int32_t bmin, bn;
};

struct RunPtrs {
    double *x, *v, *F, *E, *R, *barrier, *state, *well, *periodic, *xp, *local;
    uint64_t *cnt0, *cnt1;
    ldpcp_lane* lanes;
    uint64_t* win;       // [traj][2][wb]
    int32_t* wbase;      // [traj]
    double* frames;      // [field][frame][traj]
    const double* g;     // replay gaussians [k][traj]
    const double* f;     // replay flip uniforms [k][traj]
    const double* c;     // replay catch-up uniforms [traj][cap]
    const int8_t* binit; // replay barrier states [array][traj][bn]
    double* ticks;       // barrier ticks [traj][tick_cap]
    double* bdefault;    // state of barriers created with UNIFORMINITALSTATES [traj]
    double* events;      // event lists [traj][list][ev_cap]
};

// The run's branch flags, separate from RunConst so that a kernel can make them
// compile-time constants while the constants stay in the kernel arguments.
struct Flags {
    int model, over, tau0, replay;
};

__device__ __forceinline__ Flags run_flags(const RunConst& c) { return {c.model, c.overdamped, c.tau_zero, c.replay}; }

// Constants of IntegrateInitializeVariables (langevinDynamicsSingleBarrier.h:1520-1618,
// overdamped :2812-2824; langevinDynamicsDoubleBarrier.h:1733-1937, GetXc :1722-1732).
void make_const(const ldpcp_run_params& p, RunConst& c)
{
    memset(&c, 0, sizeof(c));
    c.model = p.model;
    c.overdamped = p.overdamped;
    c.tau_zero = (p.tau == 0.0);
    c.equal_prob = p.equal_probability;
    c.same_init = p.same_initial_barrier;
    c.b_wait = p.b_waiting;
    c.b_fpt = p.b_fpt;
    c.b_tpt = p.b_tpt;
    c.b_cg = p.b_fptcg;
    c.n_cg = p.n_cg;
    for (int j = 0; j < p.n_cg && j < LDPCP_MAX_CG; ++j) {
        c.cg_size[j] = p.cg_size[j];
        c.cg_inv[j] = 1.0 / p.cg_size[j];
    }
    c.save_all = p.save_all;
    c.randomize_v = p.randomize_v;
    c.randomize_state = p.randomize_state;
    c.wb = p.barrier_window;
    c.dt = p.timestep;
    c.minv = 1.0 / p.mass;
    c.mhalf = p.mass * 0.5;
    c.halfdt = p.timestep * 0.5;
    c.expmgdt = exp(-p.gamma * p.timestep);
    c.var = sqrt(p.kb * p.temperature / p.mass * (1.0 - exp(-2.0 * p.gamma * p.timestep)));
    c.L = p.periodic_length;
    c.K = p.eq_constant;
    if (p.equal_probability) {
        c.c_lap = -2.0 * p.timestep / p.tau;
        c.p_flip = c.tau_zero ? 0.0 : 0.5 * (1.0 - exp(-2.0 * p.timestep / p.tau));
        c.p0 = 0.5;
    } else {
        c.c_lap = -(p.eq_constant + 1.0) * p.timestep / p.tau;
        if (!c.tau_zero) {
            c.p_asym0 = p.eq_constant / (1.0 + p.eq_constant) * (1.0 - exp(c.c_lap));
            c.p_asym1 = 1.0 / (1.0 + p.eq_constant) * (1.0 - exp(c.c_lap));
        }
        c.p0 = 1.0 / (1.0 + p.eq_constant);
    }
    c.t_flip = flip_threshold(c.p_flip);
    c.t_asym0 = flip_threshold(c.p_asym0);
    c.t_asym1 = flip_threshold(c.p_asym1);
    c.p_create = 1.0 / (1.0 + p.eq_constant);
    c.Linv = 1.0 / c.L;
    if (p.model == 0) {
        c.w = p.half_width[0];
        for (int s = 0; s < 2; ++s) {
            c.H[s] = p.height[0][s];
            c.S[s] = p.height[0][s] / p.half_width[0];
        }
    } else {
        for (int b = 0; b < 2; ++b) {
            for (int s = 0; s < 2; ++s) c.Hd[b][s] = p.height[b][s];
            c.xb[b] = p.position[b];
        }
        const double x1b = p.position[0], w1 = p.half_width[0], x2b = p.position[1], w2 = p.half_width[1];
        const double y1 = p.height[0][0], y2 = p.height[1][0];
        const double x1h = x1b + w1, x2l = x2b - w2;
        const double xc = fma(x1h * y1, x2b - x2l, x2l * y2 * (x1h - x1b)) / fma(y1, x2b - x2l, y2 * (x1h - x1b));
        const double yc = (x1h - xc) / (x1h - x1b) * y1;
        c.xc = xc;
        c.lo[0] = x1b - w1;
        c.hi[0] = x1b + w1;
        c.lo[1] = x2b - w2;
        c.hi[1] = x2b + w2;
        const double xb_xl[2] = {c.xb[0] - c.lo[0], c.xb[1] - c.lo[1]};
        const double xh_xb[2] = {c.hi[0] - c.xb[0], c.hi[1] - c.xb[1]};
        for (int s = 0; s < 2; ++s) {
            const double H0 = c.Hd[0][s], H1 = c.Hd[1][s];
            c.slope[0][s][0] = H0 / xb_xl[0];
            c.slope[0][s][1] = (yc - H0) / (xc - c.xb[0]);
            c.slope[1][s][0] = (H1 - yc) / (c.xb[1] - xc);
            c.slope[1][s][1] = 0.0 - H1 / xh_xb[1];
            c.icept[0][s][0] = 0.0 - H0 * (c.lo[0] / xb_xl[0]);
            c.icept[0][s][1] = (xc * H0 - yc * c.xb[0]) / (xc - c.xb[0]);
            c.icept[1][s][0] = (yc * c.xb[1] - xc * H1) / (c.xb[1] - xc);
            c.icept[1][s][1] = H1 * (c.hi[1] / xh_xb[1]);
        }
    }
    if (p.overdamped) {
        const double mgi = 1.0 / (p.mass * p.gamma);
        c.dtgm = p.timestep * mgi;
        c.var = sqrt(p.kb * p.temperature * mgi * p.timestep * 0.5);
    }
    c.vsigma = sqrt(p.kb * p.temperature / p.mass);
    c.x0 = p.x0;
    c.v0 = p.v0;
    c.initial_state = p.initial_state;
    const double m = fmod(fmod(p.x0, p.periodic_length) + p.periodic_length, p.periodic_length);
    c.centre = (m == p.periodic_length * 0.5);
    c.save_freq = p.save_freq;
    c.nsteps_eff = ldpcp_steps_integrated(p.nsteps, p.save_freq);
    c.frames_cap = c.nsteps_eff / p.save_freq + 1;
    c.run_begin = 0;
    c.run_end = c.nsteps_eff;
    c.clock_end = p.nsteps;
    c.frame0 = 1;
    c.key0 = (uint32_t)p.seed;
    c.key1 = (uint32_t)(p.seed >> 32);
    c.init_gauss = (uint64_t)(p.randomize_v ? 1 : 0) + (uint64_t)(p.overdamped ? 1 : 0);
    c.init_flip = p.randomize_state ? 1 : 0;
    c.tick_cap = p.tick_capacity;
    c.ev_cap = p.event_capacity;
    c.n_lists = p.event_capacity ? LDPCP_LIST_CG0 + (p.b_fptcg ? p.n_cg : 0) : 0;
}

// ---------------------------------------------------------------------------
// Per-lane step pieces, each restating one reference method.
// ---------------------------------------------------------------------------
struct Lane {
    double x, v, F, E, R, barrier, state;
    double height, slope_pos;                       // single barrier
    double dh[2], ds[2][2], di[2][2];               // double barrier
    double well, periodic, well_prev, periodic_prev, xp, local, local_prev;
    bool ent_well, ent_period, ent_local;
    int wbase;  // barrier id of the window's first entry
    uint64_t cnt0, cnt1;
};

// Heights, slopes and intercepts of both barrier states, which every step
// selects from by the current state.  A selection between two kernel-argument
// fields is compiled as one load from the selected address, a memory round trip
// on the critical path of every step; each value here passes through an empty
// asm statement, so it stays an opaque register value loaded once per launch.
struct StateParams {
    double H[2], S[2], Sn[2];                   // single barrier [state]; Sn = 0 - S
    uint32_t Sm[2];                             // single barrier: the sign bit that turns S into Sn (0 for S = +0)
    double dh[2][2], ds[2][2][2], di[2][2][2];  // double barrier [state][barrier][side]
    double dfn[2][2][2];                        // double barrier force 0 - ds
    double p_flip[2];                           // per-step flip probability leaving state [state]
    uint64_t t_flip[2];                         // its flip_threshold
};

// Wave-uniform copy of a 64-bit value, so that what derives from it stays in
// scalar registers.
__device__ __forceinline__ uint64_t uniform_u64(uint64_t v)
{
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)v);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readfirstlane((int)(uint32_t)(v >> 32));
    return ((uint64_t)hi << 32) | lo;
}

// The lanes where a < b, and where a > b, or either is NaN: the compare's lane
// mask itself (one v_cmp; a ballot of a bool would first spell the bool out in
// a vector register).
__device__ __forceinline__ uint64_t lanes_ult(double a, double b) { return __builtin_amdgcn_fcmp(a, b, 12); }
__device__ __forceinline__ uint64_t lanes_ugt(double a, double b) { return __builtin_amdgcn_fcmp(a, b, 10); }

__device__ __forceinline__ double in_register(double v)
{
    asm("" : "+v"(v));
    return v;
}

__device__ __forceinline__ uint64_t in_register(uint64_t v)
{
    asm("" : "+v"(v));
    return v;
}

__device__ __forceinline__ uint32_t in_register(uint32_t v)
{
    asm("" : "+v"(v));
    return v;
}

__device__ __forceinline__ StateParams state_params(const RunConst& c, const Flags& f)
{
    StateParams sp;
    // equally probable states share one flip probability (Single.h:2010-2015, 2078-2086)
    sp.p_flip[0] = in_register(c.equal_prob ? c.p_flip : c.p_asym0);
    sp.p_flip[1] = in_register(c.equal_prob ? c.p_flip : c.p_asym1);
    sp.t_flip[0] = in_register(c.equal_prob ? c.t_flip : c.t_asym0);
    sp.t_flip[1] = in_register(c.equal_prob ? c.t_flip : c.t_asym1);
#pragma unroll
    for (int st = 0; st < 2; ++st) {
        if (f.model == 0) {
            sp.H[st] = in_register(c.H[st]);
            sp.S[st] = in_register(c.S[st]);
            sp.Sn[st] = in_register(0.0 - c.S[st]);
            // 0 - S is S with the sign bit flipped, except for S = +0 (0 - 0 = +0)
            sp.Sm[st] = in_register(__double_as_longlong(c.S[st]) == 0 ? 0u : 0x80000000u);
            continue;
        }
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            sp.dh[st][b] = in_register(c.Hd[b][st]);
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                sp.ds[st][b][side] = in_register(c.slope[b][st][side]);
                sp.di[st][b][side] = in_register(c.icept[b][st][side]);
                sp.dfn[st][b][side] = in_register(0.0 - c.slope[b][st][side]);
            }
        }
    }
    return sp;
}

// UpdateBarrierStateEtaParamter (Single.h:2154-2165; Double.h:2737-2757)
__device__ __forceinline__ void lane_param(Lane& s, const StateParams& sp, const Flags& f, bool count = true)
{
    const bool one = (s.state == 1.0);
    if (f.model == 0) {
        s.height = one ? sp.H[1] : sp.H[0];
        s.slope_pos = one ? sp.S[1] : sp.S[0];
    } else {
#pragma unroll
        for (int b = 0; b < 2; ++b) {
            s.dh[b] = one ? sp.dh[1][b] : sp.dh[0][b];
#pragma unroll
            for (int side = 0; side < 2; ++side) {
                s.ds[b][side] = one ? sp.ds[1][b][side] : sp.ds[0][b][side];
                s.di[b][side] = one ? sp.di[1][b][side] : sp.di[0][b][side];
            }
        }
    }
    if (!count) return;
    const uint64_t inc1 = one ? 1u : 0u;
    s.cnt1 += inc1;
    s.cnt0 += 1u - inc1;
}

// EnergyAndForce (Single.h:1292-1324; Double.h:1464-1522)
// force: sets F; energy: sets E and barrier (the step needs F only; E and
// barrier, functions of the same x and state, are computed when stored)
__device__ __forceinline__ void lane_energy_force(Lane& s, const RunConst& c, const Flags& f, bool force = true,
                                                  bool energy = true)
{
    if (f.model == 0) {
        const double r = s.x * c.Linv;
        const double pidx = rint(r);
        const double xp = s.x - pidx * c.L;
        const double slope_neg = 0.0 - s.slope_pos;
        const bool neg = (xp < 0.0) && (xp > -c.w);
        const bool pos = (xp < c.w) && (xp > 0.0);
        const double vneg = s.height - xp * slope_neg;
        const double vpos = s.height - xp * s.slope_pos;
        double E = neg ? vneg : 0.0;
        E = pos ? vpos : E;
        double F = neg ? slope_neg : 0.0;
        F = pos ? s.slope_pos : F;
        if (energy) {
            s.E = E;
            s.barrier = s.height;
        }
        if (force) s.F = F;
    } else {
        const double xp = s.xp;
        const bool n0 = (xp < c.xb[0]) && (xp > c.lo[0]);
        const bool p0 = (xp < c.xc) && (xp > c.xb[0]);
        const bool n1 = (xp < c.xb[1]) && (xp > c.xc);
        const bool p1 = (xp < c.hi[1]) && (xp > c.xb[1]);
        const double v00 = xp * s.ds[0][0] + s.di[0][0];
        const double v01 = xp * s.ds[0][1] + s.di[0][1];
        const double v10 = xp * s.ds[1][0] + s.di[1][0];
        const double v11 = xp * s.ds[1][1] + s.di[1][1];
        double E = n0 ? v00 : 0.0;
        E = p0 ? v01 : E;
        E = n1 ? v10 : E;
        E = p1 ? v11 : E;
        double F = n0 ? s.ds[0][0] : 0.0;
        F = p0 ? s.ds[0][1] : F;
        F = n1 ? s.ds[1][0] : F;
        F = p1 ? s.ds[1][1] : F;
        F = 0.0 - F;
        double B = n0 ? s.dh[0] : 0.0;
        B = p0 ? s.dh[0] : B;
        B = n1 ? s.dh[1] : B;
        B = p1 ? s.dh[1] : B;
        if (energy) {
            s.E = E;
            s.barrier = B;
        }
        if (force) s.F = F;
    }
}

// Single-barrier force for state st, from the slopes as separate values (selecting
// between elements of one array would let the compiler turn the selection into an
// indexed load from a stack copy of the array).  -w < xp < 0 or 0 < xp < w
// (strict, Single.h:1292-1324) is |xp| < w with xp != 0.
__device__ __forceinline__ double force_single(double x, const RunConst& c, double S0, double S1, double Sn0,
                                               double Sn1, bool st)
{
    const double r = x * c.Linv;
    const double xp = x - rint(r) * c.L;
    const bool band = (fabs(xp) < c.w) && (xp != 0.0);
    const double Sp = st ? S1 : S0;
    const double Sm = st ? Sn1 : Sn0;
    return band ? ((xp < 0.0) ? Sm : Sp) : 0.0;
}

// force_single for the offset xp = x - (periodic index) L of the state's slope
// S: S on (0, w), 0 - S on (-w, 0), 0 elsewhere; 0 - S is S with xp's sign bit
// xor-ed into its high word where Sm is the sign bit.
__device__ __forceinline__ double force_single_xp(double xp, double w, double S, uint32_t Sm)
{
    const bool band = (fabs(xp) < w) && (xp != 0.0);
    const uint64_t sb = (uint64_t)__double_as_longlong(S);
    const uint32_t hi = (uint32_t)(sb >> 32) ^ ((uint32_t)((uint64_t)__double_as_longlong(xp) >> 32) & Sm);
    const double Fs = __hiloint2double((int)hi, (int)(uint32_t)sb);
    return band ? Fs : 0.0;
}

// The force of EnergyAndForce for barrier state st, selected from that state's
// slopes directly: the same values lane_param + lane_energy_force give (the
// single barrier's slopeNeg and the double barrier's F are 0 - slope, computed
// once per launch in StateParams), without the heights and intercepts the step
// does not use.
__device__ __forceinline__ double lane_force(const Lane& s, const RunConst& c, const StateParams& sp, const Flags& f,
                                             bool st)
{
    if (f.model == 0) return force_single(s.x, c, sp.S[0], sp.S[1], sp.Sn[0], sp.Sn[1], st);
    const double xp = s.xp;
    const bool n0 = (xp < c.xb[0]) && (xp > c.lo[0]);
    const bool p0 = (xp < c.xc) && (xp > c.xb[0]);
    const bool n1 = (xp < c.xb[1]) && (xp > c.xc);
    const bool p1 = (xp < c.hi[1]) && (xp > c.xb[1]);
    double F = 0.0;  // 0 - 0 outside the barriers
    F = n0 ? (st ? sp.dfn[1][0][0] : sp.dfn[0][0][0]) : F;
    F = p0 ? (st ? sp.dfn[1][0][1] : sp.dfn[0][0][1]) : F;
    F = n1 ? (st ? sp.dfn[1][1][0] : sp.dfn[0][1][0]) : F;
    F = p1 ? (st ? sp.dfn[1][1][1] : sp.dfn[0][1][1]) : F;
    return F;
}

// Positions certainly inside a lane's cell: where rint(x Linv) == periodic,
// floor(x Linv) == well and, for the double barrier, (x - periodic L < xc) ==
// (local == 0) (lane_indices).  The cell's ends (periodic -+ 1/2) L, well L,
// (well + 1) L and periodic L + xc are moved inwards by 2^-48 of their
// magnitude (at least 16 ulps; the rounding of x Linv, x - periodic L and of
// the ends moves a true end by a few ulps) and by 2^-1000.  A step whose x
// stays in [lo, hi] changes no index; one outside takes the rare branch, which
// computes the indices.
__device__ __forceinline__ void cell_bounds(const RunConst& c, const Flags& f, double periodic, double well,
                                            double local, double& lo, double& hi)
{
    double a = fmax((periodic - 0.5) * c.L, well * c.L);
    double b = fmin((periodic + 0.5) * c.L, (well + 1.0) * c.L);
    if (f.model == 1) {
        const double xm = periodic * c.L + c.xc;
        a = local == 0.0 ? a : fmax(a, xm);
        b = local == 0.0 ? fmin(b, xm) : b;
    }
    lo = a + (fabs(a) * 0x1p-48 + 0x1p-1000);
    hi = b - (fabs(b) * 0x1p-48 + 0x1p-1000);
}

// The double barrier's force inside the barrier of a lane's local half
// (local == 0: barrier 0, whose segments are (lo[0], xb[0]) and (xb[0], xc);
// local == 1: barrier 1, (xc, xb[1]) and (xb[1], hi[1])): lane_force's four
// segment tests reduced to the two that can hold there.  a and b are the
// segments' forces 0 - slope, [state].
struct LocalBarrier {
    double lo, xb, hi, a[2], b[2];
};

__device__ __forceinline__ LocalBarrier local_barrier(const RunConst& c, const StateParams& sp, double local)
{
    const bool one = local != 0.0;
    LocalBarrier r;
    r.lo = one ? c.xc : c.lo[0];
    r.xb = one ? c.xb[1] : c.xb[0];
    r.hi = one ? c.hi[1] : c.xc;
#pragma unroll
    for (int st = 0; st < 2; ++st) {
        r.a[st] = one ? sp.dfn[st][1][0] : sp.dfn[st][0][0];
        r.b[st] = one ? sp.dfn[st][1][1] : sp.dfn[st][0][1];
    }
    return r;
}

__device__ __forceinline__ double force_double_local(double xp, const LocalBarrier& lb, bool st)
{
    const bool n = (xp < lb.xb) && (xp > lb.lo);
    const bool p = (xp < lb.hi) && (xp > lb.xb);
    double F = 0.0;
    F = n ? (st ? lb.a[1] : lb.a[0]) : F;
    F = p ? (st ? lb.b[1] : lb.b[0]) : F;
    return F;
}

// InitializePeriodicIndices / UpdatePeriodicIndices (Single.h:1838-1884; Double.h:2150-2235)
__device__ __forceinline__ void lane_indices(Lane& s, const RunConst& c, const Flags& f, bool init)
{
    const double r = s.x * c.Linv;
    const double well = floor(r), periodic = rint(r);
    if (init) {
        s.well = s.well_prev = well;
        s.periodic = s.periodic_prev = periodic;
        s.ent_well = s.ent_period = false;
    } else {
        s.well_prev = s.well;
        s.periodic_prev = s.periodic;
        s.well = well;
        s.periodic = periodic;
        s.ent_well = (s.well_prev != s.well);
        s.ent_period = (s.periodic_prev != s.periodic);
    }
    if (f.model == 1) {
        s.xp = s.x - s.periodic * c.L;
        const double local = (s.xp < c.xc) ? 0.0 : 1.0;
        if (init) {
            s.local = s.local_prev = local;
            s.ent_local = false;
        } else {
            s.local_prev = s.local;
            s.local = local;
            s.ent_local = (s.local_prev != s.local);
        }
    } else {
        s.ent_local = false;
    }
}

// ---------------------------------------------------------------------------
// Noise
// ---------------------------------------------------------------------------
struct StepNoise {
    uint64_t q;  // cached Philox pair index (+1), 0 = none
    double g1;
    uint32_t w3;
};

// replayed noise of step iStep (>= 1)
__device__ __forceinline__ void replay_noise(const RunConst& c, const Flags& f, const RunPtrs& P, uint32_t t,
                                             uint64_t iStep, double& g, double& u)
{
    const uint64_t k = iStep - 1;
    g = P.g[(c.init_gauss + k) * c.n_traj + t];
    u = f.tau0 ? 1.0 : P.f[(c.init_flip + k) * c.n_traj + t];
}

// Philox noise of step iStep (>= 1): the Gaussian and the 32-bit word of the flip
// uniform (unit_f64 of it)
__device__ __forceinline__ void step_noise(const RunConst& c, uint32_t t, uint64_t iStep, StepNoise& sn, double& g,
                                           uint32_t& fw)
{
    const uint64_t q = (iStep - 1) >> 1;
    if ((iStep - 1) & 1) {
        if (sn.q != q + 1) {
            const Philox4 w = philox4x32_10((uint32_t)q, (uint32_t)(q >> 32), c.traj_offset + t, 0u, c.key0, c.key1);
            double g0;
            box_muller_f64(w.w[0], w.w[1], g0, sn.g1);
            sn.w3 = w.w[3];
        }
        g = sn.g1;
        fw = sn.w3;
    } else {
        const Philox4 w = philox4x32_10((uint32_t)q, (uint32_t)(q >> 32), c.traj_offset + t, 0u, c.key0, c.key1);
        box_muller_f64(w.w[0], w.w[1], g, sn.g1);
        sn.w3 = w.w[3];
        sn.q = q + 1;
        fw = w.w[2];
    }
}

__device__ __forceinline__ int barrier_init_state(const RunConst& c, const Flags& f, const RunPtrs& P, uint32_t t,
                                                  int array, int id, int& err)
{
    // UserData::m_bInitialValue: the lane's BarrierState0Instance as it stood before step 0
    // (Single.h:1378-1380, 80-84): 0 for a fresh integrator, the previous trial's or the
    // restart file's initial state otherwise
    if (c.same_init) return (int)P.bdefault[t];
    if (f.replay) {
        const int k = id - c.bmin;
        if (k < 0 || k >= c.bn) {
            err = 2;
            return 0;
        }
        const int8_t v = P.binit[((uint64_t)array * c.n_traj + t) * c.bn + k];
        if (v < 0) {
            err = 2;
            return 0;
        }
        return v;
    }
    const Philox4 w = philox4x32_10((uint32_t)id, 0u, c.traj_offset + t, 3u + 256u * (uint32_t)array, c.key0, c.key1);
    return unit_f64(w.w[0]) < c.p_create ? 0 : 1;
}

// A trajectory's barrier table: the window of wb entries per barrier array
// (valid bit, state bit, 62-bit time step) from barrier id wbase on, optionally
// held in a per-trajectory LDS cache, direct-mapped with kWinSlots entries per
// array, that writes an entry back to the window when it is evicted and at the
// end of the launch (win_flush).  With the cache a hand-over between nearby
// barriers touches LDS only and never waits for global memory; every access of
// a launch goes through the same view, so the cache is the table's current
// state until it is flushed.
constexpr int kWinSlots = 16;
constexpr int kWinEmpty = INT_MIN;  // cache slot holds no barrier

struct Window {
    uint64_t* win;   // this trajectory's arrays: win[array * wb + k]
    int wb, wbase;
    bool cached;     // a constant of the kernel, so that the cache's code folds away when unused
    uint64_t* ce;    // LDS cache entries [array][slot][64 lanes]
    int* cid;        // their barrier ids (kWinEmpty: none), same layout
    uint8_t* dirty;  // 1: entry changed since it was read from the window
    uint32_t lane;
};

__device__ __forceinline__ Window make_window(const RunConst& c, const RunPtrs& P, uint32_t t, int wbase,
                                              uint64_t* ce = nullptr, int* cid = nullptr, uint8_t* dirty = nullptr)
{
    return {P.win + (uint64_t)t * 2 * c.wb, c.wb, wbase, ce != nullptr, ce, cid, dirty, threadIdx.x & 63u};
}

__device__ __forceinline__ int win_cache_slot(const Window& w, int array, int id)
{
    return (array * kWinSlots + (id & (kWinSlots - 1))) * 64 + (int)w.lane;
}

// writes the entry in cache slot sl (barrier `id` of `array`) back if it changed
__device__ __forceinline__ void win_evict(const Window& w, int array, int sl, int id)
{
    if (id != kWinEmpty && w.dirty[sl]) w.win[array * w.wb + (id - w.wbase)] = w.ce[sl];
}

__device__ __forceinline__ void win_flush(const Window& w)
{
    if (!w.cached) return;
    for (int a = 0; a < 2; ++a)
        for (int j = 0; j < kWinSlots; ++j) {
            const int sl = (a * kWinSlots + j) * 64 + (int)w.lane;
            win_evict(w, a, sl, w.cid[sl]);
        }
}

// window index of barrier `id`, or -1 with err = 3 when the trajectory has left the window
__device__ __forceinline__ int win_slot(const Window& w, int id, int& err)
{
    const int k = id - w.wbase;
    if (k < 0 || k >= w.wb) {
        err = 3;
        return -1;
    }
    return k;
}

__device__ __forceinline__ uint64_t win_get(const Window& w, int array, int id, int k)
{
    if (w.cached) {
        const int sl = win_cache_slot(w, array, id);
        const int held = w.cid[sl];
        if (held == id) return w.ce[sl];
        win_evict(w, array, sl, held);
        const uint64_t v = w.win[array * w.wb + k];
        w.cid[sl] = id;
        w.ce[sl] = v;
        w.dirty[sl] = 0;
        return v;
    }
    return w.win[array * w.wb + k];
}

__device__ __forceinline__ void win_put(const Window& w, int array, int id, int k, uint64_t v)
{
    if (w.cached) {
        const int sl = win_cache_slot(w, array, id);
        const int held = w.cid[sl];
        if (held != id) win_evict(w, array, sl, held);
        w.cid[sl] = id;
        w.ce[sl] = v;
        w.dirty[sl] = 1;
        return;
    }
    w.win[array * w.wb + k] = v;
}

// entry of barrier (array, id) at window index k, created (its state at step 0)
// if this trajectory has not touched it yet
__device__ __forceinline__ uint64_t win_load(const RunConst& c, const Flags& f, const RunPtrs& P, uint32_t t,
                                             const Window& w, int array, int id, int k, int& err)
{
    uint64_t v = win_get(w, array, id, k);
    if (!(v & kValid)) {
        v = kValid | (barrier_init_state(c, f, P, t, array, id, err) ? kState : 0ull);
        win_put(w, array, id, k, v);
    }
    return v;
}

// ---------------------------------------------------------------------------
// Statistics state machines (Single.h:2166-2592; Double.h:2758-3209)
// ---------------------------------------------------------------------------
__device__ __forceinline__ void arm_indices(int idx[2], const Lane& s)
{
    const int pidx = (int)s.periodic;
    if (s.well == s.periodic) {
        idx[0] = pidx;
        idx[1] = pidx + 1;
    } else {
        idx[0] = pidx - 1;
        idx[1] = pidx;
    }
}

// Stored state of barrier `id` of the single barrier's array, created (its
// state at step 0) if this trajectory has not touched it yet.
__device__ __forceinline__ int stored_state(const RunConst& c, const Flags& f, const RunPtrs& P, uint32_t t,
                                            const Window& w, int id, int& err)
{
    const int k = win_slot(w, id, err);
    if (k < 0) return 0;
    return (win_load(c, f, P, t, w, 0, id, k, err) & kState) ? 1 : 0;
}

// FPT1 arming (Single.h:2364-2387; at step 0 :2184-2212): the barriers either
// side of the well the particle sits in, the one under the particle with its
// current state, the other with its stored state.
__device__ __forceinline__ void fpt1_arm(const RunConst& c, const Flags& f, const RunPtrs& P, uint32_t t,
                                         const Window& w, int32_t (&init_index)[2], double (&init_state)[2],
                                         const Lane& s, int& err)
{
    const int pidx = (int)s.periodic;
    if (s.well == s.periodic) {
        init_state[0] = s.state;
        init_index[0] = pidx;
        init_state[1] = (double)stored_state(c, f, P, t, w, pidx + 1, err);
        init_index[1] = pidx + 1;
    } else {
        init_state[0] = (double)stored_state(c, f, P, t, w, pidx - 1, err);
        init_index[0] = pidx - 1;
        init_state[1] = s.state;
        init_index[1] = pidx;
    }
}

// Whether FPTCG size j divides the well index (an integer-valued double below
// 2^31 in magnitude): rint(well / size) computed with the reciprocal is exact
// when size divides well, and k * size is a multiple of size otherwise, so the
// test equals (int)well % size == 0 without an integer division.
__device__ __forceinline__ bool cg_divides(const RunConst& c, int j, double well)
{
    return rint(well * c.cg_inv[j]) * (double)c.cg_size[j] == well;
}

// Appends v as entry n of list `list` when the lists are kept.
__device__ __forceinline__ void push_event(const RunConst& c, double* ev, int list, uint64_t n, double v)
{
    if (n < c.ev_cap) ev[(uint64_t)list * c.ev_cap + n] = v;
}

__device__ __forceinline__ void stats_init(const RunConst& c, const Flags& f, const RunPtrs& P, uint32_t t,
                                           const Window& w, ldpcp_lane* o, const Lane& s, int& err)
{
    if (c.b_wait) {
        o->wt_tick_prev = 0.0;
        o->wt_passed = 0;
    }
    if (c.b_fpt) {
        o->fpt_last_tick = 0.0;
        o->fpt_to_update = c.centre ? 0 : 1;
        if (f.model == 0 && c.centre) fpt1_arm(c, f, P, t, w, o->fpt_init_index, o->fpt_init_state, s, err);
        o->fpt2_last_tick = 0.0;
        if (c.centre) {
            o->fpt2_to_update = 0;
            int idx[2];
            arm_indices(idx, s);
            o->fpt2_idx[0] = idx[0];
            o->fpt2_idx[1] = idx[1];
        } else {
            o->fpt2_to_update = 1;
        }
    }
    if (c.b_tpt) {
        o->tpt_last_tick = 0.0;
        o->tpt_to_update = c.centre ? 0 : 1;
        o->tpt2_last_tick = 0.0;
        o->tpt2_to_update = c.centre ? 0 : 1;
        o->tpt2_initial_well = (int)s.well;
    }
    if (c.b_cg) {
        for (int j = 0; j < c.n_cg; ++j) {
            o->cg_last_tick[j] = 0.0;
            o->cg_initial_well[j] = (int)s.well;
            o->cg_to_update[j] = !(c.centre && cg_divides(c, j, s.well));
        }
    }
}

// Event duration `m_TimeStep0 * iStep - lastTick` (Single.h:2398, 2453, 2509,
// 2525, 2576): a scalar statement the reference's compiler contracts into one
// fused multiply-add; the recorded event lists of the quick test match only
// this form.
__device__ __forceinline__ double duration(const RunConst& c, uint64_t iStep, double last)
{
    return fma(c.dt, (double)iStep, -last);
}

// The statistics state an event reads and updates (the FPTCG lists aside). An
// event reads it from the lane record in one batch of independent loads and
// writes it back at its end, so that it waits for the record once instead of
// once per branch.
struct EventState {
    double wt_tick_prev, wt_last, wt_sum, wt_sum2, wt_max;
    uint64_t wt_ticks, wt_n;
    int32_t wt_passed;
    double fpt_last_tick, fpt_sum;
    uint64_t fpt_n;
    int32_t fpt_to_update;
    int32_t fpt_init_index[2];
    double fpt_init_state[2];
    double fpt2_last_tick, fpt2_sum;
    uint64_t fpt2_n;
    int32_t fpt2_to_update;
    int32_t fpt2_idx[2];
    double tpt_last_tick, tpt_sum;
    uint64_t tpt_n;
    int32_t tpt_to_update;
    double tpt2_last_tick, tpt2_sum;
    uint64_t tpt2_n;
    int32_t tpt2_to_update, tpt2_initial_well;
};

#define LDPCP_EVENT_FIELDS(X)                                                                                  \
    X(wt_tick_prev) X(wt_last) X(wt_sum) X(wt_sum2) X(wt_max) X(wt_ticks) X(wt_n) X(wt_passed)                \
    X(fpt_last_tick) X(fpt_sum) X(fpt_n) X(fpt_to_update) X(fpt_init_index[0]) X(fpt_init_index[1])           \
    X(fpt_init_state[0]) X(fpt_init_state[1]) X(fpt2_last_tick) X(fpt2_sum) X(fpt2_n) X(fpt2_to_update)       \
    X(fpt2_idx[0]) X(fpt2_idx[1]) X(tpt_last_tick) X(tpt_sum) X(tpt_n) X(tpt_to_update) X(tpt2_last_tick)     \
    X(tpt2_sum) X(tpt2_n) X(tpt2_to_update) X(tpt2_initial_well)

__device__ __forceinline__ EventState load_event_state(const ldpcp_lane* o)
{
    EventState e;
#define LDPCP_LOAD(m) e.m = o->m;
    LDPCP_EVENT_FIELDS(LDPCP_LOAD)
#undef LDPCP_LOAD
    return e;
}

__device__ __forceinline__ void store_event_state(ldpcp_lane* o, const EventState& e)
{
#define LDPCP_STORE(m) o->m = e.m;
    LDPCP_EVENT_FIELDS(LDPCP_STORE)
#undef LDPCP_STORE
}

__device__ __forceinline__ void stats_event(const RunConst& c, const Flags& f, const RunPtrs& P, uint32_t traj,
                                            const Window& w, ldpcp_lane* o, const Lane& s, uint64_t iStep,
                                            double* ticks, int& err)
{
    const double t = c.dt * (double)iStep;
    double* const ev = c.ev_cap ? P.events + (uint64_t)traj * c.n_lists * c.ev_cap : nullptr;
    const bool single = (f.model == 0);
    EventState e = load_event_state(o);
    if (c.b_wait && single) {
        if (s.ent_well) {
            e.wt_tick_prev = (double)iStep;
            e.wt_passed = 1;
        } else if (s.ent_period && e.wt_passed) {
            const double tick = e.wt_tick_prev;
            if (e.wt_ticks < c.tick_cap) ticks[e.wt_ticks] = tick;
            if (e.wt_ticks > 0) {
                const double w = (tick - e.wt_last) * c.dt;
                e.wt_sum += w;
                e.wt_sum2 = fma(w, w, e.wt_sum2);
                if (e.wt_n == 0 || w > e.wt_max) e.wt_max = w;
                e.wt_n++;
            }
            e.wt_last = tick;
            e.wt_ticks++;
            e.wt_passed = 0;
        }
    }
    if (c.b_fpt) {
        if (single) {
            if (e.fpt_to_update && s.ent_period) {
                e.fpt_to_update = 0;
                e.fpt_last_tick = t;
                fpt1_arm(c, f, P, traj, w, e.fpt_init_index, e.fpt_init_state, s, err);
            }
            if (!e.fpt_to_update && s.ent_well) {
                e.fpt_to_update = 1;
                const double d = duration(c, iStep, e.fpt_last_tick);
                e.fpt_sum += d;
                if (ev) {
                    // the barrier logs (Single.h:2398-2404)
                    push_event(c, ev, LDPCP_LIST_FPT, e.fpt_n, d);
                    push_event(c, ev, LDPCP_LIST_INITIAL, e.fpt_n,
                               s.periodic == (double)e.fpt_init_index[0] ? e.fpt_init_state[0]
                                                                           : e.fpt_init_state[1]);
                    push_event(c, ev, LDPCP_LIST_INSTANT, e.fpt_n, s.state);
                }
                e.fpt_n++;
            }
        }
        if (e.fpt2_to_update && s.ent_period) {
            e.fpt2_to_update = 0;
            e.fpt2_last_tick = t;
            int idx[2];
            arm_indices(idx, s);
            e.fpt2_idx[0] = idx[0];
            e.fpt2_idx[1] = idx[1];
        }
        if (s.ent_period && !e.fpt2_to_update && ((double)e.fpt2_idx[0] != s.periodic) &&
            ((double)e.fpt2_idx[1] != s.periodic)) {
            const double d = duration(c, iStep, e.fpt2_last_tick);
            e.fpt2_sum += d;
            if (ev) push_event(c, ev, LDPCP_LIST_FPT2, e.fpt2_n, d);
            e.fpt2_n++;
            e.fpt2_last_tick = t;
            int idx[2];
            arm_indices(idx, s);
            e.fpt2_idx[0] = idx[0];
            e.fpt2_idx[1] = idx[1];
        }
    }
    if (c.b_tpt && single) {
        if (s.ent_period) {
            e.tpt_to_update = 0;
            e.tpt_last_tick = t;
        }
        if (!e.tpt_to_update && s.ent_well) {
            e.tpt_to_update = 1;
            const double d = duration(c, iStep, e.tpt_last_tick);
            e.tpt_sum += d;
            if (ev) push_event(c, ev, LDPCP_LIST_TPT, e.tpt_n, d);
            e.tpt_n++;
        }
        if (s.ent_period) {
            if (e.tpt2_to_update) {
                e.tpt2_to_update = 0;
            } else if (s.well != (double)e.tpt2_initial_well) {
                const double d = duration(c, iStep, e.tpt2_last_tick);
                e.tpt2_sum += d;
                if (ev) push_event(c, ev, LDPCP_LIST_TPT2, e.tpt2_n, d);
                e.tpt2_n++;
            }
            e.tpt2_last_tick = t;
            e.tpt2_initial_well = (int)s.well;
        }
    }
    if (c.b_cg && s.ent_period) {
        for (int j = 0; j < c.n_cg; ++j) {
            if (!cg_divides(c, j, s.well)) continue;
            if (o->cg_to_update[j]) {
                o->cg_to_update[j] = 0;
                o->cg_last_tick[j] = t;
                o->cg_initial_well[j] = (int)s.well;
            } else if (s.well != (double)o->cg_initial_well[j]) {
                const double d = duration(c, iStep, o->cg_last_tick[j]);
                o->cg_sum[j] += d;
                if (ev) push_event(c, ev, LDPCP_LIST_CG0 + j, o->cg_n[j], d);
                o->cg_n[j]++;
                o->cg_last_tick[j] = t;
                o->cg_initial_well[j] = (int)s.well;
            }
        }
    }
    store_event_state(o, e);
}

// Record (Single.h:1707-1765); frames are [field][frame][traj]
__device__ __forceinline__ void record(const RunConst& c, const Flags& f, const RunPtrs& P, uint32_t t, const Lane& s,
                                       uint64_t frame)
{
    if (frame >= c.frames_cap) return;
    const uint64_t plane = c.frames_cap * c.n_traj;
    const uint64_t o = frame * c.n_traj + t;
    P.frames[o] = s.x;
    if (c.save_all) {
        Lane e = s;
        lane_energy_force(e, c, f, false, true);
        P.frames[plane * 1 + o] = f.over ? 0.0 : s.v;
        P.frames[plane * 2 + o] = s.F;
        P.frames[plane * 3 + o] = e.E;
        P.frames[plane * 4 + o] = c.mhalf * (s.v * s.v);
        P.frames[plane * 5 + o] = s.state;
        P.frames[plane * 6 + o] = e.barrier;
    }
}

__device__ __forceinline__ void load_lane(Lane& s, const StateParams& sp, const Flags& f, const RunPtrs& P,
                                          uint32_t t)
{
    s.x = P.x[t];
    s.v = P.v[t];
    s.F = P.F[t];
    s.E = P.E[t];
    s.R = P.R[t];
    s.barrier = P.barrier[t];
    s.state = P.state[t];
    s.well = P.well[t];
    s.periodic = P.periodic[t];
    s.xp = P.xp[t];
    s.local = P.local[t];
    s.cnt0 = P.cnt0[t];
    s.cnt1 = P.cnt1[t];
    s.wbase = P.wbase[t];
    s.ent_well = s.ent_period = s.ent_local = false;
    lane_param(s, sp, f, false);  // rebuild heights and slopes from the stored state
}

__device__ __forceinline__ void store_lane(const Lane& s, const RunPtrs& P, uint32_t t)
{
    P.x[t] = s.x;
    P.v[t] = s.v;
    P.F[t] = s.F;
    P.E[t] = s.E;
    P.R[t] = s.R;
    P.barrier[t] = s.barrier;
    P.state[t] = s.state;
    P.well[t] = s.well;
    P.periodic[t] = s.periodic;
    P.xp[t] = s.xp;
    P.local[t] = s.local;
    P.cnt0[t] = s.cnt0;
    P.cnt1[t] = s.cnt1;
}

// IntegrateStep0 (Single.h:1619-1652, overdamped :2825-2832)
__global__ void __launch_bounds__(256) run_init_kernel(RunConst c, RunPtrs P)
{
    const Flags f = run_flags(c);
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= c.n_traj) return;
    ldpcp_lane* o = P.lanes + t;
    int err = 0;
    Lane s;
    s.v = s.F = s.E = s.R = s.barrier = 0.0;
    s.xp = s.local = s.local_prev = 0.0;
    s.x = c.x0;
    lane_indices(s, c, f, true);
    Philox4 w0;
    if (!f.replay) w0 = philox4x32_10(0u, 0u, c.traj_offset + t, 1u, c.key0, c.key1);
    uint64_t ng = 0, nf = 0;
    if (c.randomize_state) {
        const double u = f.replay ? P.f[t] : unit_f64(w0.w[2]);
        nf++;
        s.state = (u <= c.p0) ? 0.0 : 1.0;
    } else {
        s.state = c.initial_state;
    }
    o->state0_instance = s.state;
    s.wbase = (int)s.periodic - c.wb / 2;
    P.wbase[t] = s.wbase;
    const Window w = make_window(c, P, t, s.wbase);
    {
        const int a = f.model == 0 ? 0 : (int)s.local;
        const int k = win_slot(w, (int)s.periodic, err);
        if (k >= 0) win_put(w, a, (int)s.periodic, k, kValid | ((int)s.state ? kState : 0ull));
    }
    s.cnt0 = s.cnt1 = 0;
    lane_param(s, state_params(c, f), f);
    double gv = 0.0, gr = 0.0;
    if (!f.replay) box_muller_f64(w0.w[0], w0.w[1], gv, gr);
    if (c.randomize_v) {
        const double g = f.replay ? P.g[ng * c.n_traj + t] : gv;
        ng++;
        s.v = c.vsigma * g;
    } else {
        s.v = c.v0;
    }
    o->v0_instance = s.v;
    lane_energy_force(s, c, f);
    stats_init(c, f, P, t, w, o, s, err);
    if (f.over) {
        const double g = f.replay ? P.g[ng * c.n_traj + t] : gr;
        ng++;
        s.R = c.var * g;
    }
    record(c, f, P, t, s, 0);
    o->n_gauss = ng;
    o->n_flip = nf;
    o->error = err;
    store_lane(s, P, t);
}

// The step's rare branches, a hand-over and a statistics event, taken on the few
// steps where some trajectory of the wavefront has one (0.7 % of the steps of the
// quick test's 4-trajectory ranks).  They are inline: the step loop reaches them
// through one wave-uniform branch, and inline they use the kernel's constants
// in registers and address the LDS tables directly.

struct HandOver {
    double state;
    uint64_t n_catch;
    int err;
};

// Hand-over to the barrier just entered (Single.h:2025-2068; Double.h:2466-2510):
// the barrier left keeps its state and the step, the barrier entered takes its
// stored state and catches up on the steps since it was left.
__device__ __forceinline__ HandOver hand_over(const RunConst& c, const RunPtrs& P, const Flags& f, uint32_t t,
                                             const Window& w, uint64_t iStep, double periodic_prev, double periodic,
                                             double local_prev, double local, double state, uint64_t n_catch)
{
    HandOver r{state, n_catch, 0};
    const int ap = f.model == 0 ? 0 : (int)local_prev;
    const int ac = f.model == 0 ? 0 : (int)local;
    const int kp = win_slot(w, (int)periodic_prev, r.err);
    const int kc = win_slot(w, (int)periodic, r.err);
    if (r.err) return r;
    win_put(w, ap, (int)periodic_prev, kp, kValid | ((int)state ? kState : 0ull) | (iStep & kStepMask));
    const uint64_t ev = win_load(c, f, P, t, w, ac, (int)periodic, kc, r.err);
    const double prev_state = (ev & kState) ? 1.0 : 0.0;
    if (f.tau0) {
        r.state = prev_state;
        return r;
    }
    const double lap = (double)(iStep - (ev & kStepMask));
    double pc;
    if (c.equal_prob) pc = 0.5 * (1.0 - exp(c.c_lap * lap));
    else if (prev_state == 0.0) pc = c.K / (1.0 + c.K) * (1.0 - exp(c.c_lap * lap));
    else pc = 1.0 / (1.0 + c.K) * (1.0 - exp(c.c_lap * lap));
    double uc;
    if (f.replay) {
        if (n_catch >= c.catch_cap) {
            r.err = 4;
            return r;
        }
        uc = P.c[(uint64_t)t * c.catch_cap + n_catch];
    } else {
        // four catch-up uniforms per call: word n mod 4 of index n / 4
        const uint64_t q = n_catch >> 2;
        const Philox4 u4 = philox4x32_10((uint32_t)q, (uint32_t)(q >> 32), c.traj_offset + t, 2u, c.key0, c.key1);
        const uint32_t k = (uint32_t)n_catch & 3u;
        uc = unit_f64(k == 0 ? u4.w[0] : k == 1 ? u4.w[1] : k == 2 ? u4.w[2] : u4.w[3]);
    }
    r.n_catch = n_catch + 1;
    r.state = (uc < pc) ? ((prev_state == 1.0) ? 0.0 : 1.0) : prev_state;
    return r;
}

// stats_event of a step with ent_well or ent_period; returns its error code
__device__ __forceinline__ int stats_event_at(const RunConst& c, const RunPtrs& P, const Flags& f, uint32_t t,
                                              const Window& w, ldpcp_lane* o, uint64_t iStep, double well,
                                              double periodic, double state, bool ent_well, bool ent_period)
{
    Lane s;  // the fields the statistics read
    s.well = well;
    s.periodic = periodic;
    s.state = state;
    s.ent_well = ent_well;
    s.ent_period = ent_period;
    s.wbase = w.wbase;
    int err = 0;
    stats_event(c, f, P, t, w, o, s, iStep, P.ticks + (uint64_t)t * c.tick_cap, err);
    return err;
}

// Steps (s0, s1] of IntegrateStepEta / IntegrateStep
// (underdamped Single.h:1654-1706; overdamped :2833-2886; Double.h:1971-2024, 3475-3508).
// One instantiation per (model, integrator, TAU = 0, noise source): the flags
// become constants, so the step is straight-line code without the branches and
// constant reloads of a generic loop (a handle with few trajectories runs one
// wavefront, and its time per step is the latency of that code).
//
// Noise sources: kNoisePhilox draws Philox4x32-10 + Box-Muller inside the step
// loop; kNoiseReplay reads recorded streams; kNoiseBuffered reads the same
// Philox noise that run_noise_kernel wrote for the launch's steps, staged
// through LDS 32 steps at a time by direct global->LDS loads, so that the
// stepping wavefront's dependent chain holds the dynamics only (the noise is
// computed in parallel over all trajectory-steps beforehand).
constexpr int kNoisePhilox = 0, kNoiseReplay = 1, kNoiseBuffered = 2;
// records per LDS stage array of the buffered noise: 32 steps of 64
// trajectories and two rows of read-ahead padding
constexpr uint32_t kStageSlots = 34u * 64u;
// s_waitcnt immediate (gfx9 encoding: vmcnt bits 3:0 and 15:14, expcnt 6:4,
// lgkmcnt 11:8) that waits for vmcnt = 0 and lgkmcnt = 0 and leaves expcnt free
constexpr int kWaitVmLgkm0 = 0x0070;

// One buffered noise record per trajectory-step: Gaussian (lo, hi words) and
// the 32-bit flip word.
__device__ __forceinline__ void buffered_noise(const uint4& r, double& g, uint32_t& fw)
{
    g = __hiloint2double((int)r.y, (int)r.x);
    fw = r.z;
}

template <int MODEL, int OVER, int TAU0, int SRC>
__global__ void __launch_bounds__(256) run_step_kernel(RunConst c, RunPtrs P, uint64_t s0, uint64_t s1,
                                                       const uint4* __restrict__ noise)
{
    constexpr int REPLAY = (SRC == kNoiseReplay);
    constexpr Flags f{MODEL, OVER, TAU0, REPLAY};
    // the buffered launch (one wavefront per workgroup, few trajectories) keeps the
    // statistics record and the barrier table in LDS: an event's reads and a
    // hand-over's table lookups are on the step's critical path
    constexpr bool kLocal = (SRC == kNoiseBuffered);
    // A buffered launch keeps every lane of its wavefront: the stage loads copy
    // 64 noise records per instruction only with all 64 lanes on.  A lane
    // without a trajectory (beyond n_traj, or one whose record holds an error)
    // reads an existing trajectory's data, steps as a dead lane (nothing it
    // computes leaves its registers) and stores nothing.
    const uint32_t t_raw = blockIdx.x * blockDim.x + threadIdx.x;
    if (!kLocal && t_raw >= c.n_traj) return;
    const uint32_t t = t_raw < c.n_traj ? t_raw : c.n_traj - 1;
    ldpcp_lane* const og = P.lanes + t;
    const bool active = t_raw < c.n_traj && og->error == 0;
    if (!kLocal && !active) return;
    __shared__ ldpcp_lane lane_stats[64];
    __shared__ uint64_t win_cache[2 * kWinSlots * 64];
    __shared__ int win_cache_id[2 * kWinSlots * 64];
    __shared__ uint8_t win_cache_dirty[2 * kWinSlots * 64];
    ldpcp_lane* o = og;
    if constexpr (kLocal) {
        lane_stats[threadIdx.x] = *og;
        o = &lane_stats[threadIdx.x];
        for (int k = 0; k < 2 * kWinSlots; ++k) win_cache_id[k * 64 + threadIdx.x] = kWinEmpty;
    }
    int err = 0;
    const StateParams sp = state_params(c, f);
    Lane s;
    load_lane(s, sp, f, P, t);
    Window w = make_window(c, P, t, s.wbase);
    if constexpr (kLocal) {
        w.cached = true;
        w.ce = win_cache;
        w.cid = win_cache_id;
        w.dirty = win_cache_dirty;
    }
    uint64_t n_catch = o->n_catch;
    // frames every save_freq steps: the step of the next frame (wave-uniform, so the
    // test is a scalar compare)
    uint64_t next_save, frame;
    ldpcp_next_frame(c.run_begin, c.frame0, c.save_freq, s0, &next_save, &frame);
    // The barrier state as an integer, and the steps of this launch counted in
    // state 1 (StateCount, Single.h:2154-2165; a launch has at most 2^20 steps);
    // s.state, s.cnt0 and s.cnt1 are brought up to date where they are stored.
    // The state and the `dead` flag are plain integers in vector registers: as
    // lane masks, every join after a rare update would cost mask merges on the
    // common path.
    int st = (s.state == 1.0) ? 1 : 0;
    uint32_t c1 = 0;

    // The step loop runs all of the launch's steps for every trajectory (uniform
    // control flow).  A trajectory whose launch ends with an error stores its
    // record and state at that point and is dead: its later steps have no effect
    // outside registers.  live_m (wave-uniform) holds the lanes not dead, and
    // only their crossings take the rare branch.
    int dead = active ? 0 : 1;
    uint64_t live_m = __ballot(active);
    // (periodic index) x L of the cell the coming step starts in, the positions
    // [cell_lo, cell_hi] certainly inside that cell (cell_bounds) and, for the
    // double barrier, the barrier of its local half
    double pL = s.periodic * c.L;
    double cell_lo = 0.0, cell_hi = 0.0;
    cell_bounds(c, f, s.periodic, s.well, s.local, cell_lo, cell_hi);
    LocalBarrier lb{};
    if (f.model == 1) lb = local_barrier(c, sp, s.local);
    // The well, periodic and local indices change only on the rare branch: they
    // wait in LDS between its visits, so that no register carries them through
    // the common step.
    __shared__ double lane_cell[MODEL == 1 ? 3 : 2][256];
    lane_cell[0][threadIdx.x] = s.well;
    lane_cell[1][threadIdx.x] = s.periodic;
    if constexpr (MODEL == 1) lane_cell[2][threadIdx.x] = s.local;
    // stores the state after step `last`, with `counted` steps of the launch in the state counts
    auto finish = [&](uint64_t last, uint64_t counted) __attribute__((always_inline)) {
        s.state = st ? 1.0 : 0.0;
        s.cnt1 += c1;
        s.cnt0 += counted - c1;
        o->n_catch = n_catch;
        o->n_gauss += last - s0;
        if (!f.tau0) o->n_flip += last - s0;
        o->steps_done = (last == c.run_end) ? c.clock_end : last;  // a finished run's clock skips the dropped steps
        o->error = err;
        lane_param(s, sp, f, false);
        lane_energy_force(s, c, f, false, true);
        store_lane(s, P, t);
    };
    auto fail = [&](uint64_t iStep, uint64_t counted) __attribute__((always_inline)) {
        finish(iStep - 1, counted);
        dead = 1;
    };

    // One step with its noise: the Gaussian g and the flip uniform uw, either as
    // its 32-bit Philox word (tested against flip_threshold) or, replayed, as a double.
    // The common step is straight-line code; the hand-over and the statistics
    // event sit behind one wave-uniform test (the lanes whose indices changed),
    // taken on the few steps where some trajectory of the wavefront needs them.
    // `frames` (a compile-time flag) adds the frame test; a caller passes it only
    // for runs of steps that hold a frame step.
    // `istep` returns the step number: called only where the number is used (the rare
    // branch and the frame test), so the common step does not compute it.
    auto step = [&](auto istep, double g, auto uw, auto eq, auto frames) __attribute__((always_inline)) {
        if (f.over) {
            const double rf = c.var * g;
            s.x = fma(s.F, c.dtgm, s.x);
            s.x = s.x + rf;
            s.x = s.x + s.R;
            s.R = rf;
        } else {
            s.v = s.v + (c.minv * s.F) * c.halfdt;
            s.x = s.x + s.v * c.halfdt;
            s.v = c.expmgdt * s.v + c.var * g;
            s.x = s.x + s.v * c.halfdt;
        }
        // the indices are computed on the rare branch, which a step leaving
        // [cell_lo, cell_hi] takes (below)
        // per-step flip of the barrier under the particle (Single.h:2010-2015, 2078-2086);
        // equally probable states share one threshold, below 2^31 (p <= 1/2)
        if (!f.tau0) {
            bool flip;
            if constexpr (sizeof(uw) == sizeof(uint32_t)) {
                if constexpr (decltype(eq)::value) flip = uw < (uint32_t)c.t_flip;
                else flip = (uint64_t)uw < (st ? sp.t_flip[1] : sp.t_flip[0]);
            } else {
                if constexpr (decltype(eq)::value) flip = uw < c.p_flip;
                else flip = uw < (st ? sp.p_flip[1] : sp.p_flip[0]);
            }
            st = flip ? st ^ 1 : st;
        }
        // the force from the offset to the origin of the cell the step started in
        // (pL) and, for the double barrier, that cell's local barrier: exact while
        // the step stays in the cell; a step that leaves it takes the rare branch,
        // which recomputes F from x where the periodic or local index changed
        const double xp = s.x - pL;
        double F = f.model == 0 ? force_single_xp(xp, c.w, st ? sp.S[1] : sp.S[0], st ? sp.Sm[1] : sp.Sm[0])
                                : force_double_local(xp, lb, st != 0);
        c1 += (uint32_t)st;
        const uint64_t m_rare = (lanes_ult(s.x, cell_lo) | lanes_ugt(s.x, cell_hi)) & live_m;
        if (__builtin_expect(m_rare != 0, 0)) {
            s.well = lane_cell[0][threadIdx.x];
            s.periodic = lane_cell[1][threadIdx.x];
            if constexpr (MODEL == 1) s.local = lane_cell[2][threadIdx.x];
            lane_indices(s, c, f, false);
            lane_cell[0][threadIdx.x] = s.well;
            lane_cell[1][threadIdx.x] = s.periodic;
            if constexpr (MODEL == 1) lane_cell[2][threadIdx.x] = s.local;
            const uint64_t iStep = istep();
            const bool live = (dead == 0);
            const bool entering = f.model == 0 ? s.ent_period : s.ent_local;
            if (entering && live) {
                const HandOver r = hand_over(c, P, f, t, w, iStep, s.periodic_prev, s.periodic, s.local_prev,
                                             s.local, st ? 1.0 : 0.0, n_catch);
                n_catch = r.n_catch;
                err = r.err;
                if (err) {
                    c1 -= (uint32_t)st;  // this step is not counted
                    fail(iStep, iStep - 1 - s0);
                } else {
                    const int nst = (r.state == 1.0) ? 1 : 0;
                    c1 += (uint32_t)nst - (uint32_t)st;
                    st = nst;
                    F = lane_force(s, c, sp, f, st != 0);
                }
            }
            if ((s.ent_period || s.ent_well) && dead == 0) {
                err = stats_event_at(c, P, f, t, w, o, iStep, s.well, s.periodic, st ? 1.0 : 0.0, s.ent_well,
                                     s.ent_period);
                if (err) fail(iStep, iStep - s0);
            }
            live_m = __ballot(dead == 0);
            pL = s.periodic * c.L;
            cell_bounds(c, f, s.periodic, s.well, s.local, cell_lo, cell_hi);
            if (f.model == 1) lb = local_barrier(c, sp, s.local);
        }
        s.F = F;
        if (!f.over) s.v = s.v + (c.minv * s.F) * c.halfdt;
        if constexpr (decltype(frames)::value) {
            if (istep() == next_save) {
                if (dead == 0) {
                    s.state = st ? 1.0 : 0.0;
                    s.xp = s.x - pL;  // the double barrier's energy reads it
                    lane_param(s, sp, f, false);
                    record(c, f, P, t, s, frame);
                }
                frame++;
                next_save += c.save_freq;
            }
        }
    };

    // the buffered noise stages (declared once: `run` below has two instantiations)
    __shared__ uint4 stage_a[kStageSlots], stage_b[kStageSlots];
    auto run = [&](auto eq) __attribute__((always_inline)) {
        const uint64_t len = s1 - s0;
        if constexpr (SRC == kNoiseBuffered) {
            // Stage q of the launch (steps s0 + 1 + q L ..., L steps a stage) sits in
            // stage_a (q even) or stage_b (q odd), loaded by direct global->LDS loads
            // while the stage before it runs.  The even and odd stages are separate
            // code over separate arrays, so the compiler sees that a stage's LDS reads
            // never touch the array the next stage's loads are landing in and waits
            // for those loads only at the stage boundary.  The launch is one wavefront
            // per workgroup (64 threads), so a boundary's __syncthreads waits for this
            // wavefront alone.
            //
            // A stage holds rows of S records (one per trajectory of the wavefront),
            // row j at records j S .. j S + S - 1.  A handle of at most 64
            // trajectories is one wavefront (S = n_traj): its rows are contiguous in
            // the buffer ([step][trajectory]), so a stage is one block that each load
            // copies 64 records of (one load per 64 / S steps), and a stage holds as
            // many steps as the array does.  A wavefront of a larger handle (S = 64)
            // loads row by row, 32 steps a stage.
            const uint32_t lane = threadIdx.x;
            const bool one_wave = c.n_traj <= 64u;
            const uint32_t S = one_wave ? c.n_traj : 64u;
            // steps per stage: two rows of read-ahead padding left, an even count
            const uint32_t L = (kStageSlots / S - 2u) & ~1u;
            // the lane's record in a row: lanes beyond the wavefront's trajectories
            // (dead) read the first trajectory's
            const uint32_t slot = lane < S ? lane : 0u;
            const uint4* const wave_noise = noise + (uint64_t)blockIdx.x * 64u;
            auto load_stage = [&](uint4 (&buf)[kStageSlots], uint64_t q) {
                const uint64_t first = q * L;
                const uint64_t m = first >= len ? 0 : (len - first < L ? len - first : L);
                if (one_wave) {
                    const uint32_t count = (uint32_t)m * S;  // records of the stage, contiguous
                    const uint4* src = noise + first * S;
                    for (uint32_t k0 = 0; k0 < count; k0 += 64u) {
                        const uint32_t k = k0 + lane;
                        __builtin_amdgcn_global_load_lds((const void*)(src + (k < count ? k : 0u)),
                                                         (__attribute__((address_space(3))) void*)&buf[k0], 16, 0,
                                                         0);
                    }
                } else {
                    const uint32_t tl = t_raw < c.n_traj ? lane : 0u;  // in-bounds address for a dead lane
                    for (uint64_t j = 0; j < m; ++j)
                        __builtin_amdgcn_global_load_lds((const void*)(wave_noise + (first + j) * c.n_traj + tl),
                                                         (__attribute__((address_space(3))) void*)&buf[j * 64u], 16,
                                                         0, 0);
                }
            };
            // steps of stage q from buf, two per iteration with two record registers:
            // each step's record is read during the step before (the LDS latency overlaps
            // a step), and the values a step carries to the next alternate registers
            // instead of being copied.  A stage array has two rows beyond the stage's
            // steps, so the reads one and two rows ahead stay inside it; those rows
            // are never used.
            // rows r0 .. m - 1 of the stage (steps base + r0 + 1 .. base + m)
            auto stage_steps = [&](uint4 (&buf)[kStageSlots], uint64_t base, uint32_t r0, uint32_t m)
                                   __attribute__((always_inline)) {
                const uint4* rec = &buf[slot + r0 * S];
                uint4 ra = rec[0];
                const uint32_t n = m - r0;
                const uint64_t b0 = base + r0;
                const uint32_t pairs = n >> 1;
                for (uint32_t k = 0; k < pairs; ++k) {
                    // rows r0 + 2k and r0 + 2k + 1 (the last read of an even run is the row
                    // after it: padding row m at most)
                    const uint4 rb = rec[S];
                    double g;
                    uint32_t fw;
                    buffered_noise(ra, g, fw);
                    step([&] { return b0 + 2 * k + 1; }, g, fw, eq, std::false_type{});
                    ra = rec[2 * S];
                    buffered_noise(rb, g, fw);
                    step([&] { return b0 + 2 * k + 2; }, g, fw, eq, std::false_type{});
                    rec += 2 * S;
                }
                if (n & 1) {
                    double g;
                    uint32_t fw;
                    buffered_noise(ra, g, fw);
                    step([&] { return base + m; }, g, fw, eq, std::false_type{});
                }
            };
            // A stage runs without a frame test: its steps are split at the frame
            // steps it holds, each frame recorded after its step.
            auto run_stage = [&](uint4 (&buf)[kStageSlots], uint64_t q) __attribute__((always_inline)) {
                const uint64_t first = q * L;
                const uint32_t m = (uint32_t)(len - first < L ? len - first : L);
                const uint64_t base = uniform_u64(s0 + first);
                uint32_t r = 0;
                while (next_save <= base + m) {
                    const uint32_t e = (uint32_t)(next_save - base);
                    stage_steps(buf, base, r, e);
                    if (dead == 0) {
                        s.state = st ? 1.0 : 0.0;
                        s.xp = s.x - pL;  // the double barrier's energy reads it
                        lane_param(s, sp, f, false);
                        record(c, f, P, t, s, frame);
                    }
                    frame++;
                    next_save += c.save_freq;
                    r = e;
                }
                if (r < m) stage_steps(buf, base, r, m);
            };
            // stage q has landed (and every earlier LDS read is done): load stage q + 1
            // into the other array, whose stage q - 1 has been read
            auto boundary = [&]() {
                __builtin_amdgcn_s_waitcnt(kWaitVmLgkm0);
                __syncthreads();
            };
            load_stage(stage_a, 0);
            for (uint64_t q = 0; q * L < len; q += 2) {
                boundary();
                load_stage(stage_b, q + 1);
                run_stage(stage_a, q);
                if ((q + 1) * L >= len) break;
                boundary();
                load_stage(stage_a, q + 2);
                run_stage(stage_b, q + 1);
            }
        } else if constexpr (REPLAY) {
            // the noise of the next kAhead steps is in flight while a step runs
            // (indices clamp to s1, which the launch keeps above s0)
            constexpr int kAhead = 4;
            double gq[kAhead], uq[kAhead];
    #pragma unroll
            for (int d = 0; d < kAhead; ++d)
                replay_noise(c, f, P, t, (s0 + 1 + d < s1) ? s0 + 1 + d : s1, gq[d], uq[d]);
            for (uint64_t iStep = s0 + 1; iStep <= s1; ++iStep) {
                const double g = gq[0], u = uq[0];
    #pragma unroll
                for (int d = 0; d + 1 < kAhead; ++d) {
                    gq[d] = gq[d + 1];
                    uq[d] = uq[d + 1];
                }
                replay_noise(c, f, P, t, (iStep + kAhead < s1) ? iStep + kAhead : s1, gq[kAhead - 1], uq[kAhead - 1]);
                step([&] { return iStep; }, g, u, eq, std::true_type{});
            }
        } else {
            StepNoise sn;
            sn.q = 0;
            sn.g1 = 0.0;
            sn.w3 = 0;
            for (uint64_t iStep = s0 + 1; iStep <= s1; ++iStep) {
                double g;
                uint32_t fw;
                step_noise(c, t, iStep, sn, g, fw);
                step([&] { return iStep; }, g, fw, eq, std::true_type{});
            }
        }
    };
    // one copy of the loops per flip rule (uniform: a kernel argument)
    if (c.equal_prob) run(std::true_type{});
    else run(std::false_type{});
    s.well = lane_cell[0][threadIdx.x];
    s.periodic = lane_cell[1][threadIdx.x];
    if constexpr (MODEL == 1) {
        s.local = lane_cell[2][threadIdx.x];
        s.xp = s.x - s.periodic * c.L;
    }
    if (dead == 0) finish(s1, s1 - s0);
    if constexpr (kLocal) {
        if (active) {
            *og = *o;
            win_flush(w);
        }
    }
}

// MeanSquareDisplacementBruteForce (Single.h:24-45) per (trajectory, lag)
__global__ void __launch_bounds__(256) msd_kernel(const double* __restrict__ frames, uint64_t n_frames,
                                                  uint32_t n_traj, const uint64_t* __restrict__ lags,
                                                  uint32_t nlags, double* __restrict__ out)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const uint32_t l = blockIdx.y;
    if (t >= n_traj || l >= nlags) return;
    const int64_t i = (int64_t)lags[l];
    const int64_t n = (int64_t)n_frames;
    double acc = 0.0;
    if (n - i > 0) {
        for (int64_t k = 0; k < n - i; ++k) {
            const double d = frames[(uint64_t)(k + i) * n_traj + t] - frames[(uint64_t)k * n_traj + t];
            acc = fma(d, d, acc);  // `result += value * value`, contracted by the reference's compiler
        }
        acc /= (double)(n - i);
    }
    out[(uint64_t)t * nlags + l] = acc;
}

// The Philox noise of steps (s0, s0 + len] for every trajectory, laid out
// [step][trajectory] as buffered_noise reads it: the values step_noise
// would draw in the loop (Gaussian of the step's Box-Muller pair half, flip
// word w[2] / w[3]).
__global__ void __launch_bounds__(256) run_noise_kernel(RunConst c, uint64_t s0, uint64_t len,
                                                        uint4* __restrict__ noise)
{
    const uint64_t i = (uint64_t)blockIdx.x * blockDim.x + threadIdx.x;  // k * n_traj + t
    if (i >= len * c.n_traj) return;
    const uint64_t k = i / c.n_traj;
    const uint32_t t = (uint32_t)(i - k * c.n_traj);
    const uint64_t iStep = s0 + 1 + k;
    const uint64_t q = (iStep - 1) >> 1;
    const Philox4 w = philox4x32_10((uint32_t)q, (uint32_t)(q >> 32), c.traj_offset + t, 0u, c.key0, c.key1);
    double g0, g1;
    box_muller_f64(w.w[0], w.w[1], g0, g1);
    const bool second = (iStep - 1) & 1;
    const double g = second ? g1 : g0;
    const long long bits = __double_as_longlong(g);
    noise[k * c.n_traj + t] = make_uint4((uint32_t)bits, (uint32_t)(bits >> 32), second ? w.w[3] : w.w[2], 0u);
}

// The barrier table as the reference's restart file lists it (DataLocator.h:163-169):
// the stored state and hand-over step of barriers lo .. hi of one array, per
// trajectory.  A barrier the trajectory has not touched is created with its
// step-0 state, as the reference's table lookup creates it (DataLocator.h:100-188).
__global__ void __launch_bounds__(256) run_table_kernel(RunConst c, RunPtrs P, int array, int lo, int n_ids,
                                                        int32_t* __restrict__ states, uint64_t* __restrict__ steps)
{
    const uint32_t t = blockIdx.x * blockDim.x + threadIdx.x;
    const int k = (int)blockIdx.y;
    if (t >= c.n_traj || k >= n_ids) return;
    const Flags f = run_flags(c);
    const int id = lo + k;
    int err = 0;
    const Window w = make_window(c, P, t, P.wbase[t]);
    const int kw = win_slot(w, id, err);
    int st;
    uint64_t step = 0;
    if (kw >= 0) {
        const uint64_t v = win_load(c, f, P, t, w, array, id, kw, err);
        st = (v & kState) ? 1 : 0;
        step = v & kStepMask;
    } else {
        err = 0;  // beyond the trajectory's window: its state at creation, not stored
        st = barrier_init_state(c, f, P, t, array, id, err);
    }
    states[(uint64_t)t * n_ids + k] = st;
    steps[(uint64_t)t * n_ids + k] = step;
}

using StepKernel = void (*)(RunConst, RunPtrs, uint64_t, uint64_t, const uint4*);

template <int MODEL, int OVER, int SRC>
StepKernel step_kernel_tau(const RunConst& c)
{
    return c.tau_zero ? run_step_kernel<MODEL, OVER, 1, SRC> : run_step_kernel<MODEL, OVER, 0, SRC>;
}

template <int MODEL, int OVER>
StepKernel step_kernel_noise(const RunConst& c, int src)
{
    if (src == kNoiseReplay) return step_kernel_tau<MODEL, OVER, kNoiseReplay>(c);
    if (src == kNoiseBuffered) return step_kernel_tau<MODEL, OVER, kNoiseBuffered>(c);
    return step_kernel_tau<MODEL, OVER, kNoisePhilox>(c);
}

StepKernel step_kernel_for(const RunConst& c, int src)
{
    if (c.model == 0) return c.overdamped ? step_kernel_noise<0, 1>(c, src) : step_kernel_noise<0, 0>(c, src);
    return c.overdamped ? step_kernel_noise<1, 1>(c, src) : step_kernel_noise<1, 0>(c, src);
}

// Handles with at most this many trajectories fill a small part of the GPU, so
// each step's latency decides their time: they take the buffered noise.
constexpr uint32_t kBufferedMaxTraj = 8192;
constexpr uint64_t kBufferedLaunch = 1ull << 20;  // most steps per launch (and per noise buffer)
constexpr uint64_t kNoiseBufferBytes = 1ull << 28;  // noise buffer bound: fewer steps per launch above 16 trajectories

// Steps per launch of a buffered-noise handle of n trajectories whose runs are
// nsteps long.  Long launches keep the number of kernels a run enqueues small:
// a stream's hardware queue holds a bounded number of them, and a launch call
// that fills it waits, which would hold back the launch of the next handle.
uint64_t buffered_steps(uint64_t n, uint64_t nsteps)
{
    uint64_t k = kNoiseBufferBytes / (n * sizeof(uint4));
    if (k > kBufferedLaunch) k = kBufferedLaunch;
    if (k > nsteps) k = nsteps;
    return k ? k : 1;
}

}  // namespace

}  // namespace ldpcp

namespace ldpcp {
// A progress line printed from the handle's stream when the run reaches a step
// (LDPCP_PROGRESS_STEPS=N in the environment: a line every N steps), so that a
// long run reports what its kernels have done, not what the host enqueued.
struct ProgressNote {
    const void* handle;
    uint32_t traj_offset, n_traj;
    uint64_t step, end;
    std::chrono::steady_clock::time_point launched;
};

static void print_progress(void* p)
{
    const ProgressNote* n = (const ProgressNote*)p;
    const double sec = std::chrono::duration<double>(std::chrono::steady_clock::now() - n->launched).count();
    fprintf(stderr, "ldpcp: handle %p trajectories %u-%u reached step %llu of %llu (%.1f %%) after %.1f s\n",
            n->handle, n->traj_offset, n->traj_offset + n->n_traj - 1, (unsigned long long)n->step,
            (unsigned long long)n->end, 100.0 * (double)n->step / (double)(n->end ? n->end : 1), sec);
    fflush(stderr);
}
}  // namespace ldpcp

// A copy on the handle's own stream, waited for.  Run mode makes no call on the
// null stream: a process gets a few hardware queues (GPU_MAX_HW_QUEUES), and a
// queue taken by the null stream can leave two handles' streams sharing one, which
// runs their kernels by turns (seen as two of four concurrent ranks at half speed).
static hipError_t copy_on(hipStream_t s, void* dst, const void* src, size_t bytes, hipMemcpyKind kind)
{
    const hipError_t e = hipMemcpyAsync(dst, src, bytes, kind, s);
    return e == hipSuccess ? hipStreamSynchronize(s) : e;
}

struct ldpcp_run {
    ldpcp_run_params params;
    ldpcp::RunConst c;
    ldpcp::RunPtrs P;
    hipStream_t stream;
    void* state_block;
    void* frames_block;   // frames of a resumed run (prior frames + this run's), else null
    uint4* noise;         // buffered noise, [buffered_steps(n_traj, nsteps)][n_traj], for handles of few trajectories
    bool events_complete; // the event lists hold every event since step 0
    uint64_t n_frames;
    uint64_t step_base;   // absolute step the next integrate starts after
    uint64_t frames_prior; // frames recorded before this handle's run (a resumed image's)
    int phase;            // 0 fresh, 1 resumed (next integrate continues), 2 resumed run done
    bool integrated;
    bool pending;         // launched, not yet waited for
    int device;           // HIP device the handle's memory and stream live on
    std::vector<ldpcp::ProgressNote> progress;  // notes of the pending run's progress callbacks
};

using namespace ldpcp;

static int run_validate(const ldpcp_run_params* p, uint32_t n_traj)
{
    if (!p || n_traj == 0) {
        set_error("ldpcp_run_create: null params or zero trajectories");
        return LDPCP_EINVAL;
    }
    if (p->model != 0 && p->model != 1) {
        set_error("ldpcp_run_create: model must be 0 (single) or 1 (double), got %d", p->model);
        return LDPCP_EINVAL;
    }
    if (p->save_freq == 0 || p->timestep <= 0.0 || p->periodic_length <= 0.0 || p->mass <= 0.0) {
        set_error("ldpcp_run_create: save_freq, timestep, periodic_length and mass must be positive");
        return LDPCP_EINVAL;
    }
    if (p->overdamped && p->gamma <= 0.0) {
        set_error("ldpcp_run_create: the overdamped integrator needs gamma > 0");
        return LDPCP_EINVAL;
    }
    if (p->b_fptcg && (p->n_cg < 0 || p->n_cg > LDPCP_MAX_CG)) {
        set_error("ldpcp_run_create: n_cg must be in [0, %d]", LDPCP_MAX_CG);
        return LDPCP_EINVAL;
    }
    for (int j = 0; p->b_fptcg && j < p->n_cg; ++j)
        if (p->cg_size[j] == 0) {
            set_error("ldpcp_run_create: FPTCG size must be non-zero");
            return LDPCP_EINVAL;
        }
    if (p->barrier_window < 4) {
        set_error("ldpcp_run_create: barrier_window must be >= 4");
        return LDPCP_EINVAL;
    }
    return LDPCP_OK;
}

extern "C" int ldpcp_run_create(const ldpcp_run_params* p, uint32_t n_traj, uint32_t traj_offset, ldpcp_run** out)
{
    if (!out) {
        set_error("ldpcp_run_create: null output handle");
        return LDPCP_EINVAL;
    }
    *out = nullptr;
    int rc = run_validate(p, n_traj);
    if (rc) return rc;
    ldpcp_run* h = new ldpcp_run();
    h->params = *p;
    if (hipGetDevice(&h->device) != hipSuccess) h->device = 0;
    make_const(*p, h->c);
    h->c.n_traj = n_traj;
    h->c.traj_offset = traj_offset;
    h->n_frames = h->c.frames_cap;
    const uint64_t n = n_traj;
    const uint64_t fields = p->save_all ? kFields : 1;
    const uint64_t bytes_soa = n * (11 * sizeof(double) + 2 * sizeof(uint64_t));
    const uint64_t bytes_lanes = n * sizeof(ldpcp_lane);
    const uint64_t bytes_win = n * 2 * (uint64_t)p->barrier_window * sizeof(uint64_t);
    const uint64_t bytes_wbase = n * sizeof(int32_t);
    const uint64_t bytes_frames = fields * h->c.frames_cap * n * sizeof(double);
    const uint64_t bytes_ticks = n * p->tick_capacity * sizeof(double);
    const uint64_t bytes_bdef = n * sizeof(double);
    const uint64_t total =
        bytes_soa + bytes_lanes + bytes_win + bytes_wbase + bytes_frames + bytes_ticks + bytes_bdef + 1024;
    hipError_t e = hipMalloc(&h->state_block, total);
    if (e != hipSuccess) {
        set_error("ldpcp_run_create: hipMalloc(%llu bytes) failed: %s", (unsigned long long)total, hipGetErrorString(e));
        delete h;
        return LDPCP_ENOMEM;
    }
    char* b = (char*)h->state_block;
    double** dptrs[11] = {&h->P.x, &h->P.v, &h->P.F, &h->P.E, &h->P.R, &h->P.barrier, &h->P.state,
                          &h->P.well, &h->P.periodic, &h->P.xp, &h->P.local};
    for (int k = 0; k < 11; ++k) {
        *dptrs[k] = (double*)b;
        b += n * sizeof(double);
    }
    h->P.cnt0 = (uint64_t*)b;
    b += n * sizeof(uint64_t);
    h->P.cnt1 = (uint64_t*)b;
    b += n * sizeof(uint64_t);
    h->P.lanes = (ldpcp_lane*)b;
    b += bytes_lanes;
    h->P.win = (uint64_t*)b;
    b += bytes_win;
    h->P.wbase = (int32_t*)b;
    b += bytes_wbase;
    b = (char*)(((uintptr_t)b + 255) & ~(uintptr_t)255);
    h->P.frames = (double*)b;
    b += bytes_frames;
    h->P.ticks = (double*)b;
    b += bytes_ticks;
    h->P.bdefault = (double*)b;
    h->P.events = nullptr;
    h->P.g = h->P.f = h->P.c = nullptr;
    h->P.binit = nullptr;
    e = hipStreamCreate(&h->stream);
    if (e != hipSuccess) {
        set_error("ldpcp_run_create: hipStreamCreate failed: %s", hipGetErrorString(e));
        (void)hipFree(h->state_block);
        delete h;
        return LDPCP_EDEVICE;
    }
    // the buffered-noise scratch and the event lists are allocated here, not at
    // launch: an allocation between launches could wait for other handles' runs
    if (n <= kBufferedMaxTraj) e = hipMalloc(&h->noise, buffered_steps(n, h->c.nsteps_eff) * n * sizeof(uint4));
    if (e == hipSuccess && h->c.n_lists)
        e = hipMalloc(&h->P.events, n * (uint64_t)h->c.n_lists * h->c.ev_cap * sizeof(double));
    if (e != hipSuccess) {
        set_error("ldpcp_run_create: hipMalloc of the noise buffer or event lists failed: %s", hipGetErrorString(e));
        if (h->noise) (void)hipFree(h->noise);
        (void)hipStreamDestroy(h->stream);
        (void)hipFree(h->state_block);
        delete h;
        return LDPCP_ENOMEM;
    }
    e = hipMemsetAsync(h->P.bdefault, 0, bytes_bdef, h->stream);
    if (e != hipSuccess) {
        set_error("ldpcp_run_create: hipMemsetAsync failed: %s", hipGetErrorString(e));
        (void)hipStreamDestroy(h->stream);
        (void)hipFree(h->state_block);
        delete h;
        return LDPCP_EDEVICE;
    }
    *out = h;
    return LDPCP_OK;
}

extern "C" int ldpcp_run_set_replay(ldpcp_run* h, const double* gauss, uint64_t n_gauss, const double* flip,
                                    uint64_t n_flip, const double* catchup, uint64_t catch_cap, const int8_t* binit,
                                    int32_t bmin, int32_t bn)
{
    if (h && h->pending) {
        set_error("ldpcp_run_set_replay: the handle has a launched run; wait for it first");
        return LDPCP_EINVAL;
    }
    if (!h || !gauss || (!flip && n_flip) || (!catchup && catch_cap) || (!binit && bn)) {
        set_error("ldpcp_run_set_replay: null argument");
        return LDPCP_EINVAL;
    }
    const uint64_t n = h->c.n_traj;
    const uint64_t end = h->step_base + h->c.nsteps_eff;  // replay streams are indexed by absolute step
    const uint64_t need_g = h->c.init_gauss + end;
    const uint64_t need_f = h->c.init_flip + (h->c.tau_zero ? 0 : end);
    if (n_gauss < need_g || n_flip < need_f) {
        set_error("ldpcp_run_set_replay: streams too short (gauss %llu < %llu or flip %llu < %llu)",
                  (unsigned long long)n_gauss, (unsigned long long)need_g, (unsigned long long)n_flip,
                  (unsigned long long)need_f);
        return LDPCP_ERANGE;
    }
    const uint64_t bytes = (n_gauss + n_flip) * n * sizeof(double) + n * catch_cap * sizeof(double) +
                           2 * n * (uint64_t)bn + 256;
    void* block = nullptr;
    LDPCP_HIP_TRY(hipMalloc(&block, bytes));
    char* b = (char*)block;
    double* g = (double*)b;
    b += n_gauss * n * sizeof(double);
    double* f = (double*)b;
    b += n_flip * n * sizeof(double);
    double* c = (double*)b;
    b += n * catch_cap * sizeof(double);
    int8_t* bi = (int8_t*)b;
    hipError_t e = copy_on(h->stream, g, gauss, n_gauss * n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && n_flip) e = copy_on(h->stream, f, flip, n_flip * n * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && catch_cap)
        e = copy_on(h->stream, c, catchup, n * catch_cap * sizeof(double), hipMemcpyHostToDevice);
    if (e == hipSuccess && bn) e = copy_on(h->stream, bi, binit, 2 * n * (uint64_t)bn, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        (void)hipFree(block);
        set_error("ldpcp_run_set_replay: copy to the device failed: %s", hipGetErrorString(e));
        return LDPCP_EDEVICE;
    }
    if (h->c.replay) (void)hipFree((void*)h->P.g);  // the streams of an earlier call
    h->P.g = g;
    h->P.f = f;
    h->P.c = c;
    h->P.binit = bi;
    h->c.replay = 1;
    h->c.n_gauss = n_gauss;
    h->c.n_flip = n_flip;
    h->c.catch_cap = catch_cap;
    h->c.bmin = bmin;
    h->c.bn = bn;
    return LDPCP_OK;
}

extern "C" int ldpcp_run_launch(ldpcp_run* h, uint64_t steps_per_launch)
{
    if (!h) {
        set_error("ldpcp_run_launch: null handle");
        return LDPCP_EINVAL;
    }
    if (h->pending) {
        set_error("ldpcp_run_launch: the handle's previous run has not been waited for");
        return LDPCP_EINVAL;
    }
    if (h->phase == 2) {
        set_error("ldpcp_run_launch: this handle already continued a checkpoint; create a new handle");
        return LDPCP_EINVAL;
    }
    if (steps_per_launch == 0) steps_per_launch = 1ull << 20;
    const uint64_t n = h->c.n_traj;
    const dim3 block(256), grid((unsigned)((n + 255) / 256));
    if (h->phase == 0) {
        h->events_complete = h->c.n_lists > 0;
        LDPCP_HIP_TRY(hipMemsetAsync(h->P.lanes, 0, n * sizeof(ldpcp_lane), h->stream));
        LDPCP_HIP_TRY(hipMemsetAsync(h->P.win, 0, n * 2 * (uint64_t)h->c.wb * sizeof(uint64_t), h->stream));
        const uint64_t fields = h->params.save_all ? kFields : 1;
        LDPCP_HIP_TRY(hipMemsetAsync(h->P.frames, 0, fields * h->c.frames_cap * n * sizeof(double), h->stream));
        // unused tick slots are part of a checkpoint image: keep them zero, as a resumed
        // handle has them, so that images of equal states are equal bytes
        if (h->c.tick_cap)
            LDPCP_HIP_TRY(hipMemsetAsync(h->P.ticks, 0, n * h->c.tick_cap * sizeof(double), h->stream));
        hipLaunchKernelGGL(run_init_kernel, grid, block, 0, h->stream, h->c, h->P);
        LDPCP_HIP_TRY(hipGetLastError());
    }
    // steps (step_base, step_base + NSTEPS]: a resumed run continues the absolute step
    // count (Single.h:1420, 1471), which is also the Philox counter of its noise
    const uint64_t end = h->step_base + h->c.nsteps_eff;
    h->c.run_begin = h->step_base;
    h->c.run_end = end;
    h->c.clock_end = h->step_base + h->params.nsteps;
    h->c.frame0 = h->phase == 0 ? 1 : h->frames_prior;
    const bool buffered = !h->c.replay && n <= kBufferedMaxTraj;
    const int src = h->c.replay ? kNoiseReplay : buffered ? kNoiseBuffered : kNoisePhilox;
    const StepKernel step = step_kernel_for(h->c, src);
    if (buffered) steps_per_launch = std::min(steps_per_launch, buffered_steps(n, h->c.nsteps_eff));
    const dim3 sblock(buffered ? 64 : 256), sgrid((unsigned)((n + sblock.x - 1) / sblock.x));
    uint64_t every = 0;
    if (const char* e = getenv("LDPCP_PROGRESS_STEPS")) every = strtoull(e, nullptr, 10);
    const auto launched = std::chrono::steady_clock::now();
    h->progress.clear();
    if (every) h->progress.reserve((end - h->step_base) / every + 2);  // stable addresses for the callbacks
    for (uint64_t s0 = h->step_base; s0 < end; s0 += steps_per_launch) {
        const uint64_t s1 = (s0 + steps_per_launch < end) ? s0 + steps_per_launch : end;
        if (buffered) {
            const dim3 ngrid((unsigned)(((s1 - s0) * n + 255) / 256));
            hipLaunchKernelGGL(run_noise_kernel, ngrid, dim3(256), 0, h->stream, h->c, s0, s1 - s0, h->noise);
            LDPCP_HIP_TRY(hipGetLastError());
        }
        step<<<sgrid, sblock, 0, h->stream>>>(h->c, h->P, s0, s1, h->noise);
        LDPCP_HIP_TRY(hipGetLastError());
        if (every && (s1 / every != s0 / every || s1 == end) && h->progress.size() < h->progress.capacity()) {
            h->progress.push_back({h, h->c.traj_offset, h->c.n_traj, s1, end, launched});
            LDPCP_HIP_TRY(hipLaunchHostFunc(h->stream, print_progress, &h->progress.back()));
        }
    }
    h->pending = true;
    if (h->phase == 1) h->phase = 2;
    return LDPCP_OK;
}

extern "C" int ldpcp_run_wait(ldpcp_run* h)
{
    if (!h || !h->pending) {
        set_error("ldpcp_run_wait: null handle or nothing launched");
        return LDPCP_EINVAL;
    }
    const uint64_t n = h->c.n_traj;
    h->pending = false;
    LDPCP_HIP_TRY(hipStreamSynchronize(h->stream));
    h->integrated = true;
    // surface the first per-lane error
    std::vector<ldpcp_lane> lanes(n);
    LDPCP_HIP_TRY(copy_on(h->stream, lanes.data(), h->P.lanes, n * sizeof(ldpcp_lane), hipMemcpyDeviceToHost));
    for (uint64_t t = 0; t < n; ++t) {
        if (lanes[t].error) {
            const char* what = lanes[t].error == 2   ? "replay barrier table does not cover a touched barrier"
                               : lanes[t].error == 3 ? "trajectory left its barrier window (raise barrier_window)"
                                                     : "replay catch-up stream exhausted";
            set_error("ldpcp_run_integrate: trajectory %llu: %s", (unsigned long long)t, what);
            return LDPCP_ERANGE;
        }
    }
    return LDPCP_OK;
}

extern "C" int ldpcp_run_integrate(ldpcp_run* h, uint64_t steps_per_launch)
{
    const int rc = ldpcp_run_launch(h, steps_per_launch);
    return rc ? rc : ldpcp_run_wait(h);
}

extern "C" int ldpcp_run_lanes(ldpcp_run* h, uint32_t begin, uint32_t n, ldpcp_lane* out)
{
    if (h && h->pending) {  // results of a launched run: wait for it first
        const int rc = ldpcp_run_wait(h);
        if (rc) return rc;
    }
    if (!h || !out || (uint64_t)begin + n > h->c.n_traj) {
        set_error("ldpcp_run_lanes: bad handle, output or range");
        return LDPCP_EINVAL;
    }
    LDPCP_HIP_TRY(copy_on(h->stream, out, h->P.lanes + begin, (uint64_t)n * sizeof(ldpcp_lane), hipMemcpyDeviceToHost));
    std::vector<double> buf((uint64_t)n);
    struct Col {
        const double* src;
        size_t off;
    } cols[] = {{h->P.x, offsetof(ldpcp_lane, x)},           {h->P.v, offsetof(ldpcp_lane, v)},
                {h->P.F, offsetof(ldpcp_lane, f)},           {h->P.E, offsetof(ldpcp_lane, e)},
                {h->P.R, offsetof(ldpcp_lane, r_prev)},      {h->P.state, offsetof(ldpcp_lane, state)},
                {h->P.barrier, offsetof(ldpcp_lane, barrier)}, {h->P.well, offsetof(ldpcp_lane, well)},
                {h->P.periodic, offsetof(ldpcp_lane, periodic)}, {h->P.xp, offsetof(ldpcp_lane, xp)},
                {h->P.local, offsetof(ldpcp_lane, local_barrier)}};
    for (const Col& col : cols) {
        LDPCP_HIP_TRY(copy_on(h->stream, buf.data(), col.src + begin, (uint64_t)n * sizeof(double), hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) memcpy((char*)(out + i) + col.off, &buf[i], sizeof(double));
    }
    std::vector<uint64_t> cnt((uint64_t)n);
    for (int s = 0; s < 2; ++s) {
        LDPCP_HIP_TRY(copy_on(h->stream, cnt.data(), (s ? h->P.cnt1 : h->P.cnt0) + begin, (uint64_t)n * sizeof(uint64_t),
                                hipMemcpyDeviceToHost));
        for (uint32_t i = 0; i < n; ++i) out[i].state_count[s] = (double)cnt[i];
    }
    return LDPCP_OK;
}

extern "C" uint64_t ldpcp_run_n_frames(const ldpcp_run* h) { return h ? h->n_frames : 0; }

extern "C" int ldpcp_run_frames(ldpcp_run* h, int32_t field, uint32_t begin, uint32_t n, double* out)
{
    if (h && h->pending) {  // results of a launched run: wait for it first
        const int rc = ldpcp_run_wait(h);
        if (rc) return rc;
    }
    if (!h || !out || (uint64_t)begin + n > h->c.n_traj || field < 0 || field >= kFields ||
        (field > 0 && !h->params.save_all)) {
        set_error("ldpcp_run_frames: bad handle, field or range");
        return LDPCP_EINVAL;
    }
    const uint64_t nf = h->n_frames, nt = h->c.n_traj;
    std::vector<double> plane(nf * nt);
    LDPCP_HIP_TRY(copy_on(h->stream, plane.data(), h->P.frames + (uint64_t)field * nf * nt, nf * nt * sizeof(double),
                            hipMemcpyDeviceToHost));
    for (uint32_t i = 0; i < n; ++i)
        for (uint64_t f = 0; f < nf; ++f) out[(uint64_t)i * nf + f] = plane[f * nt + begin + i];
    return LDPCP_OK;
}

extern "C" int ldpcp_run_ticks(ldpcp_run* h, uint32_t begin, uint32_t n, double* out)
{
    if (h && h->pending) {  // results of a launched run: wait for it first
        const int rc = ldpcp_run_wait(h);
        if (rc) return rc;
    }
    if (!h || !out || (uint64_t)begin + n > h->c.n_traj || h->c.tick_cap == 0) {
        set_error("ldpcp_run_ticks: bad handle, range, or no tick capacity");
        return LDPCP_EINVAL;
    }
    LDPCP_HIP_TRY(copy_on(h->stream, out, h->P.ticks + (uint64_t)begin * h->c.tick_cap,
                            (uint64_t)n * h->c.tick_cap * sizeof(double), hipMemcpyDeviceToHost));
    return LDPCP_OK;
}

extern "C" int ldpcp_run_msd(ldpcp_run* h, const uint64_t* lags, uint32_t nlags, double* out)
{
    if (h && h->pending) {  // results of a launched run: wait for it first
        const int rc = ldpcp_run_wait(h);
        if (rc) return rc;
    }
    if (!h || !lags || !out || nlags == 0 || nlags > 65535) {
        set_error("ldpcp_run_msd: bad arguments");
        return LDPCP_EINVAL;
    }
    const uint64_t nt = h->c.n_traj;
    uint64_t* d_lags = nullptr;
    double* d_out = nullptr;
    LDPCP_HIP_TRY(hipMalloc(&d_lags, nlags * sizeof(uint64_t)));
    LDPCP_HIP_TRY(hipMalloc(&d_out, nt * nlags * sizeof(double)));
    LDPCP_HIP_TRY(copy_on(h->stream, d_lags, lags, nlags * sizeof(uint64_t), hipMemcpyHostToDevice));
    const dim3 block(256), grid((unsigned)((nt + 255) / 256), nlags);
    hipLaunchKernelGGL(msd_kernel, grid, block, 0, h->stream, h->P.frames, h->n_frames, (uint32_t)nt, d_lags, nlags,
                       d_out);
    LDPCP_HIP_TRY(hipGetLastError());
    LDPCP_HIP_TRY(hipStreamSynchronize(h->stream));
    LDPCP_HIP_TRY(copy_on(h->stream, out, d_out, nt * nlags * sizeof(double), hipMemcpyDeviceToHost));
    (void)hipFree(d_lags);
    (void)hipFree(d_out);
    return LDPCP_OK;
}

// ---------------------------------------------------------------------------
// ldpcp_run_msd without device allocation.  ldpcp_run_msd stages the lags and
// the per-trajectory results through two hipMalloc'd buffers; this variant
// stages them through `scratch`, a fixed array in the library's device image
// (each device holds its own copy for the life of the process):
//   [0, 512 KiB)   lags (up to 65535 u64)
//   [512 KiB, 8 MiB) results, n_traj * nlags doubles (at most 983040)
// Every call on a device shares that device's copy, so calls that overlap on
// one device (from several host threads) race.  Each call writes every result
// it reads back, so nothing needs clearing between calls.
// ---------------------------------------------------------------------------
namespace {
constexpr uint64_t kScratchBytes = 8ull << 20;
constexpr uint64_t kScratchLagBytes = 65536ull * sizeof(uint64_t);
constexpr uint64_t kScratchResults = (kScratchBytes - kScratchLagBytes) / sizeof(double);
constexpr int kScratchDevices = 64;
}  // namespace

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// Device address of `scratch` on device `dev` (which must be current), looked up
// once per device: the array's host-side name is not a device address.
static int scratch_address(int dev, void** out)
{
    static std::mutex mu;
    static void* addr[kScratchDevices] = {};
    if (dev < 0 || dev >= kScratchDevices) {
        set_error("ldpcp_run_msd_noalloc: device %d beyond the %d the library tracks", dev, kScratchDevices);
        return LDPCP_EINVAL;
    }
    std::lock_guard<std::mutex> lock(mu);
    if (!addr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p) {
            set_error("ldpcp_run_msd_noalloc: hipGetSymbolAddress(scratch) failed: %s", hipGetErrorString(e));
            return LDPCP_EDEVICE;
        }
        addr[dev] = p;
    }
    *out = addr[dev];
    return LDPCP_OK;
}

extern "C" int ldpcp_run_msd_noalloc(ldpcp_run* h, const uint64_t* lags, uint32_t nlags, double* out)
{
    if (!h || !lags || !out || nlags == 0 || nlags > 65535) {
        set_error("ldpcp_run_msd_noalloc: bad arguments");
        return LDPCP_EINVAL;
    }
    if (h->pending) {
        set_error("ldpcp_run_msd_noalloc: the handle has a launched run; wait for it first");
        return LDPCP_EINVAL;
    }
    const uint64_t nt = h->c.n_traj;
    if (nt * nlags > kScratchResults) {
        set_error("ldpcp_run_msd_noalloc: %llu trajectories x %u lags exceed the %llu results the fixed "
                  "scratch holds; use ldpcp_run_msd",
                  (unsigned long long)nt, nlags, (unsigned long long)kScratchResults);
        return LDPCP_EINVAL;
    }
    int cur = 0;
    LDPCP_HIP_TRY(hipGetDevice(&cur));
    if (cur != h->device) LDPCP_HIP_TRY(hipSetDevice(h->device));
    void* base = nullptr;
    int rc = scratch_address(h->device, &base);
    hipError_t e = hipSuccess;
    if (rc == LDPCP_OK) {
        uint64_t* d_lags = (uint64_t*)base;
        double* d_out = (double*)((char*)base + kScratchLagBytes);
        e = hipMemcpyAsync(d_lags, lags, nlags * sizeof(uint64_t), hipMemcpyHostToDevice, h->stream);
        if (e == hipSuccess) {
            const dim3 block(256), grid((unsigned)((nt + 255) / 256), nlags);
            hipLaunchKernelGGL(msd_kernel, grid, block, 0, h->stream, h->P.frames, h->n_frames, (uint32_t)nt,
                               (const uint64_t*)d_lags, nlags, d_out);
            e = hipGetLastError();
        }
        if (e == hipSuccess)
            e = hipMemcpyAsync(out, d_out, nt * nlags * sizeof(double), hipMemcpyDeviceToHost, h->stream);
        if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
        if (e != hipSuccess) {
            set_error("ldpcp_run_msd_noalloc: %s", hipGetErrorString(e));
            rc = LDPCP_EDEVICE;
        }
    }
    if (cur != h->device) (void)hipSetDevice(cur);
    return rc;
}

// ---------------------------------------------------------------------------
// Checkpoints: the device state of a handle as one self-describing image
// (the state TrajectoryData::WriteRestartFile saves, langevinDynamicsSingleBarrier.h:385-576,
// with running sums in place of event lists and Philox counters in place of MKL bank
// positions).
// ---------------------------------------------------------------------------
namespace {

struct CkptLayout {
    uint64_t lanes, wbase, win, frames, ticks, events, total;
};

// entries the image holds for the event lists: min(count, capacity) per list
uint64_t packed_events(const ldpcp_lane* lanes, uint64_t n, uint32_t n_lists, uint64_t cap)
{
    uint64_t k = 0;
    for (uint64_t t = 0; t < n; ++t)
        for (uint32_t l = 0; l < n_lists; ++l) {
            const uint64_t c = ldpcp_list_count(lanes + t, (int)l);
            k += c < cap ? c : cap;
        }
    return k;
}

CkptLayout ckpt_layout(uint64_t n, uint64_t wb, uint64_t fields, uint64_t n_frames, uint64_t tick_cap,
                       uint64_t n_events)
{
    CkptLayout L;
    uint64_t o = sizeof(ldpcp_checkpoint_header);
    L.lanes = o;
    o += n * sizeof(ldpcp_lane);
    L.wbase = o;
    o += (n * sizeof(int32_t) + 7) & ~7ull;
    L.win = o;
    o += n * 2 * wb * sizeof(uint64_t);
    L.frames = o;
    o += fields * n_frames * n * sizeof(double);
    L.ticks = o;
    o += n * tick_cap * sizeof(double);
    L.events = o;
    o += n_events * sizeof(double);
    L.total = o;
    return L;
}

uint64_t run_fields(const ldpcp_run* h) { return h->params.save_all ? kFields : 1; }

// the step clock a checkpoint holds: where the next run continues from (the
// reference's previous + NUMBEROFSTEPS, ldpcp_schedule.h)
uint64_t run_steps_done(const ldpcp_run* h) { return h->phase == 1 ? h->step_base : h->step_base + h->params.nsteps; }

// frames a checkpoint holds
uint64_t run_frames_done(const ldpcp_run* h) { return h->phase == 1 ? h->frames_prior : h->n_frames; }

CkptLayout run_layout(const ldpcp_run* h, uint64_t n_events)
{
    return ckpt_layout(h->c.n_traj, (uint64_t)h->c.wb, run_fields(h), run_frames_done(h), h->c.tick_cap, n_events);
}

}  // namespace

extern "C" int ldpcp_run_set_barrier_default(ldpcp_run* h, const double* state)
{
    if (h && h->pending) {
        set_error("ldpcp_run_set_barrier_default: the handle has a launched run; wait for it first");
        return LDPCP_EINVAL;
    }
    if (!h || !state) {
        set_error("ldpcp_run_set_barrier_default: null argument");
        return LDPCP_EINVAL;
    }
    for (uint32_t t = 0; t < h->c.n_traj; ++t)
        if (state[t] != 0.0 && state[t] != 1.0) {
            set_error("ldpcp_run_set_barrier_default: trajectory %u: state must be 0 or 1", t);
            return LDPCP_EINVAL;
        }
    LDPCP_HIP_TRY(copy_on(h->stream, h->P.bdefault, state, h->c.n_traj * sizeof(double), hipMemcpyHostToDevice));
    return LDPCP_OK;
}

extern "C" int ldpcp_run_checkpoint_size(ldpcp_run* h, uint64_t* bytes)
{
    if (!h || !bytes) {
        set_error("ldpcp_run_checkpoint_size: null argument");
        return LDPCP_EINVAL;
    }
    uint64_t n_events = 0;
    if (h->c.n_lists) {
        std::vector<ldpcp_lane> lanes(h->c.n_traj);
        const int rc = ldpcp_run_lanes(h, 0, h->c.n_traj, lanes.data());
        if (rc) return rc;
        n_events = packed_events(lanes.data(), h->c.n_traj, (uint32_t)h->c.n_lists, h->c.ev_cap);
    }
    *bytes = run_layout(h, n_events).total;
    return LDPCP_OK;
}

extern "C" int ldpcp_run_checkpoint(ldpcp_run* h, void* buf, uint64_t bytes)
{
    if (h && h->pending) {  // results of a launched run: wait for it first
        const int rc = ldpcp_run_wait(h);
        if (rc) return rc;
    }
    if (!h || !buf) {
        set_error("ldpcp_run_checkpoint: null argument");
        return LDPCP_EINVAL;
    }
    if (!h->integrated && h->phase != 1) {
        set_error("ldpcp_run_checkpoint: the handle holds no state yet (integrate or resume first)");
        return LDPCP_EINVAL;
    }
    const uint64_t n = h->c.n_traj, wb = (uint64_t)h->c.wb, fields = run_fields(h);
    std::vector<ldpcp_lane> lv(n);
    int rc = ldpcp_run_lanes(h, 0, (uint32_t)n, lv.data());
    if (rc) return rc;
    const uint32_t n_lists = (uint32_t)h->c.n_lists;
    const uint64_t ecap = h->c.ev_cap;
    const CkptLayout L = run_layout(h, packed_events(lv.data(), n, n_lists, ecap));
    if (bytes < L.total) {
        set_error("ldpcp_run_checkpoint: buffer of %llu bytes, need %llu", (unsigned long long)bytes,
                  (unsigned long long)L.total);
        return LDPCP_ERANGE;
    }
    char* b = (char*)buf;
    ldpcp_lane* lanes = (ldpcp_lane*)(b + L.lanes);
    memcpy(lanes, lv.data(), n * sizeof(ldpcp_lane));
    for (uint64_t t = 0; t < n; ++t) {
        if (lanes[t].error) {
            set_error("ldpcp_run_checkpoint: trajectory %llu stopped with error %d; its state is not resumable",
                      (unsigned long long)t, lanes[t].error);
            return LDPCP_ERANGE;
        }
        // the clock the state continues from (also when a run of fewer than FREQSAVE
        // steps integrated none)
        lanes[t].steps_done = run_steps_done(h);
    }
    ldpcp_checkpoint_header hd;
    memset(&hd, 0, sizeof(hd));
    hd.magic = LDPCP_CHECKPOINT_MAGIC;
    hd.header_bytes = sizeof(ldpcp_checkpoint_header);
    hd.lane_bytes = sizeof(ldpcp_lane);
    hd.n_traj = (uint32_t)n;
    hd.traj_offset = h->c.traj_offset;
    hd.model = h->params.model;
    hd.overdamped = h->params.overdamped;
    hd.barrier_window = h->c.wb;
    hd.n_fields = (int32_t)fields;
    hd.seed = h->params.seed;
    hd.steps_done = run_steps_done(h);
    hd.n_frames = run_frames_done(h);
    hd.save_freq = h->c.save_freq;
    hd.tick_capacity = h->c.tick_cap;
    hd.timestep = h->params.timestep;
    hd.bytes = L.total;
    hd.event_capacity = ecap;
    hd.n_lists = n_lists;
    bool complete = h->events_complete && n_lists;
    for (uint64_t t = 0; complete && t < n; ++t)
        for (uint32_t l = 0; l < n_lists; ++l) complete = complete && ldpcp_list_count(lanes + t, (int)l) <= ecap;
    hd.events_complete = complete ? 1u : 0u;
    memcpy(b, &hd, sizeof(hd));
    if (n_lists) {
        // the lists, packed: each trajectory's kept entries, list after list
        std::vector<double> all(n * n_lists * ecap);
        LDPCP_HIP_TRY(copy_on(h->stream, all.data(), h->P.events, all.size() * sizeof(double), hipMemcpyDeviceToHost));
        double* dst = (double*)(b + L.events);
        for (uint64_t t = 0; t < n; ++t)
            for (uint32_t l = 0; l < n_lists; ++l) {
                const uint64_t c = ldpcp_list_count(lanes + t, (int)l), keep = c < ecap ? c : ecap;
                memcpy(dst, &all[(t * n_lists + l) * ecap], keep * sizeof(double));
                dst += keep;
            }
    }
    LDPCP_HIP_TRY(copy_on(h->stream, b + L.wbase, h->P.wbase, n * sizeof(int32_t), hipMemcpyDeviceToHost));
    LDPCP_HIP_TRY(copy_on(h->stream, b + L.win, h->P.win, n * 2 * wb * sizeof(uint64_t), hipMemcpyDeviceToHost));
    // the frames recorded so far: the first n_frames rows of each field's plane
    for (uint64_t f = 0; f < fields; ++f)
        LDPCP_HIP_TRY(copy_on(h->stream, b + L.frames + f * hd.n_frames * n * sizeof(double),
                                h->P.frames + f * h->c.frames_cap * n, hd.n_frames * n * sizeof(double),
                                hipMemcpyDeviceToHost));
    if (h->c.tick_cap)
        LDPCP_HIP_TRY(copy_on(h->stream, b + L.ticks, h->P.ticks, n * h->c.tick_cap * sizeof(double), hipMemcpyDeviceToHost));
    return LDPCP_OK;
}

extern "C" int ldpcp_run_resume(ldpcp_run* h, const void* buf, uint64_t bytes)
{
    if (h && h->pending) {
        set_error("ldpcp_run_resume: the handle has a launched run; wait for it first");
        return LDPCP_EINVAL;
    }
    if (!h || !buf || bytes < sizeof(ldpcp_checkpoint_header)) {
        set_error("ldpcp_run_resume: null argument or truncated image");
        return LDPCP_EINVAL;
    }
    ldpcp_checkpoint_header hd;
    memcpy(&hd, buf, sizeof(hd));
    if (hd.magic != LDPCP_CHECKPOINT_MAGIC || hd.header_bytes != sizeof(ldpcp_checkpoint_header) ||
        hd.lane_bytes != sizeof(ldpcp_lane)) {
        set_error("ldpcp_run_resume: not a checkpoint image of this library version");
        return LDPCP_EINVAL;
    }
    const uint64_t n = h->c.n_traj, wb = (uint64_t)h->c.wb, fields = run_fields(h);
    struct Check {
        bool ok;
        const char* what;
    } checks[] = {
        {hd.n_traj == n, "trajectory count"},
        {hd.traj_offset == h->c.traj_offset, "trajectory offset"},
        {hd.model == h->params.model, "model"},
        {hd.overdamped == h->params.overdamped, "overdamped flag"},
        {hd.barrier_window == h->c.wb, "barrier window"},
        {(uint64_t)hd.n_fields == fields, "saved fields (WRITEX)"},
        {hd.seed == h->params.seed, "seed"},
        {hd.save_freq == h->c.save_freq, "save frequency"},
        {hd.timestep == h->params.timestep, "time step"},
        {hd.n_frames >= 1 && hd.n_frames <= hd.steps_done / hd.save_freq + 1, "frame count"},
    };
    for (const Check& k : checks)
        if (!k.ok) {
            set_error("ldpcp_run_resume: inconsistent %s", k.what);
            return LDPCP_EINVAL;
        }
    if (hd.n_lists && h->c.n_lists && hd.n_lists != (uint32_t)h->c.n_lists) {
        set_error("ldpcp_run_resume: inconsistent event lists (FPTCG sizes)");
        return LDPCP_EINVAL;
    }
    const ldpcp_lane* ilanes = (const ldpcp_lane*)((const char*)buf + sizeof(ldpcp_checkpoint_header));
    if (bytes < sizeof(ldpcp_checkpoint_header) + n * sizeof(ldpcp_lane)) {
        set_error("ldpcp_run_resume: truncated image");
        return LDPCP_EINVAL;
    }
    const CkptLayout L = ckpt_layout(n, wb, fields, hd.n_frames, hd.tick_capacity,
                                     packed_events(ilanes, n, hd.n_lists, hd.event_capacity));
    if (hd.bytes != L.total || bytes < L.total) {
        set_error("ldpcp_run_resume: image of %llu bytes, its header describes %llu", (unsigned long long)bytes,
                  (unsigned long long)L.total);
        return LDPCP_EINVAL;
    }
    if (h->c.replay) {
        const uint64_t end = hd.steps_done + h->c.nsteps_eff;
        if (h->c.n_gauss < h->c.init_gauss + end || h->c.n_flip < h->c.init_flip + (h->c.tau_zero ? 0 : end)) {
            set_error("ldpcp_run_resume: replay streams end before step %llu", (unsigned long long)end);
            return LDPCP_ERANGE;
        }
    }
    const char* b = (const char*)buf;
    const ldpcp_lane* lanes = (const ldpcp_lane*)(b + L.lanes);
    for (uint64_t t = 0; t < n; ++t)
        if (lanes[t].error || lanes[t].steps_done != hd.steps_done) {
            set_error("ldpcp_run_resume: trajectory %llu is not at step %llu", (unsigned long long)t,
                      (unsigned long long)hd.steps_done);
            return LDPCP_EINVAL;
        }
    // frame buffer for the prior frames and this run's
    const uint64_t cap = hd.n_frames + h->c.nsteps_eff / h->c.save_freq;
    void* fb = nullptr;
    LDPCP_HIP_TRY(hipMalloc(&fb, fields * cap * n * sizeof(double)));
    if (h->frames_block) (void)hipFree(h->frames_block);
    h->frames_block = fb;
    h->P.frames = (double*)fb;
    h->c.frames_cap = cap;
    h->n_frames = cap;
    LDPCP_HIP_TRY(hipMemsetAsync(fb, 0, fields * cap * n * sizeof(double), h->stream));
    LDPCP_HIP_TRY(hipStreamSynchronize(h->stream));
    for (uint64_t f = 0; f < fields; ++f)
        LDPCP_HIP_TRY(copy_on(h->stream, h->P.frames + f * cap * n, b + L.frames + f * hd.n_frames * n * sizeof(double),
                                hd.n_frames * n * sizeof(double), hipMemcpyHostToDevice));
    // per-trajectory state: the statistics record as is, the stepping state into its columns
    LDPCP_HIP_TRY(copy_on(h->stream, h->P.lanes, lanes, n * sizeof(ldpcp_lane), hipMemcpyHostToDevice));
    std::vector<double> col(n);
    struct Col {
        double* dst;
        size_t off;
    } cols[] = {{h->P.x, offsetof(ldpcp_lane, x)},           {h->P.v, offsetof(ldpcp_lane, v)},
                {h->P.F, offsetof(ldpcp_lane, f)},           {h->P.E, offsetof(ldpcp_lane, e)},
                {h->P.R, offsetof(ldpcp_lane, r_prev)},      {h->P.state, offsetof(ldpcp_lane, state)},
                {h->P.barrier, offsetof(ldpcp_lane, barrier)}, {h->P.well, offsetof(ldpcp_lane, well)},
                {h->P.periodic, offsetof(ldpcp_lane, periodic)}, {h->P.xp, offsetof(ldpcp_lane, xp)},
                {h->P.local, offsetof(ldpcp_lane, local_barrier)}, {h->P.bdefault, offsetof(ldpcp_lane, state0_instance)}};
    for (const Col& c : cols) {
        for (uint64_t t = 0; t < n; ++t) memcpy(&col[t], (const char*)(lanes + t) + c.off, sizeof(double));
        LDPCP_HIP_TRY(copy_on(h->stream, c.dst, col.data(), n * sizeof(double), hipMemcpyHostToDevice));
    }
    std::vector<uint64_t> cnt(n);
    for (int k = 0; k < 2; ++k) {
        for (uint64_t t = 0; t < n; ++t) cnt[t] = (uint64_t)lanes[t].state_count[k];
        LDPCP_HIP_TRY(copy_on(h->stream, k ? h->P.cnt1 : h->P.cnt0, cnt.data(), n * sizeof(uint64_t), hipMemcpyHostToDevice));
    }
    LDPCP_HIP_TRY(copy_on(h->stream, h->P.wbase, b + L.wbase, n * sizeof(int32_t), hipMemcpyHostToDevice));
    LDPCP_HIP_TRY(copy_on(h->stream, h->P.win, b + L.win, n * 2 * wb * sizeof(uint64_t), hipMemcpyHostToDevice));
    if (h->c.tick_cap) {
        // barrier ticks beyond the smaller of the two capacities stay counted but unlisted
        std::vector<double> tk(n * h->c.tick_cap, 0.0);
        const uint64_t keep = hd.tick_capacity < h->c.tick_cap ? hd.tick_capacity : h->c.tick_cap;
        for (uint64_t t = 0; t < n; ++t)
            memcpy(&tk[t * h->c.tick_cap], b + L.ticks + t * hd.tick_capacity * sizeof(double), keep * sizeof(double));
        LDPCP_HIP_TRY(copy_on(h->stream, h->P.ticks, tk.data(), tk.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    // event lists: the image's kept entries, as many as this handle keeps
    const uint64_t ecap = h->c.ev_cap;
    const uint32_t n_lists = (uint32_t)h->c.n_lists;
    bool complete = n_lists && hd.events_complete && hd.n_lists == n_lists;
    if (n_lists) {
        std::vector<double> all(n * n_lists * ecap, 0.0);
        const double* src = (const double*)(b + L.events);
        for (uint64_t t = 0; t < n; ++t)
            for (uint32_t l = 0; l < hd.n_lists; ++l) {
                const uint64_t c = ldpcp_list_count(lanes + t, (int)l);
                const uint64_t have = c < hd.event_capacity ? c : hd.event_capacity;
                const uint64_t keep = have < ecap ? have : ecap;
                complete = complete && c <= ecap;
                memcpy(&all[(t * n_lists + l) * ecap], src, keep * sizeof(double));
                src += have;
            }
        LDPCP_HIP_TRY(copy_on(h->stream, h->P.events, all.data(), all.size() * sizeof(double), hipMemcpyHostToDevice));
    }
    h->events_complete = complete;
    h->step_base = hd.steps_done;
    h->frames_prior = hd.n_frames;
    h->phase = 1;
    h->integrated = false;
    return LDPCP_OK;
}

extern "C" int ldpcp_run_barrier_table(ldpcp_run* h, int32_t array, int32_t lo, int32_t hi, int32_t* states,
                                       uint64_t* steps)
{
    if (h && h->pending) {  // results of a launched run: wait for it first
        const int rc = ldpcp_run_wait(h);
        if (rc) return rc;
    }
    if (!h || !states || !steps || hi < lo || array < 0 || array > 1 || (uint64_t)hi - lo + 1 > 65535) {
        set_error("ldpcp_run_barrier_table: bad handle, array, range or output");
        return LDPCP_EINVAL;
    }
    const uint64_t n = h->c.n_traj, ids = (uint64_t)(hi - lo) + 1;
    void* block = nullptr;
    LDPCP_HIP_TRY(hipMalloc(&block, n * ids * (sizeof(int32_t) + sizeof(uint64_t))));
    uint64_t* d_steps = (uint64_t*)block;
    int32_t* d_states = (int32_t*)(d_steps + n * ids);
    const dim3 grid((unsigned)((n + 255) / 256), (unsigned)ids);
    hipLaunchKernelGGL(run_table_kernel, grid, dim3(256), 0, h->stream, h->c, h->P, (int)array, (int)lo, (int)ids,
                       d_states, d_steps);
    hipError_t e = hipGetLastError();
    if (e == hipSuccess) e = hipStreamSynchronize(h->stream);
    if (e == hipSuccess) e = copy_on(h->stream, states, d_states, n * ids * sizeof(int32_t), hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = copy_on(h->stream, steps, d_steps, n * ids * sizeof(uint64_t), hipMemcpyDeviceToHost);
    (void)hipFree(block);
    if (e != hipSuccess) {
        set_error("ldpcp_run_barrier_table: %s", hipGetErrorString(e));
        return LDPCP_EDEVICE;
    }
    return LDPCP_OK;
}

extern "C" int ldpcp_run_destroy(ldpcp_run* h)
{
    if (!h) return LDPCP_OK;
    if (h->pending) (void)hipStreamSynchronize(h->stream);  // no buffer is freed under a running kernel
    if (h->c.replay) (void)hipFree((void*)h->P.g);
    if (h->frames_block) (void)hipFree(h->frames_block);
    if (h->noise) (void)hipFree(h->noise);
    if (h->P.events) (void)hipFree(h->P.events);
    (void)hipFree(h->state_block);
    (void)hipStreamDestroy(h->stream);
    delete h;
    return LDPCP_OK;
}
