/*
 * cpu_simd.c -- the CPU baseline: the reference's SIMD step loop for the
 * single-barrier absorbing first passage, W trajectories per vector.
 *
 * TEST AND BENCH INFRASTRUCTURE ONLY.  bench.py's cpu_baseline leg times it
 * and tests/ check it against the scalar oracle; the product never loads it.
 *
 * What it restates: one MPI rank of the reference integrates W lanes in
 * lock-step with AVX-512 (W = 8) or AVX2 (W = 4) intrinsics
 * (src/simd.h:102-182), one step being
 *   overdamped  X = fmadd(F, dt/(m gamma), X) + R_n+1 + R_n
 *               (langevinDynamicsSingleBarrier.h:2860-2886),
 *   underdamped B-A-O-A-B (Single.h:1680-1706),
 * then UpdatePeriodicIndices (Single.h:1861-1884), the per-step telegraph flip
 * over the whole vector and the per-lane hand-over loop on entering another
 * periodic cell (UpdateBarrierStateEta, Single.h:1998-2153),
 * UpdateBarrierStateEtaParamter (Single.h:2154-2165) and EnergyAndForce
 * (Single.h:1292-1324).  The statistic is the first passage the GPU sweep
 * measures: the first barrier-top crossing (the reference's first FPT1
 * sample, Single.h:2394-2411), with the transition path starting at the last
 * well-centre crossing before it.  An absorbed or capped lane takes the next
 * trajectory at the next step pair (the reference's lanes run a fixed step
 * count instead; refilling keeps every lane busy on this workload).
 *
 * Noise: Philox4x32-10 keyed by the seed with counter {step pair, trajectory,
 * role}, Box-Muller built from IEEE operations only -- the oracle's lane-mode
 * definition -- generated for all W lanes at once every second step.  The
 * reference draws MKL MT19937/ICDF banks instead (RANDOMNUMBER.cpp:157-246);
 * MKL is not available on the GPU host.  The arithmetic follows the oracle's
 * order operation by operation (build with -ffp-contract=off), so every
 * trajectory's absorption step equals the oracle's (tests/test_cpu_simd.py).
 */
#include <immintrin.h>
#include <math.h>
#include <pthread.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "ldpcp_oracle.h"

#if defined(__AVX512F__) && defined(__AVX512DQ__)
#define W 8
typedef __m512d vd;
typedef __m512i vi;
typedef __mmask8 vm;
#define V_set1(a) _mm512_set1_pd(a)
#define V_add(a, b) _mm512_add_pd(a, b)
#define V_sub(a, b) _mm512_sub_pd(a, b)
#define V_mul(a, b) _mm512_mul_pd(a, b)
#define V_div(a, b) _mm512_div_pd(a, b)
#define V_fma(a, b, c) _mm512_fmadd_pd(a, b, c)
#define V_sqrt(a) _mm512_sqrt_pd(a)
#define V_rint(a) _mm512_roundscale_pd(a, _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC)
#define V_floor(a) _mm512_roundscale_pd(a, _MM_FROUND_TO_NEG_INF | _MM_FROUND_NO_EXC)
#define V_lt(a, b) _mm512_cmp_pd_mask(a, b, _CMP_LT_OQ)
#define V_gt(a, b) _mm512_cmp_pd_mask(a, b, _CMP_GT_OQ)
#define V_ge(a, b) _mm512_cmp_pd_mask(a, b, _CMP_GE_OQ)
#define V_eq(a, b) _mm512_cmp_pd_mask(a, b, _CMP_EQ_OQ)
#define V_ne(a, b) _mm512_cmp_pd_mask(a, b, _CMP_NEQ_UQ)
#define M_and(a, b) ((vm)((a) & (b)))
#define M_or(a, b) ((vm)((a) | (b)))
#define M_bits(m) ((unsigned)(m))
/* lanes of m take a, the others b */
#define V_sel(m, a, b) _mm512_mask_blend_pd(m, b, a)
#define V_neg(a) _mm512_castsi512_pd(_mm512_xor_si512(_mm512_castpd_si512(a), _mm512_set1_epi64(INT64_MIN)))
#define I_set1(a) _mm512_set1_epi64((int64_t)(a))
#define I_add(a, b) _mm512_add_epi64(a, b)
#define I_xor(a, b) _mm512_xor_si512(a, b)
#define I_and(a, b) _mm512_and_si512(a, b)
#define I_or(a, b) _mm512_or_si512(a, b)
#define I_srl(a, n) _mm512_srli_epi64(a, n)
#define I_mul32(a, b) _mm512_mul_epu32(a, b)
#define I_as_d(a) _mm512_castsi512_pd(a)
#define D_as_i(a) _mm512_castpd_si512(a)
#define V_store(p, a) _mm512_store_pd(p, a)
#define V_load(p) _mm512_load_pd(p)
#define I_store(p, a) _mm512_store_si512((void*)(p), a)
#define I_load(p) _mm512_load_si512((const void*)(p))
#else
#define W 4
typedef __m256d vd;
typedef __m256i vi;
typedef __m256d vm;
#define V_set1(a) _mm256_set1_pd(a)
#define V_add(a, b) _mm256_add_pd(a, b)
#define V_sub(a, b) _mm256_sub_pd(a, b)
#define V_mul(a, b) _mm256_mul_pd(a, b)
#define V_div(a, b) _mm256_div_pd(a, b)
#define V_fma(a, b, c) _mm256_fmadd_pd(a, b, c)
#define V_sqrt(a) _mm256_sqrt_pd(a)
#define V_rint(a) _mm256_round_pd(a, _MM_FROUND_TO_NEAREST_INT | _MM_FROUND_NO_EXC)
#define V_floor(a) _mm256_round_pd(a, _MM_FROUND_TO_NEG_INF | _MM_FROUND_NO_EXC)
#define V_lt(a, b) _mm256_cmp_pd(a, b, _CMP_LT_OQ)
#define V_gt(a, b) _mm256_cmp_pd(a, b, _CMP_GT_OQ)
#define V_ge(a, b) _mm256_cmp_pd(a, b, _CMP_GE_OQ)
#define V_eq(a, b) _mm256_cmp_pd(a, b, _CMP_EQ_OQ)
#define V_ne(a, b) _mm256_cmp_pd(a, b, _CMP_NEQ_UQ)
#define M_and(a, b) _mm256_and_pd(a, b)
#define M_or(a, b) _mm256_or_pd(a, b)
#define M_bits(m) ((unsigned)_mm256_movemask_pd(m))
#define V_sel(m, a, b) _mm256_blendv_pd(b, a, m)
#define V_neg(a) _mm256_xor_pd(a, _mm256_set1_pd(-0.0))
#define I_set1(a) _mm256_set1_epi64x((int64_t)(a))
#define I_add(a, b) _mm256_add_epi64(a, b)
#define I_xor(a, b) _mm256_xor_si256(a, b)
#define I_and(a, b) _mm256_and_si256(a, b)
#define I_or(a, b) _mm256_or_si256(a, b)
#define I_srl(a, n) _mm256_srli_epi64(a, n)
#define I_mul32(a, b) _mm256_mul_epu32(a, b)
#define I_as_d(a) _mm256_castsi256_pd(a)
#define D_as_i(a) _mm256_castpd_si256(a)
#define V_store(p, a) _mm256_store_pd(p, a)
#define V_load(p) _mm256_load_pd(p)
#define I_store(p, a) _mm256_store_si256((__m256i*)(p), a)
#define I_load(p) _mm256_load_si256((const __m256i*)(p))
#endif

#define ALIGN __attribute__((aligned(64)))

int ldo_simd_width(void) { return W; }

/* ------------------------------------------------------------------------- */
/* Constants (IntegrateInitializeVariables, Single.h:1520-1618; overdamped    */
/* :2812-2824): the oracle's formulas for the single barrier.                 */
/* ------------------------------------------------------------------------- */
typedef struct sconst {
    double dt, minv, halfdt, expmgdt, var, dtgm, L, Linv, w, c_lap, p_flip, p_asym[2], p0, p_create, K, vsigma;
    double H[2], S[2];
    int tau_zero, equal, overdamped;
} sconst;

static void make_sconst(const ldo_params* p, sconst* c)
{
    memset(c, 0, sizeof(*c));
    c->dt = p->timestep;
    c->minv = 1.0 / p->mass;
    c->halfdt = p->timestep * 0.5;
    c->expmgdt = exp(-p->gamma * p->timestep);
    c->var = sqrt(p->kb * p->temperature / p->mass * (1.0 - exp(-2.0 * p->gamma * p->timestep)));
    c->L = p->periodic_length;
    c->Linv = 1.0 / c->L;
    c->K = p->eq_constant;
    c->tau_zero = (p->tau == 0.0);
    c->equal = p->equal_probability;
    c->overdamped = p->overdamped;
    if (p->equal_probability) {
        c->c_lap = -2.0 * p->timestep / p->tau;
        c->p_flip = c->tau_zero ? 0.0 : 0.5 * (1.0 - exp(-2.0 * p->timestep / p->tau));
        c->p0 = 0.5;
    } else {
        c->c_lap = -(p->eq_constant + 1.0) * p->timestep / p->tau;
        if (!c->tau_zero) {
            c->p_asym[0] = p->eq_constant / (1.0 + p->eq_constant) * (1.0 - exp(c->c_lap));
            c->p_asym[1] = 1.0 / (1.0 + p->eq_constant) * (1.0 - exp(c->c_lap));
        }
        c->p0 = 1.0 / (1.0 + p->eq_constant);
    }
    c->p_create = 1.0 / (1.0 + p->eq_constant);
    c->w = p->half_width[0];
    for (int s = 0; s < 2; ++s) {
        c->H[s] = p->height[0][s];
        c->S[s] = p->height[0][s] / p->half_width[0];
    }
    if (p->overdamped) {
        double mgi = 1.0 / (p->mass * p->gamma);
        c->dtgm = p->timestep * mgi;
        c->var = sqrt(p->kb * p->temperature * mgi * p->timestep * 0.5);
    }
    c->vsigma = sqrt(p->kb * p->temperature / p->mass);
}

/* ------------------------------------------------------------------------- */
/* Noise: Philox4x32-10 (Salmon et al., SC'11) for W lanes, each 32-bit word   */
/* in the low half of a 64-bit element, so that one unsigned 32x32->64         */
/* multiply serves all lanes.                                                  */
/* ------------------------------------------------------------------------- */
static void philox_scalar(const uint32_t ctr[4], uint32_t k0, uint32_t k1, uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    for (int r = 0; r < 10; ++r) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c0, p1 = (uint64_t)0xCD9E8D57u * c2;
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0, n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        c0 = n0;
        c1 = (uint32_t)p1;
        c2 = n2;
        c3 = (uint32_t)p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

static inline void philox_vec(vi c0, vi c1, vi c2, vi c3, uint32_t k0, uint32_t k1, vi out[4])
{
    const vi m0 = I_set1(0xD2511F53u), m1 = I_set1(0xCD9E8D57u), lo = I_set1(0xFFFFFFFFu);
    for (int r = 0; r < 10; ++r) {
        const vi p0 = I_mul32(c0, m0), p1 = I_mul32(c2, m1);  /* reads the low 32 bits only */
        const vi n0 = I_xor(I_xor(I_srl(p1, 32), c1), I_set1(k0));
        const vi n2 = I_xor(I_xor(I_srl(p0, 32), c3), I_set1(k1));
        c0 = n0;
        c1 = p1;  /* upper halves are never read again unmasked */
        c2 = n2;
        c3 = p0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = I_and(c0, lo);
    out[1] = I_and(c1, lo);
    out[2] = I_and(c2, lo);
    out[3] = I_and(c3, lo);
}

/* exact conversion of a 32-bit word (low half of an element) to double */
static inline vd u32_to_d(vi w)
{
    const vi magic = I_set1(0x4330000000000000ull); /* 2^52 */
    return V_sub(I_as_d(I_or(w, magic)), V_set1(4503599627370496.0));
}

/* (w + 1/2) / 2^32 */
static inline vd unit_vec(vi w) { return V_mul(V_add(u32_to_d(w), V_set1(0.5)), V_set1(2.3283064365386963e-10)); }

static inline double unit_scalar(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }

/* natural logarithm of a positive normal x, the operation sequence of
 * ldpcp_log (csrc/ldpcp_gauss.h): 2 atanh series on the mantissa */
static inline vd log_vec(vd x)
{
    const vi b = D_as_i(x);
    const vd e = u32_to_d(I_and(I_srl(b, 52), I_set1(0x7ff)));
    vd dk = V_sub(e, V_set1(1023.0));
    vd m = I_as_d(I_or(I_and(b, I_set1(0x000fffffffffffffull)), I_set1(0x3ff0000000000000ull)));
    const vm big = V_ge(m, V_set1(1.4142135623730951));
    m = V_sel(big, V_mul(m, V_set1(0.5)), m);
    dk = V_sel(big, V_add(dk, V_set1(1.0)), dk);
    const vd f = V_div(V_sub(m, V_set1(1.0)), V_add(m, V_set1(1.0)));
    const vd z = V_mul(f, f);
    vd p = V_set1(0.04);
    p = V_fma(p, z, V_set1(0.043478260869565216));
    p = V_fma(p, z, V_set1(0.047619047619047616));
    p = V_fma(p, z, V_set1(0.05263157894736842));
    p = V_fma(p, z, V_set1(0.058823529411764705));
    p = V_fma(p, z, V_set1(0.06666666666666667));
    p = V_fma(p, z, V_set1(0.07692307692307693));
    p = V_fma(p, z, V_set1(0.09090909090909091));
    p = V_fma(p, z, V_set1(0.1111111111111111));
    p = V_fma(p, z, V_set1(0.14285714285714285));
    p = V_fma(p, z, V_set1(0.2));
    p = V_fma(p, z, V_set1(0.3333333333333333));
    const vd f2 = V_add(f, f);
    const vd lnm = V_fma(V_mul(f2, z), p, f2);
    return V_fma(dk, V_set1(0.6931471803691238), V_fma(dk, V_set1(1.9082149292705877e-10), lnm));
}

/* sin and cos of 2 pi u, the operation sequence of ldpcp_sincos_turn */
static inline void sincos_vec(vd u, vd* s_out, vd* c_out)
{
    const vd n = V_rint(V_mul(V_set1(4.0), u));
    const vd r = V_sub(u, V_mul(V_set1(0.25), n));
    const vd t = V_mul(V_set1(6.283185307179586), r);
    const vd z = V_mul(t, t);
    vd ps = V_set1(-8.22063524662433e-18);
    ps = V_fma(ps, z, V_set1(2.8114572543455206e-15));
    ps = V_fma(ps, z, V_set1(-7.647163731819816e-13));
    ps = V_fma(ps, z, V_set1(1.6059043836821613e-10));
    ps = V_fma(ps, z, V_set1(-2.505210838544172e-08));
    ps = V_fma(ps, z, V_set1(2.7557319223985893e-06));
    ps = V_fma(ps, z, V_set1(-0.0001984126984126984));
    ps = V_fma(ps, z, V_set1(0.008333333333333333));
    ps = V_fma(ps, z, V_set1(-0.16666666666666666));
    const vd sn = V_fma(V_mul(t, z), ps, t);
    vd pc = V_set1(1.5619206968586225e-16);
    pc = V_fma(pc, z, V_set1(-4.779477332387385e-14));
    pc = V_fma(pc, z, V_set1(1.1470745597729725e-11));
    pc = V_fma(pc, z, V_set1(-2.08767569878681e-09));
    pc = V_fma(pc, z, V_set1(2.755731922398589e-07));
    pc = V_fma(pc, z, V_set1(-2.48015873015873e-05));
    pc = V_fma(pc, z, V_set1(0.001388888888888889));
    pc = V_fma(pc, z, V_set1(-0.041666666666666664));
    pc = V_fma(pc, z, V_set1(0.5));
    const vd cs = V_fma(V_neg(z), pc, V_set1(1.0));
    /* quadrant n mod 4, n in {0, ..., 4} */
    const vm q1 = V_eq(n, V_set1(1.0)), q2 = V_eq(n, V_set1(2.0)), q3 = V_eq(n, V_set1(3.0));
    vd s = sn, c = cs;
    s = V_sel(q1, cs, s);
    c = V_sel(q1, V_neg(sn), c);
    s = V_sel(q2, V_neg(sn), s);
    c = V_sel(q2, V_neg(cs), c);
    s = V_sel(q3, V_neg(cs), s);
    c = V_sel(q3, sn, c);
    *s_out = s;
    *c_out = c;
}

/* ------------------------------------------------------------------------- */
/* Per-lane barrier table (the reference's lazy BarrierStatus store,           */
/* DataLocator.h:100-188): before its first barrier-top crossing a trajectory  */
/* touches the barriers of the two cells its well spans, so four slots hold     */
/* every entry it can need.                                                    */
/* ------------------------------------------------------------------------- */
#define SLOTS 4
typedef struct lane_tab {
    int n;
    int id[SLOTS];
    int state[SLOTS];
    uint64_t step[SLOTS];
} lane_tab;

typedef struct lane_ctx {
    uint32_t traj;      /* trajectory index (Philox counter word 2) */
    uint64_t start;     /* pair counter at the lane's step 1 */
    uint64_t last_period_step;
    uint64_t n_catch;
    int live;
    lane_tab tab;
} lane_ctx;

/* a barrier's entry, created on first touch with its role-3 initial state */
static int tab_slot(lane_tab* t, const sconst* c, int same_initial, uint32_t key0, uint32_t key1, uint32_t traj,
                    int id, int* err)
{
    for (int k = 0; k < t->n; ++k)
        if (t->id[k] == id) return k;
    if (t->n == SLOTS) {
        *err = 1;
        return 0;
    }
    const int k = t->n++;
    t->id[k] = id;
    t->step[k] = 0;
    if (same_initial) {
        t->state[k] = 0;
    } else {
        const uint32_t ctr[4] = {(uint32_t)id, 0u, traj, 3u};
        uint32_t w[4];
        philox_scalar(ctr, key0, key1, w);
        t->state[k] = unit_scalar(w[0]) < c->p_create ? 0 : 1;
    }
    return k;
}

/* ------------------------------------------------------------------------- */
/* The group: W lanes of one parameter set stepping together.                 */
/* ------------------------------------------------------------------------- */
typedef struct group {
    vd x, v, F, E, R, state, height, slope_pos, slope_neg, well, periodic, cnt0, cnt1;
    vi q, traj;
} group;

typedef struct set_work {
    ldo_params p;
    sconst c;
    uint32_t key0, key1;
    uint32_t n_traj;             /* trajectories to hand out (0 = unbounded) */
    uint32_t next;               /* next trajectory index (atomic) */
    uint64_t* absorbed;          /* per trajectory, or NULL */
    uint64_t* tpt0;
    uint64_t nsteps;
} set_work;

typedef struct shared {
    set_work* sets;
    int n_sets;
    double seconds;              /* > 0: run until this much time has passed */
    struct timespec t0;
    uint64_t traj_steps, traj_done;  /* atomics */
    int error;
} shared;

static double elapsed(const struct timespec* t0)
{
    struct timespec t;
    clock_gettime(CLOCK_MONOTONIC, &t);
    return (double)(t.tv_sec - t0->tv_sec) + 1e-9 * (double)(t.tv_nsec - t0->tv_nsec);
}

/* step 0 of trajectory `traj` in lane i (the oracle's lane_run set-up: initial
 * indices, state draw, barrier entry, parameters, V0 and R0 draws, force) */
static void lane_start(group* g, lane_ctx* L, int i, set_work* sw, uint32_t traj, uint64_t pair)
{
    const sconst* c = &sw->c;
    const ldo_params* p = &sw->p;
    ALIGN double a[16][W];
    V_store(a[0], g->x); V_store(a[1], g->v); V_store(a[2], g->F); V_store(a[3], g->E);
    V_store(a[4], g->R); V_store(a[5], g->state); V_store(a[6], g->height); V_store(a[7], g->slope_pos);
    V_store(a[8], g->slope_neg); V_store(a[9], g->well); V_store(a[10], g->periodic);
    V_store(a[11], g->cnt0); V_store(a[12], g->cnt1);
    ALIGN uint64_t q[W], tr[W];
    I_store(q, g->q);
    I_store(tr, g->traj);

    const uint32_t ctr[4] = {0u, 0u, traj, 1u};
    uint32_t w[4];
    philox_scalar(ctr, sw->key0, sw->key1, w);
    /* Box-Muller of role 1: cos -> V0, sin -> R0 (scalar ldpcp_box_muller order) */
    ALIGN double u1[W], u2[W], gs[W], gc[W];
    for (int k = 0; k < W; ++k) {
        u1[k] = unit_scalar(w[0]);
        u2[k] = unit_scalar(w[1]);
    }
    {
        const vd r = V_sqrt(V_mul(V_set1(-2.0), log_vec(V_load(u1))));
        vd s, cc;
        sincos_vec(V_load(u2), &s, &cc);
        V_store(gc, V_mul(r, cc));
        V_store(gs, V_mul(r, s));
    }
    const double x = p->x0;
    const double r = x * c->Linv;
    const double well = floor(r), periodic = rint(r);
    double state = p->randomize_state ? (unit_scalar(w[2]) <= c->p0 ? 0.0 : 1.0) : p->initial_state;
    memset(&L->tab, 0, sizeof(L->tab));
    int err = 0;
    const int k = tab_slot(&L->tab, c, p->same_initial_barrier, sw->key0, sw->key1, traj, (int)periodic, &err);
    L->tab.state[k] = (int)state;
    const int one = (state == 1.0);
    const double height = one ? c->H[1] : c->H[0];
    const double sp = one ? c->S[1] : c->S[0];
    const double sn = 0.0 - sp;
    const double v = p->randomize_v ? c->vsigma * gc[0] : p->v0;
    /* EnergyAndForce */
    const double pidx = rint(x * c->Linv);
    const double xp = x - pidx * c->L;
    const int neg = (xp < 0.0) && (xp > -c->w), pos = (xp < c->w) && (xp > 0.0);
    double E = neg ? height - xp * sn : 0.0;
    E = pos ? height - xp * sp : E;
    double F = neg ? sn : 0.0;
    F = pos ? sp : F;
    a[0][i] = x;
    a[1][i] = v;
    a[2][i] = F;
    a[3][i] = E;
    a[4][i] = p->overdamped ? c->var * gs[0] : 0.0;
    a[5][i] = state;
    a[6][i] = height;
    a[7][i] = sp;
    a[8][i] = sn;
    a[9][i] = well;
    a[10][i] = periodic;
    a[11][i] = one ? 0.0 : 1.0;
    a[12][i] = one ? 1.0 : 0.0;
    q[i] = 0;
    tr[i] = traj;
    g->x = V_load(a[0]); g->v = V_load(a[1]); g->F = V_load(a[2]); g->E = V_load(a[3]);
    g->R = V_load(a[4]); g->state = V_load(a[5]); g->height = V_load(a[6]); g->slope_pos = V_load(a[7]);
    g->slope_neg = V_load(a[8]); g->well = V_load(a[9]); g->periodic = V_load(a[10]);
    g->cnt0 = V_load(a[11]); g->cnt1 = V_load(a[12]);
    g->q = I_load(q);
    g->traj = I_load(tr);
    L->traj = traj;
    L->start = pair;
    L->last_period_step = 0;
    L->n_catch = 0;
    L->live = 1;
}

/* the hand-over loop of UpdateBarrierStateEta for the lanes in `bits`:
 * store the state of the barrier left, take the entered barrier's state aged
 * by the steps since it was left (catch-up flip, role 2) */
static void hand_over(group* g, lane_ctx* lanes, unsigned bits, vd periodic_prev, set_work* sw, uint64_t pair,
                      int odd, int* err)
{
    const sconst* c = &sw->c;
    ALIGN double st[W], pp[W], pc_[W];
    V_store(st, g->state);
    V_store(pp, periodic_prev);
    V_store(pc_, g->periodic);
    for (int i = 0; i < W; ++i) {
        if (!(bits >> i & 1u)) continue;
        lane_ctx* L = &lanes[i];
        const uint64_t iStep = 2 * (pair - L->start) + (odd ? 1 : 2);
        const int kp = tab_slot(&L->tab, c, sw->p.same_initial_barrier, sw->key0, sw->key1, L->traj, (int)pp[i], err);
        L->tab.state[kp] = (int)st[i];
        L->tab.step[kp] = iStep;
        const int kc = tab_slot(&L->tab, c, sw->p.same_initial_barrier, sw->key0, sw->key1, L->traj, (int)pc_[i], err);
        const double prev = (double)L->tab.state[kc];
        if (c->tau_zero) {
            st[i] = prev;
        } else {
            const double lap = (double)(iStep - L->tab.step[kc]);
            double pc;
            if (c->equal) pc = 0.5 * (1.0 - exp(c->c_lap * lap));
            else if (prev == 0.0) pc = c->K / (1.0 + c->K) * (1.0 - exp(c->c_lap * lap));
            else pc = 1.0 / (1.0 + c->K) * (1.0 - exp(c->c_lap * lap));
            /* four catch-up uniforms per call: word n mod 4 of index n / 4 */
            const uint64_t q = L->n_catch >> 2;
            const uint32_t ctr[4] = {(uint32_t)q, (uint32_t)(q >> 32), L->traj, 2u};
            uint32_t w[4];
            philox_scalar(ctr, sw->key0, sw->key1, w);
            const double u = unit_scalar(w[L->n_catch & 3]);
            L->n_catch++;
            st[i] = (u < pc) ? ((prev == 1.0) ? 0.0 : 1.0) : prev;
        }
    }
    g->state = V_load(st);
}

/* one integration step of all lanes with Gaussian gn and flip uniform un;
 * returns the lanes entering another well (barrier-top crossing) and, in
 * *period, those entering another periodic cell */
static inline unsigned step(group* g, lane_ctx* lanes, set_work* sw, vd gn, vd un, uint64_t pair, int odd,
                            unsigned live, unsigned* period, int* err)
{
    const sconst* c = &sw->c;
    if (c->overdamped) {
        const vd rf = V_mul(V_set1(c->var), gn);
        g->x = V_fma(g->F, V_set1(c->dtgm), g->x);
        g->x = V_add(g->x, rf);
        g->x = V_add(g->x, g->R);
        g->R = rf;
    } else {
        const vd hdt = V_set1(c->halfdt);
        g->v = V_add(g->v, V_mul(V_mul(V_set1(c->minv), g->F), hdt));
        g->x = V_add(g->x, V_mul(g->v, hdt));
        g->v = V_add(V_mul(V_set1(c->expmgdt), g->v), V_mul(V_set1(c->var), gn));
        g->x = V_add(g->x, V_mul(g->v, hdt));
    }
    /* UpdatePeriodicIndices */
    const vd r = V_mul(g->x, V_set1(c->Linv));
    const vd well = V_floor(r), periodic = V_rint(r);
    const vm ent_well = V_ne(g->well, well), ent_period = V_ne(g->periodic, periodic);
    const vd periodic_prev = g->periodic;
    g->well = well;
    g->periodic = periodic;
    /* per-step telegraph flip of the whole vector */
    if (!c->tau_zero) {
        const vm one = V_eq(g->state, V_set1(1.0));
        vm flip;
        if (c->equal) flip = V_lt(un, V_set1(c->p_flip));
        else flip = V_lt(un, V_sel(one, V_set1(c->p_asym[1]), V_set1(c->p_asym[0])));
        g->state = V_sel(flip, V_sel(one, V_set1(0.0), V_set1(1.0)), g->state);
    }
    const unsigned per = M_bits(ent_period) & live;
    if (per) hand_over(g, lanes, per, periodic_prev, sw, pair, odd, err);
    /* UpdateBarrierStateEtaParamter */
    const vm one = V_eq(g->state, V_set1(1.0));
    g->height = V_sel(one, V_set1(c->H[1]), V_set1(c->H[0]));
    g->slope_pos = V_sel(one, V_set1(c->S[1]), V_set1(c->S[0]));
    g->slope_neg = V_sub(V_set1(0.0), g->slope_pos);
    g->cnt1 = V_sel(one, V_add(g->cnt1, V_set1(1.0)), g->cnt1);
    g->cnt0 = V_sel(one, g->cnt0, V_add(g->cnt0, V_set1(1.0)));
    /* EnergyAndForce */
    const vd xp = V_sub(g->x, V_mul(V_rint(V_mul(g->x, V_set1(c->Linv))), V_set1(c->L)));
    const vd zero = V_set1(0.0);
    const vm neg = M_and(V_lt(xp, zero), V_gt(xp, V_set1(-c->w)));
    const vm pos = M_and(V_lt(xp, V_set1(c->w)), V_gt(xp, zero));
    vd E = V_sel(neg, V_sub(g->height, V_mul(xp, g->slope_neg)), zero);
    g->E = V_sel(pos, V_sub(g->height, V_mul(xp, g->slope_pos)), E);
    const vd F = V_sel(neg, g->slope_neg, zero);
    g->F = V_sel(pos, g->slope_pos, F);
    if (!c->overdamped) g->v = V_add(g->v, V_mul(V_mul(V_set1(c->minv), g->F), V_set1(c->halfdt)));
    *period = per;
    return M_bits(ent_well) & live;
}

/* lanes absorbed at this step: record, retire */
static void retire(lane_ctx* lanes, unsigned bits, set_work* sw, uint64_t pair, int odd, shared* sh, unsigned* live)
{
    for (int i = 0; i < W; ++i) {
        if (!(bits >> i & 1u)) continue;
        lane_ctx* L = &lanes[i];
        const uint64_t iStep = 2 * (pair - L->start) + (odd ? 1 : 2);
        if (sw->absorbed && L->traj < sw->n_traj) {
            sw->absorbed[L->traj] = iStep;
            sw->tpt0[L->traj] = L->last_period_step;
        }
        __atomic_fetch_add(&sh->traj_steps, iStep, __ATOMIC_RELAXED);
        __atomic_fetch_add(&sh->traj_done, 1, __ATOMIC_RELAXED);
        L->live = 0;
        *live &= ~(1u << i);
    }
}

static void note_period(lane_ctx* lanes, unsigned bits, uint64_t pair, int odd)
{
    for (int i = 0; i < W; ++i)
        if (bits >> i & 1u) lanes[i].last_period_step = 2 * (pair - lanes[i].start) + (odd ? 1 : 2);
}

/* live lanes that reached the step cap in this pair: record as not absorbed */
static void cap_lanes(lane_ctx* lanes, set_work* sw, shared* sh, uint64_t pair, uint64_t pairs_max, uint64_t nsteps,
                      unsigned* live)
{
    for (int i = 0; i < W; ++i) {
        if (!(*live >> i & 1u) || pair - lanes[i].start + 1 != pairs_max) continue;
        if (sw->absorbed && lanes[i].traj < sw->n_traj) {
            sw->absorbed[lanes[i].traj] = 0;
            sw->tpt0[lanes[i].traj] = 0;
        }
        __atomic_fetch_add(&sh->traj_steps, nsteps, __ATOMIC_RELAXED);
        __atomic_fetch_add(&sh->traj_done, 1, __ATOMIC_RELAXED);
        lanes[i].live = 0;
        *live &= ~(1u << i);
    }
}

/* run one set's trajectories in this thread's group until the set has no
 * more to hand out (or the time is up) */
static void run_set(set_work* sw, shared* sh, int* timed_out)
{
    const sconst* c = &sw->c;
    group g;
    memset(&g, 0, sizeof(g));
    g.x = g.v = g.F = g.E = g.R = g.state = g.height = g.slope_pos = g.slope_neg = V_set1(0.0);
    g.well = g.periodic = g.cnt0 = g.cnt1 = V_set1(0.0);
    g.q = g.traj = I_set1(0);
    lane_ctx lanes[W];
    memset(lanes, 0, sizeof(lanes));
    unsigned live = 0;
    uint64_t pair = 0;
    const uint64_t nsteps = sw->nsteps;
    const uint64_t pairs_max = (nsteps + 1) / 2;  /* pair index (from the lane's start) holding step nsteps */
    int more = 1;
    uint64_t since_check = 0;
    for (;;) {
        /* refill idle lanes at a pair boundary, so every live lane's next step is odd */
        if (more && live != (1u << W) - 1u) {
            for (int i = 0; i < W && more; ++i) {
                if (live >> i & 1u) continue;
                const uint32_t t = __atomic_fetch_add(&sw->next, 1u, __ATOMIC_RELAXED);
                if (sw->n_traj && t >= sw->n_traj) {
                    more = 0;
                    break;
                }
                lane_start(&g, &lanes[i], i, sw, t, pair);
                live |= 1u << i;
            }
        }
        if (!live) break;
        /* noise for this pair: Philox role 0 at counter {q, traj, 0} */
        vi wv[4];
        philox_vec(I_and(g.q, I_set1(0xFFFFFFFFu)), I_srl(g.q, 32), g.traj, I_set1(0), sw->key0, sw->key1, wv);
        const vd r = V_sqrt(V_mul(V_set1(-2.0), log_vec(unit_vec(wv[0]))));
        vd sn, cs;
        sincos_vec(unit_vec(wv[1]), &sn, &cs);
        const vd g0 = V_mul(r, cs), g1 = V_mul(r, sn);
        const vd ua = unit_vec(wv[2]), ub = unit_vec(wv[3]);
        int err = 0;
        unsigned per;
        /* odd step 2q+1 */
        unsigned ab = step(&g, lanes, sw, g0, ua, pair, 1, live, &per, &err);
        if (ab) retire(lanes, ab, sw, pair, 1, sh, &live);
        if (per & live) note_period(lanes, per & live, pair, 1);
        /* a lane whose step 2q+1 is its last (odd cap) */
        if (nsteps & 1u) cap_lanes(lanes, sw, sh, pair, pairs_max, nsteps, &live);
        /* even step 2q+2 */
        ab = step(&g, lanes, sw, g1, ub, pair, 0, live, &per, &err);
        if (ab) retire(lanes, ab, sw, pair, 0, sh, &live);
        if (per & live) note_period(lanes, per & live, pair, 0);
        if (!(nsteps & 1u)) cap_lanes(lanes, sw, sh, pair, pairs_max, nsteps, &live);
        if (err) sh->error = 1;
        g.q = I_add(g.q, I_set1(1));
        ++pair;
        if (sh->seconds > 0.0 && ++since_check == 1024) {
            since_check = 0;
            if (elapsed(&sh->t0) >= sh->seconds) {
                /* count the steps of the trajectories still in flight */
                for (int i = 0; i < W; ++i)
                    if (live >> i & 1u) __atomic_fetch_add(&sh->traj_steps, 2 * (pair - lanes[i].start),
                                                           __ATOMIC_RELAXED);
                *timed_out = 1;
                return;
            }
        }
    }
    (void)c;
}

typedef struct thread_arg {
    shared* sh;
    int id, n_threads;
} thread_arg;

static void* worker(void* arg)
{
    thread_arg* a = (thread_arg*)arg;
    shared* sh = a->sh;
    int timed_out = 0;
    /* start on set id mod n_sets, then take the other sets' remaining work */
    for (int k = 0; k < sh->n_sets && !timed_out; ++k) {
        set_work* sw = &sh->sets[(a->id + k) % sh->n_sets];
        run_set(sw, sh, &timed_out);
        if (sh->seconds > 0.0 && !timed_out) k = -1;  /* timed mode never runs dry: keep going */
    }
    return NULL;
}

/*
 * Absorbing first passages of n_sets single-barrier parameter sets (the
 * oracle's lane mode with absorb = 1: params[s].nsteps is the cap,
 * save_freq >= 10 drops trailing steps as the oracle does).
 *
 * n_traj > 0, seconds == 0: trajectories 0 .. n_traj-1 of every set;
 *   absorbed[s * n_traj + t] = absorption step (0 = capped) and tpt0[...] =
 *   the last well-centre crossing before it, if the arrays are given.
 * seconds > 0: trajectories are handed out without end until `seconds` of
 *   wall time have passed; the steps done so far are counted.
 * Returns 0, or -1 on bad arguments / 1 if a lane's barrier table overflowed.
 */
int ldo_simd_absorb(const ldo_params* params, const uint64_t* seeds, int n_sets, uint32_t n_traj, double seconds,
                    int n_threads, uint64_t* absorbed, uint64_t* tpt0, uint64_t* traj_steps, uint64_t* traj_done,
                    double* wall)
{
    if (n_sets < 1 || n_threads < 1 || (seconds <= 0.0 && n_traj == 0)) return -1;
    for (int s = 0; s < n_sets; ++s)
        if (params[s].model != 0 || params[s].absorb != 1) return -1;
    set_work* sets = (set_work*)calloc((size_t)n_sets, sizeof(set_work));
    shared sh;
    memset(&sh, 0, sizeof(sh));
    for (int s = 0; s < n_sets; ++s) {
        sets[s].p = params[s];
        make_sconst(&params[s], &sets[s].c);
        sets[s].key0 = (uint32_t)seeds[s];
        sets[s].key1 = (uint32_t)(seeds[s] >> 32);
        sets[s].n_traj = seconds > 0.0 ? 0 : n_traj;
        sets[s].absorbed = absorbed ? absorbed + (size_t)s * n_traj : NULL;
        sets[s].tpt0 = tpt0 ? tpt0 + (size_t)s * n_traj : NULL;
        const uint64_t F = params[s].save_freq;
        sets[s].nsteps = (F >= 10) ? (params[s].nsteps / F) * F : params[s].nsteps;
        if (sets[s].nsteps == 0) {
            free(sets);
            return -1;
        }
    }
    sh.sets = sets;
    sh.n_sets = n_sets;
    sh.seconds = seconds;
    clock_gettime(CLOCK_MONOTONIC, &sh.t0);
    pthread_t* th = (pthread_t*)calloc((size_t)n_threads, sizeof(pthread_t));
    thread_arg* ta = (thread_arg*)calloc((size_t)n_threads, sizeof(thread_arg));
    for (int t = 0; t < n_threads; ++t) {
        ta[t].sh = &sh;
        ta[t].id = t;
        ta[t].n_threads = n_threads;
        pthread_create(&th[t], NULL, worker, &ta[t]);
    }
    for (int t = 0; t < n_threads; ++t) pthread_join(th[t], NULL);
    if (wall) *wall = elapsed(&sh.t0);
    if (traj_steps) *traj_steps = sh.traj_steps;
    if (traj_done) *traj_done = sh.traj_done;
    free(th);
    free(ta);
    free(sets);
    return sh.error ? 1 : 0;
}
