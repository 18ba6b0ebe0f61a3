/*
 * ldpcp_oracle.c -- CPU restatement of the reference Langevin hot path.
 *
 * TEST INFRASTRUCTURE ONLY (the checker, never the product).  Built with
 * -ffp-contract=off so that every multiply and add rounds separately, as in
 * the reference's intrinsic code; the one fused multiply-add the reference
 * asks for (_mmSIMD_fmadd_pd in the overdamped position update,
 * langevinDynamicsSingleBarrier.h:2868) is an explicit fma().
 *
 * Citations are file:line into the reference sources.
 */
#include "ldpcp_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "ldpcp_gauss.h"

/* ------------------------------------------------------------------------- */
/* Constants: IntegrateInitializeVariables (Single.h:1520-1618, overdamped     */
/* :2812-2824; Double.h:1733-1937 with GetXc :1722-1732).                      */
/* ------------------------------------------------------------------------- */
typedef struct consts {
    double dt, minv, mhalf, halfdt, expmgdt, var;   /* var: noise amplitude of the chosen integrator */
    double dtgm;                                    /* overdamped dt/(m gamma) */
    double L, Linv, halfL;
    double c_lap;                                   /* -2 dt/tau  or -(K+1) dt/tau */
    double p_flip, p_asym[2], p0, p_create;
    double K;
    /* single barrier */
    double H[2], S[2], w;
    /* double barrier */
    double Hd[2][2];                                /* [barrier][state] */
    double xb[2], lo[2], hi[2], xc, yc;
    double slope[2][2][2], icept[2][2][2];          /* [barrier][state][0=left,1=right] */
    double vsigma;
    int tau_zero;
} consts;

static void make_consts(const ldo_params* p, consts* c)
{
    memset(c, 0, sizeof(*c));
    c->dt = p->timestep;
    c->minv = 1.0 / p->mass;
    c->mhalf = p->mass * 0.5;
    c->halfdt = p->timestep * 0.5;
    c->expmgdt = exp(-p->gamma * p->timestep);
    c->var = sqrt(p->kb * p->temperature / p->mass * (1.0 - exp(-2.0 * p->gamma * p->timestep)));
    c->L = p->periodic_length;
    c->K = p->eq_constant;
    c->tau_zero = (p->tau == 0.0);
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
    /* BarrierStatus::Initialize uses K whatever EQ's flag says (Single.h:90) */
    c->p_create = 1.0 / (1.0 + p->eq_constant);
    c->halfL = c->L * 0.5;
    c->Linv = 1.0 / c->L;
    if (p->model == 0) {
        c->w = p->half_width[0];
        for (int s = 0; s < 2; ++s) {
            c->H[s] = p->height[0][s];
            c->S[s] = p->height[0][s] / p->half_width[0];
        }
    } else {
        for (int b = 0; b < 2; ++b) {
            for (int s = 0; s < 2; ++s) c->Hd[b][s] = p->height[b][s];
            c->xb[b] = p->position[b];
        }
        double x1b = p->position[0], w1 = p->half_width[0], x2b = p->position[1], w2 = p->half_width[1];
        double y1 = p->height[0][0], y2 = p->height[1][0];
        double x1l = x1b - w1, x1h = x1b + w1, x2l = x2b - w2;
        (void)x1l;
        double xc = fma(x1h * y1, x2b - x2l, x2l * y2 * (x1h - x1b)) / fma(y1, x2b - x2l, y2 * (x1h - x1b));
        double yc = (x1h - xc) / (x1h - x1b) * y1;
        c->xc = xc;
        c->yc = yc;
        c->lo[0] = x1b - w1; c->hi[0] = x1b + w1;
        c->lo[1] = x2b - w2; c->hi[1] = x2b + w2;
        double xb_xl[2] = {c->xb[0] - c->lo[0], c->xb[1] - c->lo[1]};
        double xh_xb[2] = {c->hi[0] - c->xb[0], c->hi[1] - c->xb[1]};
        for (int s = 0; s < 2; ++s) {
            double H0 = c->Hd[0][s], H1 = c->Hd[1][s];
            c->slope[0][s][0] = H0 / xb_xl[0];
            c->slope[0][s][1] = (yc - H0) / (xc - c->xb[0]);
            c->slope[1][s][0] = (H1 - yc) / (c->xb[1] - xc);
            c->slope[1][s][1] = 0.0 - H1 / xh_xb[1];
            c->icept[0][s][0] = 0.0 - H0 * (c->lo[0] / xb_xl[0]);
            c->icept[0][s][1] = (xc * H0 - yc * c->xb[0]) / (xc - c->xb[0]);
            c->icept[1][s][0] = (yc * c->xb[1] - xc * H1) / (c->xb[1] - xc);
            c->icept[1][s][1] = H1 * (c->hi[1] / xh_xb[1]);
        }
    }
    if (p->overdamped) {
        double mgi = 1.0 / (p->mass * p->gamma);
        c->dtgm = p->timestep * mgi;
        c->var = sqrt(p->kb * p->temperature * mgi * p->timestep * 0.5);
    }
    c->vsigma = sqrt(p->kb * p->temperature / p->mass);
}

/* ------------------------------------------------------------------------- */
/* One trajectory's dynamic state (a lane of TrajectoryData, Single.h:128-217) */
/* ------------------------------------------------------------------------- */
typedef struct lstate {
    double x, v, F, E, R, barrier;
    double state;
    double slope_pos, slope_neg, height;          /* single */
    double d_slope[2][2], d_icept[2][2], d_height[2]; /* double: [barrier][left/right] */
    double well, periodic, well_prev, periodic_prev, xp;
    double local, local_prev;
    int ent_well, ent_period, ent_local;
    double count[2];
} lstate;

/* UpdateBarrierStateEtaParamter (Single.h:2154-2165; Double.h:2737-2757) */
static void lane_param(lstate* s, const consts* c, int model)
{
    int one = (s->state == 1.0);
    if (model == 0) {
        s->height = one ? c->H[1] : c->H[0];
        s->slope_pos = one ? c->S[1] : c->S[0];
        s->slope_neg = 0.0 - s->slope_pos;
    } else {
        int st = one ? 1 : 0;
        for (int b = 0; b < 2; ++b) {
            s->d_height[b] = c->Hd[b][st];
            for (int side = 0; side < 2; ++side) {
                s->d_slope[b][side] = c->slope[b][st][side];
                s->d_icept[b][side] = c->icept[b][st][side];
            }
        }
    }
    if (one) s->count[1] = s->count[1] + 1.0;
    else s->count[0] = s->count[0] + 1.0;
}

/* EnergyAndForce (Single.h:1292-1324; Double.h:1464-1522) */
static void lane_energy_force(lstate* s, const consts* c, int model)
{
    if (model == 0) {
        double r = s->x * c->Linv;
        double pidx = rint(r);
        double xp = s->x - pidx * c->L;
        int neg = (xp < 0.0) && (xp > -c->w);
        int pos = (xp < c->w) && (xp > 0.0);
        double vneg = s->height - xp * s->slope_neg;
        double vpos = s->height - xp * s->slope_pos;
        double E = neg ? vneg : 0.0;
        E = pos ? vpos : E;
        double F = neg ? s->slope_neg : 0.0;
        F = pos ? s->slope_pos : F;
        s->E = E;
        s->F = F;
        s->barrier = s->height;
    } else {
        double xp = s->xp;
        int n0 = (xp < c->xb[0]) && (xp > c->lo[0]);
        int p0 = (xp < c->xc) && (xp > c->xb[0]);
        int n1 = (xp < c->xb[1]) && (xp > c->xc);
        int p1 = (xp < c->hi[1]) && (xp > c->xb[1]);
        double v00 = xp * s->d_slope[0][0] + s->d_icept[0][0];
        double v01 = xp * s->d_slope[0][1] + s->d_icept[0][1];
        double v10 = xp * s->d_slope[1][0] + s->d_icept[1][0];
        double v11 = xp * s->d_slope[1][1] + s->d_icept[1][1];
        double E = n0 ? v00 : 0.0;
        E = p0 ? v01 : E;
        E = n1 ? v10 : E;
        E = p1 ? v11 : E;
        double F = n0 ? s->d_slope[0][0] : 0.0;
        F = p0 ? s->d_slope[0][1] : F;
        F = n1 ? s->d_slope[1][0] : F;
        F = p1 ? s->d_slope[1][1] : F;
        F = 0.0 - F;
        double B = n0 ? s->d_height[0] : 0.0;
        B = p0 ? s->d_height[0] : B;
        B = n1 ? s->d_height[1] : B;
        B = p1 ? s->d_height[1] : B;
        s->E = E;
        s->F = F;
        s->barrier = B;
    }
}

/* InitializePeriodicIndices / UpdatePeriodicIndices (Single.h:1838-1884;
 * Double.h:2150-2235).  floor = _MM_FROUND_TO_NEG_INF (0x09), rint = round
 * half to even (_MM_FROUND_TO_NEAREST_INT). */
static void lane_indices(lstate* s, const consts* c, int model, int init)
{
    double r = s->x * c->Linv;
    double well = floor(r), periodic = rint(r);
    if (init) {
        s->well = s->well_prev = well;
        s->periodic = s->periodic_prev = periodic;
        s->ent_well = s->ent_period = 0;
    } else {
        s->well_prev = s->well;
        s->periodic_prev = s->periodic;
        s->well = well;
        s->periodic = periodic;
        s->ent_well = (s->well_prev != s->well);
        s->ent_period = (s->periodic_prev != s->periodic);
    }
    if (model == 1) {
        s->xp = s->x - s->periodic * c->L;
        double local = (s->xp < c->xc) ? 0.0 : 1.0;
        if (init) {
            s->local = s->local_prev = local;
            s->ent_local = 0;
        } else {
            s->local_prev = s->local;
            s->local = local;
            s->ent_local = (s->local_prev != s->local);
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Statistics state machines: InitializeStatistics / UpdateStatistics          */
/* (Single.h:2166-2592; Double.h:2758-3209).  `touch` replays the barrier      */
/* look-ups the single-barrier FPT code makes, which create barriers (and so   */
/* draw uniforms) in rank mode; lane mode passes NULL.                          */
/* ------------------------------------------------------------------------- */
typedef int (*touch_fn)(void* ctx, int lane, int id);  /* returns the barrier's stored state */

static void cap_put(double* base, uint64_t cap, uint64_t n, double v)
{
    if (base && n < cap) base[n] = v;
}

typedef struct capref {
    ldo_capture* cap;
    int lane;
    int W;
    int ncg;
} capref;

#define CAPPTR(field, cr) ((cr) && (cr)->cap && (cr)->cap->field ? (cr)->cap->field + (uint64_t)(cr)->lane * (cr)->cap->cap : NULL)

static void fpt_arm_indices(int idx[2], const lstate* s, touch_fn touch, void* tctx, int lane)
{
    int pidx = (int)s->periodic;
    if (s->well == s->periodic) {
        if (touch) { touch(tctx, lane, pidx); touch(tctx, lane, pidx + 1); }
        idx[0] = pidx;
        idx[1] = pidx + 1;
    } else {
        if (touch) { touch(tctx, lane, pidx - 1); touch(tctx, lane, pidx); }
        idx[0] = pidx - 1;
        idx[1] = pidx;
    }
}

/* FPT1 arming (Single.h:2364-2387, at step 0 :2184-2212): the barriers on
 * either side of the well the particle sits in, the one under the particle
 * with its current state, the other with its stored state (looked up, and so
 * created, in the order the reference does). */
static void fpt1_arm(ldo_lane* o, const lstate* s, touch_fn touch, void* tctx, int lane)
{
    int pidx = (int)s->periodic;
    if (s->well == s->periodic) {
        if (touch) touch(tctx, lane, pidx);
        int st = touch ? touch(tctx, lane, pidx + 1) : 0;
        o->fpt_init_state[0] = s->state;
        o->fpt_init_index[0] = pidx;
        o->fpt_init_state[1] = (double)st;
        o->fpt_init_index[1] = pidx + 1;
    } else {
        int st = touch ? touch(tctx, lane, pidx - 1) : 0;
        if (touch) touch(tctx, lane, pidx);
        o->fpt_init_state[0] = (double)st;
        o->fpt_init_index[0] = pidx - 1;
        o->fpt_init_state[1] = s->state;
        o->fpt_init_index[1] = pidx;
    }
}

static void stats_init(const ldo_params* p, ldo_lane* o, const lstate* s, touch_fn touch, void* tctx, int lane)
{
    double m = fmod(fmod(p->x0, p->periodic_length) + p->periodic_length, p->periodic_length);
    int centre = (m == p->periodic_length * 0.5);
    if (p->b_waiting) {
        o->wt_tick_prev = 0.0;
        o->wt_passed = 0;
    }
    if (p->b_fpt) {
        int idx[2];
        o->fpt_last_tick = 0.0;
        (void)idx;
        if (centre) {
            o->fpt_to_update = 0;
            if (p->model == 0) fpt1_arm(o, s, touch, tctx, lane);
        } else {
            o->fpt_to_update = 1;
            o->fpt_init_state[0] = 0.0;
        }
        o->fpt2_last_tick = 0.0;
        if (centre) {
            o->fpt2_to_update = 0;
            fpt_arm_indices(o->fpt2_idx, s, p->model == 0 ? touch : NULL, tctx, lane);
        } else {
            o->fpt2_to_update = 1;
        }
    }
    if (p->b_tpt) {
        o->tpt_last_tick = 0.0;
        o->tpt_to_update = centre ? 0 : 1;
        o->tpt2_last_tick = 0.0;
        o->tpt2_to_update = centre ? 0 : 1;
        o->tpt2_initial_well = (int)s->well;
    }
    if (p->b_fptcg) {
        for (int j = 0; j < p->n_cg; ++j) {
            o->cg_last_tick[j] = 0.0;
            o->cg_initial_well[j] = (int)s->well;
            o->cg_to_update[j] = !(centre && ((int)round(s->well) % p->cg_size[j] == 0));
        }
    }
}

/* Event durations `m_TimeStep0 * iStep - lastTick` (Single.h:2398, 2453, 2509,
 * 2525, 2576) are scalar statements that the reference's compiler contracts
 * into one fused multiply-add (icpx -fp-model=precise permits contraction
 * within a statement); the recorded event lists of the quick test match only
 * this form. */
static double ev(const ldo_params* p, uint64_t iStep, double last)
{
    return fma(p->timestep, (double)iStep, -last);
}

static void stats_update(const ldo_params* p, ldo_lane* o, const lstate* s, uint64_t iStep,
                         touch_fn touch, void* tctx, int lane, capref* cr)
{
    const double t = p->timestep * (double)iStep;
    const int single = (p->model == 0);
    if (p->b_waiting && single) {
        if (s->ent_well) {
            o->wt_tick_prev = (double)iStep;
            o->wt_passed = 1;
        } else if (s->ent_period) {
            if (o->wt_passed) {
                double tick = o->wt_tick_prev;
                cap_put(CAPPTR(ticks, cr), cr ? cr->cap->cap : 0, o->wt_ticks, tick);
                if (o->wt_ticks > 0) {
                    double w = (tick - o->wt_last) * p->timestep;
                    o->wt_sum += w;
                    o->wt_sum2 = fma(w, w, o->wt_sum2);
                    if (o->wt_n == 0 || w > o->wt_max) o->wt_max = w;
                    o->wt_n++;
                }
                o->wt_last = tick;
                o->wt_ticks++;
                o->wt_passed = 0;
            }
        }
    }
    if (p->b_fpt) {
        if (single) {
            if (o->fpt_to_update && s->ent_period) {
                o->fpt_to_update = 0;
                o->fpt_last_tick = t;
                fpt1_arm(o, s, touch, tctx, lane);
            }
            if (!o->fpt_to_update && s->ent_well) {
                o->fpt_to_update = 1;
                double v = ev(p, iStep, o->fpt_last_tick);
                cap_put(CAPPTR(fpt, cr), cr ? cr->cap->cap : 0, o->fpt_n, v);
                /* barrier logs (Single.h:2398-2404) */
                cap_put(CAPPTR(init, cr), cr ? cr->cap->cap : 0, o->fpt_n,
                        s->periodic == (double)o->fpt_init_index[0] ? o->fpt_init_state[0] : o->fpt_init_state[1]);
                cap_put(CAPPTR(inst, cr), cr ? cr->cap->cap : 0, o->fpt_n, s->state);
                o->fpt_sum += v;
                o->fpt_n++;
            }
        }
        if (o->fpt2_to_update && s->ent_period) {
            o->fpt2_to_update = 0;
            o->fpt2_last_tick = t;
            fpt_arm_indices(o->fpt2_idx, s, single ? touch : NULL, tctx, lane);
        }
        if (s->ent_period && !o->fpt2_to_update &&
            ((double)o->fpt2_idx[0] != s->periodic && (double)o->fpt2_idx[1] != s->periodic)) {
            double v = ev(p, iStep, o->fpt2_last_tick);
            cap_put(CAPPTR(fpt2, cr), cr ? cr->cap->cap : 0, o->fpt2_n, v);
            o->fpt2_sum += v;
            o->fpt2_n++;
            o->fpt2_last_tick = t;
            fpt_arm_indices(o->fpt2_idx, s, single ? touch : NULL, tctx, lane);
        }
    }
    if (p->b_tpt && single) {
        if (s->ent_period) {
            o->tpt_to_update = 0;
            o->tpt_last_tick = t;
        }
        if (!o->tpt_to_update && s->ent_well) {
            o->tpt_to_update = 1;
            double v = ev(p, iStep, o->tpt_last_tick);
            cap_put(CAPPTR(tpt, cr), cr ? cr->cap->cap : 0, o->tpt_n, v);
            o->tpt_sum += v;
            o->tpt_n++;
        }
        if (s->ent_period) {
            if (o->tpt2_to_update) {
                o->tpt2_to_update = 0;
            } else if (s->well != (double)o->tpt2_initial_well) {
                double v = ev(p, iStep, o->tpt2_last_tick);
                cap_put(CAPPTR(tpt2, cr), cr ? cr->cap->cap : 0, o->tpt2_n, v);
                o->tpt2_sum += v;
                o->tpt2_n++;
            }
            o->tpt2_last_tick = t;
            o->tpt2_initial_well = (int)s->well;
        }
    }
    if (p->b_fptcg) {
        for (int j = 0; j < p->n_cg; ++j) {
            if (s->ent_period && ((int)round(s->well) % p->cg_size[j] == 0)) {
                if (o->cg_to_update[j]) {
                    o->cg_to_update[j] = 0;
                    o->cg_last_tick[j] = t;
                    o->cg_initial_well[j] = (int)s->well;
                } else if (s->well != (double)o->cg_initial_well[j]) {
                    double v = ev(p, iStep, o->cg_last_tick[j]);
                    if (cr && cr->cap && cr->cap->cg)
                        cap_put(cr->cap->cg + ((uint64_t)j * cr->W + cr->lane) * cr->cap->cap, cr->cap->cap, o->cg_n[j], v);
                    o->cg_sum[j] += v;
                    o->cg_n[j]++;
                    o->cg_last_tick[j] = t;
                    o->cg_initial_well[j] = (int)s->well;
                }
            }
        }
    }
}

static void lane_export(ldo_lane* o, const lstate* s)
{
    o->x = s->x; o->v = s->v; o->f = s->F; o->e = s->E; o->r_prev = s->R;
    o->state = s->state; o->barrier = s->barrier;
    o->state_count[0] = s->count[0]; o->state_count[1] = s->count[1];
    o->well = s->well; o->periodic = s->periodic; o->xp = s->xp; o->local_barrier = s->local;
}

/* Record (Single.h:1707-1765): fields x, v, f, pe, ke, state, barrier. */
static void lane_record(const ldo_params* p, const consts* c, const lstate* s, double* frames,
                        uint64_t cap, uint64_t frame, uint64_t stride)
{
    if (!frames || frame >= cap) return;
    frames[frame * stride] = s->x;
    if (p->save_all) {
        uint64_t plane = cap * stride;
        frames[plane * 1 + frame * stride] = p->overdamped ? 0.0 : s->v;
        frames[plane * 2 + frame * stride] = s->F;
        frames[plane * 3 + frame * stride] = s->E;
        frames[plane * 4 + frame * stride] = c->mhalf * (s->v * s->v);
        frames[plane * 5 + frame * stride] = s->state;
        frames[plane * 6 + frame * stride] = s->barrier;
    }
}

/* ------------------------------------------------------------------------- */
/* Rank mode: W lanes sharing three banks and a lazily grown barrier table     */
/* ------------------------------------------------------------------------- */
typedef struct bank {
    double* buf;
    uint64_t size, process;
    int64_t renew;
    int which;
} bank;

typedef struct barrier_table {
    int min_id, n, cap;
    int32_t* state;   /* [n][W] */
    uint64_t* step;   /* [n][W] */
    int8_t* init;     /* [n][W] state at creation */
} barrier_table;

struct ldo_rank_state {
    ldo_params p;
    consts c;
    int W;
    ldo_fill_fn fill;
    void* ctx;
    bank g, u, us;
    barrier_table bt[2];
    double initial_value[LDO_MAX_W];  /* UserData::m_bInitialValue */
    double state0_instance[LDO_MAX_W];
    ldo_record* rec;
    uint64_t rec_flip_n[LDO_MAX_W], rec_catch_n[LDO_MAX_W], rec_gauss_n[LDO_MAX_W];
    int cur_array;
};

static void bank_renew(ldo_rank_state* s, bank* b)
{
    s->fill(s->ctx, b->which, b->buf, b->size);
    b->process = 0;
    b->renew++;
}

/* GetRandomNumberSafe (RANDOMNUMBER.cpp:230-237, 322-329) */
static void bank_vec(ldo_rank_state* s, bank* b, double* out)
{
    if (b->process + (uint64_t)s->W > b->size) bank_renew(s, b);
    memcpy(out, b->buf + b->process, sizeof(double) * s->W);
    b->process += s->W;
}

/* UniformRandomNumberSequence::GetRandomNumber (RANDOMNUMBER.cpp:129-136) */
static double bank_scalar(ldo_rank_state* s, bank* b)
{
    if (b->process == b->size) bank_renew(s, b);
    return b->buf[b->process++];
}

static void rec_flip(ldo_rank_state* s, const double* u)
{
    if (!s->rec || !s->rec->flip) return;
    for (int i = 0; i < s->W; ++i) {
        if (s->rec_flip_n[i] < s->rec->cap) s->rec->flip[(uint64_t)i * s->rec->cap + s->rec_flip_n[i]] = u[i];
        s->rec_flip_n[i]++;
    }
}

static void rec_gauss(ldo_rank_state* s, const double* g)
{
    if (!s->rec || !s->rec->gauss) return;
    for (int i = 0; i < s->W; ++i) {
        if (s->rec_gauss_n[i] < s->rec->cap) s->rec->gauss[(uint64_t)i * s->rec->cap + s->rec_gauss_n[i]] = g[i];
        s->rec_gauss_n[i]++;
    }
}

/* BarrierStatus::Initialize (Single.h:78-96) */
static void barrier_create(ldo_rank_state* s, barrier_table* t, int slot, int id)
{
    int W = s->W;
    if (s->p.same_initial_barrier) {
        for (int i = 0; i < W; ++i) {
            t->state[(uint64_t)slot * W + i] = (int)s->initial_value[i];
            t->step[(uint64_t)slot * W + i] = 0;
            t->init[(uint64_t)slot * W + i] = (int8_t)(int)s->initial_value[i];
        }
    } else {
        double u[LDO_MAX_W];
        bank_vec(s, &s->u, u);
        for (int i = 0; i < W; ++i) {
            int st = u[i] < s->c.p_create ? 0 : 1;
            t->state[(uint64_t)slot * W + i] = st;
            t->step[(uint64_t)slot * W + i] = 0;   /* UserData::m_iStep stays 0 */
            t->init[(uint64_t)slot * W + i] = (int8_t)st;
        }
    }
    if (s->rec && s->rec->binit) {
        int a = (int)(t - s->bt);
        int k = id - s->rec->bmin;
        if (k >= 0 && k < s->rec->bn)
            for (int i = 0; i < W; ++i)
                s->rec->binit[((uint64_t)a * W + i) * s->rec->bn + k] = t->init[(uint64_t)slot * W + i];
    }
}

static void table_reserve(barrier_table* t, int W, int need)
{
    if (need <= t->cap) return;
    int cap = t->cap ? t->cap : 256;
    while (cap < need) cap *= 2;
    t->state = (int32_t*)realloc(t->state, sizeof(int32_t) * (uint64_t)cap * W);
    t->step = (uint64_t*)realloc(t->step, sizeof(uint64_t) * (uint64_t)cap * W);
    t->init = (int8_t*)realloc(t->init, sizeof(int8_t) * (uint64_t)cap * W);
    t->cap = cap;
}

/* FiniteInifiniteIntegerTickArray1D::GetElementByID (DataLocator.h:121-148):
 * grows downwards or upwards, creating every missing id in ascending order. */
static int barrier_slot(ldo_rank_state* s, barrier_table* t, int id)
{
    int W = s->W;
    if (t->n == 0) {
        table_reserve(t, W, 1);
        t->min_id = id;
        t->n = 1;
        barrier_create(s, t, 0, id);
    } else if (id < t->min_id) {
        int add = t->min_id - id;
        table_reserve(t, W, t->n + add);
        memmove(t->state + (uint64_t)add * W, t->state, sizeof(int32_t) * (uint64_t)t->n * W);
        memmove(t->step + (uint64_t)add * W, t->step, sizeof(uint64_t) * (uint64_t)t->n * W);
        memmove(t->init + (uint64_t)add * W, t->init, sizeof(int8_t) * (uint64_t)t->n * W);
        int old_min = t->min_id;
        t->n += add;
        t->min_id = id;
        for (int i = id; i < old_min; ++i) barrier_create(s, t, i - id, i);
    } else if (id - t->min_id + 1 > t->n) {
        int old_end = t->n + t->min_id;
        int newn = id - t->min_id + 1;
        table_reserve(t, W, newn);
        t->n = newn;
        for (int i = old_end; i <= id; ++i) barrier_create(s, t, i - t->min_id, i);
    }
    return id - t->min_id;
}

static int rank_touch(void* ctx, int lane, int id)
{
    ldo_rank_state* s = (ldo_rank_state*)ctx;
    barrier_table* t = &s->bt[s->cur_array];
    int slot = barrier_slot(s, t, id);
    return t->state[(uint64_t)slot * s->W + lane];
}

ldo_rank_state* ldo_rank_create(const ldo_params* p, int W, ldo_fill_fn fill, void* ctx, uint64_t bank_size)
{
    if (W < 1 || W > LDO_MAX_W || !fill) return NULL;
    ldo_rank_state* s = (ldo_rank_state*)calloc(1, sizeof(*s));
    s->p = *p;
    make_consts(p, &s->c);
    s->W = W;
    s->fill = fill;
    s->ctx = ctx;
    bank* banks[3] = {&s->g, &s->u, &s->us};
    for (int k = 0; k < 3; ++k) {
        banks[k]->size = bank_size;
        banks[k]->buf = (double*)malloc(sizeof(double) * bank_size);
        banks[k]->which = k;
        banks[k]->renew = 0;
        bank_renew(s, banks[k]);   /* Initialize() renews once (RANDOMNUMBER.cpp:84-95) */
    }
    return s;
}

void ldo_rank_destroy(ldo_rank_state* s)
{
    if (!s) return;
    free(s->g.buf); free(s->u.buf); free(s->us.buf);
    for (int a = 0; a < 2; ++a) { free(s->bt[a].state); free(s->bt[a].step); free(s->bt[a].init); }
    free(s);
}

void ldo_rank_rng_progress(const ldo_rank_state* s, int64_t renew[3], uint64_t offset[3])
{
    renew[0] = s->g.renew; offset[0] = s->g.process;
    renew[1] = s->u.renew; offset[1] = s->u.process;
    renew[2] = s->us.renew; offset[2] = s->us.process;
}

int ldo_rank_barriers(const ldo_rank_state* s, int* min_id, int32_t* states, uint64_t* steps, int cap, int which)
{
    const barrier_table* t = &s->bt[which];
    *min_id = t->min_id;
    int n = t->n < cap ? t->n : cap;
    if (states) memcpy(states, t->state, sizeof(int32_t) * (uint64_t)n * s->W);
    if (steps) memcpy(steps, t->step, sizeof(uint64_t) * (uint64_t)n * s->W);
    return t->n;
}

/* Per-step telegraph update + barrier hand-over for one lane in rank mode
 * (UpdateBarrierStateEta, Single.h:2025-2068 / 2096-2145; Double.h:2466-2510;
 * UpdateBarrierState for TAU == 0, Single.h:1958-1992). */
static void rank_handover(ldo_rank_state* s, lstate* L, int i, uint64_t iStep)
{
    const int model = s->p.model;
    const consts* c = &s->c;
    int entering = model == 0 ? L->ent_period : L->ent_local;
    if (!entering) return;
    int W = s->W;
    int a_prev = model == 0 ? 0 : (int)L->local_prev;
    int a_cur = model == 0 ? 0 : (int)L->local;
    int prev_id = (int)L->periodic_prev;
    int cur_id = (int)L->periodic;
    barrier_table* tp = &s->bt[a_prev];
    int sp = barrier_slot(s, tp, prev_id);
    tp->state[(uint64_t)sp * W + i] = (int)L->state;
    tp->step[(uint64_t)sp * W + i] = iStep;
    barrier_table* tc = &s->bt[a_cur];
    int sc = barrier_slot(s, tc, cur_id);
    if (c->tau_zero) {
        L->state = (double)tc->state[(uint64_t)sc * W + i];
        return;
    }
    double prev_state = (double)tc->state[(uint64_t)sc * W + i];
    double lap = (double)(iStep - tc->step[(uint64_t)sc * W + i]);
    double pc;
    if (s->p.equal_probability) {
        pc = 0.5 * (1.0 - exp(c->c_lap * lap));
    } else if (prev_state == 0.0) {
        pc = c->K / (1.0 + c->K) * (1.0 - exp(c->c_lap * lap));
    } else {
        pc = 1.0 / (1.0 + c->K) * (1.0 - exp(c->c_lap * lap));
    }
    double u = bank_scalar(s, &s->us);
    if (s->rec && s->rec->catchup) {
        if (s->rec_catch_n[i] < s->rec->cap) s->rec->catchup[(uint64_t)i * s->rec->cap + s->rec_catch_n[i]] = u;
        s->rec_catch_n[i]++;
    }
    if (u < pc) L->state = (prev_state == 1.0) ? 0.0 : 1.0;
    else L->state = prev_state;
}

/* per-step flip of the barrier the lane sits on (Single.h:2010-2015, 2078-2086) */
static int flip_decision(const consts* c, int equal, double state, double u, double* out_state)
{
    if (equal) {
        if (u < c->p_flip) { *out_state = (state == 0.0) ? 1.0 : 0.0; return 1; }
    } else {
        int one = (state == 1.0);
        double pr = one ? c->p_asym[1] : c->p_asym[0];
        if (u < pr) { *out_state = one ? 0.0 : 1.0; return 1; }
    }
    *out_state = state;
    return 0;
}

int ldo_rank_integrate(ldo_rank_state* s, ldo_lane* lanes, double* frames, uint64_t frames_cap,
                       uint64_t* n_frames, ldo_capture* cap, ldo_record* rec)
{
    const ldo_params* p = &s->p;
    const consts* c = &s->c;
    const int W = s->W;
    const int model = p->model;
    lstate L[LDO_MAX_W];
    double vec[LDO_MAX_W];
    capref crs[LDO_MAX_W];

    s->rec = rec;
    memset(s->rec_flip_n, 0, sizeof(s->rec_flip_n));
    memset(s->rec_catch_n, 0, sizeof(s->rec_catch_n));
    memset(s->rec_gauss_n, 0, sizeof(s->rec_gauss_n));
    /* PrepareRecord -> TrajectoryData::Clear releases the barrier store (Single.h:245) */
    for (int a = 0; a < 2; ++a) s->bt[a].n = 0, s->bt[a].min_id = 0;
    /* PreIntegrate copies BarrierState0Instance into UserData before step 0 (Single.h:1378-1380) */
    for (int i = 0; i < W; ++i) s->initial_value[i] = s->state0_instance[i];
    memset(lanes, 0, sizeof(ldo_lane) * W);
    memset(L, 0, sizeof(L));
    for (int i = 0; i < W; ++i) {
        crs[i].cap = cap; crs[i].lane = i; crs[i].W = W; crs[i].ncg = p->n_cg;
    }

    /* ---- IntegrateStep0 (Single.h:1619-1652; overdamped :2825-2832) ---- */
    for (int i = 0; i < W; ++i) {
        L[i].x = p->x0;
        lane_indices(&L[i], c, model, 1);
    }
    /* InitializeBarrierState (Single.h:1885-1932; Double.h:2236-2290) */
    if (p->randomize_state) {
        bank_vec(s, &s->u, vec);
        rec_flip(s, vec);
        for (int i = 0; i < W; ++i) L[i].state = (vec[i] <= c->p0) ? 0.0 : 1.0;
    } else {
        for (int i = 0; i < W; ++i) L[i].state = p->initial_state;
    }
    for (int i = 0; i < W; ++i) {
        s->state0_instance[i] = L[i].state;
        lanes[i].state0_instance = L[i].state;
    }
    int narrays = model == 0 ? 1 : 2;
    int edge = model == 0 ? 100 : 10000;
    for (int a = 0; a < narrays; ++a) {
        barrier_slot(s, &s->bt[a], -edge);
        barrier_slot(s, &s->bt[a], edge);
    }
    for (int i = 0; i < W; ++i) {
        int a = model == 0 ? 0 : (int)L[i].local;
        int sl = barrier_slot(s, &s->bt[a], (int)L[i].periodic);
        s->bt[a].state[(uint64_t)sl * W + i] = (int)s->state0_instance[i];
    }
    for (int i = 0; i < W; ++i) {
        L[i].count[0] = L[i].count[1] = 0.0;
        lane_param(&L[i], c, model);
    }
    if (p->randomize_v) {
        bank_vec(s, &s->g, vec);
        rec_gauss(s, vec);
        for (int i = 0; i < W; ++i) L[i].v = c->vsigma * vec[i];
    } else {
        for (int i = 0; i < W; ++i) L[i].v = p->v0;
    }
    for (int i = 0; i < W; ++i) {
        lanes[i].v0_instance = L[i].v;
        lane_energy_force(&L[i], c, model);
    }
    s->cur_array = 0;
    for (int i = 0; i < W; ++i) stats_init(p, &lanes[i], &L[i], rank_touch, s, i);
    if (p->overdamped) {
        bank_vec(s, &s->g, vec);
        rec_gauss(s, vec);
        for (int i = 0; i < W; ++i) L[i].R = c->var * vec[i];
    }
    uint64_t nf = 0;
    for (int i = 0; i < W; ++i) lane_record(p, c, &L[i], frames ? frames + i : NULL, frames_cap, nf, W);
    nf++;

    /* ---- main loop (Integrate, Single.h:1403-1511) ---- */
    uint64_t F = p->save_freq;
    uint64_t nsteps = (F >= 10) ? (p->nsteps / F) * F : p->nsteps;
    for (uint64_t iStep = 1; iStep <= nsteps; ++iStep) {
        double g[LDO_MAX_W];
        bank_vec(s, &s->g, g);
        rec_gauss(s, g);
        for (int i = 0; i < W; ++i) {
            lstate* l = &L[i];
            if (p->overdamped) {
                /* Single.h:2860-2876 */
                double rf = c->var * g[i];
                l->x = fma(l->F, c->dtgm, l->x);
                l->x = l->x + rf;
                l->x = l->x + l->R;
                l->R = rf;
            } else {
                /* Single.h:1680-1693 */
                l->v = l->v + (c->minv * l->F) * c->halfdt;
                l->x = l->x + l->v * c->halfdt;
                l->v = c->expmgdt * l->v + c->var * g[i];
                l->x = l->x + l->v * c->halfdt;
            }
            lane_indices(l, c, model, 0);
        }
        if (!c->tau_zero) {
            double u[LDO_MAX_W];
            bank_vec(s, &s->u, u);
            rec_flip(s, u);
            for (int i = 0; i < W; ++i) flip_decision(c, p->equal_probability, L[i].state, u[i], &L[i].state);
        }
        for (int i = 0; i < W; ++i) rank_handover(s, &L[i], i, iStep);
        for (int i = 0; i < W; ++i) {
            lane_param(&L[i], c, model);
            lane_energy_force(&L[i], c, model);
            if (!p->overdamped) L[i].v = L[i].v + (c->minv * L[i].F) * c->halfdt;
        }
        s->cur_array = 0;
        for (int i = 0; i < W; ++i) stats_update(p, &lanes[i], &L[i], iStep, rank_touch, s, i, cap ? &crs[i] : NULL);
        if (iStep % F == 0) {
            for (int i = 0; i < W; ++i) lane_record(p, c, &L[i], frames ? frames + i : NULL, frames_cap, nf, W);
            nf++;
        }
    }
    for (int i = 0; i < W; ++i) {
        lane_export(&lanes[i], &L[i]);
        lanes[i].steps_done = nsteps;
        lanes[i].n_gauss = s->rec_gauss_n[i];
        lanes[i].n_flip = s->rec_flip_n[i];
        lanes[i].n_catch = s->rec_catch_n[i];
    }
    if (n_frames) *n_frames = nf;
    s->rec = NULL;
    return 0;
}

/* ------------------------------------------------------------------------- */
/* Lane mode                                                                   */
/* ------------------------------------------------------------------------- */
void ldo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4])
{
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; ++r) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t n0 = (uint32_t)(p1 >> 32) ^ c1 ^ k0;
        uint32_t n1 = (uint32_t)p1;
        uint32_t n2 = (uint32_t)(p0 >> 32) ^ c3 ^ k1;
        uint32_t n3 = (uint32_t)p0;
        c0 = n0; c1 = n1; c2 = n2; c3 = n3;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

uint64_t ldo_flip_threshold(double p)
{
    double P = ceil(p * 4294967296.0 - 0.5);
    if (P < 0.0) P = 0.0;
    if (P > 4294967296.0) P = 4294967296.0;
    return (uint64_t)P;
}

typedef struct lane_noise {
    int kind;  /* 0 replay, 1 philox */
    const double *gauss, *flip, *catchup;
    uint64_t n_gauss, n_flip, n_catch;
    const int8_t* binit; int bmin, bn;
    uint32_t key[2]; uint32_t traj;
    int err;
} lane_noise;

static double u32_to_unit(uint32_t w) { return ((double)w + 0.5) * 2.3283064365386963e-10; }

static void philox_role(const lane_noise* n, uint64_t idx, uint32_t role, uint32_t out[4])
{
    uint32_t ctr[4] = {(uint32_t)idx, (uint32_t)(idx >> 32), n->traj, role};
    ldo_philox4x32_10(ctr, n->key, out);
}

/* Box-Muller on two 32-bit words: r = sqrt(-2 ln u1), theta = 2 pi u2, with the
 * logarithm and sine/cosine the GPU uses (ldpcp_gauss.h: IEEE operations only, so
 * the bits agree).  This is the build's noise definition, not the reference's
 * (whose MKL ICDF streams the replay mode carries); the functions are pinned
 * against a correctly rounded reference by ldo_log / ldo_sincos_turn. */
static void box_muller(uint32_t a, uint32_t b, double* g0, double* g1)
{
    ldpcp_box_muller(a, b, g0, g1);
}

double ldo_log(double x) { return ldpcp_log(x); }

void ldo_double_apex(const ldo_params* p, double* xc, double* yc)
{
    consts c;
    make_consts(p, &c);
    *xc = c.xc;
    *yc = c.yc;
}
void ldo_sincos_turn(double u, double* s, double* c) { ldpcp_sincos_turn(u, s, c); }

/* role 1: initial draws; word 0/1 -> V0 (cos) and R0 (sin), word 2 -> state */
static double noise_init_gauss(lane_noise* n, int which, ldo_lane* o)
{
    if (n->kind == 0) {
        if (o->n_gauss >= n->n_gauss) { n->err = 1; return 0.0; }
        return n->gauss[o->n_gauss++];
    }
    uint32_t w[4];
    philox_role(n, 0, 1, w);
    double g0, g1;
    box_muller(w[0], w[1], &g0, &g1);
    o->n_gauss++;
    return which == 0 ? g0 : g1;
}

static double noise_init_state_u(lane_noise* n, ldo_lane* o)
{
    if (n->kind == 0) {
        if (o->n_flip >= n->n_flip) { n->err = 1; return 0.0; }
        return n->flip[o->n_flip++];
    }
    uint32_t w[4];
    philox_role(n, 0, 1, w);
    o->n_flip++;
    return u32_to_unit(w[2]);
}

/* role 0: one Philox call per pair of steps -> two Gaussians, two flips */
static void noise_step(lane_noise* n, uint64_t iStep, ldo_lane* o, double* g, double* u)
{
    if (n->kind == 0) {
        if (o->n_gauss >= n->n_gauss) { n->err = 1; *g = 0.0; } else *g = n->gauss[o->n_gauss++];
        if (u) {
            if (o->n_flip >= n->n_flip) { n->err = 1; *u = 1.0; } else *u = n->flip[o->n_flip++];
        }
        return;
    }
    uint64_t q = (iStep - 1) >> 1;
    int half = (int)((iStep - 1) & 1);
    uint32_t w[4];
    philox_role(n, q, 0, w);
    double g0, g1;
    box_muller(w[0], w[1], &g0, &g1);
    *g = half ? g1 : g0;
    o->n_gauss++;
    if (u) { *u = u32_to_unit(half ? w[3] : w[2]); o->n_flip++; }
}

static double noise_catch(lane_noise* n, ldo_lane* o)
{
    if (n->kind == 0) {
        if (o->n_catch >= n->n_catch) { n->err = 1; return 1.0; }
        return n->catchup[o->n_catch++];
    }
    /* four uniforms per call: the n-th hand-over takes word n mod 4 of index n / 4 */
    uint32_t w[4];
    philox_role(n, o->n_catch >> 2, 2, w);
    const double u = u32_to_unit(w[o->n_catch & 3]);
    o->n_catch++;
    return u;
}

/* state a barrier has when first touched (role 3 / replay table) */
static int noise_barrier_init(lane_noise* n, const consts* c, int same, double init_value, int array, int id)
{
    if (same) return (int)init_value;
    if (n->kind == 0) {
        int k = id - n->bmin;
        if (!n->binit || k < 0 || k >= n->bn) { n->err = 2; return 0; }
        int8_t v = n->binit[(uint64_t)array * n->bn + k];
        if (v < 0) { n->err = 2; return 0; }
        return v;
    }
    uint32_t w[4];
    philox_role(n, (uint64_t)(uint32_t)id, 3u + (uint32_t)array * 0x100u, w);
    return u32_to_unit(w[0]) < c->p_create ? 0 : 1;
}

/* Per-lane barrier store: lazily materialised entries over a growing window. */
typedef struct lane_store {
    int min_id, n, cap;
    int32_t* state; uint64_t* step; uint8_t* valid;
} lane_store;

static int store_slot(lane_store* t, int id)
{
    if (t->n == 0) {
        t->cap = 64;
        t->state = (int32_t*)calloc(t->cap, sizeof(int32_t));
        t->step = (uint64_t*)calloc(t->cap, sizeof(uint64_t));
        t->valid = (uint8_t*)calloc(t->cap, 1);
        t->min_id = id - 32;
        t->n = t->cap;
    }
    while (id < t->min_id || id >= t->min_id + t->n) {
        int ncap = t->cap * 2;
        int shift = (id < t->min_id) ? ncap - t->cap : 0;
        int32_t* st = (int32_t*)calloc(ncap, sizeof(int32_t));
        uint64_t* sp = (uint64_t*)calloc(ncap, sizeof(uint64_t));
        uint8_t* va = (uint8_t*)calloc(ncap, 1);
        memcpy(st + shift, t->state, sizeof(int32_t) * t->n);
        memcpy(sp + shift, t->step, sizeof(uint64_t) * t->n);
        memcpy(va + shift, t->valid, t->n);
        free(t->state); free(t->step); free(t->valid);
        t->state = st; t->step = sp; t->valid = va;
        t->min_id -= shift;
        t->cap = ncap;
        t->n = ncap;
    }
    return id - t->min_id;
}

static void lane_store_entry(lane_store* t, lane_noise* n, const consts* c, const ldo_params* p,
                             int array, int id, int* slot)
{
    int k = store_slot(t, id);
    if (!t->valid[k]) {
        t->valid[k] = 1;
        t->state[k] = noise_barrier_init(n, c, p->same_initial_barrier, 0.0, array, id);
        t->step[k] = 0;
    }
    *slot = k;
}

/* whether barrier `id` has an entry (has been touched) */
static int store_has(const lane_store* t, int id)
{
    if (t->n == 0) return 0;
    const int k = id - t->min_id;
    return k >= 0 && k < t->n && t->valid[k];
}

/* Test hook: lane mode appends every barrier hand-over to this trace. */
static ldo_handover* g_trace;
static uint64_t g_trace_cap, g_trace_n;

void ldo_trace_handovers(ldo_handover* buf, uint64_t cap)
{
    g_trace = buf;
    g_trace_cap = buf ? cap : 0;
    g_trace_n = 0;
}

uint64_t ldo_trace_count(void) { return g_trace_n; }

/* the statistics' barrier lookups in lane mode: the lane's own store */
typedef struct lane_touch_ctx {
    lane_store* st;
    lane_noise* n;
    const consts* c;
    const ldo_params* p;
} lane_touch_ctx;

static int lane_touch(void* ctx, int lane, int id)
{
    lane_touch_ctx* tc = (lane_touch_ctx*)ctx;
    int k;
    (void)lane;
    lane_store_entry(&tc->st[0], tc->n, tc->c, tc->p, 0, id, &k);
    return tc->st[0].state[k];
}

static int lane_run(const ldo_params* p, ldo_lane* o, lane_noise* n, double* frames, uint64_t frames_cap,
                    ldo_capture* cap, int lane_in_cap)
{
    consts cc;
    make_consts(p, &cc);
    const consts* c = &cc;
    const int model = p->model;
    lstate L;
    lane_store st[2];
    capref cr;
    memset(&L, 0, sizeof(L));
    memset(st, 0, sizeof(st));
    memset(o, 0, sizeof(*o));
    cr.cap = cap; cr.lane = lane_in_cap; cr.W = 1; cr.ncg = p->n_cg;

    L.x = p->x0;
    lane_indices(&L, c, model, 1);
    if (p->randomize_state) L.state = noise_init_state_u(n, o) <= c->p0 ? 0.0 : 1.0;
    else L.state = p->initial_state;
    o->state0_instance = L.state;
    {
        int a = model == 0 ? 0 : (int)L.local, k;
        lane_store_entry(&st[a], n, c, p, a, (int)L.periodic, &k);
        st[a].state[k] = (int)L.state;
    }
    L.count[0] = L.count[1] = 0.0;
    lane_param(&L, c, model);
    L.v = p->randomize_v ? c->vsigma * noise_init_gauss(n, 0, o) : p->v0;
    o->v0_instance = L.v;
    lane_energy_force(&L, c, model);
    lane_touch_ctx tctx = {st, n, c, p};
    touch_fn touch = (model == 0) ? lane_touch : NULL;
    stats_init(p, o, &L, touch, &tctx, 0);
    if (p->overdamped) L.R = c->var * noise_init_gauss(n, 1, o);
    uint64_t nf = 0;
    lane_record(p, c, &L, frames, frames_cap, nf++, 1);

    uint64_t F = p->save_freq;
    uint64_t nsteps = (F >= 10) ? (p->nsteps / F) * F : p->nsteps;
    uint64_t last_period_step = 0, xc_first = 0, xc_last = 0;
    /* absorbing mode 2 stops on entering a well centre outside the two cells
     * {a, a+1} the start lies between (the first FPT2 event, Single.h:2445-2489) */
    const int cell_a = (L.well == L.periodic) ? (int)L.periodic : (int)L.periodic - 1;
    uint64_t iStep;
    for (iStep = 1; iStep <= nsteps; ++iStep) {
        double g, u = 1.0;
        noise_step(n, iStep, o, &g, c->tau_zero ? NULL : &u);
        if (p->overdamped) {
            double rf = c->var * g;
            L.x = fma(L.F, c->dtgm, L.x);
            L.x = L.x + rf;
            L.x = L.x + L.R;
            L.R = rf;
        } else {
            L.v = L.v + (c->minv * L.F) * c->halfdt;
            L.x = L.x + L.v * c->halfdt;
            L.v = c->expmgdt * L.v + c->var * g;
            L.x = L.x + L.v * c->halfdt;
        }
        lane_indices(&L, c, model, 0);
        if (!c->tau_zero) flip_decision(c, p->equal_probability, L.state, u, &L.state);
        int entering = model == 0 ? L.ent_period : L.ent_local;
        if (entering) {
            int ap = model == 0 ? 0 : (int)L.local_prev, ac = model == 0 ? 0 : (int)L.local, kp, kc;
            lane_store_entry(&st[ap], n, c, p, ap, (int)L.periodic_prev, &kp);
            st[ap].state[kp] = (int)L.state;
            st[ap].step[kp] = iStep;
            const int first_touch = !store_has(&st[ac], (int)L.periodic);
            lane_store_entry(&st[ac], n, c, p, ac, (int)L.periodic, &kc);
            double prev_state = (double)st[ac].state[kc];
            ldo_handover* tr = NULL;
            if (g_trace && g_trace_n < g_trace_cap) {
                tr = &g_trace[g_trace_n];
                tr->step = iStep;
                tr->x = L.x;
                tr->left_array = ap;
                tr->left_id = (int)L.periodic_prev;
                tr->left_state = (int)L.state;
                tr->entered_array = ac;
                tr->entered_id = (int)L.periodic;
                tr->stored_state = (int)prev_state;
                tr->first_touch = first_touch;
                tr->stored_step = st[ac].step[kc];
            }
            if (g_trace) g_trace_n++;
            if (c->tau_zero) {
                L.state = prev_state;
            } else {
                double lap = (double)(iStep - st[ac].step[kc]);
                double pc;
                if (p->equal_probability) pc = 0.5 * (1.0 - exp(c->c_lap * lap));
                else if (prev_state == 0.0) pc = c->K / (1.0 + c->K) * (1.0 - exp(c->c_lap * lap));
                else pc = 1.0 / (1.0 + c->K) * (1.0 - exp(c->c_lap * lap));
                double uc = noise_catch(n, o);
                L.state = (uc < pc) ? ((prev_state == 1.0) ? 0.0 : 1.0) : prev_state;
            }
            if (tr) tr->new_state = (int)L.state;
        }
        lane_param(&L, c, model);
        lane_energy_force(&L, c, model);
        if (!p->overdamped) L.v = L.v + (c->minv * L.F) * c->halfdt;
        stats_update(p, o, &L, iStep, touch, &tctx, 0, cap ? &cr : NULL);
        if (iStep % F == 0) lane_record(p, c, &L, frames, frames_cap, nf++, 1);
        /* double barrier: crossings of the intermediate-well bottom xc (a local
         * barrier change that is not a periodic-cell change) split the
         * transition path into its barrier-1 and barrier-2 parts */
        if (model == 1 && L.ent_local && !L.ent_period) {
            xc_last = iStep;
            if (!xc_first) xc_first = iStep;
        }
        if ((p->absorb == 1 && L.ent_well) ||
            (p->absorb == 2 && L.ent_period && ((int)L.periodic - cell_a < 0 || (int)L.periodic - cell_a > 1))) {
            o->absorbed_step = iStep;
            o->tpt_start_step = last_period_step;
            o->xc_first_step = xc_first;
            o->xc_last_step = xc_last;
            break;
        }
        if (L.ent_period) {
            last_period_step = iStep;
            xc_first = 0;
        }
        if (n->err) break;
    }
    o->steps_done = iStep > nsteps ? nsteps : iStep;
    lane_export(o, &L);
    o->error = n->err;
    for (int a = 0; a < 2; ++a) { free(st[a].state); free(st[a].step); free(st[a].valid); }
    return n->err;
}

int ldo_lane_run_replay(const ldo_params* p, ldo_lane* out,
                        const double* gauss, uint64_t n_gauss,
                        const double* flip, uint64_t n_flip,
                        const double* catchup, uint64_t n_catch,
                        const int8_t* binit, int bmin, int bn,
                        double* frames_x, uint64_t frames_cap, ldo_capture* cap, int lane_in_cap)
{
    lane_noise n;
    memset(&n, 0, sizeof(n));
    n.kind = 0;
    n.gauss = gauss; n.n_gauss = n_gauss;
    n.flip = flip; n.n_flip = n_flip;
    n.catchup = catchup; n.n_catch = n_catch;
    n.binit = binit; n.bmin = bmin; n.bn = bn;
    return lane_run(p, out, &n, frames_x, frames_cap, cap, lane_in_cap);
}

int ldo_lane_run_philox(const ldo_params* p, ldo_lane* out, uint64_t seed, uint32_t traj,
                        uint64_t step_begin, uint64_t step_end,
                        double* frames_x, uint64_t frames_cap)
{
    (void)step_begin; (void)step_end;
    lane_noise n;
    memset(&n, 0, sizeof(n));
    n.kind = 1;
    n.key[0] = (uint32_t)seed;
    n.key[1] = (uint32_t)(seed >> 32);
    n.traj = traj;
    return lane_run(p, out, &n, frames_x, frames_cap, NULL, 0);
}

/* ------------------------------------------------------------------------- */
/* Runner post-processing                                                      */
/* ------------------------------------------------------------------------- */
/* MOLUTILITY::GenerateLog10Scale (MOLUTILITY.cpp:59-73) */
uint64_t ldo_log10_scale(double timestep, uint64_t n_frames, uint64_t minimal, uint64_t per_order,
                         int include_zero, uint64_t* out, uint64_t cap)
{
    const double step = 1.0 / (double)per_order;
    double max_time = timestep * (double)(n_frames - minimal);
    double t0 = log10(timestep), t1 = log10(max_time);
    uint64_t n = 0;
    if (include_zero) { if (n < cap) out[n] = 0; n++; }
    for (double t = t0; t <= t1; t += step) {
        uint64_t iv = (uint64_t)round(pow(10.0, t) / timestep);
        if (n == 0 || (n - 1 < cap && out[n - 1] != iv)) {
            if (n < cap) out[n] = iv;
            n++;
        }
    }
    return n;
}

/* MeanSquareDisplacementBruteForce (Single.h:24-45) */
void ldo_msd_brute(const double* x, uint64_t n, const uint64_t* lags, uint64_t nlags, double* out)
{
    for (uint64_t ii = 0; ii < nlags; ++ii) {
        int64_t i = (int64_t)lags[ii];
        double acc = 0.0;
        if ((int64_t)n - i > 0) {
            for (int64_t k = 0; k < (int64_t)n - i; ++k) {
                double v = x[k + i] - x[k];
                acc = fma(v, v, acc);   /* `result += value * value`, contracted */
            }
            acc /= (double)((int64_t)n - i);
        }
        out[ii] = acc;
    }
}

/* STATISTICS::AVERAGE / RMSD (STATISTICS.cpp:100-106, 122-134) */
double ldo_average(const double* v, uint64_t n)
{
    if (!n) return 0.0;
    double s = 0.0;
    for (uint64_t i = 0; i < n; ++i) s += v[i];
    return s / (double)n;
}

double ldo_rmsd(const double* v, uint64_t n)
{
    if (!n) return 0.0;
    double a = ldo_average(v, n), r = 0.0;
    for (uint64_t i = 0; i < n; ++i) r = fma(v[i] - a, v[i] - a, r);  /* contracted, as above */
    return sqrt(r / (double)n);
}
