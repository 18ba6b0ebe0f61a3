/*
 * ldpcp_oracle.h -- CPU restatement of the reference Langevin hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product path (langevindynamicsforpcp_amd) never calls it.
 *
 * Two modes are restated:
 *   rank mode  -- W lanes in lock-step sharing three noise banks, exactly as
 *                 one MPI rank of the reference (langevinDynamics*Barrier.h,
 *                 RANDOMNUMBER.cpp, DataLocator.h);
 *   lane mode  -- one trajectory with its own role-tagged noise (replayed
 *                 arrays or Philox4x32-10), which is what the GPU kernels do.
 */
#ifndef LDPCP_ORACLE_H
#define LDPCP_ORACLE_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDO_MAX_CG 16
#define LDO_MAX_W 16

typedef struct ldo_params {
    int model;                 /* 0 = single barrier, 1 = double barrier */
    int overdamped;
    double x0, v0, gamma, kb, temperature, mass;
    int randomize_v, randomize_state, same_initial_barrier;
    double periodic_length;
    double height[2][2];       /* [barrier][state]; single uses [0][*]      */
    double half_width[2];      /* single uses [0]                           */
    double position[2];        /* double only                               */
    double timestep;
    uint64_t nsteps;
    uint64_t save_freq;
    double tau;
    double initial_state;
    int equal_probability;
    double eq_constant;
    int b_waiting, b_fpt, b_tpt, b_fptcg;
    int n_cg;
    int cg_size[LDO_MAX_CG];
    int save_all;              /* WRITEX: also keep v, f, pe, ke, state, barrier frames */
    int absorb;                /* lane mode: 1 stop at the first barrier-top crossing,
                                  2 at the first non-adjacent well centre */
} ldo_params;

/* Per-trajectory result; the GPU kernels keep the same fields. */
typedef struct ldo_lane {
    double x, v, f, e, r_prev, state, barrier;
    double state_count[2];
    double well, periodic, xp, local_barrier;
    double v0_instance, state0_instance;
    /* FPT1 (well centre -> barrier top) */
    double fpt_sum, fpt_last_tick; uint64_t fpt_n; int fpt_to_update;
    /* FPT2 (well centre -> non-adjacent well centre) */
    double fpt2_sum, fpt2_last_tick; uint64_t fpt2_n; int fpt2_to_update; int fpt2_idx[2];
    /* TPT1 / TPT2 */
    double tpt_sum, tpt_last_tick; uint64_t tpt_n; int tpt_to_update;
    double tpt2_sum, tpt2_last_tick; uint64_t tpt2_n; int tpt2_to_update; int tpt2_initial_well;
    /* FPTCG(n) */
    double cg_sum[LDO_MAX_CG], cg_last_tick[LDO_MAX_CG];
    uint64_t cg_n[LDO_MAX_CG]; int cg_to_update[LDO_MAX_CG]; int cg_initial_well[LDO_MAX_CG];
    /* waiting time: barrier ticks (in steps) and running moments of their gaps */
    double wt_tick_prev; int wt_passed;
    uint64_t wt_ticks; double wt_last; double wt_sum, wt_sum2, wt_max; uint64_t wt_n;
    /* absorbing mode */
    uint64_t absorbed_step;    /* 0 = not absorbed */
    uint64_t tpt_start_step;   /* last well-centre crossing before absorption */
    uint64_t xc_first_step;    /* double barrier: first crossing of the intermediate-well
                                  bottom xc after the last well-centre crossing (0 = none) */
    uint64_t xc_last_step;     /* double barrier: last crossing of xc before absorption */
    uint64_t steps_done;
    /* noise consumption (lane mode) */
    uint64_t n_gauss, n_flip, n_catch;
    int error;
    /* FPT1 barrier log (single barrier): states and ids of the two barriers
     * around the well the last FPT1 started from */
    int fpt_init_index[2];
    double fpt_init_state[2];
} ldo_lane;

/* Optional per-lane capture of the first `cap` values of every event list. */
typedef struct ldo_capture {
    uint64_t cap;
    double *fpt, *fpt2, *tpt, *tpt2, *ticks;  /* [lane][cap] */
    double *cg;                               /* [cg][lane][cap] */
    double *inst, *init;                      /* FPT1 barrier logs, [lane][cap] */
} ldo_capture;

/* Bank refill callback: write n doubles of stream `which` (0 = SIMD
 * Gaussian, 1 = SIMD uniform, 2 = scalar uniform) into buf. */
typedef void (*ldo_fill_fn)(void* ctx, int which, double* buf, uint64_t n);

typedef struct ldo_rank_state ldo_rank_state;

/* Optional recording of what each lane consumed, in the role-tagged form the
 * lane mode replays: flip[lane][k], catch[lane][k], gauss[lane][k] and the
 * state every barrier had when it was created, binit[b][lane][id - bmin]. */
typedef struct ldo_record {
    uint64_t cap;                     /* per-lane capacity of the streams */
    double *gauss, *flip, *catchup;   /* [lane][cap] */
    int bmin, bn;                     /* barrier id window recorded */
    int8_t *binit;                    /* [barrier array][lane][bn], -1 = never created */
} ldo_record;

ldo_rank_state* ldo_rank_create(const ldo_params* p, int W, ldo_fill_fn fill, void* ctx, uint64_t bank_size);
void ldo_rank_destroy(ldo_rank_state* s);
/* Runs one trial (Integrate) of the rank; frames are [frame][lane]. */
int ldo_rank_integrate(ldo_rank_state* s, ldo_lane* lanes, double* frames_x, uint64_t frames_cap,
                       uint64_t* n_frames, ldo_capture* cap, ldo_record* rec);
/* RNG progress (renew count, bank offset) per stream, as stored in .restart. */
void ldo_rank_rng_progress(const ldo_rank_state* s, int64_t renew[3], uint64_t offset[3]);
/* Barrier table: size, min id and [id][lane] states / time steps. */
int ldo_rank_barriers(const ldo_rank_state* s, int* min_id, int32_t* states, uint64_t* steps, int cap, int which);

/* Lane mode with replayed role-tagged noise. */
int ldo_lane_run_replay(const ldo_params* p, ldo_lane* out,
                        const double* gauss, uint64_t n_gauss,
                        const double* flip, uint64_t n_flip,
                        const double* catchup, uint64_t n_catch,
                        const int8_t* binit, int bmin, int bn,
                        double* frames_x, uint64_t frames_cap, ldo_capture* cap, int lane_in_cap);
/* Lane mode with counter-based Philox4x32-10 noise keyed by (seed, traj). */
int ldo_lane_run_philox(const ldo_params* p, ldo_lane* out, uint64_t seed, uint32_t traj,
                        uint64_t step_begin, uint64_t step_end,
                        double* frames_x, uint64_t frames_cap);

void ldo_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Test hook: while a trace buffer is set, lane mode appends one record per
 * barrier hand-over (Single.h:2025-2068; Double.h:2466-2510); count() is the
 * number of hand-overs since the buffer was set, also past cap. */
typedef struct ldo_handover {
    uint64_t step;              /* iStep of the hand-over */
    double x;                   /* position after the step's move */
    int32_t left_array, left_id, left_state;       /* barrier left and the state it keeps */
    int32_t entered_array, entered_id, stored_state; /* barrier entered and its stored state */
    int32_t first_touch;        /* 1: the entered barrier was created by this hand-over */
    int32_t new_state;          /* state after the catch-up flip */
    uint64_t stored_step;       /* step the entered barrier was last left (0: never) */
} ldo_handover;
void ldo_trace_handovers(ldo_handover* buf, uint64_t cap);
uint64_t ldo_trace_count(void);

/* The Box-Muller logarithm and sine/cosine of 2 pi u shared with the GPU
 * (ldpcp_gauss.h), exported so the tests can pin them against a correctly
 * rounded reference. */
double ldo_log(double x);
/* The double barrier's intermediate-well bottom (GetXc, Double.h:1722-1732) as
 * the constants hold it. */
void ldo_double_apex(const ldo_params* p, double* xc, double* yc);
void ldo_sincos_turn(double u, double* s, double* c);
uint64_t ldo_flip_threshold(double p);

/* Post-processing helpers restated from the runner. */
uint64_t ldo_log10_scale(double timestep, uint64_t n_frames, uint64_t minimal, uint64_t per_order,
                         int include_zero, uint64_t* out, uint64_t cap);
void ldo_msd_brute(const double* x, uint64_t n, const uint64_t* lags, uint64_t nlags, double* out);
double ldo_average(const double* v, uint64_t n);
double ldo_rmsd(const double* v, uint64_t n);

#ifdef __cplusplus
}
#endif
#endif
