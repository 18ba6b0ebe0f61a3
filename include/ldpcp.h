/*
 * ldpcp.h -- C ABI of the MI355X Langevin hot-path library (libldpcp.so).
 *
 * The reference has no FFI: its runner drives the integrator object
 * LANGEVINEQUATION::LangevinEquationSIMDMultiple through C++ calls.  Each
 * entry point below names the reference interface it replaces.  All entry
 * points return 0 on success and a negative code on failure; the message is
 * then available from ldpcp_last_error() (thread-local).  Outputs are caller
 * allocated.  One handle serves one host thread; handles are independent and
 * may run concurrently on separate HIP streams.
 */
#ifndef LDPCP_H
#define LDPCP_H
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define LDPCP_MAX_CG 16

#define LDPCP_OK 0
#define LDPCP_EINVAL (-1)
#define LDPCP_EDEVICE (-2)
#define LDPCP_ENOMEM (-3)
#define LDPCP_ERANGE (-4)   /* a trajectory left its barrier window or replay stream */

/* Parameters of one reference run (the option table of
 * langevinDynamicsSingleBarrier.cpp:65-108 / ...DoubleBarrier.cpp:80-85 as
 * copied into the integrator by LangevinEquationRunnerMPI::Run,
 * langevinDynamicsSingleBarrier.h:3249-3301). */
typedef struct ldpcp_run_params {
    int32_t model;                 /* 0 single barrier, 1 double barrier */
    int32_t overdamped;
    double x0, v0, gamma, kb, temperature, mass;
    int32_t randomize_v, randomize_state, same_initial_barrier;
    double periodic_length;
    double height[2][2];           /* [barrier][state] */
    double half_width[2];
    double position[2];
    double timestep;
    uint64_t nsteps;               /* NUMBEROFSTEPS */
    uint64_t save_freq;            /* FREQSAVE */
    double tau;
    double initial_state;
    int32_t equal_probability;
    double eq_constant;
    int32_t b_waiting, b_fpt, b_tpt, b_fptcg;
    int32_t n_cg;
    int32_t cg_size[LDPCP_MAX_CG];
    int32_t save_all;              /* WRITEX: keep v, f, pe, ke, state, barrier frames */
    uint64_t seed;                 /* Philox key (free-running noise) */
    int32_t barrier_window;        /* barrier entries kept per trajectory and barrier array */
    uint64_t tick_capacity;        /* barrier-crossing ticks kept per trajectory (0 = none) */
    uint64_t event_capacity;       /* entries kept per event list per trajectory (0 = none): the
                                      FPT / FPT2 / TPT / TPT2 / FPTCG lists and the FPT1 barrier
                                      logs the reference keeps for its restart files */
} ldpcp_run_params;

/* Event lists kept per trajectory when event_capacity > 0, in this order
 * (TrajectoryData, langevinDynamicsSingleBarrier.h:160-235): FPT1, FPT2,
 * TPT1, TPT2, the FPT1 logs of the barrier crossed (m_InstantBarrierList) and
 * of the barrier left (m_InitialBarrierList), then FPTCG for each size. */
#define LDPCP_LIST_FPT 0
#define LDPCP_LIST_FPT2 1
#define LDPCP_LIST_TPT 2
#define LDPCP_LIST_TPT2 3
#define LDPCP_LIST_INSTANT 4
#define LDPCP_LIST_INITIAL 5
#define LDPCP_LIST_CG0 6

/* Per-trajectory results: the values the runner reads through
 * GetFirstPassageTimeAverage{,2}, GetTransitionPathTimeAverage{,2},
 * GetFirstPassageTimeCGAverage2, GetStateCount, GetV0Instance,
 * GetBarrierState0Instance and GetBarrierTick
 * (langevinDynamicsSingleBarrier.h:2620-2769), kept as running sums and
 * counts so the device never stores event lists. */
typedef struct ldpcp_lane {
    double x, v, f, e, r_prev, state, barrier;
    double state_count[2];
    double well, periodic, xp, local_barrier;
    double v0_instance, state0_instance;
    double fpt_sum, fpt_last_tick; uint64_t fpt_n; int32_t fpt_to_update;
    double fpt2_sum, fpt2_last_tick; uint64_t fpt2_n; int32_t fpt2_to_update; int32_t fpt2_idx[2];
    double tpt_sum, tpt_last_tick; uint64_t tpt_n; int32_t tpt_to_update;
    double tpt2_sum, tpt2_last_tick; uint64_t tpt2_n; int32_t tpt2_to_update; int32_t tpt2_initial_well;
    double cg_sum[LDPCP_MAX_CG], cg_last_tick[LDPCP_MAX_CG];
    uint64_t cg_n[LDPCP_MAX_CG]; int32_t cg_to_update[LDPCP_MAX_CG]; int32_t cg_initial_well[LDPCP_MAX_CG];
    double wt_tick_prev; int32_t wt_passed;
    uint64_t wt_ticks; double wt_last; double wt_sum, wt_sum2, wt_max; uint64_t wt_n;
    uint64_t absorbed_step, tpt_start_step, steps_done;
    uint64_t n_gauss, n_flip, n_catch;
    int32_t error;
    /* FPT1 barrier log state (m_InitialBarrierIndex / m_InitialBarrierState,
     * Single.h:2364-2387): the barriers either side of the well the last FPT1
     * started from */
    int32_t fpt_init_index[2];
    double fpt_init_state[2];
} ldpcp_lane;

typedef struct ldpcp_run ldpcp_run;

/* Replaces `new LangevinEquationSIMDMultiple` / `Overdamped...` + the
 * parameter copy + Initialize(seed) (langevinDynamicsSingleBarrier.h:3231-3306,
 * :1276-1291).  One handle holds n_traj trajectories ("lanes"); the Philox
 * counter of trajectory i is keyed by traj_offset + i, so results do not
 * depend on how trajectories are split over handles. */
int ldpcp_run_create(const ldpcp_run_params* params, uint32_t n_traj, uint32_t traj_offset, ldpcp_run** out);

/* Replay mode (no reference counterpart: the reference's own MKL banks,
 * RANDOMNUMBER.cpp:157-340, are replaced by role-tagged per-trajectory
 * streams).  gauss and flip are [k][n_traj]; catchup is [n_traj][catch_cap];
 * binit is [barrier array][n_traj][bn] holding each barrier's state at
 * creation for ids bmin .. bmin+bn-1 (-1 = never created). */
int ldpcp_run_set_replay(ldpcp_run* h, const double* gauss, uint64_t n_gauss,
                         const double* flip, uint64_t n_flip,
                         const double* catchup, uint64_t catch_cap,
                         const int8_t* binit, int32_t bmin, int32_t bn);

/* Replaces Integrate(iTrial, bTimer) (langevinDynamicsSingleBarrier.h:1386-1519):
 * step 0, then NSTEPS steps (trailing NSTEPS mod FREQSAVE steps dropped when
 * FREQSAVE >= 10), frames every FREQSAVE steps, statistics every step. */
int ldpcp_run_integrate(ldpcp_run* h, uint64_t steps_per_launch);

/* ldpcp_run_integrate in two halves: ldpcp_run_launch enqueues the run on the
 * handle's own stream and returns; ldpcp_run_wait waits for it and reports
 * per-trajectory errors.  Handles launched one after another run concurrently
 * on the GPU (the reference's MPI ranks integrate at the same time,
 * langevinDynamicsSingleBarrier.h:3429-3446).  A run is split into launches of
 * up to steps_per_launch (0: 2^20) steps; ldpcp_run_launch returns once they
 * are enqueued, waiting while the stream's hardware queue is full, so long runs
 * of several handles are launched from one thread per handle.  Reading results
 * waits first. */
int ldpcp_run_launch(ldpcp_run* h, uint64_t steps_per_launch);
int ldpcp_run_wait(ldpcp_run* h);

/* Per-trajectory results for trajectories [begin, begin+n). */
int ldpcp_run_lanes(ldpcp_run* h, uint32_t begin, uint32_t n, ldpcp_lane* out);

/* Replaces GetNumberOfFrameX and GetX/GetV/GetF/GetPE/GetKE/GetState/GetBarrier
 * (langevinDynamicsSingleBarrier.h:2629-2701).  field: 0 x, 1 v, 2 f, 3 pe,
 * 4 ke, 5 state, 6 barrier.  out is [n][n_frames]. */
uint64_t ldpcp_run_n_frames(const ldpcp_run* h);
int ldpcp_run_frames(ldpcp_run* h, int32_t field, uint32_t begin, uint32_t n, double* out);

/* Replaces GetBarrierTick (langevinDynamicsSingleBarrier.h:2702-2704): the
 * waiting-time barrier ticks (in steps) of trajectories [begin, begin+n),
 * out is [n][tick_capacity]; the count per trajectory is ldpcp_lane.wt_ticks
 * (a count above tick_capacity means the list was truncated). */
int ldpcp_run_ticks(ldpcp_run* h, uint32_t begin, uint32_t n, double* out);

/* Replaces MeanSquareDisplacementBruteForce over each lane's saved x frames
 * (langevinDynamicsSingleBarrier.h:24-45, called at :3454).  out is
 * [n_traj][nlags]: the per-lane, time-origin-averaged MSD at each lag (in
 * frames), in the reference's summation order. */
int ldpcp_run_msd(ldpcp_run* h, const uint64_t* lags, uint32_t nlags, double* out);

/* No counterpart in the reference: ldpcp_run_msd for callers that cannot
 * allocate device memory (code under HIP graph capture, tight serving loops,
 * harnesses that forbid hipMalloc).  Same arguments and results; the lags and
 * results stage through a fixed 8 MiB device array instead of two allocations,
 * so n_traj * nlags must not exceed 983040 (else LDPCP_EINVAL).  The array is
 * shared by every call on a device: calls that overlap on one device race.
 * The handle must not have a launched run pending (wait for it first). */
int ldpcp_run_msd_noalloc(ldpcp_run* h, const uint64_t* lags, uint32_t nlags, double* out);

/* Replaces UserData::m_bInitialValue (langevinDynamicsSingleBarrier.h:1378-1380,
 * used by BarrierStatus::Initialize :80-84): with UNIFORMINITALSTATES, barriers
 * created during the run take this state per trajectory ([n_traj], 0 or 1).  It
 * is 0 for a new handle; the reference's runner carries each lane's initial
 * state into the next trial, and a restart carries it from the file. */
int ldpcp_run_set_barrier_default(ldpcp_run* h, const double* state);

/* Checkpoints.  Replace TrajectoryData::WriteRestartFile / ReadRestartFile
 * (langevinDynamicsSingleBarrier.h:385-576, 577-820; the file-level
 * WriteRestartFile / ReadRestartFile :1048-1206 and the resumed Integrate
 * :1397-1400, 1420) with an image of the handle's device state.  The image is
 * ldpcp_checkpoint_header followed by
 *   ldpcp_lane     lanes[n_traj]                      (what ldpcp_run_lanes returns)
 *   int32_t        window_base[n_traj]                (padded to 8 bytes)
 *   uint64_t       window[n_traj][2][barrier_window]  (bit 63 created, bit 62 state,
 *                                                      bits 0-61 step of the last hand-over)
 *   double         frames[n_fields][n_frames][n_traj] (x; with WRITEX v, f, pe, ke, state, barrier)
 *   double         ticks[n_traj][tick_capacity]
 *   double         events: for each trajectory, for each of the n_lists lists,
 *                  its first min(count, event_capacity) entries (counts in ldpcp_lane:
 *                  fpt_n for FPT1 and both barrier logs, fpt2_n, tpt_n, tpt2_n, cg_n[j])
 * Event lists are also held as running sums and counts (ldpcp_lane), and the noise
 * position is the step count itself (the Philox counter), so a resumed run
 * continues bit for bit: NSTEPS = a + b in one run equals a run of a steps,
 * a checkpoint, and a resumed run of b steps. */
#define LDPCP_CHECKPOINT_MAGIC 0x324B43504350444Cull /* "LDPCPCK2" */
typedef struct ldpcp_checkpoint_header {
    uint64_t magic;
    uint32_t header_bytes, lane_bytes;  /* sizeof of the two structs, as written */
    uint32_t n_traj, traj_offset;
    int32_t model, overdamped;
    int32_t barrier_window, n_fields;
    uint64_t seed;
    uint64_t steps_done;                /* step clock the next run continues from: the runs' NUMBEROFSTEPS
                                           total, which exceeds the last integrated step by the trailing
                                           NUMBEROFSTEPS mod FREQSAVE steps a run drops when FREQSAVE >= 10
                                           (langevinDynamicsSingleBarrier.h:1397-1420) */
    uint64_t n_frames;                  /* frames recorded: 1 + NUMBEROFSTEPS / FREQSAVE per run, summed */
    uint64_t save_freq;
    uint64_t tick_capacity;
    double timestep;
    uint64_t bytes;                     /* size of the whole image */
    uint64_t event_capacity;
    uint32_t n_lists;                   /* event lists per trajectory (6 + FPTCG sizes, or 0) */
    uint32_t events_complete;           /* 1: the lists hold every event since step 0 */
} ldpcp_checkpoint_header;

/* Number of events list `list` holds for a trajectory (entries kept:
 * the smaller of this and event_capacity). */
static inline uint64_t ldpcp_list_count(const ldpcp_lane* l, int list)
{
    switch (list) {
    case LDPCP_LIST_FPT:
    case LDPCP_LIST_INSTANT:
    case LDPCP_LIST_INITIAL: return l->fpt_n;
    case LDPCP_LIST_FPT2: return l->fpt2_n;
    case LDPCP_LIST_TPT: return l->tpt_n;
    case LDPCP_LIST_TPT2: return l->tpt2_n;
    default: return l->cg_n[list - LDPCP_LIST_CG0];
    }
}

/* Size of the image ldpcp_run_checkpoint writes for this handle now. */
int ldpcp_run_checkpoint_size(ldpcp_run* h, uint64_t* bytes);
/* Writes the image after ldpcp_run_integrate (or after ldpcp_run_resume). */
int ldpcp_run_checkpoint(ldpcp_run* h, void* buf, uint64_t bytes);
/* Loads an image into a handle created with the same parameters, trajectory
 * count and offset (NSTEPS may differ: it is the number of further steps).  The
 * next ldpcp_run_integrate continues the image's step clock (steps_done + 1 on),
 * recording frames every FREQSAVE steps of its own after the image's;
 * ldpcp_run_n_frames then counts both.  Such a handle integrates once. */
int ldpcp_run_resume(ldpcp_run* h, const void* buf, uint64_t bytes);

/* Replaces the barrier table the reference's restart file lists
 * (FiniteInifiniteIntegerTickArray1D::Write, DataLocator.h:163-169): the
 * stored state (0/1) and last hand-over step of barriers lo .. hi of `array`
 * for every trajectory, out as [n_traj][hi - lo + 1].  Barriers a trajectory
 * has not touched are created with their step-0 state, as the reference's
 * table lookups create them. */
int ldpcp_run_barrier_table(ldpcp_run* h, int32_t array, int32_t lo, int32_t hi, int32_t* states,
                            uint64_t* steps);

/* Replaces SAFE_DELETE(leOMP) (langevinDynamicsSingleBarrier.h:3972). */
int ldpcp_run_destroy(ldpcp_run* h);

/* ------------------------------------------------------------------------- */
/* First-passage engine: absorbing runs over ragged parameter sets.             */
/* No reference counterpart (the reference runs one parameter file per MPI     */
/* job, langevinDynamicsSingleBarrier.h:3429-3589); it is the batched form of   */
/* the same integrator, stopped at the first passage event, with FPT/TPT/MSD    */
/* statistics reduced on the device.  Two front ends share it: a static sweep   */
/* (all sets admitted at once) and a request server (continuous batching).      */
/* ------------------------------------------------------------------------- */
#define LDPCP_HIST_BINS 256
#define LDPCP_MSD_LAGS 64

typedef struct ldpcp_sweep_set {
    int32_t model, overdamped;
    double x0, v0, gamma, kb, temperature, mass;
    int32_t randomize_v, randomize_state;
    double initial_state;
    double periodic_length;
    double height[2][2];
    double half_width[2];
    double position[2];
    double timestep, tau;
    int32_t equal_probability;
    double eq_constant;
    uint64_t n_traj;
    uint64_t max_steps;            /* step cap per trajectory (< 2^32) */
    uint64_t seed;                 /* Philox key */
} ldpcp_sweep_set;

/* Statistics of one parameter set.  The first passage event is the first
 * barrier-top crossing for the single barrier (the reference's first FPT1
 * sample) and the first arrival at a non-adjacent well centre for the double
 * barrier (its first FPT2 sample).  Times are in the set's time units.  The
 * duration sums are kept on the device as exact integer step sums and read
 * as dt * sum and dt^2 * sum of squares: bit-identical whatever the launch
 * shape or the other sets resident. */
typedef struct ldpcp_set_stats {
    uint64_t n_traj, n_absorbed, n_capped;
    double fpt_sum, fpt_sum2;      /* first-passage times of absorbed trajectories */
    double tpt_sum, tpt_sum2;      /* transition-path times: last well-centre crossing -> event */
    /* double barrier, per barrier: [0] last well-centre crossing -> first crossing
     * of the intermediate-well bottom xc after it; [1] last crossing of xc -> event */
    double tptb_sum[2], tptb_sum2[2];
    uint64_t traj_steps;           /* trajectory-steps integrated */
    uint64_t n_finished;           /* absorbed + capped so far */
    /* FPT histogram: 256 log10 bins over [hist_lo, hist_hi] (log10 of time);
     * absorption step k (FPT = dt k) falls in bin
     * min(255, floor((log10(dt k) - hist_lo) * 256 / (hist_hi - hist_lo))) */
    double hist_lo, hist_hi;
    uint64_t hist[LDPCP_HIST_BINS];
    uint64_t msd_step[LDPCP_MSD_LAGS]; /* lag steps (log-spaced in [1, max_steps]) */
    double msd_sum[LDPCP_MSD_LAGS];    /* sum of (x(lag) - x0)^2 over trajectories alive at the lag */
    uint64_t msd_n[LDPCP_MSD_LAGS];
    /* first absorption step of each histogram bin, 0 where no step up to the
     * cap reaches the bin: bin b holds the steps k with hist_step[b] <= k and
     * k < hist_step[b'] for the next b' > b with hist_step[b'] > hist_step[b]
     * (bins too narrow to hold an integer step repeat the next start) */
    uint64_t hist_step[LDPCP_HIST_BINS];
} ldpcp_set_stats;

/* Launch shape.  Zero fields take the defaults (block 256, 8 blocks per CU,
 * 4096 steps per launch, graph replay on). */
typedef struct ldpcp_engine_config {
    int32_t block;                 /* threads per workgroup: 64, 128 or 256 */
    int32_t blocks_per_cu;         /* persistent workgroups per compute unit */
    uint32_t steps_per_launch;
    int32_t use_graph;             /* 1: replay a captured hipGraph per launch, -1: plain launches */
} ldpcp_engine_config;

typedef struct ldpcp_sweep ldpcp_sweep;

/* precision: 64 = plain FP64 stepping (the reference arithmetic), 32 = FP32
 * stepping with a compensated (cell, offset) position and an integer clock. */
int ldpcp_sweep_create(const ldpcp_sweep_set* sets, int32_t nsets, int32_t precision, ldpcp_sweep** out);
int ldpcp_sweep_set_config(ldpcp_sweep* h, const ldpcp_engine_config* cfg);
int ldpcp_sweep_get_config(ldpcp_sweep* h, ldpcp_engine_config* cfg);
/* Autotuner: times candidate launch shapes on the sweep's own trajectories
 * (results do not depend on the shape) and keeps the fastest. */
int ldpcp_sweep_tune(ldpcp_sweep* h, ldpcp_engine_config* chosen);
/* One launch of steps_per_launch steps (0 = the configured value) over every
 * live trajectory; returns the number still live in *live. */
int ldpcp_sweep_launch(ldpcp_sweep* h, uint32_t steps_per_launch, uint64_t* live);
/* Launches until every trajectory is absorbed or capped. */
int ldpcp_sweep_run(ldpcp_sweep* h, uint32_t steps_per_launch);
int ldpcp_sweep_stats(ldpcp_sweep* h, int32_t set, ldpcp_set_stats* out);
/* Per trajectory of a set: absorption step (0 = not absorbed), last
 * well-centre crossing before it, and (double barrier) the first and last
 * crossings of xc used by tptb.  Any output but absorbed_step may be NULL. */
int ldpcp_sweep_fpt(ldpcp_sweep* h, int32_t set, uint64_t* absorbed_step, uint64_t* tpt_start, uint64_t* xc_first,
                    uint64_t* xc_last);
/* Total duration (ms) of the launches so far, timed with HIP events on the
 * handle's stream around each launch, and the launch count. */
int ldpcp_sweep_kernel_time(ldpcp_sweep* h, double* ms_total, uint64_t* launches);
int ldpcp_sweep_destroy(ldpcp_sweep* h);

/* ------------------------------------------------------------------------- */
/* Request server: continuous batching of sweep requests on one GPU.           */
/* Requests wait in FIFO order until their HBM footprint fits the budget, then */
/* join the resident trajectories between two launches; absorbed trajectories  */
/* are retired after every launch and a request completes when its last        */
/* trajectory finishes.  A request's results do not depend on what else is    */
/* resident.  One server per host thread.                                      */
/* ------------------------------------------------------------------------- */
typedef struct ldpcp_server ldpcp_server;

typedef struct ldpcp_server_config {
    uint64_t budget_bytes;         /* HBM for resident trajectory state + per-request statistics */
    int32_t precision;             /* 32 or 64 */
    int32_t max_requests;          /* resident requests at once (0 = 4096) */
    ldpcp_engine_config engine;
    int32_t keep_fpt;              /* 1: keep each completed request's per-trajectory lists (ldpcp_server_fpt) */
} ldpcp_server_config;

enum { LDPCP_REQ_QUEUED = 0, LDPCP_REQ_RUNNING = 1, LDPCP_REQ_DONE = 2 };

int ldpcp_server_create(const ldpcp_server_config* cfg, ldpcp_server** out);
/* HBM bytes a request occupies while resident. */
int ldpcp_server_footprint(const ldpcp_server* s, const ldpcp_sweep_set* set, uint64_t* bytes);
int ldpcp_server_submit(ldpcp_server* s, const ldpcp_sweep_set* set, uint64_t* request_id);
/* Admits what fits, runs up to n_launches launches (fewer if nothing is
 * live), retires finished requests.  Returns without waiting for the last
 * launch when wait == 0. */
int ldpcp_server_pump(ldpcp_server* s, uint32_t n_launches, int32_t wait);
int ldpcp_server_status(ldpcp_server* s, uint64_t request_id, int32_t* state);
/* Ids of requests completed since the last call (at most max). */
int ldpcp_server_completed(ldpcp_server* s, uint64_t* ids, int32_t max, int32_t* n);
/* Per trajectory of a completed request, in the layout of ldpcp_sweep_fpt
 * (n_traj entries each; the values equal those of the same set run alone by
 * a sweep).  Needs keep_fpt; call before ldpcp_server_result, which frees the
 * record with the lists.  xc_first / xc_last are zero for the single barrier. */
int ldpcp_server_fpt(ldpcp_server* s, uint64_t request_id, uint64_t* absorbed_step, uint64_t* tpt_start,
                     uint64_t* xc_first, uint64_t* xc_last);
/* Results of a completed request; frees its record. */
int ldpcp_server_result(ldpcp_server* s, uint64_t request_id, ldpcp_set_stats* out);
/* queued / resident requests, live trajectories, resident bytes, trajectory-steps done */
int ldpcp_server_info(ldpcp_server* s, uint64_t* queued, uint64_t* resident, uint64_t* live_traj,
                      uint64_t* resident_bytes, uint64_t* traj_steps);
int ldpcp_server_destroy(ldpcp_server* s);

/* Library-wide. */
const char* ldpcp_last_error(void);
int ldpcp_device_info(char* name, int32_t name_len, int32_t* n_cu, int64_t* hbm_bytes);

#ifdef __cplusplus
}
#endif
#endif
