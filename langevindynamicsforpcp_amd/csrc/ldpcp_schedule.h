/*
 * ldpcp_schedule.h -- step clock and frame schedule of one integrate call.
 *
 * A run continues the step clock of the run before it (Single.h:1397-1400,
 * 1420): it integrates steps base+1 .. base+nsteps_eff, where base is the
 * previous runs' NUMBEROFSTEPS total and nsteps_eff drops the trailing
 * NUMBEROFSTEPS mod FREQSAVE steps when FREQSAVE >= 10 (the reference's
 * `ngroup = m_NumberOfStep / m_SaveFreq` groups), and records a frame after
 * every FREQSAVE steps of its own (`Record(m_NumberOfStepDonePreviously +
 * iStepBase + m_SaveFreq - 1)`), i.e. at base + k FREQSAVE.  Its clock then
 * stands at base + NUMBEROFSTEPS: the dropped steps are skipped, not
 * integrated, and the next run's step numbers (event times dt*iStep, the noise
 * counters) start after them.
 *
 * Plain C++ for host and device: the kernels and a host test use the same
 * arithmetic.
 */
#ifndef LDPCP_SCHEDULE_H
#define LDPCP_SCHEDULE_H
#include <stdint.h>

#if defined(__HIPCC__)
#define LDPCP_SCHED_HD __host__ __device__ inline
#else
#define LDPCP_SCHED_HD static inline
#endif

/* Steps a run of nsteps integrates (FREQSAVE >= 10 drops the trailing partial group). */
LDPCP_SCHED_HD uint64_t ldpcp_steps_integrated(uint64_t nsteps, uint64_t save_freq)
{
    return save_freq >= 10 ? (nsteps / save_freq) * save_freq : nsteps;
}

/* The first frame after step s0 of a run that started at base (base <= s0):
 * its step, and its index among the handle's frames (frame0 = the index of the
 * run's first frame, i.e. the frames recorded before the run). */
LDPCP_SCHED_HD void ldpcp_next_frame(uint64_t base, uint64_t frame0, uint64_t save_freq, uint64_t s0,
                                     uint64_t* step, uint64_t* index)
{
    const uint64_t k = (s0 - base) / save_freq;
    *step = base + (k + 1) * save_freq;
    *index = frame0 + k;
}

#undef LDPCP_SCHED_HD
#endif
