"""The named workloads of the batched engine, as lists of parameter sets.

single_barrier   first passage over the single barrier: heights 3.6 Eb / Eb
                 (sample_input/CuIDB_single geometry: L = 5.8, w = 0.65,
                 X0 = -2.9, the well centre, at rest), dt = 1e-4,
                 tau = 1e-5, Philox seed 1000 + Eb index, Eb in {2, 3, 4, 5},
                 overdamped sets first, then underdamped; step cap 1e8.
double_barrier   the same Eb values over the sample_input/CuIDB_double
                 geometry (barriers at -1.42 / -0.56, half widths 0.70 /
                 0.65, heights 1.7 Eb / Eb and 3.6 Eb / Eb), from the
                 outer-left well centre to the first non-adjacent well centre.
ragged           256 parameter sets drawn from seed 7 (sweep.ragged_sets).
serving_requests a Poisson stream of requests (20 per second, seed 11; single
                 or double barrier, under- or overdamped, Eb ~ U[2, 6.5] with
                 ratio 3.6, 1e3 .. 1e6 trajectories log-uniform) at
                 dt = 1e-3, tau = 1e-4 and a cap of 1e5 steps per trajectory.

The serving cap is a load choice.  The stream brings ~2.9e6 trajectories per
second (mean count (1e6 - 1e3) / ln 1000 = 1.45e5 per request); at 1e5 steps
each that is at most 2.9e11 trajectory-steps per second, below what the FP32
engine integrates (~6.5e11 per second), so the queue reaches a steady state.
At the sweep's 1e8-step cap the same stream would bring ~440 times the
engine's rate and latencies would grow without bound.
"""
from .sweep import double_barrier_set, ragged_sets, sweep_set

EBS = (2.0, 3.0, 4.0, 5.0)
SWEEP_CAP = 10 ** 8

SERVING_RATE = 20.0
SERVING_SEED = 11
SERVING_CAP = 100000
SERVING_TIMESTEP = 1e-3
SERVING_TAU = 1e-4


def single_barrier(n_traj=1 << 20, ebs=EBS, max_steps=SWEEP_CAP, overdamped=(True, False)):
    out = []
    for over in overdamped:
        for eb in ebs:
            out.append(sweep_set(model=0, overdamped=over, eb=eb, ratio=3.6, n_traj=n_traj, max_steps=max_steps,
                                 seed=1000 + EBS.index(eb), x0=-2.9, periodic_length=5.8, half_width=(0.65, 0.65),
                                 timestep=1e-4, tau=1e-5, randomize_v=False, v0=0.0))
    return out


def double_barrier(n_traj=1 << 22, ebs=EBS, max_steps=SWEEP_CAP, overdamped=(True, False)):
    return [double_barrier_set(eb, overdamped=over, n_traj=n_traj, max_steps=max_steps, seed=1000 + EBS.index(eb),
                               timestep=1e-4, tau=1e-5, randomize_v=False, v0=0.0)
            for over in overdamped for eb in ebs]


def ragged(max_steps=200000, seed=7):
    return ragged_sets(seed=seed, max_steps=max_steps)


def serving_requests(seconds=300.0, rate=SERVING_RATE, seed=SERVING_SEED, max_steps=SERVING_CAP):
    from .server import poisson_requests
    return poisson_requests(rate, seconds, seed, max_steps=max_steps, timestep=SERVING_TIMESTEP, tau=SERVING_TAU)
