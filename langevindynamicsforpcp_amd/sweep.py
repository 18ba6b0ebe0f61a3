"""Host side of the batched first-passage engine (ldpcp_sweep_* in include/ldpcp.h).

A sweep admits every parameter set at once and launches until every
trajectory is absorbed or capped.  Results per trajectory depend only on its
set, index and step count, never on the launch shape or on the other sets.
"""
import ctypes as C
import math

import numpy as np

from . import _native as N


def sweep_set(model=0, overdamped=True, eb=2.0, ratio=3.6, n_traj=1 << 20, max_steps=10 ** 8, seed=1000,
              x0=-2.9, periodic_length=5.8, half_width=(0.65, 0.65), position=(0.0, 0.0), timestep=1e-4,
              tau=1e-5, gamma=1.0, kb=1.0, temperature=1.0, mass=1.0, heights=None,
              randomize_state=True, randomize_v=False, v0=0.0, equal_probability=True, eq_constant=1.0):
    """A parameter set.  By default the barrier alternates between ratio*Eb and
    Eb (BARRIERHEIGHTS ratio*Eb Eb, as in sample_input/CuIDB_single)."""
    s = N.SweepSet()
    s.model = model
    s.overdamped = int(overdamped)
    s.x0, s.v0, s.gamma, s.kb, s.temperature, s.mass = x0, v0, gamma, kb, temperature, mass
    s.randomize_v = int(randomize_v)
    s.randomize_state = int(randomize_state)
    s.initial_state = 0.0
    s.periodic_length = periodic_length
    if heights is None:
        heights = [[ratio * eb, eb], [ratio * eb, eb]]
    for b in range(2):
        for k in range(2):
            s.height[b][k] = heights[b][k]
        s.half_width[b] = half_width[b]
        s.position[b] = position[b]
    s.timestep, s.tau = timestep, tau
    s.equal_probability = int(equal_probability)
    s.eq_constant = eq_constant
    s.n_traj, s.max_steps, s.seed = int(n_traj), int(max_steps), int(seed)
    return s


def double_barrier_set(eb, overdamped=True, n_traj=1 << 22, max_steps=10 ** 8, seed=1000, timestep=1e-4, tau=1e-5,
                       ratio0=1.7, ratio1=3.6, **kw):
    """sample_input/CuIDB_double geometry: L = 5.8, barriers at -1.42 / -0.56
    with half widths 0.70 / 0.65, barrier 0 alternating between ratio0*Eb and
    Eb, barrier 1 between ratio1*Eb and Eb, X0 = -2.9 (outer-well centre)."""
    return sweep_set(model=1, overdamped=overdamped, heights=[[ratio0 * eb, eb], [ratio1 * eb, eb]],
                     n_traj=n_traj, max_steps=max_steps, seed=seed, x0=-2.9, periodic_length=5.8,
                     half_width=(0.70, 0.65), position=(-1.42, -0.56), timestep=timestep, tau=tau, **kw)


def ragged_sets(seed=7, n_sets=256, max_steps=200000, n_lo=2 ** 10, n_hi=2 ** 18, timestep=1e-4):
    """A ragged batch of parameter sets: Eb ~ U[2, 6.5]; amplitude ratio 2.7
    (sample_input/CuOPTz geometry: L = 2.1, w = 0.6, X0 = -1.05) or 3.6
    (sample_input/CuIDB_single: L = 5.8, w = 0.65, X0 = -2.9); tau
    log-uniform in [1e-3, 1e2]; gamma log-uniform in [0.1, 10]; single or
    double barrier (CuIDB_double geometry) with p = 0.5; under- or overdamped
    with p = 0.5; trajectory counts log-uniform in [n_lo, n_hi]."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n_sets):
        eb = float(rng.uniform(2.0, 6.5))
        ratio = float(rng.choice([2.7, 3.6]))
        tau = float(math.exp(rng.uniform(math.log(1e-3), math.log(1e2))))
        gamma = float(math.exp(rng.uniform(math.log(0.1), math.log(10.0))))
        model = int(rng.integers(2))
        over = bool(rng.integers(2))
        n = int(round(math.exp(rng.uniform(math.log(n_lo), math.log(n_hi)))))
        sd = int(rng.integers(1, 2 ** 62))
        if model == 1:
            s = double_barrier_set(eb, overdamped=over, n_traj=n, max_steps=max_steps, seed=sd, timestep=timestep,
                                   tau=tau, ratio1=ratio, gamma=gamma, randomize_v=not over)
        elif ratio == 2.7:
            s = sweep_set(model=0, overdamped=over, eb=eb, ratio=ratio, n_traj=n, max_steps=max_steps, seed=sd,
                          x0=-1.05, periodic_length=2.1, half_width=(0.6, 0.6), timestep=timestep, tau=tau,
                          gamma=gamma, randomize_v=not over)
        else:
            s = sweep_set(model=0, overdamped=over, eb=eb, ratio=ratio, n_traj=n, max_steps=max_steps, seed=sd,
                          x0=-2.9, periodic_length=5.8, half_width=(0.65, 0.65), timestep=timestep, tau=tau,
                          gamma=gamma, randomize_v=not over)
        out.append(s)
    return out


class Sweep:
    """Absorbing first-passage runs over a list of parameter sets on the GPU."""

    def __init__(self, sets, precision=32):
        self.sets = list(sets)
        arr = (N.SweepSet * len(self.sets))(*self.sets)
        lib = N.load()
        h = C.c_void_p()
        N.check(lib.ldpcp_sweep_create(arr, len(self.sets), int(precision), C.byref(h)), "ldpcp_sweep_create")
        self._h = h
        self.precision = precision

    def config(self):
        c = N.EngineConfig()
        N.check(N.load().ldpcp_sweep_get_config(self._h, C.byref(c)), "ldpcp_sweep_get_config")
        return c

    def set_config(self, block=0, blocks_per_cu=0, steps_per_launch=0, use_graph=0):
        c = N.EngineConfig(block, blocks_per_cu, steps_per_launch, use_graph)
        N.check(N.load().ldpcp_sweep_set_config(self._h, C.byref(c)), "ldpcp_sweep_set_config")

    def tune(self):
        """Autotune block size, persistent workgroups per CU and steps per launch
        on this sweep's own trajectories; returns the chosen config."""
        c = N.EngineConfig()
        N.check(N.load().ldpcp_sweep_tune(self._h, C.byref(c)), "ldpcp_sweep_tune")
        return c

    def launch(self, steps_per_launch=0):
        live = C.c_uint64()
        N.check(N.load().ldpcp_sweep_launch(self._h, int(steps_per_launch), C.byref(live)), "ldpcp_sweep_launch")
        return live.value

    def run(self, steps_per_launch=0):
        N.check(N.load().ldpcp_sweep_run(self._h, int(steps_per_launch)), "ldpcp_sweep_run")

    def stats(self, i):
        st = N.SetStats()
        N.check(N.load().ldpcp_sweep_stats(self._h, int(i), C.byref(st)), "ldpcp_sweep_stats")
        return st

    def paths(self, i):
        """Per trajectory: absorption step (0 = capped or live), last well-centre
        crossing before it, first / last crossing of the intermediate-well bottom
        (double barrier)."""
        n = self.sets[i].n_traj
        out = [np.zeros(n, np.uint64) for _ in range(4)]
        ptr = [a.ctypes.data_as(C.POINTER(C.c_uint64)) for a in out]
        N.check(N.load().ldpcp_sweep_fpt(self._h, int(i), *ptr), "ldpcp_sweep_fpt")
        return out

    def fpt_steps(self, i):
        a, t, _, _ = self.paths(i)
        return a, t

    def kernel_time(self):
        ms = C.c_double()
        n = C.c_uint64()
        N.check(N.load().ldpcp_sweep_kernel_time(self._h, C.byref(ms), C.byref(n)), "ldpcp_sweep_kernel_time")
        return ms.value, n.value

    def mfpt(self, i):
        st = self.stats(i)
        if st.n_absorbed == 0:
            return math.nan, math.nan
        m = st.fpt_sum / st.n_absorbed
        var = max(st.fpt_sum2 / st.n_absorbed - m * m, 0.0)
        return m, math.sqrt(var / st.n_absorbed)

    def close(self):
        if self._h is not None:
            N.load().ldpcp_sweep_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass
