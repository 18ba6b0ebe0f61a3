"""The SIMD CPU baseline (oracle/cpu_simd.c) against the scalar oracle.

The baseline restates the reference's W-lane SIMD step loop for the absorbing
first passage; bench.py times it beside the GPU.  With the oracle's Philox
noise and operation order it must reproduce every trajectory's absorption
step and transition-path start exactly, for both builds (AVX-512, 8 lanes;
AVX2, 4 lanes) wherever the host runs them, and however the trajectories are
spread over threads and lanes.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol

FAST = dict(x0=-1.05, periodic_length=2.1, half_width=[0.6, 0.6], timestep=1e-3)


def params(over, eb=1.0, ratio=3.0, cap=20000, tau=1e-3, equal=1, k=1.0, randomize_v=None, randomize_state=1,
           same=0, **kw):
    geo = dict(FAST)
    geo.update(kw)
    return ol.make_params(model=0, overdamped=int(over), heights=[[ratio * eb, eb], [ratio * eb, eb]],
                          tau=tau, nsteps=cap, save_freq=cap, randomize_v=int(not over if randomize_v is None
                                                                                 else randomize_v),
                          v0=0.0, randomize_state=randomize_state, same_initial_barrier=same,
                          equal_probability=equal, eq_constant=k, b_waiting=0, b_fpt=0, b_tpt=0, b_fptcg=0,
                          absorb=1, **geo)


def oracle_paths(oracle, p, seed, n):
    ab = np.zeros(n, dtype=np.uint64)
    tp = np.zeros(n, dtype=np.uint64)
    for t in range(n):
        out = ol.Lane()
        oracle.ldo_lane_run_philox(C.byref(p), C.byref(out), seed, t, 0, 0, None, 0)
        ab[t], tp[t] = out.absorbed_step, out.tpt_start_step
    return ab, tp


def widths():
    w = ol.host_simd_width()
    return [x for x in (8, 4) if x <= w]


CASES = {
    "over": dict(over=True),
    "under": dict(over=False),
    "under_at_rest": dict(over=False, randomize_v=0),
    "asymmetric": dict(over=True, equal=0, k=2.5),
    "tau_zero": dict(over=True, tau=0.0),
    "same_initial": dict(over=False, same=1, randomize_state=0),
    "odd_cap": dict(over=True, eb=2.0, cap=4001),
    "reference_geometry": dict(over=True, eb=0.5, ratio=3.6, x0=-2.9, periodic_length=5.8,
                               half_width=[0.65, 0.65], timestep=1e-4, tau=1e-5, cap=40000),
}


@pytest.mark.parametrize("width", widths())
@pytest.mark.parametrize("case", sorted(CASES))
def test_simd_baseline_equals_oracle(oracle, width, case):
    lib = ol.load_simd(width)
    assert lib.ldo_simd_width() == width
    p = params(**CASES[case])
    n = 96
    ref_ab, ref_tp = oracle_paths(oracle, p, 41, n)
    ab, tp, steps, done, _ = ol.simd_absorb(lib, [p], [41], n_traj=n, threads=3)
    assert done == n
    bad = [(t, int(ab[0, t]), int(ref_ab[t])) for t in range(n)
           if ab[0, t] != ref_ab[t] or tp[0, t] != ref_tp[t]]
    assert not bad, f"trajectories differ (t, simd step, oracle step): {bad[:8]}"
    cap = p.nsteps
    assert steps == int(sum(int(a) if a else cap for a in ref_ab))
    if case not in ("odd_cap", "reference_geometry"):
        assert np.count_nonzero(ref_ab) > n // 2  # the case exercises absorption
    if case == "odd_cap":
        assert np.count_nonzero(ref_ab == 0) > 0  # and the cap


@pytest.mark.parametrize("width", widths())
def test_simd_baseline_several_sets_and_threads(oracle, width):
    """Sets of different kinds in one call, more threads than sets: each set's
    trajectories still equal the oracle's."""
    lib = ol.load_simd(width)
    ps = [params(over=True, eb=1.0), params(over=False, eb=1.5), params(over=True, eb=2.0, cap=3001)]
    seeds = [7, 8, 2 ** 33 + 9]
    n = 40
    ab, tp, _, done, _ = ol.simd_absorb(lib, ps, seeds, n_traj=n, threads=7)
    assert done == 3 * n
    for s, (p, seed) in enumerate(zip(ps, seeds)):
        ref_ab, ref_tp = oracle_paths(oracle, p, seed, n)
        assert np.array_equal(ab[s], ref_ab), s
        assert np.array_equal(tp[s], ref_tp), s


@pytest.mark.parametrize("width", widths())
def test_simd_baseline_timed_mode_counts_steps(width):
    lib = ol.load_simd(width)
    p = params(over=True, eb=2.0, cap=2000)
    _, _, steps, done, wall = ol.simd_absorb(lib, [p], [3], seconds=0.3, threads=2, want_paths=False)
    assert 0.3 <= wall < 5.0
    assert done > 0 and steps >= done
