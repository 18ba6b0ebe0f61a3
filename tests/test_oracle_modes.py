"""The oracle's two modes agree, and the restatement obeys the mathematics.

* lane mode replaying the role-tagged streams a rank consumed reproduces the
  rank's trajectories bit for bit (this is the contract the GPU replay path
  is held to);
* Philox4x32-10 matches the published known-answer vectors;
* closed forms: flip probability, free-particle diffusion, velocity
  variance, stationary telegraph occupancy, continuity of the double-barrier
  potential at its segment ends.
"""
import ctypes as C
import math

import numpy as np
import pytest

import oracle_lib as ol
from streams import run_rank_recorded

CONFIGS = {
    "single_over": dict(model=0, overdamped=1, x0=-1.05, periodic_length=2.1, heights=[[3.0, 1.0], [3.0, 1.0]],
                        half_width=[0.6, 0.6], timestep=1e-3, tau=1e-3, randomize_v=1),
    "single_under": dict(model=0, overdamped=0, x0=-1.05, periodic_length=2.1, heights=[[3.0, 1.0], [3.0, 1.0]],
                         half_width=[0.6, 0.6], timestep=1e-3, tau=1e-2, randomize_v=1, gamma=2.0),
    "single_tau0": dict(model=0, overdamped=1, x0=-1.05, periodic_length=2.1, heights=[[3.0, 1.0], [3.0, 1.0]],
                        half_width=[0.6, 0.6], timestep=1e-3, tau=0.0, randomize_v=0),
    "single_eq_false": dict(model=0, overdamped=1, x0=-1.05, periodic_length=2.1,
                            heights=[[3.0, 1.0], [3.0, 1.0]], half_width=[0.6, 0.6], timestep=1e-3, tau=5e-3,
                            equal_probability=0, eq_constant=2.0),
    "single_same_init": dict(model=0, overdamped=1, x0=-1.05, periodic_length=2.1,
                             heights=[[3.0, 1.0], [3.0, 1.0]], half_width=[0.6, 0.6], timestep=1e-3, tau=1e-3,
                             same_initial_barrier=1, randomize_state=0, initial_state=1.0),
    "double_over": dict(model=1, overdamped=1, x0=-2.9, periodic_length=5.8,
                        heights=[[1.0, 0.5], [2.0, 0.5]], half_width=[0.70, 0.65], position=[-1.42, -0.56],
                        timestep=2e-3, tau=1e-3, randomize_v=1),
    "double_under": dict(model=1, overdamped=0, x0=-2.9, periodic_length=5.8,
                         heights=[[1.0, 0.5], [2.0, 0.5]], half_width=[0.70, 0.65], position=[-1.42, -0.56],
                         timestep=2e-3, tau=1e-2, randomize_v=1, gamma=3.0, temperature=2.0),
    "double_tau0": dict(model=1, overdamped=1, x0=-2.9, periodic_length=5.8,
                        heights=[[1.0, 0.5], [2.0, 0.5]], half_width=[0.70, 0.65], position=[-1.42, -0.56],
                        timestep=2e-3, tau=0.0, randomize_v=0),
}


def make(name, nsteps=60000, save_freq=100, **kw):
    cfg = dict(CONFIGS[name])
    cfg.update(kw)
    return ol.make_params(nsteps=nsteps, save_freq=save_freq, **cfg)


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_lane_replay_equals_rank(oracle, name):
    p = make(name)
    W = 4
    lanes, frames, s = run_rank_recorded(oracle, p, W, seed=7)
    if p.tau != 0.0:
        assert sum(lanes[i].n_catch for i in range(W)) > 0, "no barrier hand-over exercised"
    assert sum(lanes[i].fpt2_n + lanes[i].cg_n[0] for i in range(W)) > 0, "no statistics event exercised"
    for i in range(W):
        out = ol.Lane()
        fr = np.zeros(frames.shape[1])
        binit = np.ascontiguousarray(s["binit"][:, i, :])
        rc = oracle.ldo_lane_run_replay(C.byref(p), C.byref(out), ol.dptr(s["gauss"][i]), s["n_gauss"][i],
                                        ol.dptr(s["flip"][i]), s["n_flip"][i], ol.dptr(s["catch"][i]),
                                        s["n_catch"][i], binit.ctypes.data_as(C.POINTER(C.c_int8)), s["bmin"],
                                        binit.shape[1], ol.dptr(fr), len(fr), None, 0)
        assert rc == 0
        np.testing.assert_array_equal(fr, frames[0, :, i])
        ref = lanes[i]
        for f in ("x", "v", "f", "e", "r_prev", "state", "fpt_sum", "fpt2_sum", "tpt_sum", "tpt2_sum", "wt_sum",
                  "wt_sum2", "wt_max"):
            assert getattr(out, f) == getattr(ref, f), f
        for f in ("fpt_n", "fpt2_n", "tpt_n", "tpt2_n", "wt_ticks", "wt_n", "n_catch"):
            assert getattr(out, f) == getattr(ref, f), f
        assert list(out.state_count) == list(ref.state_count)
        assert list(out.cg_n) == list(ref.cg_n)
        assert list(out.cg_sum) == list(ref.cg_sum)


def test_philox_known_answers(oracle):
    """Random123 kat_vectors for philox4x32_10."""
    cases = [
        ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
        ((0xffffffff,) * 4, (0xffffffff,) * 2, (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
        ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
         (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
    ]
    for ctr, key, want in cases:
        c = (C.c_uint32 * 4)(*ctr)
        k = (C.c_uint32 * 2)(*key)
        o = (C.c_uint32 * 4)()
        oracle.ldo_philox4x32_10(c, k, o)
        assert tuple(o) == want


def test_flip_threshold_equals_double_compare(oracle):
    rng = np.random.default_rng(3)
    for p in [0.0, 1e-12, 0.25, 0.4999999989694238, 0.5, 0.75, 1.0 - 1e-12, 1.0]:
        T = oracle.ldo_flip_threshold(p)
        ws = np.concatenate([rng.integers(0, 2 ** 32, 2000, dtype=np.uint64),
                             np.array([max(T, 1) - 1, min(T, 2 ** 32 - 1)], np.uint64)])
        u = (ws.astype(np.float64) + 0.5) * 2.0 ** -32
        assert np.array_equal(ws < T, u < p)


def test_flip_probability_closed_form():
    """p = (1 - exp(-2 dt/tau)) / 2 (Single.h:1551): 0.4999999989694... at the
    quick test's dt = 1e-4, tau = 1e-5."""
    p = 0.5 * (1.0 - math.exp(-2.0 * 1e-4 / 1e-5))
    assert p == pytest.approx(0.5 * (1 - math.exp(-20.0)), rel=0, abs=0)
    assert 0.49999999 < p < 0.5


def _lane_philox(oracle, p, traj, nfr):
    out = ol.Lane()
    fr = np.zeros(nfr)
    oracle.ldo_lane_run_philox(C.byref(p), C.byref(out), 1234, traj, 0, 0, ol.dptr(fr), nfr)
    return out, fr


def test_free_particle_overdamped_diffusion(oracle):
    """H = 0: the MA(1) scheme (Single.h:2868-2876) has x_n - x_0 =
    sigma (g_0 + 2 g_1 + ... + 2 g_{n-1} + g_n); E[(x_n - x_0)^2] = 2 kT dt n /(m gamma) - kT dt/(m gamma)."""
    n_steps, save = 2000, 2000
    p = ol.make_params(model=0, overdamped=1, x0=-1.0, periodic_length=2.0, heights=[[0.0, 0.0], [0.0, 0.0]],
                       half_width=[0.5, 0.5], timestep=1e-3, tau=1e-2, nsteps=n_steps, save_freq=save,
                       gamma=2.0, temperature=1.5, b_waiting=0, b_fpt=0, b_tpt=0, b_fptcg=0)
    d2 = []
    for t in range(3000):
        _, fr = _lane_philox(oracle, p, t, 2)
        d2.append((fr[1] - fr[0]) ** 2)
    D = 1.5 / 2.0
    want = 2 * D * 1e-3 * n_steps - D * 1e-3
    assert np.mean(d2) == pytest.approx(want, rel=4 * math.sqrt(2.0 / 3000))


def test_underdamped_velocity_variance(oracle):
    """H = 0: the O step keeps <v^2> = kT/m (Single.h:1535-1537, 1689-1691)."""
    p = ol.make_params(model=0, overdamped=0, x0=-1.0, periodic_length=2.0, heights=[[0.0, 0.0], [0.0, 0.0]],
                       half_width=[0.5, 0.5], timestep=1e-2, tau=1e-2, nsteps=500, save_freq=500,
                       gamma=1.0, temperature=2.0, mass=0.5, randomize_v=1, b_waiting=0, b_fpt=0, b_tpt=0,
                       b_fptcg=0)
    v2 = [_lane_philox(oracle, p, t, 2)[0].v ** 2 for t in range(4000)]
    assert np.mean(v2) == pytest.approx(2.0 / 0.5, rel=4 * math.sqrt(2.0 / 4000))


def test_telegraph_stationary_occupancy(oracle):
    """Equal-probability telegraph noise spends half its time in each state;
    with EQ false and K, state 1 holds K/(1+K) (Single.h:1563-1583)."""
    for eq, K, want in ((1, 1.0, 0.5), (0, 3.0, 0.75)):
        p = ol.make_params(model=0, overdamped=1, x0=-1.0, periodic_length=2.0, heights=[[1.0, 0.5], [1.0, 0.5]],
                           half_width=[0.5, 0.5], timestep=1e-3, tau=2e-2, nsteps=20000, save_freq=20000,
                           equal_probability=eq, eq_constant=K, b_waiting=0, b_fpt=0, b_tpt=0, b_fptcg=0)
        c1 = c = 0.0
        for t in range(64):
            L, _ = _lane_philox(oracle, p, t, 2)
            c1 += L.state_count[1]
            c += L.state_count[0] + L.state_count[1]
        assert c1 / c == pytest.approx(want, abs=0.03)


def test_double_barrier_potential_is_continuous(oracle):
    """CuIDB_double geometry: E rises from 0 at x1b - w1 to H0 at x1b, falls to
    yc at xc, rises to H1 at x2b and falls to 0 at x2b + w2 (state 0)."""
    geo = dict(model=1, overdamped=1, periodic_length=5.8, heights=[[3.4, 2.0], [7.2, 2.0]],
               half_width=[0.70, 0.65], position=[-1.42, -0.56], timestep=1e-4, tau=0.0, nsteps=10,
               save_freq=10, randomize_state=0, initial_state=0.0, b_waiting=0, b_fpt=0, b_tpt=0, b_fptcg=0,
               randomize_v=0)
    x1b, w1, x2b, w2 = -1.42, 0.70, -0.56, 0.65

    def energy(x):
        p = ol.make_params(x0=x, **geo)
        out = ol.Lane()
        oracle.ldo_lane_run_replay(C.byref(p), C.byref(out), ol.dptr(np.zeros(64)), 64, ol.dptr(np.zeros(64)), 0,
                                   ol.dptr(np.zeros(1)), 0, None, 0, 0, None, 0, None, 0)
        return out

    eps = 1e-9
    for xa, want in ((x1b - w1 + eps, 0.0), (x1b - eps, 3.4), (x1b + eps, 3.4), (x2b - eps, 7.2),
                     (x2b + eps, 7.2), (x2b + w2 - eps, 0.0)):
        p = ol.make_params(x0=xa, nsteps=0, **{k: v for k, v in geo.items() if k != "nsteps"})
        out = ol.Lane()
        oracle.ldo_lane_run_replay(C.byref(p), C.byref(out), ol.dptr(np.zeros(4)), 4, ol.dptr(np.zeros(4)), 0,
                                   ol.dptr(np.zeros(1)), 0, None, 0, 0, None, 0, None, 0)
        assert out.e == pytest.approx(want, abs=1e-7), xa
    # the apex of the intermediate well joins both inner slopes
    lo = ol.make_params(x0=-1.06058, nsteps=0, **{k: v for k, v in geo.items() if k != "nsteps"})
    del energy
    out = ol.Lane()
    oracle.ldo_lane_run_replay(C.byref(lo), C.byref(out), ol.dptr(np.zeros(4)), 4, ol.dptr(np.zeros(4)), 0,
                               ol.dptr(np.zeros(1)), 0, None, 0, 0, None, 0, None, 0)
    assert 0.0 < out.e < 3.4
