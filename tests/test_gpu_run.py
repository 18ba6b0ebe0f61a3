"""GPU parity of the reference-semantics integrator (libldpcp.so, run mode).

Replay: the GPU consumes the role-tagged noise a reference-style rank (the
oracle's rank mode, fed by seeded banks) drew, and must reproduce the rank's
trajectories bit for bit: every saved frame, the end state, every statistics
sum and count.  Philox: GPU and oracle draw the same counter-based noise and
build its Gaussians from IEEE operations only (ldpcp_gauss.h), so every saved
position agrees bit for bit and event counts exactly.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from langevindynamicsforpcp_amd.integrator import LangevinEquation
from streams import run_rank_recorded
from test_oracle_modes import CONFIGS, make

pytestmark = pytest.mark.gpu


def to_engine(p, n_traj, seed=0):
    """LangevinEquation with the same parameters as an oracle Params."""
    le = LangevinEquation(n_traj, model=p.model, overdamped=bool(p.overdamped))
    le.m_X0, le.m_V0, le.m_Gamma0, le.m_kB0, le.m_T0, le.m_M0 = p.x0, p.v0, p.gamma, p.kb, p.temperature, p.mass
    le.m_TimeStep0, le.m_NumberOfStep, le.m_SaveFreq = p.timestep, p.nsteps, p.save_freq
    le.m_RandomizeVelocity, le.m_RandomizeState0 = bool(p.randomize_v), bool(p.randomize_state)
    le.m_bSameInitialBarrier = bool(p.same_initial_barrier)
    le.m_PeriodicLength0 = p.periodic_length
    le.m_BarrierHeight0 = [[p.height[b][s] for s in range(2)] for b in range(2)]
    le.m_SlopeHalfLength0 = [p.half_width[0], p.half_width[1]]
    le.m_BarrierPosition0 = [p.position[0], p.position[1]]
    le.m_InitialState0 = p.initial_state
    le.m_Tau0 = p.tau
    le.m_bEqualProbability = bool(p.equal_probability)
    le.m_EquilibriumConstant = p.eq_constant
    le.m_bWaitingTimeDistribution = bool(p.b_waiting)
    le.m_bFirstPassageTime = bool(p.b_fpt)
    le.m_bTransitionPathTime = bool(p.b_tpt)
    le.m_bFirstPassageTimeCG = bool(p.b_fptcg)
    le.m_FirstPassageTimeCGSize = [p.cg_size[j] for j in range(p.n_cg)]
    le.m_SaveAll = bool(p.save_all)
    le.Initialize(seed)
    return le


EXACT = ("x", "v", "f", "e", "r_prev", "state", "barrier", "fpt_sum", "fpt2_sum", "tpt_sum", "tpt2_sum", "wt_sum",
         "wt_sum2", "wt_max", "v0_instance", "state0_instance", "well", "periodic")
COUNTS = ("fpt_n", "fpt2_n", "tpt_n", "tpt2_n", "wt_ticks", "wt_n", "n_catch")


@pytest.mark.parametrize("name", sorted(CONFIGS))
def test_gpu_replay_bit_exact(oracle, name):
    p = make(name, save_all=1)
    W = 8
    print(f"[gpu replay] {name} W={W} nsteps={p.nsteps}", flush=True)
    lanes, frames, s = run_rank_recorded(oracle, p, W, seed=11)
    ng, nf = min(s["n_gauss"]), min(s["n_flip"])
    le = to_engine(p, W)
    le.barrier_window = 1024
    le.steps_per_launch = 7919  # odd chunking on purpose: state must survive launches
    le.SetReplay(s["gauss"][:, :ng].T, s["flip"][:, :nf].T, s["catch"], s["binit"], s["bmin"])
    le.Integrate(0)
    got = le.lanes()
    for field in range(7):
        np.testing.assert_array_equal(le.frames(field).T, frames[field], err_msg=f"frame field {field}")
    for i in range(W):
        for f in EXACT:
            assert getattr(got[i], f) == getattr(lanes[i], f), (name, i, f)
        for f in COUNTS:
            assert getattr(got[i], f) == getattr(lanes[i], f), (name, i, f)
        assert list(got[i].state_count) == list(lanes[i].state_count)
        assert list(got[i].cg_n) == list(lanes[i].cg_n)
        assert list(got[i].cg_sum) == list(lanes[i].cg_sum)
    le.close()


@pytest.mark.parametrize("name", ["single_over", "single_under", "double_over", "double_under"])
def test_gpu_philox_matches_oracle(oracle, name):
    p = make(name, nsteps=10000, save_freq=100)
    n = 64
    print(f"[gpu philox] {name} n={n}", flush=True)
    le = to_engine(p, n, seed=99)
    le.Integrate(0)
    got = le.lanes()
    fr = le.frames(0)
    for t in range(n):
        ref = ol.Lane()
        rf = np.zeros(fr.shape[1])
        oracle.ldo_lane_run_philox(C.byref(p), C.byref(ref), 99, t, 0, 0, ol.dptr(rf), len(rf))
        np.testing.assert_array_equal(fr[t], rf, err_msg=f"{name} trajectory {t}")
        for f in COUNTS:
            assert getattr(got[t], f) == getattr(ref, f), (name, t, f)
        assert got[t].state == ref.state
    le.close()


def test_gpu_msd_kernel_matches_oracle(oracle):
    p = make("single_over", nsteps=200000, save_freq=100)
    le = to_engine(p, 16, seed=5)
    le.Integrate(0)
    fr = le.frames(0)
    lags = np.zeros(256, np.uint64)
    nl = oracle.ldo_log10_scale(p.timestep * p.save_freq, fr.shape[1], 1000, 20, 1,
                                lags.ctypes.data_as(C.POINTER(C.c_uint64)), 256)
    lags = lags[:nl]
    got = le.msd(lags)
    out = np.zeros(nl)
    for t in range(16):
        x = np.ascontiguousarray(fr[t])
        oracle.ldo_msd_brute(ol.dptr(x), len(x), lags.ctypes.data_as(C.POINTER(C.c_uint64)), nl, ol.dptr(out))
        np.testing.assert_array_equal(got[t], out)
    le.close()


def test_gpu_barrier_window_overflow_is_an_error():
    p = make("single_over", nsteps=200000, save_freq=100)
    le = to_engine(p, 64, seed=3)
    le.barrier_window = 4
    from langevindynamicsforpcp_amd import LdpcpError
    with pytest.raises(LdpcpError, match="barrier window"):
        le.Integrate(0)
    le.close()


@pytest.mark.parametrize("name", ["single_over", "single_under", "double_over", "double_under", "single_tau0",
                                  "double_tau0"])
def test_gpu_buffered_noise_equals_inline_noise(name):
    """A handle of few trajectories draws its Philox noise in a separate kernel
    (staged through LDS); a large handle draws it inside the step loop.  The
    trajectories both handles share (same ids) must agree bit for bit.  The
    buffered handles cover a one-wavefront handle of 7 trajectories (stages of
    308 steps, lanes 7-63 dead), one of 64 (32-step stages) and one of 200
    (four wavefronts loading row by row, the last with 8 trajectories)."""
    p = make(name, nsteps=40000, save_freq=100, save_all=1)
    big = 8192 + 128                     # above the buffered-noise limit
    print(f"[gpu noise paths] {name}", flush=True)
    b = to_engine(p, big, seed=31)
    b.tick_capacity = 256
    b.Integrate(0)
    lb = b.lanes()
    for small in (7, 64, 200):
        print(f"[gpu noise paths] {name} buffered handle of {small}", flush=True)
        a = to_engine(p, small, seed=31)
        a.tick_capacity = 256
        a.steps_per_launch = 3001        # stages straddle launches
        a.Integrate(0)
        la = a.lanes()
        for t in range(small):
            for f, _ in type(la[t])._fields_:
                x, y = getattr(la[t], f), getattr(lb[t], f)
                assert (list(x) if hasattr(x, "__len__") else x) == (list(y) if hasattr(y, "__len__") else y), \
                    (small, t, f)
        for field in range(7):
            np.testing.assert_array_equal(a.frames(field), b.frames(field)[:small])
        a.close()
    b.close()


def test_gpu_replay_streams_can_be_replaced(oracle):
    """A second ldpcp_run_set_replay on a handle replaces the first call's
    streams (and frees them): the run follows the second set."""
    p = make("single_under", save_all=1)
    W = 8
    lanes, frames, s = run_rank_recorded(oracle, p, W, seed=23)
    ng, nf = min(s["n_gauss"]), min(s["n_flip"])
    le = to_engine(p, W)
    le.barrier_window = 1024
    le._create()
    from langevindynamicsforpcp_amd import _native as N
    lib = N.load()
    g, f = np.ascontiguousarray(s["gauss"][:, :ng].T), np.ascontiguousarray(s["flip"][:, :nf].T)
    c, b = np.ascontiguousarray(s["catch"]), np.ascontiguousarray(s["binit"], np.int8)
    dp = lambda a: a.ctypes.data_as(C.POINTER(C.c_double))  # noqa: E731
    junk = np.zeros_like(g)
    for gauss in (junk, g):
        N.check(lib.ldpcp_run_set_replay(le._h, dp(gauss), g.shape[0], dp(f), f.shape[0], dp(c), c.shape[1],
                                         b.ctypes.data_as(C.POINTER(C.c_int8)), s["bmin"], b.shape[2]),
                "ldpcp_run_set_replay")
    N.check(lib.ldpcp_run_integrate(le._h, 4096), "ldpcp_run_integrate")
    got = (N.LaneResult * W)()
    N.check(lib.ldpcp_run_lanes(le._h, 0, W, got), "ldpcp_run_lanes")
    for i in range(W):
        for fld in EXACT + COUNTS:
            assert getattr(got[i], fld) == getattr(lanes[i], fld), (i, fld)
    le.close()
