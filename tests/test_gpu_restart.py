"""Restart on the GPU: checkpoint images and resumed runs (run mode).

The reference writes its whole trajectory state at the end of a run
(WriteRestartFile, langevinDynamicsSingleBarrier.h:1048-1136) and a later run
with READRESTART continues it (ReadRestartFile :1137-1206; the resumed
Integrate skips step 0 and continues the step count, :1397-1400, 1420).  Here
the state is the library's checkpoint image and the noise position is the step
count itself, so:

* a run of a steps, a checkpoint and a resumed run of b steps reproduce one
  run of a + b steps bit for bit: every frame of every field, every per-lane
  statistic, the barrier ticks and the MSD over all frames;
* in replay mode the resumed run also reproduces the oracle's rank, so the
  continuation is pinned to the reference's semantics, not only to itself;
* UNIFORMINITALSTATES barriers created in a later trial take the previous
  trial's initial state (UserData::m_bInitialValue, :1378-1380).
"""
import ctypes as C
import math

import numpy as np
import pytest

from langevindynamicsforpcp_amd import LdpcpError
from langevindynamicsforpcp_amd import _native as N
from langevindynamicsforpcp_amd.integrator import read_checkpoint
from streams import run_rank_recorded
from test_gpu_run import COUNTS, EXACT, to_engine
from test_oracle_modes import make

pytestmark = pytest.mark.gpu

CASES = ["single_over", "single_under", "single_eq_false", "single_tau0", "double_over", "double_under"]


def lane_dict(lane):
    out = {}
    for name, _ in N.LaneResult._fields_:
        v = getattr(lane, name)
        out[name] = list(v) if hasattr(v, "__len__") else v
    return out


def engine(name, nsteps, n, seed, spl):
    le = to_engine(make(name, nsteps=nsteps, save_freq=100, save_all=1), n, seed=seed)
    le.tick_capacity = 4096
    le.event_capacity = 4096
    le.steps_per_launch = spl
    return le


@pytest.mark.parametrize("name", CASES)
def test_gpu_resume_continues_bit_exact(name):
    n, seed, a, b = 48, 21, 30000, 45000
    print(f"[gpu restart] {name} n={n} a={a} b={b}", flush=True)
    full = engine(name, a + b, n, seed, 6007)
    full.Integrate(0)
    first = engine(name, a, n, seed, 7919)
    first.Integrate(0)
    image = first.Checkpoint()
    hd = N.CheckpointHeader.from_buffer_copy(image)
    assert hd.magic == N.CHECKPOINT_MAGIC and hd.steps_done == a and hd.n_frames == a // 100 + 1
    assert hd.bytes == len(image)
    second = engine(name, b, n, seed, 5003)
    second.Resume(image)
    second.Integrate(0)

    for t, (x, y) in enumerate(zip(full.lanes(), second.lanes())):
        assert lane_dict(x) == lane_dict(y), (name, t)
    assert sum(full.lanes()[t].fpt2_n + full.lanes()[t].cg_n[0] for t in range(n)) > 0
    for field in range(7):
        np.testing.assert_array_equal(second.frames(field), full.frames(field), err_msg=f"field {field}")
    for x, y in zip(full.ticks(), second.ticks()):
        np.testing.assert_array_equal(x, y)
    lags = np.array([0, 1, 2, 5, 10, 50, 200, 600], np.uint64)
    np.testing.assert_array_equal(second.msd(lags), full.msd(lags))
    # the whole device state, event lists included, byte for byte
    assert second.Checkpoint() == full.Checkpoint()
    for le in (full, first, second):
        le.close()


def test_gpu_resume_after_dropped_steps_continues_the_reference_clock():
    """NUMBEROFSTEPS 45 with FREQSAVE 10 integrates 40 steps; the reference's
    clock, its restart header and its next run continue at step 45 (the resumed
    run integrates steps 46 .. 75, frames at 55, 65, 75; Single.h:1397-1450).
    A free overdamped particle (F = 0) with replayed Gaussians has the exact
    recurrence x <- (x + var g_k) + R, R <- var g_k with k = iStep - 1, so every
    frame of both runs is known bit for bit."""
    W = 8
    p = make("single_over", nsteps=45, save_freq=10)
    p.height[0][0] = p.height[0][1] = p.height[1][0] = p.height[1][1] = 0.0
    p.periodic_length, p.randomize_v, p.randomize_state = 1e6, 0, 0
    p.b_waiting = p.b_fpt = p.b_tpt = p.b_fptcg = 0
    print("[gpu restart] dropped trailing steps", flush=True)
    rng = np.random.default_rng(5)
    n_g = 1 + 45 + 30  # R0, then one Gaussian per step of the clock
    gauss = rng.standard_normal((n_g, W))
    flip = rng.random((45 + 30, W))
    first = to_engine(p, W)
    first.SetReplay(gauss, flip, np.zeros((W, 1)), np.full((2, W, 1), -1, np.int8), 0)
    first.Integrate(0)
    image = first.Checkpoint()
    hd = N.CheckpointHeader.from_buffer_copy(image)
    assert hd.steps_done == 45 and hd.n_frames == 5
    assert all(L.steps_done == 45 for L in first.lanes())
    p.nsteps = 30
    second = to_engine(p, W)
    second.SetReplay(gauss, flip, np.zeros((W, 1)), np.full((2, W, 1), -1, np.int8), 0)
    second.Resume(image)
    second.Integrate(0)
    assert N.load().ldpcp_run_n_frames(second._h) == 8
    assert all(L.steps_done == 75 for L in second.lanes())
    # the recurrence, step by step, in the same operation order
    var = math.sqrt(p.kb * p.temperature / (p.mass * p.gamma) * p.timestep * 0.5)
    x = np.full(W, p.x0)
    R = var * gauss[0]
    want = {0: x.copy()}
    for i in list(range(1, 41)) + list(range(46, 76)):
        rf = var * gauss[1 + i - 1]
        x = (x + rf) + R
        R = rf
        if i in (10, 20, 30, 40, 55, 65, 75):
            want[i] = x.copy()
    got = second.frames(0)
    for k, step in enumerate([0, 10, 20, 30, 40, 55, 65, 75]):
        np.testing.assert_array_equal(got[:, k], want[step], err_msg=f"frame {k} (step {step})")
    first.close()
    second.close()


def test_gpu_checkpoint_of_a_resumed_handle_is_the_image():
    n, seed = 16, 4
    first = engine("double_under", 20000, n, seed, 4096)
    first.Integrate(0)
    image = first.Checkpoint()
    second = engine("double_under", 10000, n, seed, 4096)
    second.Resume(image)
    second._create()                         # resumed, not yet integrated
    assert second.Checkpoint() == image
    second.Integrate(0)
    with pytest.raises(LdpcpError, match="already continued"):
        N.check(N.load().ldpcp_run_integrate(second._h, 4096), "ldpcp_run_integrate")
    first.close()
    second.close()


def test_gpu_resume_rejects_a_different_run():
    first = engine("single_over", 10000, 16, 5, 4096)
    first.Integrate(0)
    image = first.Checkpoint()
    other = engine("single_over", 10000, 16, 6, 4096)   # another seed: another run
    other.Resume(image)
    with pytest.raises(LdpcpError, match="inconsistent seed"):
        other.Integrate(0)
    fewer = engine("single_over", 10000, 8, 5, 4096)
    fewer.Resume(image)
    with pytest.raises(LdpcpError, match="inconsistent trajectory count"):
        fewer.Integrate(0)
    bad = bytearray(image)
    bad[0] ^= 1
    third = engine("single_over", 10000, 16, 5, 4096)
    third.Resume(bytes(bad))
    with pytest.raises(LdpcpError, match="not a checkpoint image"):
        third.Integrate(0)
    for le in (first, other, fewer, third):
        le.close()


@pytest.mark.parametrize("name", ["single_over", "double_under"])
def test_gpu_resumed_replay_matches_oracle(oracle, name):
    """Replay of a reference-style rank, split in two runs at a checkpoint."""
    W, a, b = 8, 20000, 40000
    p = make(name, nsteps=a + b, save_freq=100, save_all=1)
    print(f"[gpu restart replay] {name} W={W}", flush=True)
    lanes, frames, s = run_rank_recorded(oracle, p, W, seed=13)
    ng, nf = min(s["n_gauss"]), min(s["n_flip"])
    replay = (s["gauss"][:, :ng].T, s["flip"][:, :nf].T, s["catch"], s["binit"], s["bmin"])
    first = to_engine(make(name, nsteps=a, save_freq=100, save_all=1), W)
    first.barrier_window = 1024
    first.SetReplay(*replay)
    first.Integrate(0)
    second = to_engine(make(name, nsteps=b, save_freq=100, save_all=1), W)
    second.barrier_window = 1024
    second.SetReplay(*replay)
    second.Resume(first.Checkpoint())
    second.Integrate(0)
    got = second.lanes()
    for field in range(7):
        np.testing.assert_array_equal(second.frames(field).T, frames[field], err_msg=f"frame field {field}")
    for i in range(W):
        for f in EXACT + COUNTS:
            assert getattr(got[i], f) == getattr(lanes[i], f), (name, i, f)
        assert list(got[i].cg_sum) == list(lanes[i].cg_sum)
    first.close()
    second.close()


def test_gpu_same_initial_barrier_second_trial_matches_oracle(oracle):
    """Trial 2 of a rank with UNIFORMINITALSTATES: barriers created during the
    trial take trial 1's initial states (1 here), not 0."""
    W = 8
    p = make("single_same_init", nsteps=60000, save_freq=100, save_all=1)
    lanes, frames, s = run_rank_recorded(oracle, p, W, seed=17, trials=2)
    ng, nf = min(s["n_gauss"]), min(s["n_flip"])
    le = to_engine(p, W)
    le.barrier_window = 1024
    le.SetReplay(s["gauss"][:, :ng].T, s["flip"][:, :nf].T, s["catch"], s["binit"], s["bmin"])
    le.SetBarrierDefault(np.full(W, p.initial_state))
    le.Integrate(0)
    got = le.lanes()
    assert sum(lanes[i].n_catch for i in range(W)) > 0
    np.testing.assert_array_equal(le.frames(0).T, frames[0])
    for i in range(W):
        for f in EXACT + COUNTS:
            assert getattr(got[i], f) == getattr(lanes[i], f), (i, f)
    le.close()


@pytest.mark.parametrize("name", ["single_over", "single_under", "single_eq_false", "double_over", "double_under"])
def test_gpu_event_lists_match_oracle(oracle, name):
    """The raw event lists a restart file in the reference's layout needs
    (FPT1, FPT2, TPT1, TPT2, FPTCG and the FPT1 barrier logs of the crossed and
    the left barrier, TrajectoryData, langevinDynamicsSingleBarrier.h:160-235;
    the double barrier fills FPT2 and FPTCG, langevinDynamicsDoubleBarrier.h:2931-3209),
    kept on the device, equal the oracle's lists entry for entry (replayed noise)."""
    W, capn = 8, 8192
    p = make(name, nsteps=60000, save_freq=100)
    print(f"[gpu event lists] {name}", flush=True)
    lanes, frames, s = run_rank_recorded(oracle, p, W, seed=19, capn=capn)
    ng, nf = min(s["n_gauss"]), min(s["n_flip"])
    le = to_engine(p, W)
    le.barrier_window = 1024
    le.event_capacity = capn
    le.SetReplay(s["gauss"][:, :ng].T, s["flip"][:, :nf].T, s["catch"], s["binit"], s["bmin"])
    le.Integrate(0)
    img = read_checkpoint(le.Checkpoint())
    assert img["header"].events_complete == 1 and img["header"].n_lists == 6 + p.n_cg
    cap = s["capture"]
    names = ["fpt", "fpt2", "tpt", "tpt2", "inst", "init"]
    for i in range(W):
        got = img["events"][i]
        assert list(img["lanes"][i].fpt_init_index) == list(lanes[i].fpt_init_index)
        assert list(img["lanes"][i].fpt_init_state) == list(lanes[i].fpt_init_state)
        for k, key in enumerate(names):
            assert len(got[k]) <= capn
            np.testing.assert_array_equal(got[k], cap[key][i, :len(got[k])], err_msg=f"{key} lane {i}")
        for j in range(p.n_cg):
            np.testing.assert_array_equal(got[6 + j], cap["cg"][j, i, :len(got[6 + j])], err_msg=f"cg{j}")
    if name.startswith("single"):
        assert sum(len(img["events"][i][4]) for i in range(W)) > 0, "no FPT1 event: the barrier logs are not exercised"
    else:
        assert sum(len(img["events"][i][1]) for i in range(W)) > 0, "no FPT2 event: the lists are not exercised"
    le.close()
