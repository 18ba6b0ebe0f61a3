"""Pin the CPU oracle to the reference's own recorded quick-test run.

The fixtures in tests/golden/ were decoded from the reference's
sample_input/quick_test/executed_job (restart files and outputs); see
tests/golden/make_golden.py.  Where this container has the Intel MKL runtime
the reference linked (conda's libmkl_rt), the oracle's rank mode is driven by
the very same MT19937 / ICDF streams and must reproduce the recorded
trajectories bit for bit.
"""
import ctypes as C
import json
import os

import numpy as np
import pytest

import oracle_lib as ol

GOLD = os.path.join(os.path.dirname(__file__), "golden")


@pytest.fixture(scope="module")
def golden():
    with open(os.path.join(GOLD, "quick_test_state.json")) as f:
        state = json.load(f)
    frames = np.load(os.path.join(GOLD, "quick_test_frames.npz"))["x"]
    return state, frames


def quick_test_params(nsteps, **kw):
    """sample_input/quick_test/2.0_7.2_m1.0_1.00e-05.inp"""
    return ol.make_params(model=0, overdamped=1, x0=-2.9, randomize_v=1, periodic_length=5.8,
                          heights=[[7.2, 2.0], [7.2, 2.0]], half_width=[0.65, 0.65], timestep=1e-4,
                          nsteps=nsteps, save_freq=300000, tau=1e-5, randomize_state=1, **kw)


def test_rank_seeds_follow_glibc(golden):
    """MPIKernel::RandomSeed: srand(RANDOMSEED) then distinct rand() values
    (mpiKernel.cpp:124-138, RANDOMNUMBER.cpp:5-34) give the recorded seeds."""
    state, _ = golden
    libc = C.CDLL("libc.so.6")
    libc.srand(state["random_seed_master"])
    seeds = []
    while len(seeds) < 4:
        v = libc.rand() % 2147483647
        if v not in seeds:
            seeds.append(v)
    assert seeds == state["seeds"]


def test_log10_lags_and_msd_match_msd_log(oracle, golden):
    """GenerateLog10Scale + MeanSquareDisplacementBruteForce over the recorded
    frames reproduce msd.log (printed with %14.7E)."""
    state, frames = golden
    msd_log = np.array(state["outputs"]["msd.log"])
    lags = np.zeros(512, np.uint64)
    n = oracle.ldo_log10_scale(30.0, 10001, 1000, 40, 1, lags.ctypes.data_as(C.POINTER(C.c_uint64)), 512)
    lags = lags[:n]
    assert n == len(msd_log)
    np.testing.assert_allclose(30.0 * lags.astype(float), msd_log[:, 0], rtol=5e-8, atol=0)
    total = np.zeros(n)
    out = np.zeros(n)
    for lane in range(16):
        x = np.ascontiguousarray(frames[lane])
        oracle.ldo_msd_brute(ol.dptr(x), len(x), lags.ctypes.data_as(C.POINTER(C.c_uint64)), n, ol.dptr(out))
        total += out
    np.testing.assert_allclose(total / 16.0, msd_log[:, 1], rtol=6e-8, atol=1e-300)


def test_outputs_are_per_lane_means(golden):
    """firstPassageTime.dat holds STATISTICS::AVERAGE of each lane's event list
    (rank-major), the .log its average / population SD / count."""
    state, _ = golden
    dat = state["outputs"]["firstPassageTime.dat"]
    means = [L["mean_fpt"] for L in state["lanes"]]
    assert dat == means
    vals = dict(state["outputs"]["firstPassageTime.log"])
    assert vals["count"] == 16
    assert vals["average"] == pytest.approx(np.mean(dat), rel=1e-15)
    assert vals["sd"] == pytest.approx(np.std(dat), rel=1e-13)


needs_mkl = pytest.mark.skipif(not ol.mkl_available(), reason="MKL runtime (reference noise) not present")


def _run_rank(oracle, rank, seed, nsteps, capn=0):
    p = quick_test_params(nsteps)
    ms = ol.MklStreams([seed, seed, seed])
    st = oracle.ldo_rank_create(C.byref(p), 4, ms.fill, ms.ctx, 40000)
    lanes = (ol.Lane * 4)()
    nfr = nsteps // 300000 + 1
    frames = np.zeros((nfr, 4))
    nf = C.c_uint64()
    cap = None
    bufs = None
    if capn:
        bufs = {k: np.zeros((4, capn)) for k in ("fpt", "fpt2", "tpt", "tpt2", "ticks", "inst", "init")}
        bufs["cg"] = np.zeros((7, 4, capn))
        cap = ol.Capture(capn, *[ol.dptr(bufs[k]) for k in ("fpt", "fpt2", "tpt", "tpt2", "ticks", "cg", "inst",
                                                             "init")])
    oracle.ldo_rank_integrate(st, lanes, ol.dptr(frames), nfr, C.byref(nf), C.byref(cap) if cap else None, None)
    renew = (C.c_int64 * 3)()
    off = (C.c_uint64 * 3)()
    oracle.ldo_rank_rng_progress(st, renew, off)
    oracle.ldo_rank_destroy(st)
    ms.close()
    return lanes, frames, [(renew[k], off[k]) for k in range(3)], bufs


@needs_mkl
@pytest.mark.parametrize("rank", [0, 1, 2, 3])
def test_rank_mode_reproduces_recorded_frames(oracle, golden, rank):
    """20 saved frames (6e6 steps) of every lane equal the restart file bit for bit,
    and so do V0 and the initial barrier state drawn at step 0."""
    state, frames_ref = golden
    seed = state["seeds"][rank]
    lanes, frames, _, bufs = _run_rank(oracle, rank, seed, 20 * 300000, capn=64)
    for i in range(4):
        ref = state["lanes"][rank * 4 + i]
        assert lanes[i].v0_instance == ref["v0_instance"]
        assert lanes[i].state0_instance == ref["state0_instance"]
        np.testing.assert_array_equal(frames[:, i], frames_ref[rank * 4 + i, :21])
        # the first recorded events of every list, bit for bit
        for key, n_attr in (("fpt", "fpt_n"), ("fpt2", "fpt2_n"), ("tpt", "tpt_n"), ("tpt2", "tpt2_n"),
                            ("inst", "fpt_n"), ("init", "fpt_n")):
            k = min(int(getattr(lanes[i], n_attr)), len(ref[key + "_first"]), 64)
            assert k > 0, key
            np.testing.assert_array_equal(bufs[key][i, :k], ref[key + "_first"][:k])


@needs_mkl
@pytest.mark.slow
@pytest.mark.parametrize("rank", [0, 1, 2, 3])
def test_rank_mode_full_run_matches_restart(oracle, golden, rank):
    """The whole 3e9-step quick test: every frame, event count, end state and
    RNG progress counter of the restart file (several minutes per rank)."""
    state, frames_ref = golden
    seed = state["seeds"][rank]
    lanes, frames, rng, _ = _run_rank(oracle, rank, seed, 3000000000)
    assert [list(r) for r in rng] == state["ranks"][rank]["rng"]
    for i in range(4):
        ref = state["lanes"][rank * 4 + i]
        np.testing.assert_array_equal(frames[:, i], frames_ref[rank * 4 + i])
        assert lanes[i].x == ref["x"]
        assert [lanes[i].state_count[0], lanes[i].state_count[1]] == ref["state_count"]
        assert (lanes[i].fpt_n, lanes[i].fpt2_n, lanes[i].tpt_n, lanes[i].tpt2_n, lanes[i].wt_ticks) == (
            ref["fpt_n"], ref["fpt2_n"], ref["tpt_n"], ref["tpt2_n"], ref["ticks_n"])
        assert [lanes[i].cg_n[j] for j in range(7)] == ref["cg_n"]
        assert lanes[i].fpt_sum / lanes[i].fpt_n == ref["mean_fpt"]
        assert list(lanes[i].fpt_init_state) == ref["fpt_init_state"]
        assert list(lanes[i].fpt_init_index) == ref["fpt_init_index"]
