"""End-to-end runs of the drop-in executables on the GPU.

The expected output values are recomputed with the CPU oracle (lane mode,
same Philox keys: rank r's glibc seed, trajectory trial*W + lane) and the
runner's reductions (rank-major per-lane means, MSD summed lane by lane,
state counts); file layouts are checked against the reference's recorded
quick-test outputs.
"""
import ctypes as C
import json
import os
import re
import subprocess

import numpy as np
import pytest

import oracle_lib as ol

pytestmark = pytest.mark.gpu

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
BIN = os.path.join(REPO, "langevindynamicsforpcp_amd", "bin")
GOLD = os.path.join(os.path.dirname(__file__), "golden")

SINGLE_INP = """\
OUTPUTBASE           out
OVERDAMPED           {over}
X0                   -1.05
V0                   1.0
GAMMA                1.0
KB                   1.0
TEMPERATURE          1.0
MASS                 1.0
RANDOMIZEV           true
PERIODICLENGTH       2.1
BARRIERHEIGHTS       3.0 1.0
BARRIERHALFWIDTH     0.6
TIMESTEP             0.001
NUMBEROFSTEPS        200000
NUMBEROFPREVIOUSSTEPS 0
NUMBEROFRUNS         2
FREQSAVE             100
FREQSTAT             1
DISTRIBUTIONV0       true
DISTRIBUTIONS0       true
TAU                  0.001
INITIALSTATE         0
RANDOMIZESTATES      true
UNIFORMINITALSTATES  false
EQ                   true 1.0
WAITINGTIME          true
FPT                  true
TPT                  true
FPTCG                true 2
FPTCG                true 3
FPTCG                true 10
READRESTART          false
WRITERESTART         false
INPUTRESTART         none
OUTPUTRESTART        none
WRITEX               true
NUMBEROFLOGSAMPLES   20
RANDOMSEED           123
PRINTTIMER           false
"""


def rank_seeds(master, ranks):
    libc = C.CDLL("libc.so.6")
    libc.srand(master)
    out = []
    while len(out) < ranks:
        v = libc.rand() % 2147483647
        if v not in out:
            out.append(v)
    return out


def oracle_lanes(oracle, p, seed, trials, W):
    res = []
    nfr = p.nsteps // p.save_freq + 1
    for i in range(trials):
        for ch in range(W):
            out = ol.Lane()
            fr = np.zeros(nfr)
            oracle.ldo_lane_run_philox(C.byref(p), C.byref(out), seed, i * W + ch, 0, 0, ol.dptr(fr), nfr)
            res.append((out, fr))
    return res


@pytest.mark.parametrize("over", ["true", "false"])
def test_single_barrier_executable_matches_oracle(oracle, tmp_path, over):
    (tmp_path / "run.inp").write_text(SINGLE_INP.format(over=over))
    ranks, W, trials = 3, 4, 2
    print(f"[gpu exe] single over={over}", flush=True)
    r = subprocess.run([os.path.join(BIN, "langevinDynamicsSingleBarrier"), "-i", "run.inp", "--ranks", str(ranks),
                        "--simd-width", str(W)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    p = ol.make_params(model=0, overdamped=1 if over == "true" else 0, x0=-1.05, periodic_length=2.1,
                       heights=[[3.0, 1.0], [3.0, 1.0]], half_width=[0.6, 0.6], timestep=1e-3, tau=1e-3,
                       nsteps=200000, save_freq=100, randomize_v=1, cg=[2, 3, 10])
    seeds = rank_seeds(123, ranks)
    lanes = []
    for rk in range(ranks):
        lanes += oracle_lanes(oracle, p, seeds[rk], trials, W)
    mean = lambda s, n: s / n if n else 0.0  # noqa: E731
    exp = {
        "firstPassageTime.dat": [mean(L.fpt_sum, L.fpt_n) for L, _ in lanes],
        "firstPassageTime2.dat": [mean(L.fpt2_sum, L.fpt2_n) for L, _ in lanes],
        "transitionPathTime.dat": [mean(L.tpt_sum, L.tpt_n) for L, _ in lanes],
        "transitionPathTime2.dat": [mean(L.tpt2_sum, L.tpt2_n) for L, _ in lanes],
        "firstPassageTimeCG2_0000_0002.dat": [mean(L.cg_sum[0], L.cg_n[0]) for L, _ in lanes],
        "firstPassageTimeCG2_0002_0010.dat": [mean(L.cg_sum[2], L.cg_n[2]) for L, _ in lanes],
    }
    for name, want in exp.items():
        got = np.fromfile(tmp_path / f"out_{name}", "<f8")
        np.testing.assert_array_equal(got, want, err_msg=name)
    # msd.log: lags from GenerateLog10Scale, per-lane time-origin MSD summed rank-major / nrun
    lags = np.zeros(512, np.uint64)
    nl = oracle.ldo_log10_scale(0.1, 2001, 1000, 20, 1, lags.ctypes.data_as(C.POINTER(C.c_uint64)), 512)
    lags = lags[:nl]
    total = np.zeros(nl)
    out = np.zeros(nl)
    for _, fr in lanes:
        oracle.ldo_msd_brute(ol.dptr(fr), len(fr), lags.ctypes.data_as(C.POINTER(C.c_uint64)), nl, ol.dptr(out))
        total += out
    msd = np.loadtxt(tmp_path / "out_msd.log", comments="#")
    np.testing.assert_allclose(msd[:, 0], 0.1 * lags, rtol=1e-7)
    np.testing.assert_allclose(msd[:, 1], total / len(lanes), rtol=1e-6)
    # state counts and initial-state histogram are exact
    s0 = sum(L.state_count[0] for L, _ in lanes)
    s1 = sum(L.state_count[1] for L, _ in lanes)
    sc = [ln for ln in (tmp_path / "out_stateCount.log").read_text().splitlines() if ln.startswith("# state")]
    assert sc[0].split()[3] == f"{s0:.0f}" and sc[1].split()[3] == f"{s1:.0f}"
    st0 = np.loadtxt(tmp_path / "out_State0Distribution.log", comments="#")
    assert st0[:, 3].tolist() == [sum(1 for L, _ in lanes if L.state0_instance == 0.0),
                                  sum(1 for L, _ in lanes if L.state0_instance == 1.0)]
    # one trajectory log per (trial, lane, rank), frames first, then the event ticks
    logs = sorted(f for f in os.listdir(tmp_path) if "_trajectory_" in f)
    assert len(logs) == ranks * trials * W
    body = np.loadtxt(tmp_path / logs[0], comments="#")
    assert body.shape == (2001, 8)
    np.testing.assert_allclose(body[:, 1], lanes[0][1], rtol=1e-6, atol=1e-6)
    # waiting-time moments
    wt = [ln for ln in (tmp_path / "out_waitingTime.log").read_text().splitlines() if ln.startswith("# Nwt")]
    assert int(wt[0].split()[-1]) == sum(L.wt_n for L, _ in lanes)


def test_header_layout_matches_reference(tmp_path):
    """The PrintParameter block of our .log files has the reference's lines in
    the reference's order (values differ only where the run differs)."""
    with open(os.path.join(GOLD, "quick_test_headers.json")) as f:
        ref = json.load(f)["firstPassageTime.log"]
    inp = open(os.path.join(GOLD, "quick_test.inp")).read()
    inp = inp.replace("3000000000", "3000000").replace("FREQSAVE             300000", "FREQSAVE             300")
    inp = inp.replace("WRITEX               true", "WRITEX               false")
    inp = inp.replace("WRITERESTART         true", "WRITERESTART         false")
    (tmp_path / "q.inp").write_text(inp)
    r = subprocess.run([os.path.join(BIN, "langevinDynamicsSingleBarrier"), "-i", "q.inp", "--ranks", "4",
                        "--simd-width", "4"], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    assert (tmp_path / "m1.0_t1.00e-05_0_firstPassageTime.log").exists(), r.stderr[-2000:]
    got = (tmp_path / "m1.0_t1.00e-05_0_firstPassageTime.log").read_text().splitlines()
    varying = ("Save frequency", "Save time step", "Number of step ", "Simulation time", "Number of traj",
               "Memory usage", "# average", "# sd ", "Write trajectories", "Write restart")
    assert len(got) == len(ref)
    number = re.compile(r"[-+]?[0-9][-+0-9.eE]*")
    for a, b in zip(got, ref):
        if any(v in b for v in varying):
            # same label and layout, different value
            assert number.sub("N", a.split("=")[0]) == number.sub("N", b.split("=")[0])
        else:
            assert a == b


DOUBLE_INP = """\
OUTPUTBASE           dbl
OVERDAMPED           {over}
X0                   -2.9
V0                   1.0
GAMMA                1.0
KB                   1.0
TEMPERATURE          1.0
MASS                 1.0
RANDOMIZEV           true
PERIODICLENGTH       5.8
BARRIERHEIGHTS0      1.0 0.5
BARRIERHEIGHTS1      2.0 0.5
BARRIERHALFWIDTH0    0.70
BARRIERHALFWIDTH1    0.65
BARRIERPOSITION0     -1.42
BARRIERPOSITION1     -0.56
TIMESTEP             0.002
NUMBEROFSTEPS        200000
NUMBEROFPREVIOUSSTEPS 0
NUMBEROFRUNS         1
FREQSAVE             100
FREQSTAT             1
DISTRIBUTIONV0       false
DISTRIBUTIONS0       true
TAU                  0.001
INITIALSTATE         0
RANDOMIZESTATES      true
UNIFORMINITALSTATES  false
EQ                   true 1.0
WAITINGTIME          true
FPT                  true
TPT                  true
FPTCG                true 2
FPTCG                true 3
READRESTART          false
WRITERESTART         false
INPUTRESTART         none
OUTPUTRESTART        none
WRITEX               false
NUMBEROFLOGSAMPLES   20
RANDOMSEED           77
PRINTTIMER           false
"""


@pytest.mark.parametrize("over", ["true", "false"])
def test_double_barrier_executable_matches_oracle(oracle, tmp_path, over):
    """langevinDynamicsDoubleBarrier end to end: FPT2 and FPTCG per-lane means
    (the statistics the double-barrier code keeps), FPT1 / TPT outputs zero,
    state counts and MSD against the oracle with the same Philox keys."""
    (tmp_path / "run.inp").write_text(DOUBLE_INP.format(over=over))
    ranks, W = 2, 8
    print(f"[gpu exe] double over={over}", flush=True)
    r = subprocess.run([os.path.join(BIN, "langevinDynamicsDoubleBarrier"), "-i", "run.inp", "--ranks", str(ranks),
                        "--simd-width", str(W)], cwd=tmp_path, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stderr[-2000:]
    p = ol.make_params(model=1, overdamped=1 if over == "true" else 0, x0=-2.9, periodic_length=5.8,
                       heights=[[1.0, 0.5], [2.0, 0.5]], half_width=[0.70, 0.65], position=[-1.42, -0.56],
                       timestep=2e-3, tau=1e-3, nsteps=200000, save_freq=100, randomize_v=1, cg=[2, 3])
    seeds = rank_seeds(77, ranks)
    lanes = []
    for rk in range(ranks):
        lanes += oracle_lanes(oracle, p, seeds[rk], 1, W)
    mean = lambda s, n: s / n if n else 0.0  # noqa: E731
    assert sum(L.fpt2_n for L, _ in lanes) > 0, "no FPT2 events: the case does not exercise the statistics"
    exp = {
        "firstPassageTime2.dat": [mean(L.fpt2_sum, L.fpt2_n) for L, _ in lanes],
        "firstPassageTimeCG2_0000_0002.dat": [mean(L.cg_sum[0], L.cg_n[0]) for L, _ in lanes],
        "firstPassageTimeCG2_0001_0003.dat": [mean(L.cg_sum[1], L.cg_n[1]) for L, _ in lanes],
    }
    for name, want in exp.items():
        got = np.fromfile(tmp_path / f"dbl_{name}", "<f8")
        np.testing.assert_array_equal(got, want, err_msg=name)
    for name in ("firstPassageTime.dat", "transitionPathTime.dat", "transitionPathTime2.dat"):
        got = np.fromfile(tmp_path / f"dbl_{name}", "<f8")
        assert got.shape == (ranks * W,) and not got.any(), name
    s0 = sum(L.state_count[0] for L, _ in lanes)
    s1 = sum(L.state_count[1] for L, _ in lanes)
    sc = [ln for ln in (tmp_path / "dbl_stateCount.log").read_text().splitlines() if ln.startswith("# state")]
    assert sc[0].split()[3] == f"{s0:.0f}" and sc[1].split()[3] == f"{s1:.0f}"
    lags = np.zeros(512, np.uint64)
    nl = oracle.ldo_log10_scale(0.2, 2001, 1000, 20, 1, lags.ctypes.data_as(C.POINTER(C.c_uint64)), 512)
    lags = lags[:nl]
    total = np.zeros(nl)
    out = np.zeros(nl)
    for _, fr in lanes:
        oracle.ldo_msd_brute(ol.dptr(fr), len(fr), lags.ctypes.data_as(C.POINTER(C.c_uint64)), nl, ol.dptr(out))
        total += out
    msd = np.loadtxt(tmp_path / "dbl_msd.log", comments="#")
    np.testing.assert_allclose(msd[:, 0], 0.2 * lags, rtol=1e-6)
    np.testing.assert_allclose(msd[:, 1], total / len(lanes), rtol=1e-6, atol=6e-8)


def _outputs(d):
    """Every output file of a run, header lines (`# key = value`) and the
    timing log left out: numbers only."""
    out = {}
    for name in sorted(os.listdir(d)):
        if not name.startswith("out_") or name.endswith("_time.log"):
            continue
        if name.endswith(".dat"):
            out[name] = (d / name).read_bytes()
        else:
            out[name] = [ln for ln in (d / name).read_text().splitlines() if not (ln.startswith("#") and "=" in ln)]
    return out


@pytest.mark.parametrize("model", ["single", "double"])
def test_executable_restart_continues_exactly(tmp_path, model):
    """WRITERESTART after 100000 steps, then READRESTART for 100000 more
    (NUMBEROFPREVIOUSSTEPS 100000), gives every output file of one
    200000-step run: means, logs, MSD over all frames, trajectory logs."""
    base = (SINGLE_INP.format(over="true") if model == "single" else DOUBLE_INP.format(over="false"))
    base = re.sub(r"(?m)^OUTPUTBASE .*$", "OUTPUTBASE out", base)
    base = re.sub(r"(?m)^FREQSAVE .*$", "FREQSAVE 10", base)
    base = re.sub(r"(?m)^NUMBEROFRUNS .*$", "NUMBEROFRUNS 2", base)
    exe = os.path.join(BIN, "langevinDynamicsSingleBarrier" if model == "single" else "langevinDynamicsDoubleBarrier")

    def run(d, **keys):
        text = base
        for k, v in keys.items():
            text = re.sub(rf"(?m)^{k} .*$", f"{k} {v}", text)
        (d / "run.inp").write_text(text)
        print(f"[gpu exe restart] {model} {d.name} {keys}", flush=True)
        r = subprocess.run([exe, "-i", "run.inp", "--ranks", "2", "--simd-width", "4"], cwd=d, capture_output=True,
                           text=True, timeout=300)
        assert r.returncode == 0 and "Error" not in r.stderr, r.stderr[-2000:]

    full, split = tmp_path / "full", tmp_path / "split"
    full.mkdir()
    split.mkdir()
    run(full, NUMBEROFSTEPS=200000)
    run(split, NUMBEROFSTEPS=100000, WRITERESTART="true", OUTPUTRESTART="rs")
    restarts = sorted(f for f in os.listdir(split) if f.endswith(".restart"))
    assert restarts == [f"rs_{r:05d}_{t:05d}.restart" for r in range(2) for t in range(2)]
    for f in os.listdir(split):
        if f.startswith("out_"):
            os.remove(split / f)
    run(split, NUMBEROFSTEPS=100000, NUMBEROFPREVIOUSSTEPS=100000, READRESTART="true", INPUTRESTART="rs")
    want, got = _outputs(full), _outputs(split)
    assert sorted(got) == sorted(want)
    for name in want:
        assert got[name] == want[name], name


def test_executable_restart_checks_settings(tmp_path):
    """A restart file written under other settings is refused with the
    reference's message, and the run ends (exit status 0, as the reference)."""
    base = re.sub(r"(?m)^FREQSAVE .*$", "FREQSAVE 10", SINGLE_INP.format(over="true"))
    base = re.sub(r"(?m)^NUMBEROFRUNS .*$", "NUMBEROFRUNS 1", base)
    base = re.sub(r"(?m)^WRITEX .*$", "WRITEX false", base)
    exe = os.path.join(BIN, "langevinDynamicsSingleBarrier")
    first = re.sub(r"(?m)^WRITERESTART .*$", "WRITERESTART true", base).replace("OUTPUTRESTART        none",
                                                                             "OUTPUTRESTART        rs")
    first = re.sub(r"(?m)^NUMBEROFSTEPS .*$", "NUMBEROFSTEPS 100000", first)
    (tmp_path / "a.inp").write_text(first)
    r = subprocess.run([exe, "-i", "a.inp", "--ranks", "1", "--simd-width", "4"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0 and (tmp_path / "rs_00000_00000.restart").exists(), r.stderr[-2000:]
    second = re.sub(r"(?m)^READRESTART .*$", "READRESTART true", base).replace("INPUTRESTART         none",
                                                                           "INPUTRESTART         rs")
    second = re.sub(r"(?m)^TIMESTEP .*$", "TIMESTEP 0.002", second)
    (tmp_path / "b.inp").write_text(second)
    r = subprocess.run([exe, "-i", "b.inp", "--ranks", "1", "--simd-width", "4"], cwd=tmp_path, capture_output=True,
                       text=True, timeout=300)
    assert r.returncode == 0
    assert "Error: inconsistent time step0" in r.stderr and "Error: can't read restart file on rank" in r.stderr


def test_executable_continues_a_reference_restart(tmp_path):
    """READRESTART of the reference program's own restart files: the quick
    test's state after 3e9 steps (4 ranks x 4 lanes: positions, states, state
    counts, barrier tables, R_prev and all 10001 x frames, from the fixtures)
    with event lists drawn here, written in the reference's layout.  With
    NUMBEROFSTEPS 0 every output is the imported state's: the per-lane means of
    the lists, the reference's own msd.log and stateCount.log; with 600000 more
    steps the run continues from it (state counts grow by exactly the steps)."""
    import sys
    sys.path.insert(0, GOLD)
    import restart_decode as R
    st = json.load(open(os.path.join(GOLD, "quick_test_state.json")))
    frames_x = np.load(os.path.join(GOLD, "quick_test_frames.npz"))["x"]
    rng = np.random.default_rng(5)
    cg_sizes = [2, 3, 4, 5, 10, 100, 1000]

    def seq_mean(v):
        s = 0.0
        for x in v:
            s += x
        return s / len(v) if len(v) else 0.0

    expect = {k: [] for k in ("fpt", "fpt2", "tpt", "tpt2")}
    expect_cg = [[] for _ in cg_sizes]
    spacings = []
    for rank in range(4):
        rk = st["ranks"][rank]
        lanes = []
        for i in range(4):
            w = st["lanes"][rank * 4 + i]
            L = {"x": w["x"], "v": 0.0, "f": w["f"], "e": w["e"], "eta": 0.0,
                 "barrier": 7.2 if w["state"] == 0.0 else 2.0,
                 "traj_x": frames_x[rank * 4 + i], "current_state": w["state"], "previous_state": w["state"],
                 "slope_pos": 0.0, "slope_neg": 0.0, "current_height": 0.0,
                 "state_count0": w["state_count"][0], "state_count1": w["state_count"][1],
                 "well": w["well"], "periodic": w["periodic"], "well_prev": w["well"], "periodic_prev": w["periodic"],
                 "entering_well": 0, "entering_period": 0, "v0_instance": w["v0_instance"], "eta0_instance": 0.0,
                 "state0_instance": w["state0_instance"], "well_stored": 0.0, "periodic_stored": 0.0,
                 "well_prev_stored": 0.0, "periodic_prev_stored": 0.0,
                 "tick_prev": 0.0, "barrier_passed": 0,
                 "barrier_ticks": np.sort(rng.integers(0, 3 * 10**9, w["ticks_n"])).astype(np.float64),
                 "instant_barrier": [], "initial_barrier": [], "fpt_last_tick": w["fpt_last_tick"],
                 "fpt_to_update": w["fpt_to_update"], "fpt_init_state": (0.0, 0.0), "fpt_init_index": (0, 0),
                 "fpt2_last_tick": w["fpt2_last_tick"], "fpt2_to_update": 0, "fpt2_index": tuple(w["fpt2_index"]),
                 "tpt_last_tick": 0.0, "tpt_to_update": 1, "tpt2_last_tick": 0.0, "tpt2_to_update": 1,
                 "tpt2_initial_well": int(w["well"]), "cg_sizes": cg_sizes, "cg": []}
            for k in ("traj_v", "traj_f", "traj_pe", "traj_ke", "traj_eta", "traj_state", "traj_barrier"):
                L[k] = []
            for k in ("fpt", "fpt2", "tpt", "tpt2"):
                L[k] = rng.exponential(40.0, w[k + "_n"])
                expect[k].append(seq_mean(L[k]))
            for j, n in enumerate(w["cg_n"]):
                lst = rng.exponential(300.0, n)
                L["cg"].append({"list": lst, "last_tick": 0.0, "to_update": 0, "initial_well": int(w["well"])})
                expect_cg[j].append(seq_mean(lst))
            spacings += list(np.diff(L["barrier_ticks"]) * 1e-4)
            lanes.append(L)
        header = {"signature": 2345, "simd_width": 4, "x0": -2.9, "v0": 1.0, "gamma": 1.0, "kb": 1.0,
                  "temperature": 1.0, "mass": 1.0, "timestep": 1e-4, "nsteps_total": 3 * 10**9, "waiting_time": 1,
                  "stat_freq": 1, "save_freq": 300000, "save_flags": [1, 0, 0, 0], "randomize_v": 1,
                  "randomize_eta": 1, "periodic_length": 5.8, "h0": 7.2, "h1": 2.0, "initial_state": 0.0,
                  "randomize_state": 1, "same_initial_barrier": 0, "q": 0.2, "tau": 1e-5, "eta0": 0.0,
                  "save_flags2": [0, 0, 0], "rank": rank, "rng": [(0, 0)] * 3}
        R.encode({"header": header, "lanes": lanes, "barrier_state": np.array(rk["barrier_state"], np.int32),
                  "barrier_step": np.array(rk["barrier_step"], np.uint64), "barrier_min_id": rk["barrier_min_id"],
                  "r_prev": np.array(rk["r_prev"])}, str(tmp_path / f"ref_{rank:05d}_00000.restart"))

    base = open(os.path.join(GOLD, "quick_test.inp")).read()
    for k, v in dict(OUTPUTBASE="out", WRITEX="false", READRESTART="true", INPUTRESTART="ref",
                     WRITERESTART="false", NUMBEROFPREVIOUSSTEPS="3000000000").items():
        base = re.sub(rf"(?m)^{k} .*$", f"{k} {v}", base)
    exe = os.path.join(BIN, "langevinDynamicsSingleBarrier")

    def run(nsteps):
        (tmp_path / "run.inp").write_text(re.sub(r"(?m)^NUMBEROFSTEPS .*$", f"NUMBEROFSTEPS {nsteps}", base))
        print(f"[gpu exe reference restart] NUMBEROFSTEPS {nsteps}", flush=True)
        r = subprocess.run([exe, "-i", "run.inp", "--ranks", "4", "--simd-width", "4"], cwd=tmp_path,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0 and "Error" not in r.stderr, r.stderr[-2000:]

    run(0)
    for k, name in (("fpt", "firstPassageTime"), ("fpt2", "firstPassageTime2"), ("tpt", "transitionPathTime"),
                    ("tpt2", "transitionPathTime2")):
        assert np.fromfile(tmp_path / f"out_{name}.dat", "<f8").tolist() == expect[k], name
    for j, size in enumerate(cg_sizes):
        got = np.fromfile(tmp_path / f"out_firstPassageTimeCG2_{j:04d}_{size:04d}.dat", "<f8")
        assert got.tolist() == expect_cg[j], size
    msd = np.loadtxt(tmp_path / "out_msd.log", comments="#")
    np.testing.assert_allclose(msd, np.array(st["outputs"]["msd.log"]), rtol=2e-7, atol=0)
    sc = [ln for ln in (tmp_path / "out_stateCount.log").read_text().splitlines() if ln.startswith("# state")]
    assert [ln.split() for ln in sc] == [ln.split() for ln in st["outputs"]["stateCount.log"]]
    wt = [ln for ln in (tmp_path / "out_waitingTime.log").read_text().splitlines() if ln.startswith("# Nwt")]
    assert int(wt[0].split()[-1]) == len(spacings)

    run(600000)
    total = sum(sum(int(x) for x in ln.split()[3:4]) for ln in
                [ln for ln in (tmp_path / "out_stateCount.log").read_text().splitlines() if ln.startswith("# state")])
    assert total == sum(int(ln.split()[3]) for ln in st["outputs"]["stateCount.log"]) + 16 * 600000


def test_executable_writes_the_reference_restart_layout(tmp_path):
    """WRITERESTART of the single-barrier program writes the reference's own
    layout (TrajectoryData::WriteRestartFile, langevinDynamicsSingleBarrier.h:385-576):
    it decodes with the decoder pinned on the reference's recorded files, every
    lane's event lists average to the .dat values the run wrote, its x frames
    end at its position; this program's exact image follows the end marker.
    Cut after the end marker, the file alone restarts the run: with
    NUMBEROFSTEPS 0 the outputs are the writing run's."""
    import sys
    sys.path.insert(0, GOLD)
    import restart_decode as R
    base = SINGLE_INP.format(over="true")
    for k, v in dict(WRITERESTART="true", OUTPUTRESTART="rs", WRITEX="false", NUMBEROFRUNS="2").items():
        base = re.sub(rf"(?m)^{k} .*$", f"{k} {v}", base)
    ranks, W, trials = 2, 4, 2
    exe = os.path.join(BIN, "langevinDynamicsSingleBarrier")
    (tmp_path / "run.inp").write_text(base)
    print("[gpu exe reference layout] write", flush=True)
    r = subprocess.run([exe, "-i", "run.inp", "--ranks", str(ranks), "--simd-width", str(W)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "Error" not in r.stderr, r.stderr[-2000:]
    dat = {k: np.fromfile(tmp_path / f"out_{k}.dat", "<f8") for k in
           ("firstPassageTime", "firstPassageTime2", "transitionPathTime", "transitionPathTime2",
            "firstPassageTimeCG2_0000_0002")}

    def mean(v):
        s = 0.0
        for x in v:
            s += x
        return s / len(v) if len(v) else 0.0

    for rk in range(ranks):
        for tr in range(trials):
            path = tmp_path / f"rs_{rk:05d}_{tr:05d}.restart"
            d = R.decode(str(path))
            h = d["header"]
            assert (h["signature"], h["simd_width"], h["rank"], h["nsteps_total"]) == (2345, W, rk, 200000)
            assert d["trailer"] is not None
            for ch, L in enumerate(d["lanes"]):
                g = rk * trials * W + tr * W + ch      # rank-major, then trial, then lane
                assert mean(L["fpt"]) == dat["firstPassageTime"][g]
                assert mean(L["fpt2"]) == dat["firstPassageTime2"][g]
                assert mean(L["tpt"]) == dat["transitionPathTime"][g]
                assert mean(L["tpt2"]) == dat["transitionPathTime2"][g]
                assert mean(L["cg"][0]["list"]) == dat["firstPassageTimeCG2_0000_0002"][g]
                assert len(L["instant_barrier"]) == len(L["initial_barrier"]) == len(L["fpt"])
                assert set(np.unique(L["instant_barrier"])) <= {0.0, 1.0}
                assert len(L["traj_x"]) == 200000 // 100 + 1 and L["traj_x"][-1] == L["x"]
                assert L["state_count0"] + L["state_count1"] == 200001
            assert d["barrier_min_id"] <= -100 and d["barrier_state"].shape[0] >= 201
            assert set(np.unique(d["barrier_state"])) <= {0, 1}
    # the reference layout alone: cut the file after the end marker
    for f in os.listdir(tmp_path):
        if f.endswith(".restart"):
            d = R.decode(str(tmp_path / f))
            data = (tmp_path / f).read_bytes()
            (tmp_path / f).write_bytes(data[:len(data) - len(d["trailer"]) - 16])
            assert R.decode(str(tmp_path / f))["trailer"] is None
    before = {k: v.copy() for k, v in dat.items()}
    msd_before = np.loadtxt(tmp_path / "out_msd.log", comments="#")
    text = base
    for k, v in dict(WRITERESTART="false", READRESTART="true", INPUTRESTART="rs", NUMBEROFSTEPS="0",
                     NUMBEROFPREVIOUSSTEPS="200000").items():
        text = re.sub(rf"(?m)^{k} .*$", f"{k} {v}", text)
    (tmp_path / "run.inp").write_text(text)
    print("[gpu exe reference layout] read", flush=True)
    r = subprocess.run([exe, "-i", "run.inp", "--ranks", str(ranks), "--simd-width", str(W)], cwd=tmp_path,
                       capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "Error" not in r.stderr, r.stderr[-2000:]
    for k, v in before.items():
        np.testing.assert_array_equal(np.fromfile(tmp_path / f"out_{k}.dat", "<f8"), v, err_msg=k)
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "out_msd.log", comments="#"), msd_before)


@pytest.mark.parametrize("over", ["true", "false"])
def test_double_barrier_executable_writes_the_reference_restart_layout(tmp_path, over):
    """WRITERESTART of the double-barrier program writes the reference's own
    layout (TrajectoryData::WriteRestartFile, langevinDynamicsDoubleBarrier.h:414-642):
    per lane xp = x - periodic*L and the local-barrier index (xp < xc -> 0,
    Double.h:2198-2224) besides the single barrier's record, two barrier
    tables spanning at least the ids -10000 .. 10000 the reference creates at
    step 0 (:2250-2255), the marker 9948, R_prev when overdamped.  Every
    lane's FPT2 and FPTCG lists average to the .dat values the run wrote and
    its x frames end at its position.  Cut after the end marker, the file
    alone restarts the run: with NUMBEROFSTEPS 0 the outputs are the writing
    run's; 100000 further steps add exactly that many steps to the state
    counts."""
    import sys
    sys.path.insert(0, GOLD)
    import restart_decode as R
    base = DOUBLE_INP.format(over=over)
    for k, v in dict(WRITERESTART="true", OUTPUTRESTART="rs", NUMBEROFRUNS="2").items():
        base = re.sub(rf"(?m)^{k} .*$", f"{k} {v}", base)
    ranks, W, trials = 2, 4, 2
    overdamped = over == "true"
    exe = os.path.join(BIN, "langevinDynamicsDoubleBarrier")

    def run(text, what):
        (tmp_path / "run.inp").write_text(text)
        print(f"[gpu exe double reference layout] over={over} {what}", flush=True)
        r = subprocess.run([exe, "-i", "run.inp", "--ranks", str(ranks), "--simd-width", str(W)], cwd=tmp_path,
                           capture_output=True, text=True, timeout=300)
        assert r.returncode == 0 and "Error" not in r.stderr, r.stderr[-2000:]

    run(base, "write")
    names = ("firstPassageTime2", "firstPassageTimeCG2_0000_0002", "firstPassageTimeCG2_0001_0003")
    dat = {k: np.fromfile(tmp_path / f"dbl_{k}.dat", "<f8") for k in names}

    def mean(v):
        s = 0.0
        for x in v:
            s += x
        return s / len(v) if len(v) else 0.0

    counted = 0.0
    n_fpt2 = 0
    for rk in range(ranks):
        for tr in range(trials):
            d = R.decode(str(tmp_path / f"rs_{rk:05d}_{tr:05d}.restart"), overdamped=overdamped, model=1)
            h = d["header"]
            assert (h["signature"], h["simd_width"], h["rank"], h["nsteps_total"]) == (2345, W, rk, 200000)
            assert (h["h00"], h["h01"], h["h10"], h["h11"], h["position0"], h["position1"]) == \
                (1.0, 0.5, 2.0, 0.5, -1.42, -0.56)
            assert d["trailer"] is not None
            xc = h["xc"]
            for ch, L in enumerate(d["lanes"]):
                g = rk * trials * W + tr * W + ch      # rank-major, then trial, then lane
                assert mean(L["fpt2"]) == dat["firstPassageTime2"][g]
                assert mean(L["cg"][0]["list"]) == dat["firstPassageTimeCG2_0000_0002"][g]
                assert mean(L["cg"][1]["list"]) == dat["firstPassageTimeCG2_0001_0003"][g]
                n_fpt2 += len(L["fpt2"])
                assert L["xp"] == L["x"] - L["periodic"] * 5.8
                assert L["local"] == (0.0 if L["xp"] < xc else 1.0) == L["local_stored"]
                assert len(L["traj_x"]) == 200000 // 100 + 1 and L["traj_x"][-1] == L["x"]
                assert L["state_count0"] + L["state_count1"] == 200001
                s = int(L["current_state"])
                assert (L["current_height0"], L["current_height1"]) == ((1.0, 0.5)[s], (2.0, 0.5)[s])
                counted += L["state_count0"] + L["state_count1"]
            for a in ("", "1"):
                st = d["barrier_state" + a]
                assert d["barrier_min_id" + a] <= -10000 and st.shape[0] >= 20001
                assert set(np.unique(st)) <= {0, 1}
            assert (d["r_prev"] is not None) == overdamped
    assert n_fpt2 > 0, "no FPT2 events: the case does not exercise the lists"
    # the reference layout alone: cut every file after its end marker
    for f in os.listdir(tmp_path):
        if f.endswith(".restart"):
            d = R.decode(str(tmp_path / f), overdamped=overdamped, model=1)
            data = (tmp_path / f).read_bytes()
            (tmp_path / f).write_bytes(data[:len(data) - len(d["trailer"]) - 16])
            assert R.decode(str(tmp_path / f), overdamped=overdamped, model=1)["trailer"] is None
    msd_before = np.loadtxt(tmp_path / "dbl_msd.log", comments="#")
    resumed = base
    for k, v in dict(WRITERESTART="false", READRESTART="true", INPUTRESTART="rs", NUMBEROFSTEPS="0",
                     NUMBEROFPREVIOUSSTEPS="200000").items():
        resumed = re.sub(rf"(?m)^{k} .*$", f"{k} {v}", resumed)
    run(resumed, "read, 0 steps")
    for k, v in dat.items():
        np.testing.assert_array_equal(np.fromfile(tmp_path / f"dbl_{k}.dat", "<f8"), v, err_msg=k)
    np.testing.assert_array_equal(np.loadtxt(tmp_path / "dbl_msd.log", comments="#"), msd_before)
    run(re.sub(r"(?m)^NUMBEROFSTEPS .*$", "NUMBEROFSTEPS 100000", resumed), "read, 100000 steps")
    sc = [ln for ln in (tmp_path / "dbl_stateCount.log").read_text().splitlines() if ln.startswith("# state")]
    assert float(sc[0].split()[3]) + float(sc[1].split()[3]) == counted + ranks * trials * W * 100000
