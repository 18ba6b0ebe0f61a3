"""First-passage engine on the GPU: MSD sampling, launch-shape and batching
invariance, graph replay, and the request server (continuous batching).

* MSD at the 64 log-spaced lags of the FP64 engine equals the sum over the
  oracle's trajectories (same Philox noise, x saved every step) of
  (x(lag) - x0)^2 over trajectories not absorbed before the lag.
* Per-trajectory results do not depend on the launch shape (block size,
  persistent workgroups, steps per launch, graph replay or plain launches),
  on the autotuner, or on which other sets share the launches (wavefronts that
  straddle sets of different kinds included).
* A server request's results are the same alone and in a mix of requests
  admitted at different launches; the HBM budget bounds what is resident.
"""
import ctypes as C

import numpy as np
import pytest

import oracle_lib as ol
from langevindynamicsforpcp_amd import _native as N
from langevindynamicsforpcp_amd.server import Server
from langevindynamicsforpcp_amd.sweep import Sweep, sweep_set

pytestmark = pytest.mark.gpu

FAST = dict(x0=-1.05, periodic_length=2.1, half_width=(0.6, 0.6), timestep=1e-3, tau=1e-3)
DOUBLE = dict(x0=-2.9, periodic_length=5.8, half_width=(0.70, 0.65), position=(-1.42, -0.56), timestep=2e-3,
              tau=1e-3)


def mixed_sets(seed0=40, cap=60000):
    """Four kinds, odd sizes, so that wavefronts straddle sets and kinds."""
    return [
        sweep_set(model=0, overdamped=True, eb=1.0, ratio=3.0, n_traj=37, max_steps=cap, seed=seed0, **FAST),
        sweep_set(model=1, overdamped=False, heights=[[1.0, 0.5], [2.0, 0.5]], n_traj=101, max_steps=cap,
                  seed=seed0 + 1, randomize_v=True, **DOUBLE),
        sweep_set(model=0, overdamped=False, eb=0.8, ratio=3.0, n_traj=5, max_steps=cap, seed=seed0 + 2,
                  randomize_v=True, **FAST),
        sweep_set(model=1, overdamped=True, heights=[[1.0, 0.5], [2.0, 0.5]], n_traj=70, max_steps=cap,
                  seed=seed0 + 3, **DOUBLE),
        sweep_set(model=0, overdamped=True, eb=1.5, ratio=3.0, n_traj=300, max_steps=cap, seed=seed0 + 4, **FAST),
    ]


def run_sweep(sets, precision=32, config=None, tune=False, K=0):
    sw = Sweep(sets, precision=precision)
    if config:
        sw.set_config(**config)
    if tune:
        sw.tune()
    sw.run(K)
    out = [(sw.paths(i), sw.stats(i)) for i in range(len(sets))]
    sw.close()
    return out


def assert_same_set(a, b, exact_sums=False):
    (pa, sa), (pb, sb) = a, b
    for x, y in zip(pa, pb):
        assert np.array_equal(x, y)
    for f in ("n_traj", "n_absorbed", "n_capped", "n_finished", "traj_steps"):
        assert getattr(sa, f) == getattr(sb, f), f
    assert list(sa.hist) == list(sb.hist)
    assert list(sa.msd_n) == list(sb.msd_n)
    assert list(sa.msd_step) == list(sb.msd_step)
    # the duration sums are integer step sums on the device: exact in any order
    for f in ("fpt_sum", "fpt_sum2", "tpt_sum", "tpt_sum2"):
        assert getattr(sa, f) == getattr(sb, f), f
    for k in range(2):
        assert sa.tptb_sum[k] == sb.tptb_sum[k] and sa.tptb_sum2[k] == sb.tptb_sum2[k]
    np.testing.assert_allclose(np.array(sa.msd_sum), np.array(sb.msd_sum), rtol=1e-9, atol=0)


def test_msd_matches_oracle_frames(oracle):
    cap = 3000
    s = sweep_set(model=0, overdamped=False, eb=1.0, ratio=3.0, n_traj=64, max_steps=cap, seed=77,
                  randomize_v=True, **FAST)
    print("[gpu engine] msd vs oracle", flush=True)
    (paths, st), = run_sweep([s], precision=64)
    ab = paths[0]
    lags = np.array(st.msd_step, dtype=np.int64)
    valid = lags > 0
    p = ol.make_params(model=s.model, overdamped=s.overdamped, x0=s.x0, v0=s.v0, gamma=s.gamma, kb=s.kb,
                       temperature=s.temperature, mass=s.mass, randomize_v=s.randomize_v,
                       randomize_state=s.randomize_state, periodic_length=s.periodic_length,
                       heights=[[s.height[b][k] for k in range(2)] for b in range(2)],
                       half_width=[s.half_width[0], s.half_width[1]], position=[s.position[0], s.position[1]],
                       timestep=s.timestep, nsteps=cap, save_freq=1, tau=s.tau,
                       equal_probability=s.equal_probability, eq_constant=s.eq_constant, b_waiting=0, b_fpt=0,
                       b_tpt=0, b_fptcg=0, absorb=1)
    want = np.zeros(64)
    want_n = np.zeros(64, np.int64)
    for t in range(s.n_traj):
        out = ol.Lane()
        fr = np.zeros(cap + 1)
        oracle.ldo_lane_run_philox(C.byref(p), C.byref(out), s.seed, t, 0, 0, ol.dptr(fr), cap + 1)
        assert out.absorbed_step == ab[t]
        for k in np.nonzero(valid)[0]:
            lag = lags[k]
            if out.absorbed_step == 0 or out.absorbed_step >= lag:
                want[k] += (fr[lag] - s.x0) ** 2
                want_n[k] += 1
    assert list(st.msd_n) == list(want_n)
    np.testing.assert_allclose(np.array(st.msd_sum), want, rtol=1e-12, atol=1e-300)
    assert want_n[valid].min() >= 1


@pytest.mark.parametrize("precision", [32, 64])
def test_launch_shape_and_graph_do_not_change_results(precision):
    sets = mixed_sets()
    print(f"[gpu engine] launch shapes precision={precision}", flush=True)
    base = run_sweep(sets, precision)
    plain = run_sweep(sets, precision, config=dict(block=64, blocks_per_cu=3, steps_per_launch=777, use_graph=-1))
    graph = run_sweep(sets, precision, config=dict(block=128, blocks_per_cu=5, steps_per_launch=5000, use_graph=1))
    tuned = run_sweep(sets, precision, tune=True)
    for i in range(len(sets)):
        assert_same_set(base[i], plain[i])
        assert_same_set(base[i], graph[i])
        assert_same_set(base[i], tuned[i])
        assert base[i][1].n_finished == sets[i].n_traj


def test_sets_alone_and_together_agree():
    sets = mixed_sets(seed0=90)
    print("[gpu engine] alone vs together", flush=True)
    together = run_sweep(sets, 32)
    for i, s in enumerate(sets):
        alone = run_sweep([s], 32)
        assert_same_set(together[i], alone[0])


def test_server_results_same_alone_and_in_mix():
    sets = mixed_sets(seed0=120, cap=40000)
    print("[gpu engine] server mix", flush=True)
    srv = Server(budget_bytes=256 << 20, precision=32, steps_per_launch=1500, keep_fpt=True)
    ids = []
    # requests join at different launches while others are resident
    for k, s in enumerate(sets):
        ids.append(srv.submit(s))
        srv.pump(k + 1)
    n_of = dict(zip(ids, (s.n_traj for s in sets)))
    done, paths = {}, {}
    while len(done) < len(ids):
        srv.pump(4, wait=True)
        for rid in srv.completed():
            paths[rid] = srv.paths(rid, n_of[rid])
            done[rid] = srv.result(rid)
    for s, rid in zip(sets, ids):
        (p_alone, st_alone), = run_sweep([s], 32)
        # per-trajectory lists: absorption step, TPT start, xc crossings
        for x, y in zip(paths[rid], p_alone):
            assert np.array_equal(x, y)
        assert np.count_nonzero(paths[rid][0]) == st_alone.n_absorbed
        st = done[rid]
        for f in ("n_traj", "n_absorbed", "n_capped", "n_finished", "traj_steps"):
            assert getattr(st, f) == getattr(st_alone, f), f
        assert list(st.hist) == list(st_alone.hist)
        assert list(st.msd_n) == list(st_alone.msd_n)
        assert st.fpt_sum == st_alone.fpt_sum and st.fpt_sum2 == st_alone.fpt_sum2
        assert st.tpt_sum == st_alone.tpt_sum and st.tpt_sum2 == st_alone.tpt_sum2
        np.testing.assert_allclose(np.array(st.msd_sum), np.array(st_alone.msd_sum), rtol=1e-9)
    info = srv.info()
    assert info["resident"] == 0 and info["queued"] == 0
    # the lists go with the record: gone once the result is read
    with pytest.raises(N.LdpcpError):
        srv.paths(ids[0], n_of[ids[0]])
    srv.close()


def test_server_budget_bounds_resident_state():
    sets = [sweep_set(model=0, overdamped=True, eb=1.0, ratio=3.0, n_traj=20000, max_steps=20000, seed=300 + i,
                      **FAST) for i in range(5)]
    probe = Server(budget_bytes=64 << 20, precision=32)
    one = probe.footprint(sets[0])
    probe.close()
    budget = int(2.5 * one)  # two requests fit, not three
    srv = Server(budget_bytes=budget, precision=32, max_requests=8, steps_per_launch=2000)
    print(f"[gpu engine] server budget={budget} footprint={one}", flush=True)
    ids = [srv.submit(s) for s in sets]
    seen_queue = False
    done = {}
    while len(done) < len(ids):
        srv.pump(2, wait=True)
        info = srv.info()
        assert info["resident_bytes"] <= budget
        assert info["resident"] <= 2
        seen_queue |= info["queued"] > 0
        for rid in srv.completed():
            with pytest.raises(N.LdpcpError):  # the server keeps no per-trajectory lists
                srv.paths(rid, 20000)
            done[rid] = srv.result(rid)
    assert seen_queue
    for s, rid in zip(sets, ids):
        assert done[rid].n_finished == s.n_traj
    with pytest.raises(N.LdpcpError):
        srv.submit(sweep_set(model=0, n_traj=10 ** 6, max_steps=1000, **FAST))
    srv.close()


def test_ragged_batch_fp32_statistics_match_fp64():
    """A ragged batch of sets of every kind (single / double barrier, under /
    overdamped, Eb in [2, 6.5], tau over five decades, gamma over two): per set,
    the FP32 engine's absorbed fraction and MFPT agree with the FP64 engine's
    within 5 standard errors (same noise, so usually far closer)."""
    import math

    from langevindynamicsforpcp_amd.sweep import ragged_sets
    sets = ragged_sets(seed=7, n_sets=24, n_lo=256, n_hi=4096, max_steps=60000, timestep=1e-3)
    print("[gpu engine] ragged fp32 vs fp64", flush=True)
    res = {}
    for prec in (32, 64):
        sw = Sweep(sets, precision=prec)
        sw.tune()
        sw.run()
        res[prec] = [sw.stats(i) for i in range(len(sets))]
        sw.close()
    checked = 0
    for s, a, b in zip(sets, res[32], res[64]):
        assert a.n_finished == b.n_finished == s.n_traj
        n = s.n_traj
        pb = b.n_absorbed / n
        se = math.sqrt(max(pb * (1 - pb), 1.0 / n) / n)
        assert abs(a.n_absorbed / n - pb) <= 5 * se + 1e-12
        if a.n_absorbed >= 200 and b.n_absorbed >= 200:
            ma, mb = a.fpt_sum / a.n_absorbed, b.fpt_sum / b.n_absorbed
            sd = math.sqrt(max(b.fpt_sum2 / b.n_absorbed - mb * mb, 0.0) / b.n_absorbed)
            assert abs(ma - mb) <= 5 * sd + 1e-9 * mb
            checked += 1
    assert checked >= 4


def test_underdamped_32_trajectories_match_oracle(oracle):
    """The reference's CPU-sized case: 32 trajectories (4 ranks x 8 lanes) of
    the single-barrier underdamped model (barrier 7.2 / 2.0, L = 5.8,
    w = 0.65, dt = 1e-4, tau = 1e-5), at rest in the left well centre, run
    to first passage over the barrier or 1e7 steps, Philox seed 1, FP64
    engine.  Absorption step and TPT start of every trajectory equal the
    oracle's; the 256-bin FPT histogram holds every FPT in its bin; MSD at the
    lags up to 1e5 steps equals the oracle's frames."""
    cap, frames_cap = 10 ** 7, 100001
    s = sweep_set(model=0, overdamped=False, eb=2.0, ratio=3.6, n_traj=32, max_steps=cap, seed=1, x0=-2.9,
                  periodic_length=5.8, half_width=(0.65, 0.65), timestep=1e-4, tau=1e-5, randomize_v=False, v0=0.0)
    print("[gpu engine] 32 underdamped trajectories, barrier 7.2/2.0", flush=True)
    (paths, st), = run_sweep([s], precision=64)
    ab, tp = paths[0], paths[1]
    p = ol.make_params(model=0, overdamped=0, x0=-2.9, v0=0.0, randomize_v=0, periodic_length=5.8,
                       heights=[[7.2, 2.0], [7.2, 2.0]], half_width=[0.65, 0.65], timestep=1e-4, nsteps=cap,
                       save_freq=1, tau=1e-5, b_waiting=0, b_fpt=0, b_tpt=0, b_fptcg=0, absorb=1)
    lags = np.array(st.msd_step, dtype=np.int64)
    use = (lags > 0) & (lags < frames_cap)
    want = np.zeros(64)
    want_n = np.zeros(64, np.int64)
    fpts = []
    for t in range(32):
        out = ol.Lane()
        fr = np.zeros(frames_cap)
        oracle.ldo_lane_run_philox(C.byref(p), C.byref(out), 1, t, 0, 0, ol.dptr(fr), frames_cap)
        assert out.absorbed_step == ab[t], t
        if ab[t]:
            assert out.tpt_start_step == tp[t], t
            fpts.append(ab[t] * 1e-4)
        for k in np.nonzero(use)[0]:
            if out.absorbed_step == 0 or out.absorbed_step >= lags[k]:
                want[k] += (fr[lags[k]] - s.x0) ** 2
                want_n[k] += 1
    assert st.n_absorbed == len(fpts) and st.n_absorbed + st.n_capped == 32
    assert st.fpt_sum == pytest.approx(sum(fpts), rel=1e-12)
    assert list(np.array(st.msd_n)[use]) == list(want_n[use])
    np.testing.assert_allclose(np.array(st.msd_sum)[use], want[use], rtol=1e-12)
    # histogram: log10 bins over [log10 dt, log10 (dt cap)]
    edges = np.linspace(st.hist_lo, st.hist_hi, 257)
    hist = np.array(st.hist)
    assert hist.sum() == len(fpts)
    for f in fpts:
        b = int(np.searchsorted(edges, np.log10(f), side="right") - 1)
        near = [k for k in range(max(b - 1, 0), min(b + 2, 256)) if hist[k] > 0]
        assert near, f
