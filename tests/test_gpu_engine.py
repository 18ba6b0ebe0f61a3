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
    rel = 0 if exact_sums else 1e-12
    for f in ("fpt_sum", "fpt_sum2", "tpt_sum", "tpt_sum2"):
        assert getattr(sa, f) == pytest.approx(getattr(sb, f), rel=rel, abs=1e-300), f
    for k in range(2):
        assert sa.tptb_sum[k] == pytest.approx(sb.tptb_sum[k], rel=rel, abs=1e-300)
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
    srv = Server(budget_bytes=256 << 20, precision=32, steps_per_launch=1500)
    ids = []
    # requests join at different launches while others are resident
    for k, s in enumerate(sets):
        ids.append(srv.submit(s))
        srv.pump(k + 1)
    done = {}
    while len(done) < len(ids):
        srv.pump(4, wait=True)
        for rid in srv.completed():
            done[rid] = srv.result(rid)
    for s, rid in zip(sets, ids):
        (_, st_alone), = run_sweep([s], 32)
        st = done[rid]
        for f in ("n_traj", "n_absorbed", "n_capped", "n_finished", "traj_steps"):
            assert getattr(st, f) == getattr(st_alone, f), f
        assert list(st.hist) == list(st_alone.hist)
        assert list(st.msd_n) == list(st_alone.msd_n)
        assert st.fpt_sum == pytest.approx(st_alone.fpt_sum, rel=1e-12)
        assert st.tpt_sum == pytest.approx(st_alone.tpt_sum, rel=1e-12)
        np.testing.assert_allclose(np.array(st.msd_sum), np.array(st_alone.msd_sum), rtol=1e-9)
    info = srv.info()
    assert info["resident"] == 0 and info["queued"] == 0
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
            done[rid] = srv.result(rid)
    assert seen_queue
    for s, rid in zip(sets, ids):
        assert done[rid].n_finished == s.n_traj
    with pytest.raises(N.LdpcpError):
        srv.submit(sweep_set(model=0, n_traj=10 ** 6, max_steps=1000, **FAST))
    srv.close()
