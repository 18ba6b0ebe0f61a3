"""Request server: continuous batching of sweep requests (ldpcp_server_* in include/ldpcp.h).

Requests wait in FIFO order until their HBM footprint fits the budget, then
join the trajectories already resident between two launches; a request
completes when its last trajectory is absorbed or capped.  Results of a
request are the same alone and in any mix (its noise is keyed by its own seed,
trajectory index and step).

``serve`` drives a server in real time against a list of timed arrivals and
measures each request's latency (completion time minus arrival time).
"""
import ctypes as C
import math
import time

import numpy as np

from . import _native as N
from .sweep import Sweep, sweep_set


class Server:
    def __init__(self, budget_bytes, precision=32, max_requests=0, block=0, blocks_per_cu=0, steps_per_launch=0,
                 use_graph=0, keep_fpt=False):
        cfg = N.ServerConfig()
        cfg.budget_bytes = int(budget_bytes)
        cfg.precision = int(precision)
        cfg.max_requests = int(max_requests)
        cfg.engine = N.EngineConfig(block, blocks_per_cu, steps_per_launch, use_graph)
        cfg.keep_fpt = 1 if keep_fpt else 0
        h = C.c_void_p()
        N.check(N.load().ldpcp_server_create(C.byref(cfg), C.byref(h)), "ldpcp_server_create")
        self._h = h

    def footprint(self, s):
        b = C.c_uint64()
        N.check(N.load().ldpcp_server_footprint(self._h, C.byref(s), C.byref(b)), "ldpcp_server_footprint")
        return b.value

    def submit(self, s):
        rid = C.c_uint64()
        N.check(N.load().ldpcp_server_submit(self._h, C.byref(s), C.byref(rid)), "ldpcp_server_submit")
        return rid.value

    def pump(self, n_launches=1, wait=False):
        N.check(N.load().ldpcp_server_pump(self._h, int(n_launches), 1 if wait else 0), "ldpcp_server_pump")

    def status(self, rid):
        st = C.c_int32()
        N.check(N.load().ldpcp_server_status(self._h, int(rid), C.byref(st)), "ldpcp_server_status")
        return st.value

    def completed(self):
        buf = (C.c_uint64 * 256)()
        n = C.c_int32()
        out = []
        while True:
            N.check(N.load().ldpcp_server_completed(self._h, buf, 256, C.byref(n)), "ldpcp_server_completed")
            out += list(buf[:n.value])
            if n.value < 256:
                return out

    def paths(self, rid, n_traj):
        """Per trajectory of a completed request (server made with keep_fpt):
        absorption step (0 = capped), last well-centre crossing before it,
        first / last crossing of the intermediate-well bottom (double barrier).
        Read before result(), which frees the request's record."""
        out = [np.zeros(int(n_traj), np.uint64) for _ in range(4)]
        ptr = [a.ctypes.data_as(C.POINTER(C.c_uint64)) for a in out]
        N.check(N.load().ldpcp_server_fpt(self._h, int(rid), *ptr), "ldpcp_server_fpt")
        return out

    def result(self, rid):
        st = N.SetStats()
        N.check(N.load().ldpcp_server_result(self._h, int(rid), C.byref(st)), "ldpcp_server_result")
        return st

    def info(self):
        v = [C.c_uint64() for _ in range(5)]
        N.check(N.load().ldpcp_server_info(self._h, *[C.byref(x) for x in v]), "ldpcp_server_info")
        keys = ("queued", "resident", "live_traj", "resident_bytes", "traj_steps")
        return dict(zip(keys, (x.value for x in v)))

    def close(self):
        if self._h is not None:
            N.load().ldpcp_server_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_requests(rate, duration, seed, n_min=1000, n_max=1000000, eb_lo=2.0, eb_hi=6.5, ratio=3.6,
                     max_steps=100000, timestep=1e-3, tau=1e-4):
    """Timed sweep requests: Poisson arrivals at `rate` per second over
    `duration` seconds; each request draws single or double barrier, under- or
    overdamped, Eb ~ U[eb_lo, eb_hi] (heights ratio*Eb / Eb) and a trajectory
    count log-uniform in [n_min, n_max]."""
    rng = np.random.default_rng(seed)
    t, out = 0.0, []
    while True:
        t += rng.exponential(1.0 / rate)
        if t >= duration:
            return out
        model = int(rng.integers(2))
        over = bool(rng.integers(2))
        eb = float(rng.uniform(eb_lo, eb_hi))
        n = int(round(math.exp(rng.uniform(math.log(n_min), math.log(n_max)))))
        seed_r = int(rng.integers(1, 2 ** 62))
        if model == 0:
            s = sweep_set(model=0, overdamped=over, eb=eb, ratio=ratio, n_traj=n, max_steps=max_steps, seed=seed_r,
                          x0=-2.9, periodic_length=5.8, half_width=(0.65, 0.65), timestep=timestep, tau=tau,
                          randomize_v=not over)
        else:
            # sample_input/CuIDB_double geometry, both barriers at ratio*Eb / Eb
            s = sweep_set(model=1, overdamped=over, heights=[[ratio * eb, eb], [ratio * eb, eb]], n_traj=n,
                          max_steps=max_steps, seed=seed_r, x0=-2.9, periodic_length=5.8, half_width=(0.70, 0.65),
                          position=(-1.42, -0.56), timestep=timestep, tau=tau, randomize_v=not over)
        out.append((t, s))


def serve(server, requests, launches_per_pump=1, idle_sleep=5e-4, on_done=None):
    """Runs `server` in real time against `requests` [(arrival seconds, set)].
    Returns per-request arrival / completion times (seconds from the start),
    results and the wall time; the latency clock starts at the arrival."""
    n = len(requests)
    arrival = np.array([a for a, _ in requests])
    done_at = np.full(n, np.nan)
    results = [None] * n
    rid_to_i = {}
    t0 = time.perf_counter()
    i = 0
    n_done = 0
    while n_done < n:
        now = time.perf_counter() - t0
        while i < n and arrival[i] <= now:
            rid_to_i[server.submit(requests[i][1])] = i
            i += 1
        if len(rid_to_i) > n_done:
            server.pump(launches_per_pump, wait=False)
        elif i < n:
            time.sleep(min(idle_sleep, max(0.0, arrival[i] - now)))
        for rid in server.completed():
            k = rid_to_i[rid]
            done_at[k] = time.perf_counter() - t0
            results[k] = server.result(rid)
            n_done += 1
            if on_done:
                on_done(k, results[k])
        if i == n and len(rid_to_i) == n and n_done < n:
            # everything submitted: keep pumping until the last completion is seen
            server.pump(launches_per_pump, wait=True)
    wall = time.perf_counter() - t0
    return {"arrival": arrival, "done": done_at, "latency": done_at - arrival, "results": results, "wall": wall}


def one_at_a_time(requests, precision=64, steps_per_launch=0, limit_s=None):
    """The same requests on the plain engine, one job at a time in arrival
    order: each request's service time is measured on the GPU, and latencies
    follow from the arrival times (a job starts when it has arrived and the
    previous one has finished).  With limit_s, stops after the request that
    takes the total service time past it."""
    service = []
    results = []
    for _, s in requests:
        if limit_s is not None and sum(service) > limit_s:
            break
        t = time.perf_counter()
        sw = Sweep([s], precision=precision)
        sw.run(steps_per_launch)
        results.append(sw.stats(0))
        sw.close()
        service.append(time.perf_counter() - t)
    arrival = np.array([a for a, _ in requests[:len(service)]])
    done = np.zeros(len(service))
    free = 0.0
    for k, a in enumerate(arrival):
        start = max(a, free)
        free = start + service[k]
        done[k] = free
    return {"arrival": arrival, "done": done, "latency": done - arrival, "service": np.array(service),
            "results": results, "wall": float(done[-1]) if len(done) else 0.0}


def percentiles(lat):
    lat = np.asarray(lat, float)
    return {f"p{q}": float(np.percentile(lat, q)) for q in (50, 95, 99)}
