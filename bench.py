"""Throughput of the Langevin first-passage hot path on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--serve-seconds S]

Workload: the single-barrier model with 2^20 trajectories per parameter set,
Eb in {2, 3, 4, 5} with the barrier alternating between 3.6 Eb and Eb
(sample_input/CuIDB_single geometry: L = 5.8, w = 0.65, dt = 1e-4,
tau = 1e-5), underdamped and overdamped -- 8 sets, 2^23 trajectories in one
ragged batch -- starting at rest in the well centre X0 = -2.9, Philox seed
1000 + Eb index, step cap 1e8, absorbed trajectories retired after every
launch.  A bench "step" is one launch (steps_per_launch integration steps of
every live trajectory).  The FP32+compensated engine is timed (value) next to
the plain FP64 engine, both with the autotuned launch shape.  The roofline
carries the census ops of the timed launches and, from a rocprofv3 --pmc pass
run first, the VALU instructions the kernel actually issues.  Afterwards the
single-barrier Eb = 2 and Eb = 3 sets and the double barrier's Eb = 2 sets run
to completion in both arithmetics and their mean first-passage and transition
path times are compared.

Double barrier: 8 sets (Eb in {2,3,4,5} x {over, under}, sample_input/CuIDB_double
geometry) of 2^22 trajectories, FP32 throughput.  Ragged: 256 parameter sets
drawn from seed 7 (sweep.ragged_sets) run to completion in FP32 and FP64,
per-set statistics compared.

Serving: a Poisson stream of sweep requests (20 per second for --serve-seconds,
default 300, seed 11; single or double barrier, under- or overdamped,
Eb ~ U[2, 6.5] with ratio 3.6, 1e3 to 1e6 trajectories log-uniform, dt = 1e-3,
tau = 1e-4, cap 1e5 steps) into the FP32 request server with a 200 GB HBM
budget; p50 / p95 / p99 request latency and the aggregate trajectory-steps/s
are reported next to the plain FP64 engine running the same requests one at a
time.

CPU baseline: the reference's SIMD step loop restated with AVX-512 (or AVX2)
intrinsics (oracle/cpu_simd.c), one thread per host core, on a bounded sample
of the same 8 sets.
"""
import argparse
import json
import math
import os
import subprocess
import sys
import time

REPO = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, REPO)

# VALU lane-op rates of MI355X: single issue, 256 CU x 64 lanes/cycle x 2.4 GHz;
# the published 157.3 TFLOP/s FP32 vector peak is twice that (packed or
# dual-issued FMA lane-ops, 2 flops each)
P_ISSUE = 256 * 64 * 2.4e9
P_VALU = 157.3e12 / 2
# algorithmic VALU ops per trajectory-step (Philox4x32-10 40, Box-Muller 12,
# flip 2, and the integrator, indices, telegraph, force and events)
OPS_PER_STEP = {(0, True): 93, (0, False): 102, (1, True): 115, (1, False): 124}
EBS = (2.0, 3.0, 4.0, 5.0)
STEPS_PER_LAUNCH = 4096
COPY_BYTES = 1 << 28  # 256 MiB read + 256 MiB written per calibration copy


def workload_sets(n_traj, ebs=EBS):
    from langevindynamicsforpcp_amd import workloads as WL
    return WL.single_barrier(n_traj=n_traj, ebs=ebs)


def total_steps(sw, n):
    return sum(sw.stats(i).traj_steps for i in range(n))


def time_engine(precision, n_traj, steps, warmup, spl, tune=True):
    from langevindynamicsforpcp_amd.sweep import Sweep
    sets = workload_sets(n_traj)
    sw = Sweep(sets, precision=precision)
    cfg = sw.tune() if tune else None
    if spl:
        c = sw.config()
        sw.set_config(block=c.block, blocks_per_cu=c.blocks_per_cu, steps_per_launch=spl, use_graph=c.use_graph)
    for _ in range(warmup):
        sw.launch()
    per_set0 = [sw.stats(i).traj_steps for i in range(len(sets))]
    ms0, n0 = sw.kernel_time()
    t0 = time.perf_counter()
    for _ in range(steps):
        sw.launch()
    wall = time.perf_counter() - t0
    per_set1 = [sw.stats(i).traj_steps for i in range(len(sets))]
    ms1, n1 = sw.kernel_time()
    d = [b - a for a, b in zip(per_set0, per_set1)]
    ops = sum(di * OPS_PER_STEP[(s.model, bool(s.overdamped))] for di, s in zip(d, sets)) / max(sum(d), 1)
    c = sw.config()
    res = {"traj_steps": sum(d), "total_traj_steps": sum(per_set1), "total_launches": n1,
           "wall_s": wall, "kernel_ms": ms1 - ms0, "launches": n1 - n0, "ops_per_step": ops,
           "config": {"block": c.block, "blocks_per_cu": c.blocks_per_cu, "steps_per_launch": c.steps_per_launch,
                      "use_graph": c.use_graph, "tuned": cfg is not None}}
    sw.close()
    return res


def mfpt_sets():
    """The sets whose mean first-passage time is run to completion in both
    arithmetics: single barrier Eb = 2 and 3, over- and underdamped (the timed
    workload's geometry), and the double barrier's Eb = 2 sets (CuIDB_double
    geometry); trajectory counts sized so that every set finishes (absorbed or
    capped at 1e8 steps) within the bench's time."""
    from langevindynamicsforpcp_amd import workloads as WL
    sets = []
    for eb, over, n in ((2.0, True, 1 << 20), (2.0, False, 1 << 18), (3.0, True, 1 << 18), (3.0, False, 1 << 16)):
        s = workload_sets(n, ebs=(eb,))[0 if over else 1]
        sets.append(("single", s))
    for over, n in ((True, 1 << 18), (False, 1 << 16)):
        sets.append(("double", WL.double_barrier(n_traj=n, ebs=(2.0,), overdamped=(over,))[0]))
    return sets


def mfpt_check(precision, limit_s):
    """Runs mfpt_sets() in one sweep until every trajectory is absorbed or
    capped (or `limit_s` of wall time has passed); per set the MFPT, its
    standard error, the mean transition-path time and, for the double barrier,
    the mean transition-path time through each barrier."""
    from langevindynamicsforpcp_amd.sweep import Sweep
    named = mfpt_sets()
    sets = [s for _, s in named]
    sw = Sweep(sets, precision=precision)
    sw.tune()
    t0 = time.perf_counter()
    live = 1
    while live and time.perf_counter() - t0 < limit_s:
        live = sw.launch()
    wall = time.perf_counter() - t0
    out = []
    for i, (kind, s) in enumerate(named):
        m, e = sw.mfpt(i)
        st = sw.stats(i)
        n = max(st.n_absorbed, 1)
        r = {"model": kind, "eb": s.height[0][1] if kind == "single" else s.height[1][1],
             "overdamped": bool(s.overdamped), "n_traj": s.n_traj, "mfpt": m, "stderr": e,
             "n_absorbed": st.n_absorbed, "n_capped": st.n_capped,
             "unfinished": s.n_traj - st.n_absorbed - st.n_capped, "tpt_mean": st.tpt_sum / n}
        if kind == "double":
            r["tpt_barrier0_mean"] = st.tptb_sum[0] / n
            r["tpt_barrier1_mean"] = st.tptb_sum[1] / n
        out.append(r)
    sw.close()
    return out, wall


def compare_mfpt(f32, f64, t32, t64):
    """Per set: FP32 MFPT (and TPT means) against FP64's, relative and in
    units of the FP64 estimate's standard error."""
    fin = {"sets": [], "seconds": {"f32": t32, "f64": t64}, "cap_steps": 10 ** 8}
    for a, b in zip(f32, f64):
        r = {"model": a["model"], "eb": a["eb"], "overdamped": a["overdamped"], "n_traj": a["n_traj"],
             "mfpt_f32": a["mfpt"], "mfpt_f64": b["mfpt"], "rel_err": abs(a["mfpt"] - b["mfpt"]) / b["mfpt"],
             "z": abs(a["mfpt"] - b["mfpt"]) / b["stderr"] if b["stderr"] > 0 else None,
             "stderr_f64": b["stderr"], "tpt_f32": a["tpt_mean"], "tpt_f64": b["tpt_mean"],
             "capped": [a["n_capped"], b["n_capped"]], "unfinished": [a["unfinished"], b["unfinished"]]}
        if a["model"] == "double":
            for k in ("tpt_barrier0_mean", "tpt_barrier1_mean"):
                r[k] = [a[k], b[k]]
        fin["sets"].append(r)
    fin["max_rel_err"] = max(r["rel_err"] for r in fin["sets"])
    fin["all_finished"] = all(r["unfinished"] == [0, 0] for r in fin["sets"])
    return fin


def double_barrier(n_traj, steps, warmup):
    """Double-barrier sets (sample_input/CuIDB_double geometry: barrier 0
    between 1.7 Eb and Eb, barrier 1 between 3.6 Eb and Eb), Eb in {2,3,4,5}
    x {overdamped, underdamped}, n_traj each, FP32: throughput over `steps`
    launches after the autotuner and `warmup` launches, and the per-barrier
    transition paths of the trajectories absorbed by then."""
    from langevindynamicsforpcp_amd import workloads as WL
    from langevindynamicsforpcp_amd.sweep import Sweep
    sets = WL.double_barrier(n_traj=n_traj)
    sw = Sweep(sets, precision=32)
    c = sw.tune()
    for _ in range(warmup):
        sw.launch()
    s0 = [sw.stats(i).traj_steps for i in range(len(sets))]
    t0 = time.perf_counter()
    for _ in range(steps):
        sw.launch()
    wall = time.perf_counter() - t0
    st = [sw.stats(i) for i in range(len(sets))]
    sw.close()
    d = [b.traj_steps - a for a, b in zip(s0, st)]
    ops = sum(di * OPS_PER_STEP[(1, bool(x.overdamped))] for di, x in zip(d, sets)) / max(sum(d), 1)
    # statistics of this leg would be truncated after a few launches: the
    # double barrier's completed MFPT and per-barrier TPT are in the mfpt leg
    per_set = [{"eb": x.height[1][1], "overdamped": bool(x.overdamped), "absorbed_so_far": y.n_absorbed}
               for x, y in zip(sets, st)]
    return {"n_traj_per_set": n_traj, "sets": len(sets), "traj_steps_per_s": sum(d) / wall, "wall_s": wall,
            "ops_per_traj_step": ops, "launch": {"block": c.block, "blocks_per_cu": c.blocks_per_cu,
                                                  "steps_per_launch": c.steps_per_launch},
            "per_set": per_set}


def ragged(max_steps=200000, seed=7):
    """256 ragged parameter sets (sweep.ragged_sets, seed 7) run to completion
    (absorbed or capped at max_steps) in FP32 and in FP64, each with the
    autotuned launch shape; per set, the absorbed fraction and the MFPT of the
    absorbed trajectories are compared between the two."""
    import numpy as np
    from langevindynamicsforpcp_amd.sweep import Sweep, ragged_sets
    sets = ragged_sets(seed=seed, max_steps=max_steps)
    out, st = {}, {}
    for prec in (32, 64):
        sw = Sweep(sets, precision=prec)
        c = sw.tune()
        t0 = time.perf_counter()
        sw.run()
        wall = time.perf_counter() - t0
        st[prec] = [sw.stats(i) for i in range(len(sets))]
        sw.close()
        steps = sum(x.traj_steps for x in st[prec])
        out[f"f{prec}"] = {"traj_steps": steps, "wall_s_after_tune": wall,
                           "launch": {"block": c.block, "blocks_per_cu": c.blocks_per_cu,
                                      "steps_per_launch": c.steps_per_launch}}
    z_frac, rel_mfpt, z_mfpt = [], [], []
    for x, a, b in zip(sets, st[32], st[64]):
        n = x.n_traj
        pa, pb = a.n_absorbed / n, b.n_absorbed / n
        se = math.sqrt(max(pb * (1 - pb), 1.0 / n) / n)
        z_frac.append(abs(pa - pb) / se)
        if a.n_absorbed >= 1000 and b.n_absorbed >= 1000:
            ma, mb = a.fpt_sum / a.n_absorbed, b.fpt_sum / b.n_absorbed
            var = max(b.fpt_sum2 / b.n_absorbed - mb * mb, 0.0)
            rel_mfpt.append(abs(ma - mb) / mb)
            z_mfpt.append(abs(ma - mb) / max(math.sqrt(var / b.n_absorbed), 1e-300))
    return {"sets": len(sets), "trajectories": sum(x.n_traj for x in sets), "max_steps": max_steps, "seed": seed,
            **out, "fp32_over_fp64_wall": out["f64"]["wall_s_after_tune"] / out["f32"]["wall_s_after_tune"],
            "check_vs_fp64": {"max_z_absorbed_fraction": float(max(z_frac)),
                              "sets_with_mfpt": len(rel_mfpt),
                              "max_rel_err_mfpt": float(max(rel_mfpt)) if rel_mfpt else None,
                              "median_rel_err_mfpt": float(np.median(rel_mfpt)) if rel_mfpt else None,
                              "max_z_mfpt": float(max(z_mfpt)) if z_mfpt else None}}


# Requests of the serving stream (workloads.serving_requests): dt and tau of
# the single-barrier sets scaled by 10 (same dt/tau = 10), a step cap of 1e5
# steps = 100 time units.  At 20 requests/s and 1.45e5 trajectories per request
# on average (log-uniform in [1e3, 1e6]) the stream offers 2.9e6
# trajectories/s: a cap of 1e5 steps keeps the offered work (<= 2.9e11
# trajectory-steps/s) below the FP32 engine's rate, where the sweep's cap of
# 1e8 steps would offer ~400x what the device can do, and the latency
# percentiles would then measure an ever-growing queue.


def serving(seconds, budget_gb, rate=20.0, seed=11, fp64_limit_s=30.0, engine_rate=None):
    import numpy as np
    from langevindynamicsforpcp_amd import server as S
    from langevindynamicsforpcp_amd import workloads as WL
    reqs = WL.serving_requests(seconds, rate=rate, seed=seed)
    srv = S.Server(budget_bytes=int(budget_gb * 1e9), precision=32)
    # warm the graphs and kernels on a throwaway request before the clock starts
    srv.submit(reqs[0][1])
    while not srv.completed():
        srv.pump(1, wait=True)
    last = [time.perf_counter()]

    def progress(k, _):
        if time.perf_counter() - last[0] > 30.0:  # a line every ~30 s on long streams
            last[0] = time.perf_counter()
            print(f"[bench] serving: request {k + 1} of {len(reqs)} done", file=sys.stderr, flush=True)

    res = S.serve(srv, reqs, on_done=progress)
    info = srv.info()
    srv.close()
    steps = sum(r.traj_steps for r in res["results"])
    out = {"requests": len(reqs), "rate_per_s": rate, "duration_s": seconds, "seed": seed,
           "budget_gb": budget_gb, "wall_s": res["wall"], "traj_steps": steps,
           "traj_steps_per_s": steps / res["wall"], "latency_s": S.percentiles(res["latency"]),
           "latency_mean_s": float(np.mean(res["latency"])), "left_resident": info["resident"],
           "request": {"timestep": WL.SERVING_TIMESTEP, "tau": WL.SERVING_TAU, "max_steps": WL.SERVING_CAP,
                       "trajectories": "log-uniform in [1e3, 1e6]", "eb": "U[2, 6.5], ratio 3.6",
                       "cap_note": "cap 1e5 steps (100 time units at dt 1e-3): the stream offers at most "
                                   "2.9e6 trajectories/s x 1e5 steps = 2.9e11 trajectory-steps/s, under the FP32 "
                                   "engine's rate; the 8-set sweep's 1e8 cap would offer ~400x the device's rate"},
           "absorbed_fraction": float(sum(r.n_absorbed for r in res["results"]) /
                                      max(sum(r.n_absorbed + r.n_capped for r in res["results"]), 1))}
    if engine_rate:
        out["served_over_engine_peak"] = out["traj_steps_per_s"] / engine_rate
    # the plain FP64 engine, one job at a time, on the requests its time limit allows
    S.one_at_a_time(reqs[:1], precision=64)  # warm-up
    one = S.one_at_a_time(reqs, precision=64, limit_s=fp64_limit_s)
    n = len(one["latency"])
    steps64 = sum(r.traj_steps for r in one["results"])
    out["fp64_one_at_a_time"] = {"requests_run": n, "service_s": float(one["service"].sum()),
                                 "traj_steps_per_s": steps64 / max(float(one["service"].sum()), 1e-9),
                                 "latency_s": S.percentiles(one["latency"]) if n else None,
                                 "fp32_served_latency_s_same_requests": S.percentiles(res["latency"][:n]) if n else None}
    return out


def run_mode_c1():
    """The drop-in executable on the 32-trajectory case (tests/golden/c1_32traj.inp:
    underdamped single barrier 7.2/2.0, X0=-2.9 at rest, seed 1, 4 ranks x 8 lanes,
    1e7 steps, every statistic on): FP64 reference arithmetic, the ranks'
    handles running concurrently.  The integrate time is the one the program
    writes to c1_time.log."""
    import re
    import shutil
    import tempfile
    exe = os.path.join(REPO, "langevindynamicsforpcp_amd", "bin", "langevinDynamicsSingleBarrier")
    d = tempfile.mkdtemp(prefix="ldpcp_c1_")
    try:
        shutil.copyfile(os.path.join(REPO, "tests", "golden", "c1_32traj.inp"), os.path.join(d, "c1.inp"))
        t0 = time.perf_counter()
        r = subprocess.run([exe, "-i", "c1.inp", "--ranks", "4", "--simd-width", "8"], cwd=d, capture_output=True,
                           text=True, timeout=600)
        wall = time.perf_counter() - t0
        if r.returncode != 0:
            raise RuntimeError(r.stderr[-2000:])
        log = open(os.path.join(d, "c1_time.log")).read()
        integ = float(re.search(r"Integrate\s+([0-9.]+) s", log).group(1))
    finally:
        shutil.rmtree(d, ignore_errors=True)
    steps = 32 * 10**7
    return {"workload": "drop-in executable, single barrier underdamped, 4 ranks x 8 lanes x 1e7 steps, FP64",
            "integrate_s": integ, "wall_s": wall, "traj_steps_per_s": steps / integ,
            "reference_published_traj_steps_per_s": 1.12e8,
            "reference_note": "the reference's quick test on 4 CPU cores (16 trajectories x 3e9 overdamped steps in "
                              "7.13 min); it publishes no figure for this case"}


def host_cores():
    """Host cores this process may use: the affinity mask, the cgroup CPU
    quota where one is set, and at most 16 (one GPU's share of a shared host)."""
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    try:
        quota, period = open("/sys/fs/cgroup/cpu.max").read().split()[:2]
        if quota != "max":
            n = min(n, max(1, int(float(quota) / float(period))))
    except (OSError, ValueError):
        pass
    return max(1, min(16, n))


def cpu_model():
    try:
        for line in open("/proc/cpuinfo"):
            if line.startswith("model name"):
                return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(seconds=10.0, cap=200000):
    """The reference's SIMD CPU path restated (oracle/cpu_simd.c: W = 8 AVX-512
    or W = 4 AVX2 lanes per vector, the reference's step, telegraph and
    hand-over loop, Philox noise), one thread per host core, on a bounded
    sample of the same workload: the 8 parameter sets, threads spread over
    them, trajectories absorbed or capped at `cap` steps, ~`seconds` of wall
    time."""
    sys.path.insert(0, os.path.join(REPO, "tests"))
    import oracle_lib as ol
    lib = ol.load_simd()
    threads = host_cores()
    params, seeds = [], []
    for s in workload_sets(1):
        params.append(ol.make_params(model=0, overdamped=int(s.overdamped), x0=s.x0, v0=s.v0, randomize_v=0,
                                     periodic_length=s.periodic_length,
                                     heights=[[s.height[0][0], s.height[0][1]]] * 2,
                                     half_width=[s.half_width[0], s.half_width[1]], timestep=s.timestep, tau=s.tau,
                                     nsteps=cap, save_freq=cap, b_waiting=0, b_fpt=0, b_tpt=0, b_fptcg=0, absorb=1))
        seeds.append(s.seed)
    _, _, steps, done, wall = ol.simd_absorb(lib, params, seeds, seconds=seconds, threads=threads,
                                             want_paths=False)
    width = lib.ldo_simd_width()
    return {"value": steps / wall, "unit": "traj-steps/s", "cores": threads, "kind": "port",
            "per_core": steps / wall / threads, "simd_width": width,
            "isa": "AVX-512" if width == 8 else "AVX2+FMA", "cpu_model": cpu_model(),
            "host_cpus_visible": os.cpu_count(), "trajectories_finished": done,
            "sample": f"the reference's SIMD step loop restated (oracle/cpu_simd.c, FP64, {width} lanes per vector, "
                      f"Philox + IEEE Box-Muller noise, bit-exact with the oracle per trajectory), {threads} threads "
                      f"for {wall:.1f} s over the 8 parameter sets, trajectories absorbed or capped at {cap} steps",
            "reference_published_per_core": 2.8e7}


def pmc_child(args):
    """Counter-run body: four launches of the timed workload (FP32 engine, the
    default launch shape with --steps-per-launch steps), their trajectory-step
    counts written to --pmc-out, then a plain device-to-device copy of known
    size (the runtime's copy kernel) that calibrates the traffic counters."""
    import ctypes as C
    from langevindynamicsforpcp_amd.sweep import Sweep
    sets = workload_sets(args.n_traj)
    sw = Sweep(sets, precision=32)
    c = sw.config()
    sw.set_config(block=c.block, blocks_per_cu=c.blocks_per_cu, steps_per_launch=args.steps_per_launch,
                  use_graph=c.use_graph)
    per_launch, prev = [], total_steps(sw, len(sets))
    for _ in range(4):
        sw.launch()
        cur = total_steps(sw, len(sets))
        per_launch.append(cur - prev)
        prev = cur
    c = sw.config()
    sw.close()
    if args.pmc_out:
        with open(args.pmc_out, "w") as f:
            json.dump({"traj_steps_per_launch": per_launch,
                       "launch": {"block": c.block, "blocks_per_cu": c.blocks_per_cu,
                                  "steps_per_launch": c.steps_per_launch}}, f)
    hip = None
    for name in ("libamdhip64.so", "libamdhip64.so.7", "/opt/rocm/lib/libamdhip64.so"):
        try:
            hip = C.CDLL(name)
            break
        except OSError:
            continue
    if hip is None:
        raise RuntimeError("libamdhip64 not found")
    n = C.c_size_t(COPY_BYTES)
    src, dst = C.c_void_p(), C.c_void_p()
    for ptr in (src, dst):
        if hip.hipMalloc(C.byref(ptr), n) != 0:
            raise RuntimeError("hipMalloc failed")
    hip.hipMemset(src, 1, n)
    for _ in range(3):
        if hip.hipMemcpy(dst, src, n, 3) != 0:  # hipMemcpyDeviceToDevice
            raise RuntimeError("hipMemcpy failed")
    hip.hipDeviceSynchronize()
    hip.hipFree(src)
    hip.hipFree(dst)


def read_counter_csv(out_dir, counter):
    """{kernel name: [per-dispatch value in dispatch order]} from a rocprofv3 csv."""
    import csv
    import glob
    rows = []
    for path in glob.glob(os.path.join(out_dir, "**", "*counter_collection.csv"), recursive=True):
        with open(path) as f:
            for r in csv.DictReader(f):
                if r.get("Counter_Name") == counter:
                    rows.append((int(r["Dispatch_Id"]), r["Kernel_Name"], float(r["Counter_Value"])))
    rows.sort()
    out = {}
    for _, name, v in rows:
        out.setdefault(name, []).append(v)
    return out


def run_pmc(args, counters, name, timeout=240):
    """One rocprofv3 --pmc pass of the counter-run body; returns
    ({counter: {kernel: [per-dispatch values]}}, the body's launch record), or
    None if rocprofv3 is missing or the pass fails.  The raw csv stays under
    --pmc-dir when one is given."""
    import shutil
    import tempfile
    prof = shutil.which("rocprofv3")
    if prof is None:
        return None
    keep = bool(args.pmc_dir)
    d = os.path.join(args.pmc_dir, name) if keep else tempfile.mkdtemp(prefix="ldpcp_pmc_")
    os.makedirs(d, exist_ok=True)
    rec = os.path.join(d, "launches.json")
    cmd = [prof, "--pmc", *counters, "--kernel-include-regex", "engine_step|copyBuffer", "-d", d, "-o", "pmc",
           "-f", "csv", "--", sys.executable, os.path.abspath(__file__), "--pmc-child", "--pmc-out", rec,
           "--n-traj", str(args.n_traj), "--steps-per-launch", str(args.steps_per_launch or STEPS_PER_LAUNCH)]
    try:
        pr = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True, timeout=timeout,
                            check=False)
    except subprocess.TimeoutExpired:
        print(f"# pmc {name}: run timed out", file=sys.stderr, flush=True)
        return None
    vals = {c: read_counter_csv(d, c) for c in counters}
    try:
        launches = json.load(open(rec))
    except (OSError, ValueError):
        launches = None
    if not keep:
        shutil.rmtree(d, ignore_errors=True)
    n_step = [len([v for k, vs in vals[c].items() if "engine_step" in k for v in vs]) for c in counters]
    if launches is None or min(n_step) < 4:
        print(f"# pmc {name}: {n_step} step dispatches (rc {pr.returncode}); output tail:\n{pr.stdout[-1500:]}",
              file=sys.stderr, flush=True)
        return None
    return vals, launches


def step_values(vals, counter):
    """The engine step kernel's per-dispatch values of one counter."""
    return [v for k, vs in vals[counter].items() if "engine_step" in k for v in vs]


def pmc_traffic(args):
    """HBM bytes per launch of the FP32 step kernel from two rocprofv3 --pmc runs
    (FETCH_SIZE, then WRITE_SIZE, KiB per dispatch), each scaled by what the same
    counter reads for a plain copy of known size.  None if a run fails."""
    res = {}
    for counter in ("FETCH_SIZE", "WRITE_SIZE"):
        r = run_pmc(args, [counter], counter.lower())
        if r is None:
            return None
        vals, launches = r
        step = step_values(vals, counter)
        # the runtime's copy kernel also serves the engine's small host copies:
        # the calibration copies are the last three dispatches
        copy = [v for k, vs in vals[counter].items() if "copyBuffer" in k for v in vs][-3:]
        if not copy:
            return None
        factor = (COPY_BYTES / 1024.0) / (sum(copy) / len(copy))
        res[counter] = {"kib_per_launch": sum(step[-3:]) / 3, "copy_factor": factor,
                        "bytes_per_launch": sum(step[-3:]) / 3 * 1024.0 * factor}
    # what the launch must move: every live trajectory's state read once and
    # written once (FP32 single barrier: 56 B read -- live-list entry, set,
    # 5 state words, flags, 4 barrier steps, catch-up count, TPT start -- and
    # 48 B written)
    n_live = len(workload_sets(1)) * args.n_traj
    res["algorithmic_bytes_per_launch"] = {"read": 56 * n_live, "write": 48 * n_live}
    return res


VALU_COUNTERS = ("SQ_INSTS_VALU", "SQ_INSTS_SALU", "SQ_ACTIVE_INST_VALU", "SQ_ACTIVE_INST_VALU2",
                 "SQ_THREAD_CYCLES_VALU", "SQ_BUSY_CYCLES", "GRBM_GUI_ACTIVE")


def pmc_valu(args):
    """Instruction counts of the FP32 step kernel (one rocprofv3 --pmc pass):
    VALU and SALU wave-instructions per wavefront-step (64 trajectory-steps),
    the share of VALU issue cycles in which a second VALU instruction was
    active, and the share of lanes active, over the last three of four
    launches.  None if the pass fails."""
    r = run_pmc(args, list(VALU_COUNTERS), "valu")
    if r is None:
        return None
    vals, launches = r
    steps = sum(launches["traj_steps_per_launch"][-3:])
    tot = {c: sum(step_values(vals, c)[-3:]) for c in VALU_COUNTERS}
    wave_steps = steps / 64.0
    return {"valu_insts_per_wave_step": tot["SQ_INSTS_VALU"] / wave_steps,
            "salu_insts_per_wave_step": tot["SQ_INSTS_SALU"] / wave_steps,
            "valu2_over_valu_active": tot["SQ_ACTIVE_INST_VALU2"] / max(tot["SQ_ACTIVE_INST_VALU"], 1.0),
            "lanes_active": tot["SQ_THREAD_CYCLES_VALU"] / max(64.0 * tot["SQ_ACTIVE_INST_VALU"], 1.0),
            "traj_steps_counted": steps, "launch": launches["launch"],
            "per_dispatch": {c: step_values(vals, c)[-3:] for c in VALU_COUNTERS}}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=10)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--steps-per-launch", type=int, default=0, help="0 = the autotuner's choice")
    ap.add_argument("--n-traj", type=int, default=1 << 20, help="trajectories per parameter set")
    ap.add_argument("--serve-seconds", type=float, default=300.0)
    ap.add_argument("--serve-fp64-seconds", type=float, default=30.0,
                    help="service time given to the one-at-a-time FP64 comparison")
    ap.add_argument("--mfpt-limit", type=float, default=120.0,
                    help="wall-time limit of each arithmetic's MFPT run to completion")
    ap.add_argument("--serve-budget-gb", type=float, default=200.0)
    ap.add_argument("--no-finish", action="store_true", help="skip the MFPT check")
    ap.add_argument("--no-tune", action="store_true", help="default launch shape instead of the autotuner's")
    ap.add_argument("--no-serve", action="store_true")
    ap.add_argument("--no-double", action="store_true", help="skip the double-barrier sets (2^22 trajectories each)")
    ap.add_argument("--no-ragged", action="store_true", help="skip the 256 ragged parameter sets")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-traffic", action="store_true", help="skip the two rocprofv3 --pmc runs")
    ap.add_argument("--no-run-mode", action="store_true", help="skip the drop-in executable's 32-trajectory case")
    ap.add_argument("--pmc-child", action="store_true", help=argparse.SUPPRESS)
    ap.add_argument("--pmc-out", default="", help=argparse.SUPPRESS)
    ap.add_argument("--pmc-dir", default="", help="keep the counter passes' raw csv under this directory")
    ap.add_argument("--no-valu", action="store_true", help="skip the rocprofv3 --pmc instruction-count pass")
    args = ap.parse_args()
    if args.pmc_child:
        if not args.steps_per_launch:
            args.steps_per_launch = STEPS_PER_LAUNCH
        pmc_child(args)
        return
    world = int(os.environ.get("WORLD_SIZE", "1"))
    if args.gpus != 1 or world != 1:
        print("bench.py measures one GPU: the library drives a single MI355X; "
              "--gpus above 1 is not supported", file=sys.stderr)
        sys.exit(2)

    # counter runs first, while this process has not opened the GPU
    traffic = None if args.no_traffic else pmc_traffic(args)
    print(f"# traffic {traffic}", file=sys.stderr, flush=True)
    valu = None if args.no_valu else pmc_valu(args)
    print(f"# valu {valu}", file=sys.stderr, flush=True)

    from langevindynamicsforpcp_amd import _native
    lib = _native.load()
    import ctypes as C
    name = C.create_string_buffer(64)
    ncu = C.c_int32()
    hbm = C.c_int64()
    _native.check(lib.ldpcp_device_info(name, 64, C.byref(ncu), C.byref(hbm)), "ldpcp_device_info")

    r32 = time_engine(32, args.n_traj, args.steps, args.warmup, args.steps_per_launch, tune=not args.no_tune)
    print(f"# fp32 {r32}", file=sys.stderr, flush=True)
    r64 = time_engine(64, args.n_traj, args.steps, args.warmup, args.steps_per_launch, tune=not args.no_tune)
    print(f"# fp64 {r64}", file=sys.stderr, flush=True)
    v32 = r32["traj_steps"] / r32["wall_s"]
    v64 = r64["traj_steps"] / r64["wall_s"]
    k32 = r32["traj_steps"] / (r32["kernel_ms"] * 1e-3)
    k64 = r64["traj_steps"] / (r64["kernel_ms"] * 1e-3)

    fin = None
    if not args.no_finish:
        f32, t32 = mfpt_check(32, args.mfpt_limit)
        print(f"# mfpt f32 {t32:.1f} s", file=sys.stderr, flush=True)
        f64, t64 = mfpt_check(64, args.mfpt_limit)
        fin = compare_mfpt(f32, f64, t32, t64)
        print(f"# mfpt {fin}", file=sys.stderr, flush=True)

    dbl = None
    if not args.no_double:
        dbl = double_barrier(4 * args.n_traj, 3, 1)
        print(f"# double barrier {dbl}", file=sys.stderr, flush=True)
    rag = None
    if not args.no_ragged:
        rag = ragged()
        print(f"# ragged {rag}", file=sys.stderr, flush=True)

    run_mode = None
    if not args.no_run_mode:
        run_mode = run_mode_c1()
        print(f"# run mode {run_mode}", file=sys.stderr, flush=True)

    serve = None
    if not args.no_serve:
        serve = serving(args.serve_seconds, args.serve_budget_gb, fp64_limit_s=args.serve_fp64_seconds,
                        engine_rate=v32)
        print(f"# serving {serve}", file=sys.stderr, flush=True)

    if not args.steps_per_launch:
        args.steps_per_launch = r32["config"]["steps_per_launch"]
    ops = r32["ops_per_step"]
    issued = None
    if valu:
        # one wave64 VALU instruction per SIMD per 4 cycles is single issue
        single = ncu.value * 4 * 2.4e9 / 4
        rate = k32 / 64.0 * valu["valu_insts_per_wave_step"]
        issued = {"valu_insts_per_wave_step": valu["valu_insts_per_wave_step"],
                  "salu_insts_per_wave_step": valu["salu_insts_per_wave_step"],
                  "valu_wave_insts_per_s": rate, "frac_of_single_issue": rate / single,
                  "frac_of_co_issue": rate / (2.0 * single),
                  "valu2_over_valu_active": valu["valu2_over_valu_active"], "lanes_active": valu["lanes_active"],
                  "census_ops_per_issued_valu": ops / valu["valu_insts_per_wave_step"],
                  "counter_launch": valu["launch"], "counters": list(VALU_COUNTERS),
                  "note": "rocprofv3 --pmc pass of the same FP32 engine (default launch shape, last 3 of 4 "
                          "launches) before the GPU is opened here: VALU wave-instructions per wavefront-step "
                          "(64 trajectory-steps) times the timed kernel rate; single issue = CUs x 4 SIMDs x "
                          "2.4 GHz / 4 cycles per wave64 instruction, co-issue = twice that"}
    achieved = k32 * ops / 1e12
    peak = P_VALU / 1e12
    line = {
        "metric": "traj-steps/s/GPU FP32 vs FP64 at <=1% MFPT error; p50/p99 sweep latency",
        "value": v32,
        "unit": "traj-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": 1e3 * r32["wall_s"] / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (counter-based Philox noise, no input files)",
        "config": {"workload": "single-barrier first passage, 8 sets x 2^20 trajectories: Eb in {2,3,4,5} "
                               "(heights 3.6Eb/Eb) x {overdamped, underdamped}, L=5.8, w=0.65, dt=1e-4, "
                               "tau=1e-5, X0=-2.9 at rest, cap 1e8 steps, Philox seed 1000+Eb index",
                   "n_traj_per_set": args.n_traj, "launch": r32["config"], "launch_fp64": r64["config"],
                   "device": name.value.decode(), "compute_units": ncu.value},
        "fp64_value": v64,
        "fp32_over_fp64": v32 / v64,
        "kernel_traj_steps_per_s": {"f32": k32, "f64": k64},
        "roofline": {"bound": "valu", "achieved": achieved, "peak": peak, "unit": "Tops/s",
                     "frac": achieved / peak,
                     "frac_of_single_issue": achieved / (P_ISSUE / 1e12),
                     "traffic": (traffic["FETCH_SIZE"]["bytes_per_launch"] + traffic["WRITE_SIZE"]["bytes_per_launch"])
                     if traffic else None,
                     "traffic_detail": traffic,
                     "avg_launch_us": 1e3 * r32["kernel_ms"] / max(r32["launches"], 1),
                     "traj_steps_per_launch": r32["traj_steps"] / max(r32["launches"], 1),
                     "ops_per_traj_step": ops,
                     "issued": issued,
                     "note": "VALU bound (no MFMA; HBM traffic per trajectory-step ~0). achieved = census VALU "
                             "lane-ops per trajectory-step (FMA = 1 op; Philox 40, Box-Muller 12, flip 2, rest "
                             "39-72 by kind) x trajectory-steps / launch time (HIP events around each launch: "
                             "step, scan and scatter kernels). peak = the published 157.3 TFLOP/s FP32 vector "
                             "peak in FMA lane-ops (78.6 T/s), the dual-issue/packed ceiling; "
                             "frac_of_single_issue is against 256 CU x 64 lanes x 2.4 GHz (39.3 T/s)"},
    }
    if fin:
        line["mfpt"] = fin
    if dbl:
        line["double_barrier"] = dbl
    if rag:
        line["ragged"] = rag
    if run_mode:
        line["run_mode"] = run_mode
    if serve:
        line["serving"] = serve
    if not args.no_cpu_baseline:
        line["cpu_baseline"] = cpu_baseline()
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()
