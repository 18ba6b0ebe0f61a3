"""Write the golden fixtures of the quick test from the reference's own recorded run.

    python tests/golden/make_golden.py        # needs /root/reference (build container only)

Source: /root/reference/sample_input/quick_test/executed_job/ -- the outputs
the reference recorded for sample_input/quick_test/2.0_7.2_m1.0_1.00e-05.inp
(4 MPI ranks x 4 AVX2 lanes, overdamped single barrier, 3e9 steps).  The
fixtures are data only: decoded restart-file state, output-file values and
the per-rank seeds.  Nothing of the reference's code is copied.

Outputs (tests/golden/):
  quick_test_frames.npz   x frames 0..10000 of the 16 lanes, [rank*4+lane][frame]
  quick_test_state.json   per-lane end state, event counts and means, RNG
                          progress, barrier tables, output-file values
  quick_test_headers.json text of some recorded output files (layout checks)
  quick_test.inp          the quick test's input file
"""
import json
import os
import re
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import restart_decode  # noqa: E402

JOB = "/root/reference/sample_input/quick_test/executed_job"
BASE = "m1.0_t1.00e-05_0"


def read_log_values(path):
    """'# average / # sd / # count' blocks of an FPT/TPT .log file."""
    vals = []
    with open(path) as f:
        for line in f:
            m = re.match(r"#\s+(average|sd|count)\s+(\S+)", line)
            if m:
                vals.append((m.group(1), float(m.group(2))))
    return vals


def main():
    frames = []
    lanes = []
    ranks = []
    for rank in range(4):
        d = restart_decode.decode(os.path.join(JOB, f"m1.0_t1.00e-05_restart_0_{rank:05d}_00000.restart"))
        ranks.append({
            "rank": rank,
            "rng": d["header"]["rng"],
            "barrier_min_id": d["barrier_min_id"],
            "barrier_n": int(d["barrier_state"].shape[0]),
            "barrier_state": d["barrier_state"].tolist(),
            "barrier_step": [[int(v) for v in row] for row in d["barrier_step"]],
            "r_prev": d["r_prev"].tolist(),
        })
        for i, L in enumerate(d["lanes"]):
            frames.append(L["traj_x"])
            lanes.append({
                "rank": rank, "lane": i,
                "x": L["x"], "f": L["f"], "e": L["e"], "state": L["current_state"],
                "state_count": [L["state_count0"], L["state_count1"]],
                "well": L["well"], "periodic": L["periodic"],
                "v0_instance": L["v0_instance"], "state0_instance": L["state0_instance"],
                "fpt_n": len(L["fpt"]), "fpt_sum": float(np.sum(L["fpt"])), "fpt_first": L["fpt"][:20].tolist(),
                "fpt2_n": len(L["fpt2"]), "fpt2_first": L["fpt2"][:20].tolist(),
                "tpt_n": len(L["tpt"]), "tpt_first": L["tpt"][:20].tolist(),
                "tpt2_n": len(L["tpt2"]), "tpt2_first": L["tpt2"][:20].tolist(),
                "ticks_n": len(L["barrier_ticks"]), "ticks_first": L["barrier_ticks"][:20].tolist(),
                "fpt_last_tick": L["fpt_last_tick"], "fpt_to_update": L["fpt_to_update"],
                "fpt2_last_tick": L["fpt2_last_tick"], "fpt2_index": list(L["fpt2_index"]),
                "cg_n": [len(c["list"]) for c in L["cg"]],
                "cg_first": [c["list"][:10].tolist() for c in L["cg"]],
                "cg_sizes": L["cg_sizes"].tolist(),
                # FPT1 barrier logs: states of the barrier crossed / under the particle
                "inst_first": L["instant_barrier"][:20].tolist(), "init_first": L["initial_barrier"][:20].tolist(),
                "inst_n": len(L["instant_barrier"]), "init_n": len(L["initial_barrier"]),
                "inst_sum": float(sum(L["instant_barrier"].tolist())),
                "init_sum": float(sum(L["initial_barrier"].tolist())),
                "fpt_init_state": list(L["fpt_init_state"]), "fpt_init_index": list(L["fpt_init_index"]),
                # per-lane means exactly as the runner computes them (STATISTICS::AVERAGE)
                "mean_fpt": float(sum(L["fpt"].tolist()) / len(L["fpt"])) if len(L["fpt"]) else 0.0,
            })
    frames = np.array(frames)
    np.savez_compressed(os.path.join(HERE, "quick_test_frames.npz"), x=frames)
    outputs = {}
    for name in ("firstPassageTime", "firstPassageTime2", "transitionPathTime", "transitionPathTime2"):
        outputs[name + ".dat"] = np.fromfile(os.path.join(JOB, f"{BASE}_{name}.dat"), "<f8").tolist()
        outputs[name + ".log"] = read_log_values(os.path.join(JOB, f"{BASE}_{name}.log"))
    for j, n in enumerate([2, 3, 4, 5, 10, 100, 1000]):
        outputs[f"firstPassageTimeCG2_{j:04d}_{n:04d}.dat"] = np.fromfile(
            os.path.join(JOB, f"{BASE}_firstPassageTimeCG2_{j:04d}_{n:04d}.dat"), "<f8").tolist()
    outputs["firstPassageTimeCG2.log"] = read_log_values(os.path.join(JOB, f"{BASE}_firstPassageTimeCG2.log"))
    msd = np.loadtxt(os.path.join(JOB, f"{BASE}_msd.log"), comments="#")
    outputs["msd.log"] = msd.tolist()
    with open(os.path.join(JOB, f"{BASE}_stateCount.log")) as f:
        outputs["stateCount.log"] = [line.strip() for line in f if line.startswith("# state")]
    with open(os.path.join(JOB, f"{BASE}_waitingTime.log")) as f:
        outputs["waitingTime.log"] = [line.strip() for line in f if line.startswith("# R") or line.startswith("# Nwt")]
    with open(os.path.join(JOB, f"{BASE}_time.log")) as f:
        seeds = [int(m.group(1)) for m in re.finditer(r"Seed\s+(\d+)", f.read())]
    state = {"source": "sample_input/quick_test/executed_job", "seeds": seeds, "random_seed_master": 123,
             "ranks": ranks, "lanes": lanes, "outputs": outputs}
    with open(os.path.join(HERE, "quick_test_state.json"), "w") as f:
        json.dump(state, f, separators=(",", ":"))
    # text of a few output files (headers and bodies), for layout checks
    headers = {}
    for name in ("firstPassageTime.log", "msd.log", "stateCount.log", "waitingTime.log", "State0Distribution.log"):
        with open(os.path.join(JOB, f"{BASE}_{name}")) as f:
            headers[name] = f.read().splitlines()[:200]
    with open(os.path.join(HERE, "quick_test_headers.json"), "w") as f:
        json.dump(headers, f, indent=0)
    # the quick test's input file itself
    with open("/root/reference/sample_input/quick_test/2.0_7.2_m1.0_1.00e-05.inp") as f:
        inp = f.read()
    with open(os.path.join(HERE, "quick_test.inp"), "w") as f:
        f.write(inp)
    print("frames", frames.shape, "lanes", len(lanes))


if __name__ == "__main__":
    main()
