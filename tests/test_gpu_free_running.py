"""Free-running run mode against the oracle driven by an independent generator.

The replay tests pin the run mode's arithmetic bit for bit on recorded noise.
Here each side draws its own noise: the GPU run mode its Philox4x32-10 streams,
the oracle's rank (the reference's bank-and-lane consumption order,
RANDOMNUMBER.cpp:157-340) numpy's PCG64 normal and uniform banks.  The two are
different realisations of the same process, so their statistics must agree:

* the pooled first-passage times (FPT1, the barrier-top arrivals from a well
  centre, Single.h:2356-2411) pass a two-sample Kolmogorov-Smirnov test at
  p > 0.01, and the MFPTs differ by less than 1.96 combined standard errors
  (the GPU's MFPT lies in the oracle's 95 % confidence interval, widened by
  the GPU estimate's own error);
* the transition-path times (TPT1, :2491-2530) pass the same KS test and
  their means lie within the same 95 % bound;
* the second-neighbour first passages (FPT2, :2413-2489) pass the KS test and
  their means differ by less than 3 combined standard errors (a secondary
  statistic of the same runs);
* the MSD over the saved frames agrees within 3 standard errors at every lag
  the reference's runner evaluates (GenerateLog10Scale, MOLUTILITY.cpp:59-73,
  20 lags per decade over the 2001 frames);
* each lane's fraction of steps with the barrier in state 1 sits at the
  telegraph's stationary 1/2 (EQ true): their mean is within 4 standard
  errors of it.

The double barrier keeps FPT2 and FPTCG only; there FPT2 takes the MFPT test
and the first FPTCG list the secondary one.  These are the north-star's
free-running criteria (the reference's 95 % CI and a KS p > 0.01), with the
oracle standing in for the reference program.
"""
import ctypes as C
import math

import numpy as np
import pytest

from langevindynamicsforpcp_amd.integrator import read_checkpoint
from streams import run_rank_recorded
from test_gpu_run import to_engine
from test_oracle_modes import make

pytestmark = pytest.mark.gpu

def _compare(a, b, what, z_max):
    from scipy.stats import ks_2samp
    assert len(a) > 1000 and len(b) > 1000, (what, len(a), len(b))
    p = ks_2samp(a, b).pvalue
    se = math.sqrt(a.var() / len(a) + b.var() / len(b))
    print(f"[free running] {what}: n {len(a)} / {len(b)}, mean {a.mean():.5f} / {b.mean():.5f}, "
          f"z {(a.mean() - b.mean()) / se:+.2f}, KS p {p:.3f}", flush=True)
    assert p > 0.01, what
    assert abs(a.mean() - b.mean()) < z_max * se, what


# lists compared per model: (name, GPU event list, oracle capture key and list index, z bound);
# the double barrier keeps FPT2 and FPTCG only (Double.h:2974-3006, 3103-3154)
COMPARED = {
    0: [("fpt", 0, ("fpt", None), 1.96), ("fpt2", 1, ("fpt2", None), 3.0), ("tpt", 2, ("tpt", None), 1.96)],
    1: [("fpt2", 1, ("fpt2", None), 1.96), ("fptcg0", 6, ("cg", 0), 3.0)],
}


def _lane_count(lane, key, j):
    return int(lane.cg_n[j]) if key == "cg" else int(getattr(lane, key + "_n"))


@pytest.mark.parametrize("name", ["single_over", "single_under", "double_over", "double_under"])
def test_gpu_free_running_statistics_match_oracle(oracle, name):
    nsteps, capn, W, ranks, n_gpu = 200000, 8192, 16, 32, 4096
    p = make(name, nsteps=nsteps, save_freq=100)
    lists = COMPARED[p.model]
    print(f"[free running] {name}: oracle {ranks} x {W} lanes, GPU {n_gpu} trajectories, {nsteps} steps", flush=True)
    ref = {k: [] for k, *_ in lists}
    ref_x = []
    for r in range(ranks):
        lanes, frames, s = run_rank_recorded(oracle, p, W, seed=5000 + 13 * r, cap=8, capn=capn)
        for k, _, (key, j), _ in lists:
            src = s["capture"][key] if j is None else s["capture"][key][j]
            ref[k] += [src[i, :min(_lane_count(lanes[i], key, j or 0), capn)] for i in range(W)]
        ref_x.append(frames[0].T)
    ref = {k: np.concatenate(v) for k, v in ref.items()}
    ref_x = np.concatenate(ref_x)        # [lane][frame]

    le = to_engine(p, n_gpu, seed=99)
    le.barrier_window = 4096
    le.event_capacity = capn
    le.Integrate(0)
    img = read_checkpoint(le.Checkpoint())
    for k, gl, _, z in lists:
        got = np.concatenate([img["events"][i][gl] for i in range(n_gpu)])
        _compare(got, ref[k], k, z)

    # MSD per trajectory over the saved frames (MeanSquareDisplacementBruteForce,
    # Single.h:24-45): the mean over trajectories within 3 standard errors at
    # every lag of the runner's log10 scale (lag 0 excluded)
    n_frames = ref_x.shape[1]
    buf = np.zeros(512, np.uint64)
    nl = oracle.ldo_log10_scale(p.timestep * p.save_freq, n_frames, 1000, 20, 0,
                                buf.ctypes.data_as(C.POINTER(C.c_uint64)), len(buf))
    lags = buf[:nl]
    assert nl > 40 and lags[0] == 1 and lags[-1] <= n_frames - 1
    m_gpu = le.msd(lags)
    m_ref = np.array([[np.mean((x[l:] - x[:-l]) ** 2) for l in lags.astype(int)] for x in ref_x])
    zs = []
    for k, l in enumerate(lags):
        a, b = m_gpu[:, k], m_ref[:, k]
        se = math.sqrt(a.var() / len(a) + b.var() / len(b))
        zs.append((a.mean() - b.mean()) / se)
    worst = int(np.argmax(np.abs(zs)))
    print(f"[free running] msd at {nl} lags: max |z| {abs(zs[worst]):.2f} at lag {int(lags[worst])}", flush=True)
    assert max(abs(z) for z in zs) < 3.0, (int(lags[worst]), zs[worst])

    # telegraph state: fraction of counted steps in state 1 (EQ true: 1/2)
    lanes = le.lanes()
    frac = np.array([x.state_count[1] / (x.state_count[0] + x.state_count[1]) for x in lanes])
    assert abs(frac.mean() - 0.5) < 4 * frac.std(ddof=1) / math.sqrt(len(frac)) + 1e-3, frac.mean()
    le.close()
