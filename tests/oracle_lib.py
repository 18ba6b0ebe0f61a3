"""ctypes view of the CPU oracle (oracle/ldpcp_oracle.c) for the tests.

The oracle is the checker: tests call it next to the product and compare.
"""
import ctypes as C
import os
import subprocess

import numpy as np

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE_DIR = os.path.join(REPO, "oracle")
ORACLE_SO = os.path.join(ORACLE_DIR, "_build", "libldpcp_oracle.so")
MKL_SO = os.path.join(ORACLE_DIR, "_build", "libldpcp_mkl_streams.so")
MKL_RT = "/opt/conda/lib/libmkl_rt.so.1"

MAX_CG = 16
MAX_W = 16


class Params(C.Structure):
    _fields_ = [
        ("model", C.c_int), ("overdamped", C.c_int),
        ("x0", C.c_double), ("v0", C.c_double), ("gamma", C.c_double), ("kb", C.c_double),
        ("temperature", C.c_double), ("mass", C.c_double),
        ("randomize_v", C.c_int), ("randomize_state", C.c_int), ("same_initial_barrier", C.c_int),
        ("periodic_length", C.c_double),
        ("height", (C.c_double * 2) * 2),
        ("half_width", C.c_double * 2),
        ("position", C.c_double * 2),
        ("timestep", C.c_double),
        ("nsteps", C.c_uint64), ("save_freq", C.c_uint64),
        ("tau", C.c_double), ("initial_state", C.c_double),
        ("equal_probability", C.c_int), ("eq_constant", C.c_double),
        ("b_waiting", C.c_int), ("b_fpt", C.c_int), ("b_tpt", C.c_int), ("b_fptcg", C.c_int),
        ("n_cg", C.c_int), ("cg_size", C.c_int * MAX_CG),
        ("save_all", C.c_int), ("absorb", C.c_int),
    ]


class Lane(C.Structure):
    _fields_ = [
        ("x", C.c_double), ("v", C.c_double), ("f", C.c_double), ("e", C.c_double),
        ("r_prev", C.c_double), ("state", C.c_double), ("barrier", C.c_double),
        ("state_count", C.c_double * 2),
        ("well", C.c_double), ("periodic", C.c_double), ("xp", C.c_double), ("local_barrier", C.c_double),
        ("v0_instance", C.c_double), ("state0_instance", C.c_double),
        ("fpt_sum", C.c_double), ("fpt_last_tick", C.c_double), ("fpt_n", C.c_uint64), ("fpt_to_update", C.c_int),
        ("fpt2_sum", C.c_double), ("fpt2_last_tick", C.c_double), ("fpt2_n", C.c_uint64),
        ("fpt2_to_update", C.c_int), ("fpt2_idx", C.c_int * 2),
        ("tpt_sum", C.c_double), ("tpt_last_tick", C.c_double), ("tpt_n", C.c_uint64), ("tpt_to_update", C.c_int),
        ("tpt2_sum", C.c_double), ("tpt2_last_tick", C.c_double), ("tpt2_n", C.c_uint64),
        ("tpt2_to_update", C.c_int), ("tpt2_initial_well", C.c_int),
        ("cg_sum", C.c_double * MAX_CG), ("cg_last_tick", C.c_double * MAX_CG),
        ("cg_n", C.c_uint64 * MAX_CG), ("cg_to_update", C.c_int * MAX_CG), ("cg_initial_well", C.c_int * MAX_CG),
        ("wt_tick_prev", C.c_double), ("wt_passed", C.c_int),
        ("wt_ticks", C.c_uint64), ("wt_last", C.c_double), ("wt_sum", C.c_double), ("wt_sum2", C.c_double),
        ("wt_max", C.c_double), ("wt_n", C.c_uint64),
        ("absorbed_step", C.c_uint64), ("tpt_start_step", C.c_uint64),
        ("xc_first_step", C.c_uint64), ("xc_last_step", C.c_uint64), ("steps_done", C.c_uint64),
        ("n_gauss", C.c_uint64), ("n_flip", C.c_uint64), ("n_catch", C.c_uint64),
        ("error", C.c_int),
        ("fpt_init_index", C.c_int * 2), ("fpt_init_state", C.c_double * 2),
    ]


class Capture(C.Structure):
    _fields_ = [("cap", C.c_uint64),
                ("fpt", C.POINTER(C.c_double)), ("fpt2", C.POINTER(C.c_double)),
                ("tpt", C.POINTER(C.c_double)), ("tpt2", C.POINTER(C.c_double)),
                ("ticks", C.POINTER(C.c_double)), ("cg", C.POINTER(C.c_double)),
                ("inst", C.POINTER(C.c_double)), ("init", C.POINTER(C.c_double))]


class Record(C.Structure):
    _fields_ = [("cap", C.c_uint64),
                ("gauss", C.POINTER(C.c_double)), ("flip", C.POINTER(C.c_double)),
                ("catchup", C.POINTER(C.c_double)),
                ("bmin", C.c_int), ("bn", C.c_int), ("binit", C.POINTER(C.c_int8))]


FILL_FN = C.CFUNCTYPE(None, C.c_void_p, C.c_int, C.POINTER(C.c_double), C.c_uint64)


def build_oracle():
    """Compile the oracle (and, where an MKL runtime exists, the MKL feeder)."""
    subprocess.check_call(["make", "-s", "-C", ORACLE_DIR])


def load():
    if not os.path.exists(ORACLE_SO):
        build_oracle()
    lib = C.CDLL(ORACLE_SO)
    P, D, U64, I = C.c_void_p, C.POINTER(C.c_double), C.c_uint64, C.c_int
    lib.ldo_rank_create.restype = P
    lib.ldo_rank_create.argtypes = [C.POINTER(Params), I, C.c_void_p, P, U64]
    lib.ldo_rank_destroy.argtypes = [P]
    lib.ldo_rank_integrate.argtypes = [P, C.POINTER(Lane), D, U64, C.POINTER(U64), C.POINTER(Capture),
                                       C.POINTER(Record)]
    lib.ldo_rank_rng_progress.argtypes = [P, C.POINTER(C.c_int64), C.POINTER(U64)]
    lib.ldo_rank_barriers.restype = I
    lib.ldo_rank_barriers.argtypes = [P, C.POINTER(I), C.POINTER(C.c_int32), C.POINTER(U64), I, I]
    lib.ldo_lane_run_replay.argtypes = [C.POINTER(Params), C.POINTER(Lane), D, U64, D, U64, D, U64,
                                        C.POINTER(C.c_int8), I, I, D, U64, C.POINTER(Capture), I]
    lib.ldo_lane_run_philox.argtypes = [C.POINTER(Params), C.POINTER(Lane), U64, C.c_uint32, U64, U64, D, U64]
    lib.ldo_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
    lib.ldo_trace_handovers.argtypes = [C.c_void_p, U64]
    lib.ldo_trace_count.restype = U64
    lib.ldo_double_apex.argtypes = [C.POINTER(Params), D, D]
    lib.ldo_log.restype = C.c_double
    lib.ldo_log.argtypes = [C.c_double]
    lib.ldo_sincos_turn.argtypes = [C.c_double, D, D]
    lib.ldo_flip_threshold.restype = U64
    lib.ldo_flip_threshold.argtypes = [C.c_double]
    lib.ldo_log10_scale.restype = U64
    lib.ldo_log10_scale.argtypes = [C.c_double, U64, U64, U64, I, C.POINTER(U64), U64]
    lib.ldo_msd_brute.argtypes = [D, U64, C.POINTER(U64), U64, D]
    lib.ldo_average.restype = C.c_double
    lib.ldo_average.argtypes = [D, U64]
    lib.ldo_rmsd.restype = C.c_double
    lib.ldo_rmsd.argtypes = [D, U64]
    return lib


def dptr(a):
    return a.ctypes.data_as(C.POINTER(C.c_double))


def make_params(**kw):
    """Params with the reference runner's defaults for the hard-wired keys."""
    p = Params()
    defaults = dict(model=0, overdamped=1, x0=-2.9, v0=1.0, gamma=1.0, kb=1.0, temperature=1.0, mass=1.0,
                    randomize_v=1, randomize_state=1, same_initial_barrier=0, periodic_length=5.8,
                    timestep=1e-4, nsteps=10000, save_freq=10, tau=1e-5, initial_state=0.0,
                    equal_probability=1, eq_constant=1.0, b_waiting=1, b_fpt=1, b_tpt=1, b_fptcg=1,
                    save_all=0, absorb=0)
    heights = kw.pop("heights", [[7.2, 2.0], [7.2, 2.0]])
    half_width = kw.pop("half_width", [0.65, 0.65])
    position = kw.pop("position", [0.0, 0.0])
    cg = kw.pop("cg", [2, 3, 4, 5, 10, 100, 1000])
    defaults.update(kw)
    for k, v in defaults.items():
        setattr(p, k, v)
    for b in range(2):
        for s in range(2):
            p.height[b][s] = heights[b][s]
        p.half_width[b] = half_width[b]
        p.position[b] = position[b]
    p.n_cg = len(cg)
    for j, n in enumerate(cg):
        p.cg_size[j] = n
    return p


class MklStreams:
    """MKL MT19937 streams for one reference rank (container-only)."""

    def __init__(self, seeds):
        if not os.path.exists(MKL_SO):
            build_oracle()
        self.lib = C.CDLL(MKL_SO)
        self.lib.ldm_available.restype = C.c_int
        self.lib.ldm_available.argtypes = [C.c_char_p]
        if not self.lib.ldm_available(MKL_RT.encode()):
            raise RuntimeError("MKL runtime not available")
        self.lib.ldm_create.restype = C.c_void_p
        self.lib.ldm_create.argtypes = [C.c_uint, C.c_uint, C.c_uint]
        self.lib.ldm_destroy.argtypes = [C.c_void_p]
        self.lib.ldm_fill_ptr.restype = C.c_void_p
        self.ctx = self.lib.ldm_create(*[s & 0xFFFFFFFF for s in seeds])
        self.fill = self.lib.ldm_fill_ptr()

    def close(self):
        if self.ctx:
            self.lib.ldm_destroy(self.ctx)
            self.ctx = None


def mkl_available():
    return os.path.exists(MKL_RT)


SIMD_SO = {8: os.path.join(ORACLE_DIR, "_build", "libldpcp_simd512.so"),
           4: os.path.join(ORACLE_DIR, "_build", "libldpcp_simd256.so")}


def host_simd_width():
    """8 where the host has AVX-512F and -DQ, 4 where it has AVX2 and FMA, else 0."""
    try:
        flags = set()
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    flags = set(line.split(":", 1)[1].split())
                    break
    except OSError:
        return 0
    if {"avx512f", "avx512dq"} <= flags:
        return 8
    if {"avx2", "fma"} <= flags:
        return 4
    return 0


def load_simd(width=None):
    """The SIMD CPU baseline (oracle/cpu_simd.c) built for `width` lanes
    (default: the widest the host runs)."""
    width = width or host_simd_width()
    if width not in SIMD_SO:
        raise RuntimeError("the host has neither AVX-512 nor AVX2+FMA")
    if not os.path.exists(SIMD_SO[width]):
        build_oracle()
    lib = C.CDLL(SIMD_SO[width])
    U64P = C.POINTER(C.c_uint64)
    lib.ldo_simd_width.restype = C.c_int
    lib.ldo_simd_absorb.restype = C.c_int
    lib.ldo_simd_absorb.argtypes = [C.POINTER(Params), U64P, C.c_int, C.c_uint32, C.c_double, C.c_int,
                                    U64P, U64P, U64P, U64P, C.POINTER(C.c_double)]
    return lib


def simd_absorb(lib, params, seeds, n_traj=0, seconds=0.0, threads=1, want_paths=True):
    """Run ldo_simd_absorb over parameter sets `params` (list of Params) with
    Philox seeds `seeds`; returns (absorbed[set][traj], tpt0[...], traj_steps,
    traj_done, wall)."""
    n = len(params)
    P = (Params * n)(*params)
    S = (C.c_uint64 * n)(*seeds)
    ab = np.zeros((n, max(n_traj, 1)), dtype=np.uint64)
    tp = np.zeros_like(ab)
    steps, done, wall = C.c_uint64(), C.c_uint64(), C.c_double()
    u64p = C.POINTER(C.c_uint64)
    rc = lib.ldo_simd_absorb(P, S, n, n_traj, seconds, threads,
                             ab.ctypes.data_as(u64p) if want_paths and n_traj else None,
                             tp.ctypes.data_as(u64p) if want_paths and n_traj else None,
                             C.byref(steps), C.byref(done), C.byref(wall))
    if rc != 0:
        raise RuntimeError(f"ldo_simd_absorb returned {rc}")
    return ab, tp, steps.value, done.value, wall.value


def step_sums(dt, n):
    """(sum, sum of squares) of durations n (in steps) as the engine reports
    them: dt * sum(n) and dt^2 * sum(n^2), the integer sums rounded once."""
    n = [int(x) for x in n]
    s1, s2 = sum(n), sum(x * x for x in n)
    return dt * float(s1), dt * dt * (float(s2 >> 64) * 18446744073709551616.0 + float(s2 & (2 ** 64 - 1)))
