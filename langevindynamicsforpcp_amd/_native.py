"""ctypes binding of libldpcp.so (the C ABI in include/ldpcp.h).

The library is built in-tree by ``python -m langevindynamicsforpcp_amd.build``.
There is no fallback: if the library is missing or no GPU is usable, every
call raises.
"""
import ctypes as C
import os

PKG = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG, "lib", "libldpcp.so")

MAX_CG = 16
HIST_BINS = 256
MSD_LAGS = 64


class LdpcpError(RuntimeError):
    """Raised when a C-ABI call returns a non-zero status."""


class RunParams(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("overdamped", C.c_int32),
        ("x0", C.c_double), ("v0", C.c_double), ("gamma", C.c_double), ("kb", C.c_double),
        ("temperature", C.c_double), ("mass", C.c_double),
        ("randomize_v", C.c_int32), ("randomize_state", C.c_int32), ("same_initial_barrier", C.c_int32),
        ("periodic_length", C.c_double),
        ("height", (C.c_double * 2) * 2),
        ("half_width", C.c_double * 2),
        ("position", C.c_double * 2),
        ("timestep", C.c_double),
        ("nsteps", C.c_uint64), ("save_freq", C.c_uint64),
        ("tau", C.c_double), ("initial_state", C.c_double),
        ("equal_probability", C.c_int32), ("eq_constant", C.c_double),
        ("b_waiting", C.c_int32), ("b_fpt", C.c_int32), ("b_tpt", C.c_int32), ("b_fptcg", C.c_int32),
        ("n_cg", C.c_int32), ("cg_size", C.c_int32 * MAX_CG),
        ("save_all", C.c_int32),
        ("seed", C.c_uint64),
        ("barrier_window", C.c_int32),
        ("tick_capacity", C.c_uint64),
        ("event_capacity", C.c_uint64),
    ]


class LaneResult(C.Structure):
    _fields_ = [
        ("x", C.c_double), ("v", C.c_double), ("f", C.c_double), ("e", C.c_double),
        ("r_prev", C.c_double), ("state", C.c_double), ("barrier", C.c_double),
        ("state_count", C.c_double * 2),
        ("well", C.c_double), ("periodic", C.c_double), ("xp", C.c_double), ("local_barrier", C.c_double),
        ("v0_instance", C.c_double), ("state0_instance", C.c_double),
        ("fpt_sum", C.c_double), ("fpt_last_tick", C.c_double), ("fpt_n", C.c_uint64), ("fpt_to_update", C.c_int32),
        ("fpt2_sum", C.c_double), ("fpt2_last_tick", C.c_double), ("fpt2_n", C.c_uint64),
        ("fpt2_to_update", C.c_int32), ("fpt2_idx", C.c_int32 * 2),
        ("tpt_sum", C.c_double), ("tpt_last_tick", C.c_double), ("tpt_n", C.c_uint64), ("tpt_to_update", C.c_int32),
        ("tpt2_sum", C.c_double), ("tpt2_last_tick", C.c_double), ("tpt2_n", C.c_uint64),
        ("tpt2_to_update", C.c_int32), ("tpt2_initial_well", C.c_int32),
        ("cg_sum", C.c_double * MAX_CG), ("cg_last_tick", C.c_double * MAX_CG),
        ("cg_n", C.c_uint64 * MAX_CG), ("cg_to_update", C.c_int32 * MAX_CG), ("cg_initial_well", C.c_int32 * MAX_CG),
        ("wt_tick_prev", C.c_double), ("wt_passed", C.c_int32),
        ("wt_ticks", C.c_uint64), ("wt_last", C.c_double), ("wt_sum", C.c_double), ("wt_sum2", C.c_double),
        ("wt_max", C.c_double), ("wt_n", C.c_uint64),
        ("absorbed_step", C.c_uint64), ("tpt_start_step", C.c_uint64), ("steps_done", C.c_uint64),
        ("n_gauss", C.c_uint64), ("n_flip", C.c_uint64), ("n_catch", C.c_uint64),
        ("error", C.c_int32),
        ("fpt_init_index", C.c_int32 * 2), ("fpt_init_state", C.c_double * 2),
    ]


class SweepSet(C.Structure):
    _fields_ = [
        ("model", C.c_int32), ("overdamped", C.c_int32),
        ("x0", C.c_double), ("v0", C.c_double), ("gamma", C.c_double), ("kb", C.c_double),
        ("temperature", C.c_double), ("mass", C.c_double),
        ("randomize_v", C.c_int32), ("randomize_state", C.c_int32),
        ("initial_state", C.c_double),
        ("periodic_length", C.c_double),
        ("height", (C.c_double * 2) * 2),
        ("half_width", C.c_double * 2),
        ("position", C.c_double * 2),
        ("timestep", C.c_double), ("tau", C.c_double),
        ("equal_probability", C.c_int32), ("eq_constant", C.c_double),
        ("n_traj", C.c_uint64), ("max_steps", C.c_uint64), ("seed", C.c_uint64),
    ]


class SetStats(C.Structure):
    _fields_ = [
        ("n_traj", C.c_uint64), ("n_absorbed", C.c_uint64), ("n_capped", C.c_uint64),
        ("fpt_sum", C.c_double), ("fpt_sum2", C.c_double),
        ("tpt_sum", C.c_double), ("tpt_sum2", C.c_double),
        ("tptb_sum", C.c_double * 2), ("tptb_sum2", C.c_double * 2),
        ("traj_steps", C.c_uint64), ("n_finished", C.c_uint64),
        ("hist_lo", C.c_double), ("hist_hi", C.c_double),
        ("hist", C.c_uint64 * HIST_BINS),
        ("msd_step", C.c_uint64 * MSD_LAGS),
        ("msd_sum", C.c_double * MSD_LAGS),
        ("msd_n", C.c_uint64 * MSD_LAGS),
        ("hist_step", C.c_uint64 * HIST_BINS),
    ]


class EngineConfig(C.Structure):
    _fields_ = [("block", C.c_int32), ("blocks_per_cu", C.c_int32), ("steps_per_launch", C.c_uint32),
                ("use_graph", C.c_int32)]


class ServerConfig(C.Structure):
    _fields_ = [("budget_bytes", C.c_uint64), ("precision", C.c_int32), ("max_requests", C.c_int32),
                ("engine", EngineConfig), ("keep_fpt", C.c_int32)]


REQ_QUEUED, REQ_RUNNING, REQ_DONE = 0, 1, 2

CHECKPOINT_MAGIC = 0x324B43504350444C  # "LDPCPCK2"


class CheckpointHeader(C.Structure):
    _fields_ = [("magic", C.c_uint64), ("header_bytes", C.c_uint32), ("lane_bytes", C.c_uint32),
                ("n_traj", C.c_uint32), ("traj_offset", C.c_uint32), ("model", C.c_int32),
                ("overdamped", C.c_int32), ("barrier_window", C.c_int32), ("n_fields", C.c_int32),
                ("seed", C.c_uint64), ("steps_done", C.c_uint64), ("n_frames", C.c_uint64),
                ("save_freq", C.c_uint64), ("tick_capacity", C.c_uint64), ("timestep", C.c_double),
                ("bytes", C.c_uint64), ("event_capacity", C.c_uint64), ("n_lists", C.c_uint32),
                ("events_complete", C.c_uint32)]

# (name, restype, argtypes) for every entry point of include/ldpcp.h
_P = C.c_void_p
_SIGNATURES = [
    ("ldpcp_run_create", C.c_int, [C.POINTER(RunParams), C.c_uint32, C.c_uint32, C.POINTER(_P)]),
    ("ldpcp_run_set_replay", C.c_int, [_P, C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_double), C.c_uint64,
                                       C.POINTER(C.c_double), C.c_uint64, C.POINTER(C.c_int8), C.c_int32,
                                       C.c_int32]),
    ("ldpcp_run_integrate", C.c_int, [_P, C.c_uint64]),
    ("ldpcp_run_launch", C.c_int, [_P, C.c_uint64]),
    ("ldpcp_run_wait", C.c_int, [_P]),
    ("ldpcp_run_lanes", C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(LaneResult)]),
    ("ldpcp_run_n_frames", C.c_uint64, [_P]),
    ("ldpcp_run_frames", C.c_int, [_P, C.c_int32, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]),
    ("ldpcp_run_ticks", C.c_int, [_P, C.c_uint32, C.c_uint32, C.POINTER(C.c_double)]),
    ("ldpcp_run_msd", C.c_int, [_P, C.POINTER(C.c_uint64), C.c_uint32, C.POINTER(C.c_double)]),
    ("ldpcp_run_msd_noalloc", C.c_int, [_P, C.POINTER(C.c_uint64), C.c_uint32, C.POINTER(C.c_double)]),
    ("ldpcp_run_set_barrier_default", C.c_int, [_P, C.POINTER(C.c_double)]),
    ("ldpcp_run_checkpoint_size", C.c_int, [_P, C.POINTER(C.c_uint64)]),
    ("ldpcp_run_checkpoint", C.c_int, [_P, C.c_void_p, C.c_uint64]),
    ("ldpcp_run_resume", C.c_int, [_P, C.c_void_p, C.c_uint64]),
    ("ldpcp_run_barrier_table", C.c_int, [_P, C.c_int32, C.c_int32, C.c_int32, C.POINTER(C.c_int32),
                                          C.POINTER(C.c_uint64)]),
    ("ldpcp_run_destroy", C.c_int, [_P]),
    ("ldpcp_sweep_create", C.c_int, [C.POINTER(SweepSet), C.c_int32, C.c_int32, C.POINTER(_P)]),
    ("ldpcp_sweep_launch", C.c_int, [_P, C.c_uint32, C.POINTER(C.c_uint64)]),
    ("ldpcp_sweep_run", C.c_int, [_P, C.c_uint32]),
    ("ldpcp_sweep_stats", C.c_int, [_P, C.c_int32, C.POINTER(SetStats)]),
    ("ldpcp_sweep_set_config", C.c_int, [_P, C.POINTER(EngineConfig)]),
    ("ldpcp_sweep_get_config", C.c_int, [_P, C.POINTER(EngineConfig)]),
    ("ldpcp_sweep_tune", C.c_int, [_P, C.POINTER(EngineConfig)]),
    ("ldpcp_sweep_fpt", C.c_int, [_P, C.c_int32, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                  C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("ldpcp_sweep_kernel_time", C.c_int, [_P, C.POINTER(C.c_double), C.POINTER(C.c_uint64)]),
    ("ldpcp_sweep_destroy", C.c_int, [_P]),
    ("ldpcp_server_create", C.c_int, [C.POINTER(ServerConfig), C.POINTER(_P)]),
    ("ldpcp_server_footprint", C.c_int, [_P, C.POINTER(SweepSet), C.POINTER(C.c_uint64)]),
    ("ldpcp_server_submit", C.c_int, [_P, C.POINTER(SweepSet), C.POINTER(C.c_uint64)]),
    ("ldpcp_server_pump", C.c_int, [_P, C.c_uint32, C.c_int32]),
    ("ldpcp_server_status", C.c_int, [_P, C.c_uint64, C.POINTER(C.c_int32)]),
    ("ldpcp_server_completed", C.c_int, [_P, C.POINTER(C.c_uint64), C.c_int32, C.POINTER(C.c_int32)]),
    ("ldpcp_server_fpt", C.c_int, [_P, C.c_uint64] + [C.POINTER(C.c_uint64)] * 4),
    ("ldpcp_server_result", C.c_int, [_P, C.c_uint64, C.POINTER(SetStats)]),
    ("ldpcp_server_info", C.c_int, [_P, C.POINTER(C.c_uint64), C.POINTER(C.c_uint64), C.POINTER(C.c_uint64),
                                    C.POINTER(C.c_uint64), C.POINTER(C.c_uint64)]),
    ("ldpcp_server_destroy", C.c_int, [_P]),
    ("ldpcp_last_error", C.c_char_p, []),
    ("ldpcp_device_info", C.c_int, [C.c_char_p, C.c_int32, C.POINTER(C.c_int32), C.POINTER(C.c_int64)]),
]

_lib = None


def symbols():
    """Names of every entry point the binding expects the library to export."""
    return [s[0] for s in _SIGNATURES]


def load(path=LIB_PATH):
    """Load libldpcp.so; raises if it has not been built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(path):
        raise LdpcpError(f"{path} is missing: build it with `python -m langevindynamicsforpcp_amd.build`")
    lib = C.CDLL(path)
    for name, res, args in _SIGNATURES:
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    _lib = lib
    return lib


def check(rc, what):
    if rc != 0:
        msg = load().ldpcp_last_error().decode(errors="replace")
        raise LdpcpError(f"{what} failed ({rc}): {msg}")
