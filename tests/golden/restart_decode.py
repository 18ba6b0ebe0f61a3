"""Decoder for the reference's binary ``.restart`` files (test infrastructure).

Layout follows langevinDynamicsSingleBarrier.h:1048-1136 (header),
:385-576 (TrajectoryData::WriteRestartFile), DataLocator.h:163-169
(barrier table) and :2776-2785 (overdamped R_prev tail).  All values are
little-endian: int = 4 B, size_t / long long = 8 B, double = 8 B.
"""
import struct
import numpy as np


class _Reader:
    def __init__(self, buf):
        self.buf = buf
        self.pos = 0

    def i32(self):
        v = struct.unpack_from("<i", self.buf, self.pos)[0]
        self.pos += 4
        return v

    def u64(self):
        v = struct.unpack_from("<Q", self.buf, self.pos)[0]
        self.pos += 8
        return v

    def i64(self):
        v = struct.unpack_from("<q", self.buf, self.pos)[0]
        self.pos += 8
        return v

    def f64(self):
        v = struct.unpack_from("<d", self.buf, self.pos)[0]
        self.pos += 8
        return v

    def vec(self, n, dtype="<f8"):
        a = np.frombuffer(self.buf, dtype=dtype, count=n, offset=self.pos).copy()
        self.pos += a.nbytes
        return a

    def sized_vec(self):
        return self.vec(self.u64())


def read_header(r, model=0):
    """The settings header (Single.h:1054-1101; the double barrier writes four
    heights, the barrier positions and the apex xc, Double.h:1232-1241)."""
    h = {}
    h["signature"] = r.i32()
    h["simd_width"] = r.i32()
    for k in ("x0", "v0", "gamma", "kb", "temperature", "mass", "timestep"):
        h[k] = r.f64()
    h["nsteps_total"] = r.u64()
    h["waiting_time"] = r.i32()
    h["stat_freq"] = r.u64()
    h["save_freq"] = r.u64()
    h["save_flags"] = [r.i32() for _ in range(4)]
    h["randomize_v"] = r.i32()
    h["randomize_eta"] = r.i32()
    if model == 0:
        keys = ("periodic_length", "h0", "h1", "initial_state")
    else:
        keys = ("periodic_length", "h00", "h01", "h10", "h11", "initial_state", "position0", "position1", "xc")
    for k in keys:
        h[k] = r.f64()
    h["randomize_state"] = r.i32()
    h["same_initial_barrier"] = r.i32()
    for k in ("q", "tau", "eta0"):
        h[k] = r.f64()
    h["save_flags2"] = [r.i32() for _ in range(3)]
    h["rank"] = r.i32()
    h["rng"] = [(r.i64(), r.u64()) for _ in range(3)]
    return h


def decode(path, overdamped=True, tpt=True, fptcg=True):
    """Decode one restart file into a dict of header fields and per-lane data."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    h = read_header(r)
    W = h["simd_width"]
    assert r.i32() == 9998
    b_wt = r.i32()
    b_fpt = r.i32()
    lanes = []
    for _ in range(W):
        L = {}
        for k in ("x", "v", "f", "e", "eta", "barrier"):
            L[k] = r.f64()
        for k in ("traj_x", "traj_v", "traj_f", "traj_pe", "traj_ke", "traj_eta", "traj_state", "traj_barrier"):
            L[k] = r.sized_vec()
        for k in ("current_state", "previous_state", "slope_pos", "slope_neg", "current_height",
                  "state_count0", "state_count1", "well", "periodic", "well_prev", "periodic_prev"):
            L[k] = r.f64()
        L["entering_well"] = r.i32()
        L["entering_period"] = r.i32()
        for k in ("v0_instance", "eta0_instance", "state0_instance", "well_stored", "periodic_stored",
                  "well_prev_stored", "periodic_prev_stored"):
            L[k] = r.f64()
        r.i32(); r.i32()
        if b_wt:
            L["tick_prev"] = r.f64()
            L["barrier_passed"] = r.i32()
            L["barrier_ticks"] = r.sized_vec()
        if b_fpt:
            L["fpt"] = r.sized_vec()
            L["instant_barrier"] = r.sized_vec()
            L["initial_barrier"] = r.sized_vec()
            L["fpt_last_tick"] = r.f64()
            L["fpt_to_update"] = r.i32()
            L["fpt_init_state"] = (r.f64(), r.f64())
            L["fpt_init_index"] = (r.i32(), r.i32())
            L["fpt2"] = r.sized_vec()
            L["fpt2_last_tick"] = r.f64()
            L["fpt2_to_update"] = r.i32()
            L["fpt2_index"] = (r.i32(), r.i32())
        if tpt:
            L["tpt"] = r.sized_vec()
            L["tpt_last_tick"] = r.f64()
            L["tpt_to_update"] = r.i32()
            L["tpt2"] = r.sized_vec()
            L["tpt2_last_tick"] = r.f64()
            L["tpt2_to_update"] = r.i32()
            L["tpt2_initial_well"] = r.i32()
        if fptcg:
            n = r.u64()
            L["cg_sizes"] = r.vec(n, "<i4")
            L["cg"] = []
            for _ in range(n):
                lst = r.sized_vec()
                L["cg"].append({"list": lst, "last_tick": r.f64(), "to_update": r.i32(), "initial_well": r.i32()})
        lanes.append(L)
    nb = r.u64()
    states = np.zeros((nb, W), np.int32)
    steps = np.zeros((nb, W), np.uint64)
    for b in range(nb):
        for i in range(W):
            states[b, i] = r.i32()
            steps[b, i] = r.u64()
    min_id = r.i32()
    tail = None
    if overdamped:
        tail = r.vec(W)
    assert r.i32() == 9990
    assert r.pos == len(r.buf), (r.pos, len(r.buf))
    return {"header": h, "lanes": lanes, "barrier_state": states, "barrier_step": steps,
            "barrier_min_id": min_id, "r_prev": tail}
