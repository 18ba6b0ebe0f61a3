"""Decoder for the reference's binary ``.restart`` files (test infrastructure).

Layout follows langevinDynamicsSingleBarrier.h:1048-1136 (header),
:385-576 (TrajectoryData::WriteRestartFile), DataLocator.h:163-169
(barrier table) and :2776-2785 (overdamped R_prev tail).  The double-barrier
program (model=1) writes langevinDynamicsDoubleBarrier.h:1202-1294 (header
with four heights, the barrier positions and xc) and :414-642 (per lane also
xp, two slopes, intercepts and heights each, the local-barrier indices; then
two barrier tables and the marker 9948), R_prev after it (:3391-3400).  All
values are little-endian: int = 4 B, size_t / long long = 8 B, double = 8 B.

The single-barrier layout is pinned by the reference's recorded quick-test
files, which decode and re-encode byte for byte; no double-barrier restart
file ships with the reference, so model=1 follows the source alone.
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


# per-lane scalars between the frame lists and the state counts
_PARAMS = {0: ("current_state", "previous_state", "slope_pos", "slope_neg", "current_height"),
           1: ("current_state", "previous_state", "slope_pos0", "slope_pos1", "slope_neg0", "slope_neg1",
               "intercept_pos0", "intercept_pos1", "intercept_neg0", "intercept_neg1",
               "current_height0", "current_height1")}
_LEAD = {0: ("x", "v", "f", "e", "eta", "barrier"), 1: ("x", "v", "f", "e", "eta", "barrier", "xp")}
_LOCAL = (("local", "f64"), ("local_prev", "f64"), ("entering_local", "i32"),
          ("local_stored", "f64"), ("local_prev_stored", "f64"), ("entering_local_stored", "i32"))


def _read_table(r, W):
    nb = r.u64()
    states = np.zeros((nb, W), np.int32)
    steps = np.zeros((nb, W), np.uint64)
    for b in range(nb):
        for i in range(W):
            states[b, i] = r.i32()
            steps[b, i] = r.u64()
    return states, steps, r.i32()


def decode(path, overdamped=True, tpt=True, fptcg=True, model=0):
    """Decode one restart file into a dict of header fields and per-lane data.
    For the double barrier (model=1) the second barrier table is returned as
    barrier_state1 / barrier_step1 / barrier_min_id1."""
    with open(path, "rb") as f:
        r = _Reader(f.read())
    h = read_header(r, model)
    W = h["simd_width"]
    assert r.i32() == 9998
    b_wt = r.i32()
    b_fpt = r.i32()
    lanes = []
    for _ in range(W):
        L = {}
        for k in _LEAD[model]:
            L[k] = r.f64()
        for k in ("traj_x", "traj_v", "traj_f", "traj_pe", "traj_ke", "traj_eta", "traj_state", "traj_barrier"):
            L[k] = r.sized_vec()
        for k in _PARAMS[model] + ("state_count0", "state_count1", "well", "periodic", "well_prev", "periodic_prev"):
            L[k] = r.f64()
        L["entering_well"] = r.i32()
        L["entering_period"] = r.i32()
        if model == 1:
            for k, kind in _LOCAL:
                L[k] = getattr(r, kind)()
        for k in ("v0_instance", "eta0_instance", "state0_instance", "well_stored", "periodic_stored",
                  "well_prev_stored", "periodic_prev_stored"):
            L[k] = r.f64()
        L["entering_well_stored"] = r.i32()
        L["entering_period_stored"] = r.i32()
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
    states, steps, min_id = _read_table(r, W)
    extra = {}
    if model == 1:
        s1, p1, m1 = _read_table(r, W)
        extra = {"barrier_state1": s1, "barrier_step1": p1, "barrier_min_id1": m1}
        assert r.i32() == 9948
    tail = None
    if overdamped:
        tail = r.vec(W)
    assert r.i32() == 9990
    trailer = None
    if r.pos < len(r.buf):
        # files of the MI355X build carry their own checkpoint image after the end marker
        assert r.u64() == 0x324B43504350444C
        n = r.u64()
        trailer = r.buf[r.pos:r.pos + n]
        r.pos += n
    assert r.pos == len(r.buf), (r.pos, len(r.buf))
    return {"header": h, "lanes": lanes, "barrier_state": states, "barrier_step": steps,
            "barrier_min_id": min_id, "r_prev": tail, "trailer": trailer, **extra}


class _Writer:
    def __init__(self):
        self.parts = []

    def i32(self, v):
        self.parts.append(struct.pack("<i", int(v)))

    def u64(self, v):
        self.parts.append(struct.pack("<Q", int(v)))

    def i64(self, v):
        self.parts.append(struct.pack("<q", int(v)))

    def f64(self, v):
        self.parts.append(struct.pack("<d", float(v)))

    def vec(self, a, dtype="<f8"):
        self.parts.append(np.ascontiguousarray(a, dtype=dtype).tobytes())

    def sized_vec(self, a):
        self.u64(len(a))
        self.vec(a)


def _write_table(w, states, steps, min_id):
    w.u64(states.shape[0])
    for b in range(states.shape[0]):
        for i in range(states.shape[1]):
            w.i32(states[b, i])
            w.u64(steps[b, i])
    w.i32(min_id)


def encode(d, path, overdamped=True, tpt=True, fptcg=True, model=0):
    """The inverse of decode: writes the dict decode returns (header, lanes,
    barrier table(s), R_prev) as a restart file of the reference program, for
    tests that need such a file on a machine without the reference's own
    files."""
    w = _Writer()
    h = d["header"]
    w.i32(h["signature"])
    w.i32(h["simd_width"])
    for k in ("x0", "v0", "gamma", "kb", "temperature", "mass", "timestep"):
        w.f64(h[k])
    w.u64(h["nsteps_total"])
    w.i32(h["waiting_time"])
    w.u64(h["stat_freq"])
    w.u64(h["save_freq"])
    for v in h["save_flags"]:
        w.i32(v)
    w.i32(h["randomize_v"])
    w.i32(h["randomize_eta"])
    if model == 0:
        keys = ("periodic_length", "h0", "h1", "initial_state")
    else:
        keys = ("periodic_length", "h00", "h01", "h10", "h11", "initial_state", "position0", "position1", "xc")
    for k in keys:
        w.f64(h[k])
    w.i32(h["randomize_state"])
    w.i32(h["same_initial_barrier"])
    for k in ("q", "tau", "eta0"):
        w.f64(h[k])
    for v in h["save_flags2"]:
        w.i32(v)
    w.i32(h["rank"])
    for renew, offset in h["rng"]:
        w.i64(renew)
        w.u64(offset)
    w.i32(9998)
    b_wt = int(h["waiting_time"])
    b_fpt = int(d.get("b_fpt", 1))
    w.i32(b_wt)
    w.i32(b_fpt)
    for L in d["lanes"]:
        for k in _LEAD[model]:
            w.f64(L[k])
        for k in ("traj_x", "traj_v", "traj_f", "traj_pe", "traj_ke", "traj_eta", "traj_state", "traj_barrier"):
            w.sized_vec(L[k])
        for k in _PARAMS[model] + ("state_count0", "state_count1", "well", "periodic", "well_prev", "periodic_prev"):
            w.f64(L[k])
        w.i32(L["entering_well"])
        w.i32(L["entering_period"])
        if model == 1:
            for k, kind in _LOCAL:
                getattr(w, kind)(L[k])
        for k in ("v0_instance", "eta0_instance", "state0_instance", "well_stored", "periodic_stored",
                  "well_prev_stored", "periodic_prev_stored"):
            w.f64(L[k])
        w.i32(L.get("entering_well_stored", 0))
        w.i32(L.get("entering_period_stored", 0))
        if b_wt:
            w.f64(L["tick_prev"])
            w.i32(L["barrier_passed"])
            w.sized_vec(L["barrier_ticks"])
        if b_fpt:
            w.sized_vec(L["fpt"])
            w.sized_vec(L["instant_barrier"])
            w.sized_vec(L["initial_barrier"])
            w.f64(L["fpt_last_tick"])
            w.i32(L["fpt_to_update"])
            w.f64(L["fpt_init_state"][0])
            w.f64(L["fpt_init_state"][1])
            w.i32(L["fpt_init_index"][0])
            w.i32(L["fpt_init_index"][1])
            w.sized_vec(L["fpt2"])
            w.f64(L["fpt2_last_tick"])
            w.i32(L["fpt2_to_update"])
            w.i32(L["fpt2_index"][0])
            w.i32(L["fpt2_index"][1])
        if tpt:
            w.sized_vec(L["tpt"])
            w.f64(L["tpt_last_tick"])
            w.i32(L["tpt_to_update"])
            w.sized_vec(L["tpt2"])
            w.f64(L["tpt2_last_tick"])
            w.i32(L["tpt2_to_update"])
            w.i32(L["tpt2_initial_well"])
        if fptcg:
            w.u64(len(L["cg_sizes"]))
            w.vec(L["cg_sizes"], "<i4")
            for cg in L["cg"]:
                w.sized_vec(cg["list"])
                w.f64(cg["last_tick"])
                w.i32(cg["to_update"])
                w.i32(cg["initial_well"])
    _write_table(w, d["barrier_state"], d["barrier_step"], d["barrier_min_id"])
    if model == 1:
        _write_table(w, d["barrier_state1"], d["barrier_step1"], d["barrier_min_id1"])
        w.i32(9948)
    if overdamped:
        w.vec(d["r_prev"])
    w.i32(9990)
    with open(path, "wb") as f:
        f.write(b"".join(w.parts))
