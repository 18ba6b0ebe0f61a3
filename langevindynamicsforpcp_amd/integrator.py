"""Host mirror of the reference integrator interface, backed by libldpcp.so.

``LangevinEquation`` plays the role of
LANGEVINEQUATION::LangevinEquationSIMDMultiple / OverdampedLangevinEquationSIMDMultiple
(langevinDynamicsSingleBarrier.h:823-2888): the runner sets the public
parameters, calls ``Initialize(seed)`` and ``Integrate(iTrial)`` and reads the
results through the same getters.  A handle holds ``n_traj`` trajectories
(the reference's SIMD lanes of every rank) and integrates them on the GPU.
"""
import ctypes as C

import numpy as np

from . import _native as N


class LangevinEquation:
    """One GPU handle of ``n_traj`` independent trajectories."""

    def __init__(self, n_traj, model=0, overdamped=False, traj_offset=0):
        self.n_traj = int(n_traj)
        self.traj_offset = int(traj_offset)
        self.model = model
        self.m_Overdamped = bool(overdamped)
        # reference defaults (LangevinEquationSIMDMultiple constructor, Single.h:963-1032)
        self.m_X0 = 0.0
        self.m_V0 = 1.0
        self.m_Gamma0 = 1.0
        self.m_kB0 = 1.0
        self.m_T0 = 1.0
        self.m_M0 = 1.0
        self.m_TimeStep0 = 0.1
        self.m_NumberOfStep = 10000
        self.m_SaveFreq = 1
        self.m_RandomizeVelocity = False
        self.m_RandomizeState0 = True
        self.m_bSameInitialBarrier = False
        self.m_PeriodicLength0 = 5.0
        self.m_BarrierHeight0 = [[2.0, 6.0], [2.0, 6.0]]
        self.m_SlopeHalfLength0 = [2.5, 2.5]
        self.m_BarrierPosition0 = [0.0, 0.0]
        self.m_InitialState0 = 0.0
        self.m_Tau0 = 20.0
        self.m_bEqualProbability = True
        self.m_EquilibriumConstant = 1.0
        self.m_bWaitingTimeDistribution = False
        self.m_bFirstPassageTime = False
        self.m_bTransitionPathTime = False
        self.m_bFirstPassageTimeCG = False
        self.m_FirstPassageTimeCGSize = []
        self.m_SaveAll = False
        self.barrier_window = 1024
        self.tick_capacity = 0
        self.event_capacity = 0
        self.steps_per_launch = 1 << 20
        self._h = None
        self._seed = 0
        self._replay = None
        self._resume = None
        self._bdefault = None
        self._lanes = None
        self._frames = {}

    # -- set-up --------------------------------------------------------------
    def params(self):
        p = N.RunParams()
        p.model = self.model
        p.overdamped = int(self.m_Overdamped)
        p.x0, p.v0, p.gamma, p.kb = self.m_X0, self.m_V0, self.m_Gamma0, self.m_kB0
        p.temperature, p.mass = self.m_T0, self.m_M0
        p.randomize_v = int(self.m_RandomizeVelocity)
        p.randomize_state = int(self.m_RandomizeState0)
        p.same_initial_barrier = int(self.m_bSameInitialBarrier)
        p.periodic_length = self.m_PeriodicLength0
        for b in range(2):
            for s in range(2):
                p.height[b][s] = self.m_BarrierHeight0[b][s]
            p.half_width[b] = self.m_SlopeHalfLength0[b]
            p.position[b] = self.m_BarrierPosition0[b]
        p.timestep = self.m_TimeStep0
        p.nsteps = int(self.m_NumberOfStep)
        p.save_freq = int(self.m_SaveFreq)
        p.tau = self.m_Tau0
        p.initial_state = self.m_InitialState0
        p.equal_probability = int(self.m_bEqualProbability)
        p.eq_constant = self.m_EquilibriumConstant
        p.b_waiting = int(self.m_bWaitingTimeDistribution)
        p.b_fpt = int(self.m_bFirstPassageTime)
        p.b_tpt = int(self.m_bTransitionPathTime)
        p.b_fptcg = int(self.m_bFirstPassageTimeCG)
        p.n_cg = len(self.m_FirstPassageTimeCGSize)
        for j, n in enumerate(self.m_FirstPassageTimeCGSize):
            p.cg_size[j] = n
        p.save_all = int(self.m_SaveAll)
        p.seed = self._seed & 0xFFFFFFFFFFFFFFFF
        p.barrier_window = self.barrier_window
        p.tick_capacity = self.tick_capacity
        p.event_capacity = self.event_capacity
        return p

    def Initialize(self, seed):
        """Initialize(int seed) (Single.h:1276-1291): keys the Philox streams."""
        self._seed = int(seed)

    def SetReplay(self, gauss, flip, catchup, binit, bmin):
        """Role-tagged noise replay: gauss/flip [k][traj], catchup [traj][cap],
        binit [array][traj][bn] (int8, -1 = never created)."""
        self._replay = (np.ascontiguousarray(gauss, np.float64), np.ascontiguousarray(flip, np.float64),
                        np.ascontiguousarray(catchup, np.float64), np.ascontiguousarray(binit, np.int8), int(bmin))

    def SetBarrierDefault(self, states):
        """UserData::m_bInitialValue (Single.h:1378-1380): the state barriers created
        under UNIFORMINITALSTATES take, per trajectory (the previous trial's
        GetBarrierState0Instance in the reference's trial loop)."""
        self._bdefault = np.ascontiguousarray(states, np.float64)

    def Resume(self, image):
        """ReadRestartFile (Single.h:1137-1206): the next Integrate continues the
        checkpointed run for m_NumberOfStep further steps."""
        self._resume = bytes(image)

    def Checkpoint(self):
        """The state WriteRestartFile saves (Single.h:1048-1136), as the library's
        checkpoint image (bytes)."""
        lib = N.load()
        n = C.c_uint64()
        N.check(lib.ldpcp_run_checkpoint_size(self._h, C.byref(n)), "ldpcp_run_checkpoint_size")
        buf = C.create_string_buffer(n.value)
        N.check(lib.ldpcp_run_checkpoint(self._h, buf, n.value), "ldpcp_run_checkpoint")
        return buf.raw

    def _create(self):
        lib = N.load()
        if self._h is not None:
            lib.ldpcp_run_destroy(self._h)
            self._h = None
        h = C.c_void_p()
        p = self.params()
        N.check(lib.ldpcp_run_create(C.byref(p), self.n_traj, self.traj_offset, C.byref(h)), "ldpcp_run_create")
        self._h = h
        if self._replay is not None:
            g, f, c, b, bmin = self._replay
            N.check(lib.ldpcp_run_set_replay(
                h, g.ctypes.data_as(C.POINTER(C.c_double)), g.shape[0],
                f.ctypes.data_as(C.POINTER(C.c_double)), f.shape[0] if f.size else 0,
                c.ctypes.data_as(C.POINTER(C.c_double)), c.shape[1] if c.size else 0,
                b.ctypes.data_as(C.POINTER(C.c_int8)), bmin, b.shape[2] if b.size else 0), "ldpcp_run_set_replay")
        if self._bdefault is not None:
            N.check(lib.ldpcp_run_set_barrier_default(h, self._bdefault.ctypes.data_as(C.POINTER(C.c_double))),
                    "ldpcp_run_set_barrier_default")
        if self._resume is not None:
            img = C.create_string_buffer(self._resume, len(self._resume))
            N.check(lib.ldpcp_run_resume(h, img, len(self._resume)), "ldpcp_run_resume")

    def Integrate(self, iTrial=0, bTimer=False):
        """Integrate(iTrial) (Single.h:1386-1519) on the GPU."""
        self._create()
        lib = N.load()
        N.check(lib.ldpcp_run_integrate(self._h, int(self.steps_per_launch)), "ldpcp_run_integrate")
        self._lanes = None
        self._frames = {}
        return True

    def close(self):
        if self._h is not None:
            N.load().ldpcp_run_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- results -------------------------------------------------------------
    def lanes(self):
        if self._lanes is None:
            out = (N.LaneResult * self.n_traj)()
            N.check(N.load().ldpcp_run_lanes(self._h, 0, self.n_traj, out), "ldpcp_run_lanes")
            self._lanes = out
        return self._lanes

    def frames(self, field=0):
        if field not in self._frames:
            lib = N.load()
            nf = lib.ldpcp_run_n_frames(self._h)
            out = np.zeros((self.n_traj, nf))
            N.check(lib.ldpcp_run_frames(self._h, field, 0, self.n_traj, out.ctypes.data_as(C.POINTER(C.c_double))),
                    "ldpcp_run_frames")
            self._frames[field] = out
        return self._frames[field]

    def ticks(self):
        """Barrier ticks per trajectory (GetBarrierTick), truncated to tick_capacity."""
        cap = self.tick_capacity
        out = np.zeros((self.n_traj, cap))
        N.check(N.load().ldpcp_run_ticks(self._h, 0, self.n_traj, out.ctypes.data_as(C.POINTER(C.c_double))),
                "ldpcp_run_ticks")
        lanes = self.lanes()
        return [out[i, :min(cap, lanes[i].wt_ticks)] for i in range(self.n_traj)]

    def msd(self, lags):
        lags = np.ascontiguousarray(lags, np.uint64)
        out = np.zeros((self.n_traj, len(lags)))
        N.check(N.load().ldpcp_run_msd(self._h, lags.ctypes.data_as(C.POINTER(C.c_uint64)), len(lags),
                                       out.ctypes.data_as(C.POINTER(C.c_double))), "ldpcp_run_msd")
        return out

    # getters named as in Single.h:2620-2769
    def GetX(self, index):
        return self.frames(0)[index]

    def GetNumberOfFrameX(self, index):
        return self.frames(0).shape[1]

    def GetV0Instance(self, index):
        return self.lanes()[index].v0_instance

    def GetBarrierState0Instance(self, index):
        return self.lanes()[index].state0_instance

    @staticmethod
    def _avg(s, n):
        return s / n if n else 0.0

    def GetFirstPassageTimeAverage(self, index):
        L = self.lanes()[index]
        return self._avg(L.fpt_sum, L.fpt_n)

    def GetFirstPassageTimeAverage2(self, index):
        L = self.lanes()[index]
        return self._avg(L.fpt2_sum, L.fpt2_n)

    def GetTransitionPathTimeAverage(self, index):
        L = self.lanes()[index]
        return self._avg(L.tpt_sum, L.tpt_n)

    def GetTransitionPathTimeAverage2(self, index):
        L = self.lanes()[index]
        return self._avg(L.tpt2_sum, L.tpt2_n)

    def GetFirstPassageTimeCGAverage2(self, index, size_index):
        L = self.lanes()[index]
        return self._avg(L.cg_sum[size_index], L.cg_n[size_index])

    def GetStateCount(self, index, state):
        return self.lanes()[index].state_count[state]


def read_checkpoint(image):
    """The parts of a checkpoint image (ldpcp_run_checkpoint; layout in
    include/ldpcp.h) as arrays: header, lanes, window base and entries,
    frames [field][frame][trajectory], ticks [trajectory][capacity] and the
    event lists, events[trajectory][list] (the kept entries)."""
    buf = bytes(image)
    hd = N.CheckpointHeader.from_buffer_copy(buf)
    n, wb, nf, cap = hd.n_traj, hd.barrier_window, hd.n_frames, hd.tick_capacity
    o = C.sizeof(N.CheckpointHeader)
    lanes = (N.LaneResult * n).from_buffer_copy(buf[o:o + n * C.sizeof(N.LaneResult)])
    o += n * C.sizeof(N.LaneResult)
    wbase = np.frombuffer(buf, np.int32, n, o)
    o += (n * 4 + 7) // 8 * 8
    win = np.frombuffer(buf, np.uint64, n * 2 * wb, o).reshape(n, 2, wb)
    o += n * 2 * wb * 8
    frames = np.frombuffer(buf, np.float64, hd.n_fields * nf * n, o).reshape(hd.n_fields, nf, n)
    o += hd.n_fields * nf * n * 8
    ticks = np.frombuffer(buf, np.float64, n * cap, o).reshape(n, cap)
    o += n * cap * 8
    counts = {0: "fpt_n", 1: "fpt2_n", 2: "tpt_n", 3: "tpt2_n", 4: "fpt_n", 5: "fpt_n"}
    events = []
    for t in range(n):
        row = []
        for k in range(hd.n_lists):
            c = getattr(lanes[t], counts[k]) if k in counts else lanes[t].cg_n[k - 6]
            keep = min(c, hd.event_capacity)
            row.append(np.frombuffer(buf, np.float64, keep, o).copy())
            o += keep * 8
        events.append(row)
    if o != len(buf) or o != hd.bytes:
        raise ValueError(f"checkpoint image of {len(buf)} bytes, parts end at {o}")
    return {"header": hd, "lanes": lanes, "wbase": wbase, "window": win, "frames": frames, "ticks": ticks,
            "events": events}
