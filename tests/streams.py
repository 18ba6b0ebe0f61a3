"""Synthetic noise banks and role-tagged replay streams for the parity tests.

The rank-mode oracle is fed by a seeded numpy generator (not MKL) and records
what every lane consumed; the GPU replay path and the lane-mode oracle must
then reproduce the rank-mode trajectories exactly.
"""
import ctypes as C

import numpy as np

import oracle_lib as ol


class NumpyBanks:
    def __init__(self, seed):
        self.rng = [np.random.Generator(np.random.PCG64(seed + k)) for k in range(3)]
        self.cb = ol.FILL_FN(self._fill)

    def _fill(self, ctx, which, buf, n):
        arr = np.ctypeslib.as_array(buf, shape=(n,))
        if which == 0:
            arr[:] = self.rng[0].standard_normal(n)
        else:
            arr[:] = self.rng[which].random(n)


def run_rank_recorded(oracle, p, W, seed, bmin=-300, bn=600, cap=None, bank=4096, trials=1, capn=0):
    """Run one rank of W lanes; return lanes, frames [frame][lane] and the
    role-tagged per-lane streams.  With trials > 1 the rank first runs
    trials - 1 unrecorded trials (its streams and its lanes' initial barrier
    states carry over, as in the reference's trial loop) and the results are
    the last trial's."""
    nsteps_eff = (p.nsteps // p.save_freq) * p.save_freq if p.save_freq >= 10 else p.nsteps
    cap = cap or (nsteps_eff + 8)
    banks = NumpyBanks(seed)
    st = oracle.ldo_rank_create(C.byref(p), W, C.cast(banks.cb, C.c_void_p), None, bank)
    if not st:
        raise ValueError(f"oracle rank of {W} lanes refused (1 to 16 lanes)")
    lanes = (ol.Lane * W)()
    nfr = nsteps_eff // p.save_freq + 1
    fields = 7 if p.save_all else 1
    frames = np.zeros((fields, nfr, W))
    nf = C.c_uint64()
    for _ in range(trials - 1):
        oracle.ldo_rank_integrate(st, lanes, ol.dptr(frames), nfr, C.byref(nf), None, None)
    rec_g = np.zeros((W, cap))
    rec_f = np.zeros((W, cap))
    rec_c = np.zeros((W, cap))
    binit = np.full((2, W, bn), -1, np.int8)
    rec = ol.Record(cap, ol.dptr(rec_g), ol.dptr(rec_f), ol.dptr(rec_c), bmin, bn,
                    binit.ctypes.data_as(C.POINTER(C.c_int8)))
    capture, cbufs = None, None
    if capn:
        # the raw event lists of the recorded trial, first capn entries of each
        cbufs = {k: np.zeros((W, capn)) for k in ("fpt", "fpt2", "tpt", "tpt2", "ticks", "inst", "init")}
        cbufs["cg"] = np.zeros((max(p.n_cg, 1), W, capn))
        capture = ol.Capture(capn, *[ol.dptr(cbufs[k]) for k in ("fpt", "fpt2", "tpt", "tpt2", "ticks", "cg", "inst",
                                                                  "init")])
    oracle.ldo_rank_integrate(st, lanes, ol.dptr(frames), nfr, C.byref(nf), C.byref(capture) if capn else None,
                              C.byref(rec))
    oracle.ldo_rank_destroy(st)
    streams = {
        "gauss": rec_g, "flip": rec_f, "catch": rec_c, "binit": binit, "bmin": bmin,
        "n_gauss": [lanes[i].n_gauss for i in range(W)], "n_flip": [lanes[i].n_flip for i in range(W)],
        "n_catch": [lanes[i].n_catch for i in range(W)],
        "capture": cbufs,
    }
    return lanes, frames, streams
