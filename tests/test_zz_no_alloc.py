"""ldpcp_run_msd_noalloc: the MSD entry point that allocates no device memory.

It must return exactly what ldpcp_run_msd returns for the same handle and
lags, call after call (the fixed scratch array keeps the previous call's
contents, so a second, smaller call checks that nothing stale leaks into its
results), and refuse inputs beyond its fixed size with a clear error.
"""
import ctypes as C

import numpy as np
import pytest

from langevindynamicsforpcp_amd import _native as N
from langevindynamicsforpcp_amd.integrator import LangevinEquation

pytestmark = pytest.mark.gpu

SCRATCH_RESULTS = 983040  # n_traj * nlags the 8 MiB scratch holds


def _handle(n_traj, nsteps=4000, save=10, seed=5):
    le = LangevinEquation(n_traj, model=0, overdamped=True)
    le.m_X0, le.m_PeriodicLength0 = -1.05, 2.1
    le.m_BarrierHeight0 = [[3.0, 1.0], [3.0, 1.0]]
    le.m_SlopeHalfLength0 = [0.6, 0.6]
    le.m_TimeStep0, le.m_Tau0, le.m_NumberOfStep, le.m_SaveFreq = 1e-3, 1e-3, nsteps, save
    le.m_RandomizeVelocity = True
    le.Initialize(seed)
    le.Integrate(0)
    return le


def _noalloc(le, lags):
    lags = np.ascontiguousarray(lags, np.uint64)
    out = np.full((le.n_traj, len(lags)), np.nan)
    rc = N.load().ldpcp_run_msd_noalloc(le._h, lags.ctypes.data_as(C.POINTER(C.c_uint64)), len(lags),
                                        out.ctypes.data_as(C.POINTER(C.c_double)))
    return rc, out


def test_noalloc_msd_equals_msd_on_consecutive_calls():
    print("[no-alloc msd] 300 trajectories, then 37", flush=True)
    big = _handle(300)
    lags_big = np.arange(0, 421, 2, dtype=np.uint64)  # 401 frames: lags past the last frame give 0
    rc, got = _noalloc(big, lags_big)
    assert rc == 0, N.load().ldpcp_last_error()
    want = big.msd(lags_big)
    np.testing.assert_array_equal(got, want)
    assert np.all(got[:, lags_big > 400] == 0.0) and np.any(got > 0)
    # second call: fewer trajectories and fewer lags; the scratch still holds the
    # first call's larger results
    small = _handle(37, nsteps=2000, save=20, seed=9)
    lags_small = np.array([0, 1, 3, 7, 50, 99, 100, 200], np.uint64)
    rc, got = _noalloc(small, lags_small)
    assert rc == 0, N.load().ldpcp_last_error()
    np.testing.assert_array_equal(got, small.msd(lags_small))
    # and the first handle again, after the second
    rc, got = _noalloc(big, lags_big[:17])
    assert rc == 0
    np.testing.assert_array_equal(got, want[:, :17])
    big.close()
    small.close()


def test_noalloc_msd_rejects_inputs_beyond_its_scratch():
    print("[no-alloc msd] size limit", flush=True)
    le = _handle(8192, nsteps=200, save=10)
    lib = N.load()
    fits = SCRATCH_RESULTS // 8192
    rc, got = _noalloc(le, np.arange(fits, dtype=np.uint64))
    assert rc == 0, lib.ldpcp_last_error()
    np.testing.assert_array_equal(got, le.msd(np.arange(fits, dtype=np.uint64)))
    rc, _ = _noalloc(le, np.arange(fits + 1, dtype=np.uint64))
    assert rc == -1  # LDPCP_EINVAL
    assert b"scratch" in lib.ldpcp_last_error()
    le.close()
