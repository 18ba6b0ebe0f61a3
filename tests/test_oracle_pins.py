"""Pins for the oracle functions that no recorded reference output covers.

The reference ships outputs of one workload only (the single-barrier,
overdamped quick test), so these functions are pinned by what the mathematics
fixes instead, each test chosen so that a plausible slip (a dropped term, a
swapped operand, a wrong index or placement) fails it:

* the Box-Muller logarithm and sine/cosine shared with the GPU, against a
  correctly rounded evaluation (mpmath);
* GetXc (langevinDynamicsDoubleBarrier.h:1722-1732): the apex is the exact
  intersection of the two inner slopes, and the potential is continuous there
  and at every segment end;
* the local-barrier hand-over (Double.h:2435-2736; single barrier
  Single.h:1998-2153): barrier identity from the position, the stored state
  carried unchanged without flips, and the catch-up flip frequency against
  1/2 (1 - exp(-2 dn dt / tau)) and the per-step flip frequency against the
  same law over the steps between hand-overs (binned chi-square);
* the underdamped B-A-O-A-B step (Single.h:1680-1706): the exact solution of
  the discrete map under a constant force, which fixes where the half kicks
  and half drifts sit, and the exact variance of the free-particle
  displacement from the discrete Ornstein-Uhlenbeck recursion, which fixes
  where the noise enters;
* the absorbing mode: its absorption step is the first FPT1 (single) / FPT2
  (double) event of the reference's non-absorbing statistics
  (Single.h:2394-2411, 2445-2489) on the same noise, and its transition-path
  start gives the first TPT1 event (Single.h:2497-2516).
"""
import ctypes as C
import math
from fractions import Fraction

import numpy as np
import pytest

import oracle_lib as ol

mpmath = pytest.importorskip("mpmath")


def ulps(got, want):
    return abs(got - want) / math.ulp(want) if want != 0.0 else abs(got) / math.ulp(1e-300)


def fma(a, b, c):
    """Correctly rounded a*b + c."""
    return float(Fraction(a) * Fraction(b) + Fraction(c))


# ---------------------------------------------------------------------------
# Box-Muller functions
# ---------------------------------------------------------------------------
def _philox_units(n, seed):
    rng = np.random.default_rng(seed)
    w = np.concatenate([rng.integers(0, 2 ** 32, n, dtype=np.uint64),
                        np.array([0, 1, 2, 2 ** 29, 2 ** 30, 2 ** 31 - 1, 2 ** 31, 3 * 2 ** 30,
                                  2 ** 32 - 2, 2 ** 32 - 1], np.uint64)])
    return (w.astype(np.float64) + 0.5) * 2.0 ** -32


def test_log_is_within_one_ulp(oracle):
    mpmath.mp.prec = 120
    worst = 0.0
    for u in _philox_units(4000, 1):
        worst = max(worst, ulps(oracle.ldo_log(float(u)), float(mpmath.log(mpmath.mpf(float(u))))))
    rng = np.random.default_rng(2)
    for x in np.exp(rng.uniform(-700.0, 700.0, 1000)):
        worst = max(worst, ulps(oracle.ldo_log(float(x)), float(mpmath.log(mpmath.mpf(float(x))))))
    for x in [1.0, 2.0, 0.5, math.sqrt(2.0), math.nextafter(math.sqrt(2.0), 0.0), 1e-300, 1e300]:
        worst = max(worst, ulps(oracle.ldo_log(x), float(mpmath.log(mpmath.mpf(x)))))
    assert oracle.ldo_log(1.0) == 0.0
    assert worst <= 1.0, worst


def test_sincos_turn_is_within_two_ulps(oracle):
    mpmath.mp.prec = 120
    s, c = C.c_double(), C.c_double()
    worst = 0.0
    us = np.concatenate([_philox_units(4000, 3), [0.0, 0.125, 0.25, 0.375, 0.5, 0.625, 0.75, 0.875, 1.0]])
    for u in us:
        oracle.ldo_sincos_turn(float(u), C.byref(s), C.byref(c))
        a = 2 * mpmath.pi * mpmath.mpf(float(u))
        rs, rc = float(mpmath.sin(a)), float(mpmath.cos(a))
        if u in (0.0, 0.25, 0.5, 0.75, 1.0):
            # quadrant rotation: exact zeros and ones where the turn is a multiple of 1/4
            assert abs(s.value) in (0.0, 1.0) and abs(c.value) in (0.0, 1.0)
            assert s.value == round(rs) and c.value == round(rc)
            continue
        worst = max(worst, ulps(s.value, rs), ulps(c.value, rc))
    assert worst <= 2.0, worst


# ---------------------------------------------------------------------------
# GetXc and the double-barrier potential
# ---------------------------------------------------------------------------
CUIDB_DOUBLE = [(2.0, 3.40, 7.20), (3.0, 5.10, 10.80), (4.0, 6.80, 14.40), (5.0, 8.50, 18.00)]


def _double_params(h0, h1, eb, **kw):
    return ol.make_params(model=1, heights=[[h0, eb], [h1, eb]], half_width=[0.70, 0.65],
                          position=[-1.42, -0.56], periodic_length=5.8, **kw)


def _energy(oracle, p, x, state):
    """E(x) from the oracle's own potential (step 0 of a run at x0 = x)."""
    q = ol.Params.from_buffer_copy(p)
    q.x0 = x
    q.nsteps = 0
    q.save_freq = 1
    q.randomize_state = 0
    q.initial_state = float(state)
    q.tau = 0.0
    q.b_waiting = q.b_fpt = q.b_tpt = q.b_fptcg = 0
    out = ol.Lane()
    oracle.ldo_lane_run_philox(C.byref(q), C.byref(out), 1, 0, 0, 0, None, 0)
    return out.e, out.f


@pytest.mark.parametrize("eb,h0,h1", CUIDB_DOUBLE)
def test_double_apex_is_the_slope_intersection(oracle, eb, h0, h1):
    p = _double_params(h0, h1, eb)
    xc, yc = C.c_double(), C.c_double()
    oracle.ldo_double_apex(C.byref(p), C.byref(xc), C.byref(yc))
    xc, yc = xc.value, yc.value
    x1b, w1, x2b, w2 = -1.42, 0.70, -0.56, 0.65
    x1h, x2l = x1b + w1, x2b - w2
    # exact intersection of the line through (x1b, y1), (x1h, 0) with the line
    # through (x2l, 0), (x2b, y2), in rationals over the double inputs
    F = Fraction
    X1b, X1h, X2b, X2l, Y1, Y2 = F(x1b), F(x1h), F(x2b), F(x2l), F(h0), F(h1)
    s1 = -Y1 / (X1h - X1b)
    s2 = Y2 / (X2b - X2l)
    X = (s1 * X1h - s2 * X2l) / (s1 - s2)
    Yx = s1 * (X - X1h)
    assert ulps(xc, float(X)) <= 4
    assert ulps(yc, float(Yx)) <= 16  # yc inherits xc's rounding through the slope
    assert x1b < xc < x2b
    # Each segment is evaluated as xp * slope + intercept (Double.h:1498-1513),
    # whose two terms are far larger than E near the apex (the intercept of the
    # barrier-1 rise is ~13 for E ~ 1.65): continuity holds to a few ulp of the
    # larger term, which bounds the rounding of that sum.
    def term_scale(h_top, x_top, x):
        slope = abs(F(h_top) - F(yc)) / abs(F(x_top) - F(xc))
        return float(abs(slope * F(x)) + abs(F(h_top)) + slope * abs(F(x_top)))

    # E(xc-) and E(xc+) equal yc in either barrier state: the state-1 lines are
    # re-sloped to the same apex (Double.h:1852-1899)
    for st in (0, 1):
        hb = ([h0, eb][st], [h1, eb][st])
        e_lo, _ = _energy(oracle, p, math.nextafter(xc, -math.inf), st)
        e_hi, _ = _energy(oracle, p, math.nextafter(xc, math.inf), st)
        assert abs(e_lo - yc) <= 4 * math.ulp(term_scale(hb[0], x1b, xc)), (st, e_lo, yc)
        assert abs(e_hi - yc) <= 4 * math.ulp(term_scale(hb[1], x2b, xc)), (st, e_hi, yc)
        # the barrier tops and the outer feet
        for x, want, k in [(x1b, hb[0], 0), (x2b, hb[1], 1), (x1b - w1, 0.0, 0), (x2b + w2, 0.0, 1)]:
            e_a, _ = _energy(oracle, p, math.nextafter(x, -math.inf), st)
            e_b, _ = _energy(oracle, p, math.nextafter(x, math.inf), st)
            tol = 4 * math.ulp(term_scale(hb[k], [x1b, x2b][k], x))
            assert abs(e_a - want) <= tol and abs(e_b - want) <= tol, (st, x, e_a, e_b, want)
        # strict compares: the potential is 0 exactly at the segment ends
        # (langevinDynamicsDoubleBarrier.h:1498-1513)
        for x in (x1b, xc, x2b):
            assert _energy(oracle, p, x, st)[0] == 0.0
        # force = -dE/dx on each segment, from two points inside it
        for a, b in [(x1b - w1, x1b), (x1b, xc), (xc, x2b), (x2b, x2b + w2)]:
            xa, xb = a + 0.3 * (b - a), a + 0.7 * (b - a)
            (ea, fa), (eb_, fb) = _energy(oracle, p, xa, st), _energy(oracle, p, xb, st)
            assert fa == fb
            assert fa == pytest.approx(-(eb_ - ea) / (xb - xa), rel=1e-9)


def test_double_apex_contraction_choice(oracle):
    """GetXc's numerator and denominator are each `a*b + c*d` statements, which
    the reference's compiler contracts (icpx -fp-model=precise keeps
    contraction on; the recorded event lists show it for the event durations).
    Clang-based compilers fuse the left product of such a sum when both sides
    are products; the oracle and the kernels take that form.  No recorded output
    can tell the forms apart (no double-barrier output ships); they differ by at
    most an ulp or two in xc, far below any statistic's resolution."""
    x1b, w1, x2b, w2 = -1.42, 0.70, -0.56, 0.65
    x1h, x2l = x1b + w1, x2b - w2
    for eb, h0, h1 in CUIDB_DOUBLE:
        p = _double_params(h0, h1, eb)
        xc = C.c_double()
        oracle.ldo_double_apex(C.byref(p), C.byref(xc), C.byref(C.c_double()))
        left = fma(x1h * h0, x2b - x2l, x2l * h1 * (x1h - x1b)) / fma(h0, x2b - x2l, h1 * (x1h - x1b))
        plain = (x1h * h0 * (x2b - x2l) + x2l * h1 * (x1h - x1b)) / (h0 * (x2b - x2l) + h1 * (x1h - x1b))
        right = fma(x2l * h1, x1h - x1b, x1h * h0 * (x2b - x2l)) / fma(h1, x1h - x1b, h0 * (x2b - x2l))
        assert xc.value == left
        assert ulps(plain, left) <= 4 and ulps(right, left) <= 4


# ---------------------------------------------------------------------------
# Barrier hand-over
# ---------------------------------------------------------------------------
HANDOVER_DTYPE = np.dtype([("step", np.uint64), ("x", np.float64), ("left_array", np.int32),
                           ("left_id", np.int32), ("left_state", np.int32), ("entered_array", np.int32),
                           ("entered_id", np.int32), ("stored_state", np.int32), ("first_touch", np.int32),
                           ("new_state", np.int32), ("stored_step", np.uint64)])


def _trace(oracle, p, seed, traj, cap=400000):
    buf = np.zeros(cap, HANDOVER_DTYPE)
    oracle.ldo_trace_handovers(buf.ctypes.data, cap)
    out = ol.Lane()
    oracle.ldo_lane_run_philox(C.byref(p), C.byref(out), seed, traj, 0, 0, None, 0)
    n = oracle.ldo_trace_count()
    oracle.ldo_trace_handovers(None, 0)
    assert 0 < n <= cap
    return buf[:n], out


def _handover_params(model, tau, over=True):
    # underdamped paths are smooth and recross a point far less often: run longer
    n = 200000 if over else 2000000
    common = dict(overdamped=int(over), timestep=1e-3, nsteps=n, save_freq=n, tau=tau,
                  b_waiting=0, b_fpt=0, b_tpt=0, b_fptcg=0, randomize_v=int(not over))
    if model == 1:
        return _double_params(0.9, 1.0, 0.5, x0=-2.9, **common)
    return ol.make_params(model=0, heights=[[1.5, 0.5], [1.5, 0.5]], half_width=[0.6, 0.6],
                          periodic_length=2.1, x0=-1.05, **common)


def _check_timeline(tr, p, xc):
    """Barrier identity from the position; the left barrier is the one entered
    last; the stored state and step of the entered barrier are those it was
    left with."""
    L = p.periodic_length
    cell = np.rint(tr["x"] * (1.0 / L))
    assert np.array_equal(cell.astype(np.int32), tr["entered_id"])
    if p.model == 1:
        xp = tr["x"] - cell * L
        assert np.array_equal(np.where(xp < xc, 0, 1), tr["entered_array"])
    else:
        assert np.all(tr["entered_array"] == 0) and np.all(tr["left_array"] == 0)
    assert np.array_equal(tr["left_id"][1:], tr["entered_id"][:-1])
    assert np.array_equal(tr["left_array"][1:], tr["entered_array"][:-1])
    last = {}
    for r in tr:
        last[(int(r["left_array"]), int(r["left_id"]))] = (int(r["left_state"]), int(r["step"]))
        key = (int(r["entered_array"]), int(r["entered_id"]))
        if key in last:
            assert not r["first_touch"]
            assert (int(r["stored_state"]), int(r["stored_step"])) == last[key]
        else:
            # untouched before (the start barrier is stored at step 0 and never a first touch here)
            assert r["stored_step"] == 0


@pytest.mark.parametrize("model", [0, 1])
def test_handover_without_flips_restores_the_stored_state(oracle, model):
    p = _handover_params(model, tau=1e300)
    xc = C.c_double()
    if model == 1:
        oracle.ldo_double_apex(C.byref(p), C.byref(xc), C.byref(C.c_double()))
    for traj in range(3):
        tr, _ = _trace(oracle, p, 11, traj)
        assert len(tr) > 200
        _check_timeline(tr, p, xc.value)
        # no per-step flips, no catch-up flips: the state entered is the state stored,
        # and the state carried to the next hand-over is unchanged
        assert np.array_equal(tr["new_state"], tr["stored_state"])
        assert np.array_equal(tr["left_state"][1:], tr["new_state"][:-1])


def _chi2_binned(dn, flipped, p_of, edges):
    """Chi-square of the flip counts per bin of dn against the sum of the
    model probabilities (binomial variance); returns (chi2, dof)."""
    chi2, dof = 0.0, 0
    for lo, hi in zip(edges[:-1], edges[1:]):
        m = (dn >= lo) & (dn < hi)
        if m.sum() < 30:
            continue
        pr = p_of(dn[m])
        e, var = pr.sum(), (pr * (1 - pr)).sum()
        if var < 5:
            continue
        chi2 += (flipped[m].sum() - e) ** 2 / var
        dof += 1
    return chi2, dof


@pytest.mark.parametrize("model,over", [(0, True), (1, True), (1, False)])
def test_handover_catch_up_and_step_flips_follow_the_telegraph_law(oracle, model, over):
    tau = 0.05
    p = _handover_params(model, tau=tau, over=over)
    dt = p.timestep
    xc = C.c_double()
    if model == 1:
        oracle.ldo_double_apex(C.byref(p), C.byref(xc), C.byref(C.c_double()))
    parts = []
    for traj in range(6):
        tr, _ = _trace(oracle, p, 23, traj)
        _check_timeline(tr, p, xc.value)
        parts.append(tr)
    law = lambda n: 0.5 * (1.0 - np.exp(-2.0 * n.astype(float) * dt / tau))  # noqa: E731
    edges = np.unique(np.round(np.geomspace(1, 2e5, 40)))
    # catch-up flip of the barrier entered, against the steps since it was left
    dn = np.concatenate([(t["step"] - t["stored_step"])[t["first_touch"] == 0] for t in parts])
    fl = np.concatenate([(t["new_state"] != t["stored_state"])[t["first_touch"] == 0] for t in parts])
    chi2, dof = _chi2_binned(dn, fl, law, edges)
    assert dof >= 5 and chi2 < dof + 5 * math.sqrt(2 * dof), (chi2, dof)
    # per-step flips of the barrier under the particle between two hand-overs
    dn2 = np.concatenate([(t["step"][1:] - t["step"][:-1]) for t in parts])
    fl2 = np.concatenate([(t["left_state"][1:] != t["new_state"][:-1]) for t in parts])
    chi2, dof = _chi2_binned(dn2, fl2, law, edges)
    assert dof >= 5 and chi2 < dof + 5 * math.sqrt(2 * dof), (chi2, dof)
    # a wrong lap (counted from step 0) fails the same test where short laps are
    # common (overdamped paths recross within a few steps; underdamped laps are
    # mostly long enough that both laws are 1/2)
    if over:
        dn_wrong = np.concatenate([t["step"][t["first_touch"] == 0] for t in parts])
        chi2w, dofw = _chi2_binned(dn_wrong, fl, law, edges)
        assert chi2w > dofw + 5 * math.sqrt(2 * max(dofw, 1))


# ---------------------------------------------------------------------------
# Underdamped B-A-O-A-B
# ---------------------------------------------------------------------------
@pytest.mark.parametrize("x0,sign", [(1.0, 1.0), (-1.0, -1.0)])
def test_underdamped_constant_force_follows_the_discrete_map(oracle, x0, sign):
    """T = 0 on one slope of a wide barrier: the force is the constant +-S and
    the B-A-O-A-B map (v += a h; x += v h; v *= c; x += v h; v += a h, h = dt/2)
    solves to u_n = u* + c^n (u_0 - u*), u = v + a h, u* = 2 a h / (1 - c),
    x_n = x_0 + h (1 + c) [n u* + (u_0 - u*)(1 - c^n)/(1 - c)]."""
    H, w, dt, gamma, m, v0 = 4.0, 400.0, 1e-2, 0.5, 2.0, 0.2
    n_steps, save = 10000, 100
    p = ol.make_params(model=0, overdamped=0, x0=x0, v0=v0, gamma=gamma, mass=m, kb=1.0, temperature=0.0,
                       periodic_length=1000.0, heights=[[H, H], [H, H]], half_width=[w, w], timestep=dt,
                       tau=0.0, randomize_v=0, randomize_state=0, initial_state=0.0, nsteps=n_steps,
                       save_freq=save, b_waiting=0, b_fpt=0, b_tpt=0, b_fptcg=0)
    out = ol.Lane()
    fr = np.zeros(n_steps // save + 1)
    oracle.ldo_lane_run_philox(C.byref(p), C.byref(out), 1, 0, 0, 0, ol.dptr(fr), len(fr))
    mpmath.mp.prec = 100
    mp = mpmath.mpf
    a = sign * mp(H) / mp(w) / mp(m)
    h, c = mp(dt) / 2, mpmath.exp(-mp(gamma) * mp(dt))
    u0 = mp(v0) + a * h
    us = 2 * a * h / (1 - c)
    for k, xk in enumerate(fr):
        n = k * save
        want = mp(x0) + h * (1 + c) * (n * us + (u0 - us) * (1 - c ** n) / (1 - c))
        assert xk == pytest.approx(float(want), rel=1e-11, abs=1e-12), k
    vn = us + c ** n_steps * (u0 - us) - a * h
    assert out.v == pytest.approx(float(vn), rel=1e-11)


def test_underdamped_free_displacement_variance(oracle):
    """H = 0 from a stationary velocity: x_{k+1} = x_k + h(1+c) v_k + h s g_k,
    v_{k+1} = c v_k + s g_k with s^2 = kT/m (1 - c^2); Var(x_n - x_0) follows
    exactly from the AR(1) covariances, at every n.  Placing the O step before a
    full drift (x += dt v_{k+1}) changes the one-step variance by 10 %."""
    kT, m, gamma, dt, n = 1.5, 0.5, 1.0, 0.2, 12
    p = ol.make_params(model=0, overdamped=0, x0=-1.0, gamma=gamma, mass=m, kb=1.0, temperature=kT,
                       periodic_length=1e6, heights=[[0.0, 0.0], [0.0, 0.0]], half_width=[0.5, 0.5],
                       timestep=dt, tau=1.0, randomize_v=1, nsteps=n, save_freq=1,
                       b_waiting=0, b_fpt=0, b_tpt=0, b_fptcg=0)
    N = 40000
    d = np.empty((N, n))
    fr = np.zeros(n + 1)
    out = ol.Lane()
    for t in range(N):
        oracle.ldo_lane_run_philox(C.byref(p), C.byref(out), 77, t, 0, 0, ol.dptr(fr), n + 1)
        d[t] = fr[1:] - fr[0]
    h, c = dt / 2, math.exp(-gamma * dt)
    v2 = kT / m
    s = math.sqrt(v2 * (1 - c * c))

    def var_of(A, B, k):
        """Var of sum_i (A v_i + B g_i), i < k, with v_{i+1} = c v_i + s g_i stationary."""
        vv = sum(v2 * c ** abs(i - j) for i in range(k) for j in range(k))
        vg = sum(s * c ** (i - 1 - j) for i in range(k) for j in range(i))
        return A * A * vv + B * B * k + 2 * A * B * vg

    for k in range(1, n + 1):
        var = var_of(h * (1 + c), h * s, k)
        se = var * math.sqrt(2.0 / N)
        assert abs(d[:, k - 1].var() - var) < 4 * se, (k, d[:, k - 1].var(), var, se)
    # the alternative placement predicts a one-step variance this sample excludes
    var1, var1_alt = var_of(h * (1 + c), h * s, 1), var_of(dt * c, dt * s, 1)
    assert abs(d[:, 0].var() - var1_alt) > 8 * var1 * math.sqrt(2.0 / N)


# ---------------------------------------------------------------------------
# Absorbing mode = first event of the reference statistics
# ---------------------------------------------------------------------------
def _absorb_params(model, over):
    common = dict(overdamped=int(over), timestep=1e-3, tau=1e-3, randomize_v=int(not over), b_waiting=0,
                  b_fptcg=0, save_freq=1)
    if model == 0:
        return ol.make_params(model=0, x0=-1.05, periodic_length=2.1, heights=[[3.0, 1.0], [3.0, 1.0]],
                              half_width=[0.6, 0.6], **common)
    return _double_params(1.0, 2.0, 0.5, x0=-2.9, **dict(common, timestep=2e-3))


@pytest.mark.parametrize("model,over", [(0, True), (0, False), (1, True), (1, False)])
def test_absorbed_step_is_the_first_reference_event(oracle, model, over):
    p = _absorb_params(model, over)
    dt = p.timestep
    checked = 0
    for traj in range(48):
        q = ol.Params.from_buffer_copy(p)
        q.absorb, q.nsteps, q.b_fpt, q.b_tpt = model + 1, 400000, 0, 0
        a = ol.Lane()
        oracle.ldo_lane_run_philox(C.byref(q), C.byref(a), 31, traj, 0, 0, None, 0)
        A = a.absorbed_step
        if A == 0:
            continue
        checked += 1
        q.absorb, q.b_fpt, q.b_tpt = 0, 1, 1
        before, at = ol.Lane(), ol.Lane()
        q.nsteps = A - 1
        oracle.ldo_lane_run_philox(C.byref(q), C.byref(before), 31, traj, 0, 0, None, 0)
        q.nsteps = A
        oracle.ldo_lane_run_philox(C.byref(q), C.byref(at), 31, traj, 0, 0, None, 0)
        if model == 0:
            assert before.fpt_n == 0 and at.fpt_n == 1
            assert at.fpt_sum == dt * A  # fma(dt, A, -0) is the rounded product
            assert before.tpt_n == 0 and at.tpt_n == 1
            assert at.tpt_sum == fma(dt, A, -(dt * a.tpt_start_step))
        else:
            assert before.fpt2_n == 0 and at.fpt2_n == 1
            assert at.fpt2_sum == dt * A
        # the trajectory itself is the same noise path in both modes
        assert at.x == a.x and at.state == a.state
    assert checked >= 40
