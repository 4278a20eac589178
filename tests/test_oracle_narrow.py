"""Pins of the oracle's narrow-state dSB forms (X and Y as fixed-point codes
of width w = 12, 10 or 8 bits): known answers of the rounding rule, the
code-grid invariants, the Y range bound of the dynamics, the distance to the
FP32 step, the exact kick against scipy, and a solve that must find a
bipartite graph's maximum cut.

Each check targets a plausible mistake: a wrong scale or shift, rounding half
away from zero or truncation instead of round-to-nearest-even, a negative
zero that flips a sign, rounding before the update instead of after it, a Y
code range that clamps live values, a lost wall code, the widths' constants
swapped.
"""
import ctypes as C

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
from test_oracle_pins import random_problem

# w: X scale shift xs = w - 2 (codes +-2^xs, wall +-(2^xs + 1)), Y code limit
WIDTHS = {12: dict(xs=10, ym=2047), 10: dict(xs=8, ym=511), 8: dict(xs=6, ym=127)}


@pytest.mark.parametrize("w,dt,want", [
    (12, 0.5, 8), (12, 0.1, 6), (12, 1.0, 9), (12, 2.0, 9), (12, 0.05, 5), (12, 0.3, 8), (12, 0.2, 7),
    (10, 0.5, 6), (10, 0.1, 4), (10, 1.0, 7), (10, 2.0, 7), (10, 0.05, 3), (10, 0.3, 6), (10, 0.2, 5),
    (8, 0.5, 4), (8, 0.1, 2), (8, 1.0, 5), (8, 2.0, 5), (8, 0.05, 1), (8, 0.3, 4), (8, 0.2, 3)])
def test_shift_known_answers(w, dt, want):
    # largest s in [0, w - 1] with ym 2^-s >= 2/dt + dt, worked by hand:
    # w 12, dt 0.5: ymax 4.5, 2047/256 = 8.0 >= 4.5 > 2047/512; dt 0.1: 20.1 -> 64;
    # dt 0.3: 6.97 <= 2047/256; dt 0.2: 10.2 -> 128.
    # w 10, dt 0.5: 511/64 = 7.98 >= 4.5 > 511/128; dt 0.1: 511/16 = 31.9 >= 20.1 > 511/32;
    # dt 1 and 2: 511/128 = 3.99 >= 3.0 > 511/256; dt 0.05: 511/8 = 63.9 >= 40.05 > 511/16;
    # dt 0.3: 6.97 <= 511/64; dt 0.2: 511/32 = 15.97 >= 10.2 > 511/64.
    # w 8, dt 0.5: 127/16 = 7.9 >= 4.5 > 127/32; dt 0.1: 127/4 = 31.75 >= 20.1 > 127/8;
    # dt 1 and 2: 3.0 <= 127/32 = 3.97; dt 0.05: 40.05 <= 63.5; dt 0.2: 10.2 <= 15.9
    s = orc.qn_shift(dt, w)
    assert s == want
    ym = WIDTHS[w]["ym"]
    ymax = 2.0 / dt + dt
    assert ym * 2.0 ** -s >= ymax and (s == w - 1 or ym * 2.0 ** -(s + 1) < ymax)
    if w == 12:
        assert orc.q12_shift(dt) == s


@pytest.mark.parametrize("w", [12, 10, 8])
def test_shift_rejects_nonpositive_dt(w):
    assert orc.qn_shift(0.0, w) == -1 and orc.qn_shift(-0.5, w) == -1


def test_shift_rejects_unknown_width():
    for w in (11, 9, 16, 0):
        assert orc.qn_shift(0.5, w) == -1


@pytest.mark.parametrize("w", [12, 10, 8])
def test_rounding_known_answers(w):
    xs, ym = WIDTHS[w]["xs"], WIDTHS[w]["ym"]
    top = 2 ** xs
    sy = {12: 8, 10: 6, 8: 4}[w]
    x = np.array([2.5, 3.5, -2.5, -0.3, 0.4, top, top + 0.0001 * top / 1024, -1.5 * top, top - 0.4, -0.0],
                 np.float64) / top
    y = np.array([0.5, 1.5, -0.5, 0.0, 100.0, -100.0, 3.0, 3.0, 2.0, 0.7], np.float64) / 2.0 ** sy
    X, Y = x.astype(np.float32), y.astype(np.float32)
    assert X[6] > 1.0  # just beyond the wall in float32
    orc.qn_round(X, Y, sy, w)
    xq = X.astype(np.float64) * top
    yq = Y.astype(np.float64) * 2.0 ** sy
    # round half to even; tiny negatives become +0 (sgn +1); |x| > 1 -> wall code, y = 0
    np.testing.assert_array_equal(xq, [2, 4, -2, 0, 0, top, top + 1, -top - 1, top, 0])
    np.testing.assert_array_equal(yq, [0, 2, 0, 0, 100, -100, 0, 0, 2, 1])
    assert not np.signbit(X[3]) and not np.signbit(X[9])
    # the Y clamp: +-ym codes
    X2, Y2 = np.zeros(2, np.float32), np.array([1e4, -1e4], np.float32)
    orc.qn_round(X2, Y2, sy, w)
    np.testing.assert_array_equal(Y2.astype(np.float64) * 2.0 ** sy, [ym, -ym])


def _trajectory(w, n=120, A=64, dt=0.5, seed=1):
    pb, M = random_problem(n, 0.06, seed, weights="int")
    X, Y = orc.init_state(seed, orc.MAXCUT, n, A)
    sy = orc.qn_shift(dt, w)
    orc.qn_round(X, Y, sy, w)
    orc.prestep_qn(X, Y, dt, sy, w)
    return pb, M, X, Y, sy


@pytest.mark.parametrize("w", [12, 10, 8])
def test_code_grid_and_wall_invariants(w):
    xs, ym = WIDTHS[w]["xs"], WIDTHS[w]["ym"]
    pb, M, X, Y, sy = _trajectory(w)
    dt, T = 0.5, 300
    xi0 = orc.xi0_maxcut(pb.n, pb.val64)
    walls = 0
    for t0 in range(0, T, 50):
        orc.advance_qn(pb, X, Y, t0, 50, T, dt, xi0, sy, w)
        xq = X.astype(np.float64) * 2.0 ** xs
        yq = Y.astype(np.float64) * 2.0 ** sy
        assert np.array_equal(xq, np.round(xq)) and np.abs(xq).max() <= 2 ** xs + 1
        assert np.array_equal(yq, np.round(yq)) and np.abs(yq).max() <= ym
        wall = np.abs(xq) == 2 ** xs + 1
        assert (Y[wall] == 0).all()
        assert not np.signbit(X[xq == 0]).any()
        walls += int(wall.sum())
    assert walls > 0  # the run reaches the walls, so the wall path is exercised


@pytest.mark.parametrize("w", [12, 10, 8])
def test_y_range_bound_holds_on_live_elements(w):
    # |y| <= 2/dt + dt while |x| <= 1, so the code clamp never binds a live
    # element: the largest |y| seen stays below the code range.
    ym = WIDTHS[w]["ym"]
    for dt in (0.5, 0.3, 1.0):
        pb, M, X, Y, sy = _trajectory(w, dt=dt, seed=4)
        T = 400
        xi0 = orc.xi0_maxcut(pb.n, pb.val64)
        ymax = 0.0
        for t0 in range(0, T, 20):
            orc.advance_qn(pb, X, Y, t0, 20, T, dt, xi0, sy, w)
            live = np.abs(X) <= 1
            ymax = max(ymax, float(np.abs(Y[live]).max()))
        assert ymax <= 2 / dt + dt < ym * 2.0 ** -sy


@pytest.mark.parametrize("w", [12, 10, 8])
def test_one_step_within_half_a_code_of_fp32(w):
    # From the same on-grid state, the narrow step equals the FP32 step
    # rounded once at the end: within half a code of X and of Y (and the wall
    # code wherever FP32 leaves the box).
    xs = WIDTHS[w]["xs"]
    pb, M, X, Y, sy = _trajectory(w, seed=7)
    xi0 = orc.xi0_maxcut(pb.n, pb.val64)
    for t in (0, 137, 299):
        c = orc.sched(orc.MAXCUT, t, 300, 0.5, xi0)
        Y32 = Y.copy()
        X32 = orc.step(pb, orc.MAXCUT, orc.DSB, X, Y32, c)
        Yq = Y.copy()
        Xq = orc.step_qn(pb, X, Yq, c, sy, w)
        out = np.abs(X32) > 1
        assert (np.abs(Xq[out]) == np.float32((2 ** xs + 1) / 2 ** xs)).all()
        assert (np.sign(Xq[out]) == np.sign(X32[out])).all()
        live = ~out
        assert np.abs(Xq[live] - X32[live]).max() <= 2.0 ** -(xs + 1)
        assert np.abs(Yq[live] - Y32[live]).max() <= 2.0 ** -(sy + 1)


@pytest.mark.parametrize("w,sy", [(12, 6), (10, 4), (8, 2)])
def test_kick_is_exact_minus_J_sgn_x(w, sy):
    # c = (0, 0, 1, 0), y = 0: the new Y is the force F = -J sgn(x) exactly
    # (integers within the code range), with sgn(+0) = +1 for zero codes.
    xs, ym = WIDTHS[w]["xs"], WIDTHS[w]["ym"]
    pb, M = random_problem(41, 0.15, 3, symmetric=False, weights="int")
    rng = np.random.default_rng(5)
    X = (rng.integers(-2 ** xs, 2 ** xs + 1, (41, 6)) / 2 ** xs).astype(np.float32)
    X[::5] = 0.0
    Y = np.zeros_like(X)
    Xn = orc.step_qn(pb, X, Y, np.array([0.0, 0.0, 1.0, 0.0]), sy, w)
    F = -(sp.csr_matrix(M) @ np.where(np.signbit(X), -1.0, 1.0))
    assert np.abs(F).max() < ym * 2.0 ** -sy
    np.testing.assert_array_equal(Y, F)
    np.testing.assert_array_equal(Xn, X)


def test_q12_wrappers_are_width_12():
    pb, M, X, Y, sy = _trajectory(12, seed=9)
    X2, Y2 = X.copy(), Y.copy()
    xi0 = orc.xi0_maxcut(pb.n, pb.val64)
    orc.advance_q12(pb, X, Y, 0, 30, 300, 0.5, xi0, sy)
    orc.advance_qn(pb, X2, Y2, 0, 30, 300, 0.5, xi0, sy, 12)
    assert np.array_equal(X, X2) and np.array_equal(Y, Y2)
    Y3 = Y.copy()
    cc = np.ascontiguousarray(orc.sched(orc.MAXCUT, 5, 300, 0.5, xi0))
    Xa = np.zeros_like(X)
    orc.lib().orc_step_q12(orc._p(pb.rp), orc._p(pb.col), orc._p(pb.val32), C.c_int64(pb.n),
                           C.c_int64(X.shape[1]), orc._p(X), orc._p(Xa), orc._p(Y3), orc._p(cc),
                           C.c_int(sy))
    Y4 = Y.copy()
    Xb = orc.step_qn(pb, X, Y4, cc, sy, 12)
    assert np.array_equal(Xa, Xb) and np.array_equal(Y3, Y4)


@pytest.mark.parametrize("w", [12, 10, 8])
def test_narrow_dsb_finds_bipartite_max_cut(w):
    # An even cycle with +1 weights is bipartite, so its maximum cut takes
    # every edge; the narrow form must find it as the FP32 form does.
    n = 40
    rows = np.arange(n)
    M = np.zeros((n, n))
    M[rows, (rows + 1) % n] = 1
    M = M + M.T
    csr = sp.csr_matrix(M)
    pb = orc.Problem(csr.indptr, csr.indices, csr.data, S=n)
    X, Y = orc.init_state(3, orc.MAXCUT, n, 64)
    xi0 = orc.xi0_maxcut(n, pb.val64)
    r = orc.solve_qn(pb, X, Y, 2000, 0, P=100, dt=0.5, xi0=xi0, w=w)
    cuts = orc.cut(pb, r["cur"])
    assert cuts.max() == n


# ---------------------------------------------------------------- 8-bit, stochastic Y rounding

def test_q8s_step_rounds_y_up_with_the_probability_of_its_fraction():
    # One step from the same on-grid state as the FP32 step: X rounds to
    # nearest even exactly as in the 8-bit form; each Y code is the floor or
    # the ceiling of the FP32 step's y 2^sy, and the share rounded up tracks
    # the fraction (rounding the dithered value to nearest instead of down
    # would round up about half a code too often, nearest-even rounding of y
    # itself never for f < 1/2).
    pb, M, X, Y, sy = _trajectory(8, n=300, A=256, seed=11)
    xi0 = orc.xi0_maxcut(pb.n, pb.val64)
    t = 57
    c = orc.sched(orc.MAXCUT, t, 300, 0.5, xi0)
    Y32 = Y.copy()
    X32 = orc.step(pb, orc.MAXCUT, orc.DSB, X, Y32, c)
    Yq = Y.copy()
    Xq = orc.step_qn(pb, X, Yq, c, sy, 8)
    Ys = Y.copy()
    Xs = orc.step_qn(pb, X, Ys, c, sy, 8, sr=True, t=t)
    assert np.array_equal(Xs, Xq)
    live = np.abs(X32) <= 1
    v = Y32[live].astype(np.float64) * 2.0 ** sy
    code = Ys[live].astype(np.float64) * 2.0 ** sy
    assert np.array_equal(code, np.round(code))
    fl = np.floor(v)
    assert ((code == fl) | (code == fl + 1)).all()
    f = v - fl
    band = (f > 0.2) & (f < 0.3)
    assert band.sum() > 2000
    up = (code[band] == fl[band] + 1).mean()
    assert 0.2 < up < 0.3, up
    # unbiased on average, where nearest-even rounding is too (the dither
    # only changes which way each element goes)
    assert abs((code - v).mean()) < 0.01
    assert (Ys[~live] == 0).all()


def test_q8s_dither_is_a_function_of_step_node_and_agent():
    pb, M, X, Y, sy = _trajectory(8, n=120, A=300, seed=12)
    c = orc.sched(orc.MAXCUT, 10, 300, 0.5, 0.3)
    outs = []
    for t in (10, 10, 11):
        Yt = Y.copy()
        orc.step_qn(pb, X, Yt, c, sy, 8, sr=True, t=t)
        outs.append(Yt)
    assert np.array_equal(outs[0], outs[1])  # deterministic
    assert not np.array_equal(outs[0], outs[2])  # another step index, another dither


def test_q8s_kick_is_exact_and_invariants_hold():
    # c = (0, 0, 1, 0), y = 0: y 2^sy is an integer, and a dither below half a
    # code rounds it back to itself, so Y = F exactly
    pb, M = random_problem(41, 0.15, 3, symmetric=False, weights="int")
    rng = np.random.default_rng(5)
    X = (rng.integers(-64, 65, (41, 6)) / 64).astype(np.float32)
    Y = np.zeros_like(X)
    orc.step_qn(pb, X, Y, np.array([0.0, 0.0, 1.0, 0.0]), 2, 8, sr=True, t=3)
    F = -(sp.csr_matrix(M) @ np.where(np.signbit(X), -1.0, 1.0))
    np.testing.assert_array_equal(Y, F)
    # codes on the grid, walls with Y = 0, live |y| within the bound, over a run
    pb, M, X, Y, sy = _trajectory(8, seed=13)
    xi0 = orc.xi0_maxcut(pb.n, pb.val64)
    ymax, walls = 0.0, 0
    for t0 in range(0, 400, 50):
        orc.advance_qn(pb, X, Y, t0, 50, 400, 0.5, xi0, sy, 8, sr=True)
        xq = X.astype(np.float64) * 64
        yq = Y.astype(np.float64) * 2.0 ** sy
        assert np.array_equal(xq, np.round(xq)) and np.abs(xq).max() <= 65
        assert np.array_equal(yq, np.round(yq)) and np.abs(yq).max() <= 127
        wall = np.abs(xq) == 65
        assert (Y[wall] == 0).all()
        # a zero code is +0: its sign gathers as +1 (a -0 would flip it)
        assert not np.signbit(X[xq == 0]).any() and not np.signbit(Y[yq == 0]).any()
        walls += int(wall.sum())
        ymax = max(ymax, float(np.abs(Y[np.abs(X) <= 1]).max()))
    assert walls > 0 and ymax <= 2 / 0.5 + 0.5 + 2.0 ** -sy


def test_q8s_finds_bipartite_max_cut():
    n = 40
    rows = np.arange(n)
    M = np.zeros((n, n))
    M[rows, (rows + 1) % n] = 1
    M = M + M.T
    csr = sp.csr_matrix(M)
    pb = orc.Problem(csr.indptr, csr.indices, csr.data, S=n)
    X, Y = orc.init_state(3, orc.MAXCUT, n, 64)
    r = orc.solve_qn(pb, X, Y, 2000, 0, P=100, dt=0.5, xi0=orc.xi0_maxcut(n, pb.val64), w=8,
                     sr=True)
    assert orc.cut(pb, r["cur"]).max() == n
