"""Pins of the oracle's narrow-state dSB form (12-bit fixed-point X and Y):
known answers of the rounding rule, the code-grid invariants, the Y range
bound of the dynamics, the distance to the FP32 step, the exact kick against
scipy, and a solve that must find a bipartite graph's maximum cut.

Each check targets a plausible mistake: a wrong scale or shift, rounding half
away from zero or truncation instead of round-to-nearest-even, a negative
zero that flips a sign, rounding before the update instead of after it, a Y
code range that clamps live values, a lost wall code.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
from test_oracle_pins import random_problem


@pytest.mark.parametrize("dt,want", [(0.5, 8), (0.1, 6), (1.0, 9), (2.0, 9), (0.05, 5), (0.3, 8), (0.2, 7)])
def test_shift_known_answers(dt, want):
    # largest s in [0, 11] with 2047 2^-s >= 2/dt + dt, worked by hand:
    # dt 0.5: ymax 4.5, 2047/256 = 8.0 >= 4.5 > 2047/512; dt 0.1: 20.1 -> 64;
    # dt 0.3: 6.97 <= 2047/256; dt 0.2: 10.2 -> 128
    s = orc.q12_shift(dt)
    assert s == want
    ymax = 2.0 / dt + dt
    assert 2047 * 2.0 ** -s >= ymax and (s == 11 or 2047 * 2.0 ** -(s + 1) < ymax)


def test_shift_rejects_nonpositive_dt():
    assert orc.q12_shift(0.0) == -1 and orc.q12_shift(-0.5) == -1


def test_rounding_known_answers():
    sy = 8
    x = np.array([2.5, 3.5, -2.5, -0.3, 0.4, 1024.0, 1024.0001, -1500.0, 1023.6, -0.0],
                 np.float64) / 1024
    y = np.array([0.5, 1.5, -0.5, 0.0, 100.0, -100.0, 3.0, 3.0, 2.0, 0.7], np.float64) / 256
    X, Y = x.astype(np.float32), y.astype(np.float32)
    orc.q12_round(X, Y, sy)
    xq = X.astype(np.float64) * 1024
    yq = Y.astype(np.float64) * 256
    # round half to even; tiny negatives become +0 (sgn +1); |x| > 1 -> wall code, y = 0
    np.testing.assert_array_equal(xq, [2, 4, -2, 0, 0, 1024, 1025, -1025, 1024, 0])
    np.testing.assert_array_equal(yq, [0, 2, 0, 0, 100, -100, 0, 0, 2, 1])
    assert not np.signbit(X[3]) and not np.signbit(X[9])
    # the Y clamp: +-2047 codes
    X2, Y2 = np.zeros(2, np.float32), np.array([9.0, -9.0], np.float32)
    orc.q12_round(X2, Y2, sy)
    np.testing.assert_array_equal(Y2.astype(np.float64) * 256, [2047, -2047])


def _trajectory(n=120, A=64, steps=300, dt=0.5, seed=1):
    pb, M = random_problem(n, 0.06, seed, weights="int")
    X, Y = orc.init_state(seed, orc.MAXCUT, n, A)
    sy = orc.q12_shift(dt)
    orc.q12_round(X, Y, sy)
    orc.prestep_q12(X, Y, dt, sy)
    return pb, M, X, Y, sy


def test_code_grid_and_wall_invariants():
    pb, M, X, Y, sy = _trajectory()
    dt, T = 0.5, 300
    xi0 = orc.xi0_maxcut(pb.n, pb.val64)
    walls = 0
    for t0 in range(0, T, 50):
        orc.advance_q12(pb, X, Y, t0, 50, T, dt, xi0, sy)
        xq = X.astype(np.float64) * 1024
        yq = Y.astype(np.float64) * 2.0 ** sy
        assert np.array_equal(xq, np.round(xq)) and np.abs(xq).max() <= 1025
        assert np.array_equal(yq, np.round(yq)) and np.abs(yq).max() <= 2047
        wall = np.abs(xq) == 1025
        assert (Y[wall] == 0).all()
        assert not np.signbit(X[xq == 0]).any()
        walls += int(wall.sum())
    assert walls > 0  # the run reaches the walls, so the wall path is exercised


def test_y_range_bound_holds_on_live_elements():
    # |y| <= 2/dt + dt while |x| <= 1, so the 2047-code clamp never binds a
    # live element: the largest |y| seen stays below the code range.
    for dt in (0.5, 0.3, 1.0):
        pb, M, X, Y, sy = _trajectory(dt=dt, seed=4)
        T = 400
        xi0 = orc.xi0_maxcut(pb.n, pb.val64)
        ymax = 0.0
        for t0 in range(0, T, 20):
            orc.advance_q12(pb, X, Y, t0, 20, T, dt, xi0, sy)
            live = np.abs(X) <= 1
            ymax = max(ymax, float(np.abs(Y[live]).max()))
        assert ymax <= 2 / dt + dt < 2047 * 2.0 ** -sy


def test_one_step_within_half_a_code_of_fp32():
    # From the same on-grid state, the narrow step equals the FP32 step
    # rounded once at the end: within half a code of X and of Y (and the wall
    # code wherever FP32 leaves the box).
    pb, M, X, Y, sy = _trajectory(seed=7)
    xi0 = orc.xi0_maxcut(pb.n, pb.val64)
    for t in (0, 137, 299):
        c = orc.sched(orc.MAXCUT, t, 300, 0.5, xi0)
        Y32 = Y.copy()
        X32 = orc.step(pb, orc.MAXCUT, orc.DSB, X, Y32, c)
        Yq = Y.copy()
        Xq = np.zeros_like(X)
        lib = orc.lib()
        import ctypes as C
        cc = np.ascontiguousarray(c, np.float64)
        lib.orc_step_q12(orc._p(pb.rp), orc._p(pb.col), orc._p(pb.val32), C.c_int64(pb.n),
                         C.c_int64(X.shape[1]), orc._p(X), orc._p(Xq), orc._p(Yq), orc._p(cc),
                         C.c_int(sy))
        out = np.abs(X32) > 1
        assert (np.abs(Xq[out]) == np.float32(1025 / 1024)).all()
        assert (np.sign(Xq[out]) == np.sign(X32[out])).all()
        live = ~out
        assert np.abs(Xq[live] - X32[live]).max() <= 2.0 ** -11
        assert np.abs(Yq[live] - Y32[live]).max() <= 2.0 ** -(sy + 1)


def test_kick_is_exact_minus_J_sgn_x():
    # c = (0, 0, 1, 0), y = 0: the new Y is the force F = -J sgn(x) exactly
    # (integers within the code range), with sgn(+0) = +1 for zero codes.
    pb, M = random_problem(41, 0.15, 3, symmetric=False, weights="int")
    rng = np.random.default_rng(5)
    X = (rng.integers(-1024, 1025, (41, 6)) / 1024).astype(np.float32)
    X[::5] = 0.0
    Y = np.zeros_like(X)
    sy = 6  # code range +-31.98 holds every |F| here (checked below)
    import ctypes as C
    cc = np.array([0.0, 0.0, 1.0, 0.0])
    Xn = np.zeros_like(X)
    orc.lib().orc_step_q12(orc._p(pb.rp), orc._p(pb.col), orc._p(pb.val32), C.c_int64(41),
                           C.c_int64(6), orc._p(X), orc._p(Xn), orc._p(Y), orc._p(cc), C.c_int(sy))
    F = -(sp.csr_matrix(M) @ np.where(np.signbit(X), -1.0, 1.0))
    assert np.abs(F).max() < 2047 * 2.0 ** -sy
    np.testing.assert_array_equal(Y, F)
    np.testing.assert_array_equal(Xn, X)


def test_narrow_dsb_finds_bipartite_max_cut():
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
    r = orc.solve_q12(pb, X, Y, 2000, 0, P=100, dt=0.5, xi0=xi0)
    cuts = orc.cut(pb, r["cur"])
    assert cuts.max() == n
