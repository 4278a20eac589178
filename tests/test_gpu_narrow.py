"""GPU parity of the narrow-state dSB forms (SIS_FORM_Q12 / SIS_FORM_Q10 /
SIS_FORM_Q8: X and Y stored as 12-, 10- or 8-bit fixed-point codes, FP32
arithmetic, exact J sgn(X) from packed sign words) against the oracle's
restatement of the same forms (orc_*_qn with width w = 12 / 10 / 8).

Bar: bit-exact.  The codes, the FP32 update on the decoded values and the
rounding are deterministic, so the device's X, Y codes, spins, locks, step
counts and cuts equal the oracle's exactly, for every width, tile, agent count
(padded to 256 per quad), launch mode and window setting.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
import sis

pytestmark = pytest.mark.gpu

FORMW = {12: sis.FORM_Q12, 10: sis.FORM_Q10, 8: sis.FORM_Q8}
XS = {12: 10, 10: 8, 8: 6}  # X code scale 2^xs; the wall code is +-(2^xs + 1)
VG = 1 << 17  # staged gathers: a wavefront's rows' entries fetched at once into LDS
TILES = [4 | 1 << 8, 4 | 2 << 8, 4 | 4 << 8, 4 | 1 << 8 | 1 << 16, 4 | 2 << 8 | 1 << 16,
         4 | 4 << 8 | 1 << 16, 4 | 4 << 8 | VG, 4 | 8 << 8 | VG | 1 << 16, 4 | 16 << 8 | VG,
         4 | 32 << 8 | VG | 1 << 16]


def problem(n=203, p=0.04, weights=sis.W_INT3, seed=12):
    J = sis.Csr.gen_er(n, p, weights, seed)
    rp, col, val = J.arrays()
    return J, orc.Problem(rp, col, val, S=J.sum)


def seeded(n, A, seed, scale=0.1):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-scale, scale, (n, A)).astype(np.float32)
    Y = rng.uniform(-scale, scale, (n, A)).astype(np.float32)
    return X, Y


@pytest.mark.parametrize("w", [12, 10, 8])
def test_init_random_equals_rounded_oracle_stream(w):
    J, _ = problem(70)
    for A in (64, 300):
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=0.5, form=FORMW[w]))
        s.init_random(9)
        X, Y = orc.init_state(9, orc.MAXCUT, J.n, A)
        orc.qn_round(X, Y, orc.qn_shift(0.5, w), w)
        x, y = s.get_state()
        assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))


@pytest.mark.parametrize("w", [12, 10, 8])
@pytest.mark.parametrize("dt", [0.5, 0.1])
def test_set_state_rounds_like_the_oracle(w, dt):
    # values off the grid, halfway cases, beyond the wall, negative zero and
    # Y beyond the code range: the device's codes are the oracle's rounding
    J, _ = problem(50)
    A = 130
    top = 2.0 ** XS[w]
    sy = orc.qn_shift(dt, w)
    X, Y = seeded(J.n, A, 3, scale=1.5)
    X[0, :8] = np.array([2.5, 3.5, -2.5, -0.3, top, -0.0, top - 0.5, -top - 0.5]) / top
    Y[1, :4] = [100.0, -100.0, 0.5 / 2.0 ** sy, 1.5 / 2.0 ** sy]
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=FORMW[w]))
    s.set_state(X.astype(np.float64), Y.astype(np.float64))
    orc.qn_round(X, Y, sy, w)
    x, y = s.get_state()
    assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))
    assert not np.signbit(x[x == 0]).any()


@pytest.mark.parametrize("w", [12, 10, 8])
@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("agents", [96, 256, 300])
def test_trajectory_bit_exact(w, tile, agents):
    # 203 rows (not a multiple of the rows per wavefront), integer weights,
    # a state that reaches the walls; two advances cross p = 1 at t = 390
    J, pb = problem()
    T, dt = 400, 0.5
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=agents, dt=dt, form=FORMW[w], tile=tile))
    assert s.p.tile == tile
    sy = orc.qn_shift(dt, w)
    X, Y = seeded(J.n, agents, 13, scale=0.9)
    s.set_state(X.astype(np.float64), Y.astype(np.float64))
    orc.qn_round(X, Y, sy, w)
    s.prestep()
    orc.prestep_qn(X, Y, dt, sy, w)
    x, y = s.get_state()
    assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))
    for t0, n in ((0, 150), (150, 250)):
        s.advance(t0, n, T)
        orc.advance_qn(pb, X, Y, t0, n, T, dt, s.p.xi0, sy, w)
        x, y = s.get_state()
        assert np.array_equal(x, X.astype(np.float64)), (w, tile, agents, t0 + n)
        assert np.array_equal(y, Y.astype(np.float64)), (w, tile, agents, t0 + n)
    assert (np.abs(X) > 1).any()  # the wall code occurred


@pytest.mark.parametrize("w,dt", [(8, 0.016), (8, 1.0), (8, 2.0), (10, 0.004), (10, 2.0),
                                  (12, 0.001), (12, 2.0)])
def test_trajectory_bit_exact_at_the_y_range_edges(w, dt):
    # The step encodes Y without the clamp the oracle applies: exact while
    # the code range holds 2/dt + dt.  dt 0.016 (8-bit), 0.004 (10-bit) and
    # 0.001 (12-bit) sit just above the refused range (shift 0: the bound
    # 125.0 / 500.0 / 2000.0 is within one code range of 127 / 511 / 2047);
    # dt 1 and 2 give large steps.  A
    # strong coupling and a state near the walls push |y| up.
    J, pb = problem(203, 0.04, sis.W_INT3, 19)
    A, T = 128, 300
    sy = orc.qn_shift(dt, w)
    assert sy >= 0 and (2 ** (w - 1) - 1) * 2.0 ** -sy >= 2 / dt + dt
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=FORMW[w], xi0=0.9))
    X, Y = seeded(J.n, A, 41, scale=0.99)
    s.set_state(X.astype(np.float64), Y.astype(np.float64))
    orc.qn_round(X, Y, sy, w)
    ymax = 0.0
    for t0, n in ((0, 120), (120, 180)):
        s.advance(t0, n, T)
        orc.advance_qn(pb, X, Y, t0, n, T, dt, 0.9, sy, w)
        x, y = s.get_state()
        assert np.array_equal(x, X.astype(np.float64)), (w, dt, t0 + n)
        assert np.array_equal(y, Y.astype(np.float64)), (w, dt, t0 + n)
        ymax = max(ymax, float(np.abs(Y[np.abs(X) <= 1]).max()))
    # live |y| within the bound (plus half a code of rounding), inside the range
    assert ymax <= 2 / dt + dt + 2.0 ** -(sy + 1) and ymax < (2 ** (w - 1) - 1) * 2.0 ** -sy


@pytest.mark.parametrize("w", [12, 10, 8])
@pytest.mark.parametrize("tile", [4 | 4 << 8 | VG, 4 | 8 << 8 | VG | 1 << 16, 4 | 32 << 8 | VG])
@pytest.mark.parametrize("agents", [96, 300])
def test_unit_weight_waves_beside_real_weight_waves(w, tile, agents):
    # +-1 weights except in rows 40..47 (2 and -0.75) and row 150 (an exact
    # zero): waves whose staged weights are all +-1 count sign bits, the
    # others add +-|v|; both equal the oracle's CSR-order FP32 sums
    J0 = sis.Csr.gen_er(203, 0.04, sis.W_PM1, 12)
    rp, col, val = J0.arrays()
    rows = np.repeat(np.arange(203), np.diff(rp))
    val = val.astype(np.float64).copy()
    sel = (rows >= 40) & (rows < 48)
    val[sel] = np.where(val[sel] > 0, 2.0, -0.75)
    val[np.flatnonzero(rows == 150)[:1]] = 0.0
    J = sis.Csr.from_coo(203, rows, col, val, mirror=False)
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    T, dt = 400, 0.5
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=agents, dt=dt, form=FORMW[w], tile=tile))
    assert s.p.tile == tile
    sy = orc.qn_shift(dt, w)
    X, Y = seeded(J.n, agents, 17, scale=0.9)
    s.set_state(X.astype(np.float64), Y.astype(np.float64))
    orc.qn_round(X, Y, sy, w)
    for t0, n in ((0, 150), (150, 250)):
        s.advance(t0, n, T)
        orc.advance_qn(pb, X, Y, t0, n, T, dt, s.p.xi0, sy, w)
        x, y = s.get_state()
        assert np.array_equal(x, X.astype(np.float64)), (w, tile, agents, t0 + n)
        assert np.array_equal(y, Y.astype(np.float64)), (w, tile, agents, t0 + n)
    assert (np.abs(X) > 1).any()


@pytest.mark.parametrize("w", [12, 10, 8])
@pytest.mark.parametrize("tile", [4 | 8 << 8 | VG, 4 | 16 << 8 | VG | 1 << 16, 4 | 4 << 8 | VG])
@pytest.mark.parametrize("unit", [False, True])
def test_staged_gathers_hub_and_empty_rows(w, tile, unit):
    # 150 rows: a hub joined to rows 1..140 (its wave exceeds the staging
    # capacity and takes the scalar path), a sparse tail among rows 1..129,
    # rows 141..149 without entries (waves whose rows hold no entry at all),
    # real weights or +-1
    rng = np.random.default_rng(31)
    rows = [0] * 140 + list(rng.integers(1, 130, 80))
    cols = list(range(1, 141)) + list(rng.integers(1, 130, 80))
    keep = [(r, c) for r, c in zip(rows, cols) if r != c and r < 130 and c < 141]
    vals = rng.normal(size=len(keep))
    if unit:
        vals = np.sign(vals)
    J = sis.Csr.from_coo(150, [r for r, _ in keep], [c for _, c in keep], vals)
    rp, col, val = J.arrays()
    deg = np.diff(rp)
    assert deg[0] > 128 and (deg[141:] == 0).all()
    pb = orc.Problem(rp, col, val, S=J.sum)
    A, T, dt = 256, 300, 0.5
    sy = orc.qn_shift(dt, w)
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=FORMW[w], tile=tile, xi0=0.2))
    assert s.p.tile == tile
    X, Y = seeded(J.n, A, 32, scale=0.5)
    s.set_state(X.astype(np.float64), Y.astype(np.float64))
    orc.qn_round(X, Y, sy, w)
    s.advance(0, 251, T)
    orc.advance_qn(pb, X, Y, 0, 251, T, dt, 0.2, sy, w)
    x, y = s.get_state()
    assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))


@pytest.mark.parametrize("w", [12, 10, 8])
@pytest.mark.parametrize("uw,T", [(1, 6000), (0, 2345)])
@pytest.mark.parametrize("graph", [0, 1, 2])
def test_full_run_matches_oracle(w, uw, T, graph):
    J, pb = problem(120, 0.06, sis.W_PM1, 5)
    A, dt = 100, 0.5
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=FORMW[w], use_graph=graph))
    runs = 2 if graph else 1
    for _ in range(runs):  # graph replays across runs must agree with eager
        s.init_random(21)
        info = s.run(T, uw)
        X, Y = orc.init_state(21, orc.MAXCUT, J.n, A)
        r = orc.solve_qn(pb, X, Y, T, uw, P=100, K=25, dt=dt, xi0=s.p.xi0, w=w)
        assert info["steps_executed"] == r["steps"]
        x, y = s.get_state()
        assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))
        assert np.array_equal(s.spins(sis.SPINS_CURRENT), r["cur"])
        assert np.array_equal(s.spins(sis.SPINS_FINAL), r["fin"])
        st, lk, _ = s.agent_state()
        assert np.array_equal(st, r["stab"]) and np.array_equal(lk, r["bif"])
        cuts, best = s.cuts()
        want = orc.cut(pb, r["cur"])
        assert np.array_equal(cuts, want) and best == int(np.argmax(want))
        assert info["all_bifurcated"] == int(uw == 1 and r["bif"].all())


@pytest.mark.parametrize("w", [12, 10, 8])
def test_autotune_keeps_state_and_results(w):
    J, pb = problem(300, 0.03, sis.W_PM1, 14)
    A, dt = 256, 0.5
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=FORMW[w]))
    s.init_random(15)
    x0, y0 = s.get_state()
    tile, ms = s.autotune(4)
    assert ms > 0 and s.p.tile == tile and tile & 0xff == 4 and tile >> 18 == 0
    x, y = s.get_state()
    assert np.array_equal(x, x0) and np.array_equal(y, y0)
    info = s.run(700, 0)
    X, Y = orc.init_state(15, orc.MAXCUT, J.n, A)
    r = orc.solve_qn(pb, X, Y, 700, 0, P=100, dt=dt, xi0=s.p.xi0, w=w)
    assert info["steps_executed"] == r["steps"]
    assert np.array_equal(s.get_state()[0], X.astype(np.float64))
    assert np.array_equal(s.cuts()[0], orc.cut(pb, r["cur"]))


@pytest.mark.parametrize("w", [12, 10, 8])
def test_full_size_on_agent_sample_and_exact_cut(w):
    """Random 4-regular N = 2^20, 256 agents (the benchmark graph): 30 steps
    after the pre-step, every 8th agent bit-exact with the oracle; after a
    300-step run the device cut of the returned spins equals an exact
    integer host evaluation."""
    J = sis.Csr.gen_regular(1 << 20, 4, sis.W_PM1, 2024)
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    A, T, dt = 256, 20000, 0.5
    sy = orc.qn_shift(dt, w)
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=FORMW[w]))
    s.init_random(77)
    s.prestep()
    s.advance(0, 30, T)
    cols = np.arange(0, A, 8)
    X, Y = orc.init_state(77, orc.MAXCUT, J.n, A)
    X, Y = np.ascontiguousarray(X[:, cols]), np.ascontiguousarray(Y[:, cols])
    orc.qn_round(X, Y, sy, w)
    orc.prestep_qn(X, Y, dt, sy, w)
    orc.advance_qn(pb, X, Y, 0, 30, T, dt, s.p.xi0, sy, w)
    x, y = s.get_state()
    assert np.array_equal(x[:, cols], X.astype(np.float64))
    assert np.array_equal(y[:, cols], Y.astype(np.float64))
    s.init_random(78)
    assert s.run(300, 0)["steps_executed"] == 300
    cuts, _ = s.cuts()
    M = sp.csr_matrix((val, col, rp), shape=(J.n, J.n))
    s64 = s.spins().astype(np.int64)
    H = 0.5 * np.einsum("ia,ia->a", s64, M @ s64)
    assert np.array_equal(cuts, -0.5 * (H - J.sum))


@pytest.mark.parametrize("w", [12, 10, 8])
def test_rejected_where_it_does_not_apply(w):
    J, _ = problem(60)
    # dt below the Y code range's bound (127 < 2/dt + dt for 8 bits, 511 for
    # 10, 2047 for 12)
    tiny = {8: 0.015, 10: 0.0039, 12: 0.0009}[w]
    bad = [dict(dynamics=sis.BSB), dict(fused=0), dict(dt=0.0), dict(split=8), dict(split=-1),
           dict(tile=1 << 20), dict(dt=tiny)]
    for extra in bad:
        params = dict(problem=sis.MAXCUT, agents=64, dt=0.5, form=FORMW[w])
        params.update(extra)
        with pytest.raises(sis.SisError):
            sis.Solver(J, params)
    t = sis.Tsp.generate(5, 100, 0, 1)
    Jt, h = t.qubo()
    with pytest.raises(sis.SisError):
        sis.Solver(Jt, dict(problem=sis.TSP, agents=64, form=FORMW[w], dynamics=sis.DSB, ncities=5),
                   field=h)
