"""GPU parity of the 8-bit form with stochastic Y rounding (SIS_FORM_Q8S)
against the oracle's restatement (orc_*_q8s: the dither byte of a hash of
step index, node and agent, Y code = nearest-even integer to
fma(y, 2^sy, d)).

Bar: bit-exact.  The dither is a deterministic function of (step, node,
agent), so the device's codes, spins, locks, step counts and cuts equal the
oracle's exactly for every tile, agent count, launch mode and row order.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
import sis

pytestmark = pytest.mark.gpu

VG = 1 << 17
TILES = [4 | 1 << 8, 4 | 2 << 8 | 1 << 16, 4 | 4 << 8, 4 | 4 << 8 | VG, 4 | 8 << 8 | VG | 1 << 16,
         4 | 16 << 8 | VG, 4 | 32 << 8 | VG | 1 << 16]


def problem(n=203, p=0.04, weights=sis.W_INT3, seed=12):
    J = sis.Csr.gen_er(n, p, weights, seed)
    rp, col, val = J.arrays()
    return J, orc.Problem(rp, col, val, S=J.sum)


def test_init_and_set_state_round_to_nearest_even():
    # the initial state and set_state use the 8-bit form's rounding
    J, _ = problem(70)
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=300, dt=0.5, form=sis.FORM_Q8S))
    s.init_random(9)
    X, Y = orc.init_state(9, orc.MAXCUT, J.n, 300)
    orc.qn_round(X, Y, orc.qn_shift(0.5, 8), 8)
    x, y = s.get_state()
    assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))


@pytest.mark.parametrize("tile", TILES)
@pytest.mark.parametrize("agents", [96, 300])
@pytest.mark.parametrize("weights", [sis.W_INT3, sis.W_PM1])
def test_trajectory_bit_exact(tile, agents, weights):
    # 203 rows, a state that reaches the walls; two advances cross p = 1
    J, pb = problem(weights=weights)
    T, dt = 400, 0.5
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=agents, dt=dt, form=sis.FORM_Q8S, tile=tile))
    assert s.p.tile == tile
    sy = orc.qn_shift(dt, 8)
    rng = np.random.default_rng(13)
    X = rng.uniform(-0.9, 0.9, (J.n, agents)).astype(np.float32)
    Y = rng.uniform(-0.9, 0.9, (J.n, agents)).astype(np.float32)
    s.set_state(X.astype(np.float64), Y.astype(np.float64))
    orc.qn_round(X, Y, sy, 8)
    s.prestep()
    orc.prestep_qn(X, Y, dt, sy, 8)
    for t0, n in ((0, 150), (150, 250)):
        s.advance(t0, n, T)
        orc.advance_qn(pb, X, Y, t0, n, T, dt, s.p.xi0, sy, 8, sr=True)
        x, y = s.get_state()
        assert np.array_equal(x, X.astype(np.float64)), (tile, agents, t0 + n)
        assert np.array_equal(y, Y.astype(np.float64)), (tile, agents, t0 + n)
    assert (np.abs(X) > 1).any()


def test_differs_from_nearest_even_rounding():
    # the same start, the 8-bit form without the dither: another trajectory
    J, pb = problem()
    outs = []
    for form in (sis.FORM_Q8, sis.FORM_Q8S):
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=256, dt=0.5, form=form))
        s.init_random(4)
        s.advance(0, 50, 400)
        outs.append(s.get_state()[1])
    assert not np.array_equal(outs[0], outs[1])


@pytest.mark.parametrize("uw,T", [(1, 6000), (0, 2345)])
@pytest.mark.parametrize("graph", [0, 1, 2])
def test_full_run_matches_oracle(uw, T, graph):
    J, pb = problem(120, 0.06, sis.W_PM1, 5)
    A, dt = 100, 0.5
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORM_Q8S, use_graph=graph))
    for _ in range(2 if graph else 1):  # graph replays across runs agree with eager
        s.init_random(21)
        info = s.run(T, uw)
        X, Y = orc.init_state(21, orc.MAXCUT, J.n, A)
        r = orc.solve_qn(pb, X, Y, T, uw, P=100, K=25, dt=dt, xi0=s.p.xi0, w=8, sr=True)
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


@pytest.mark.parametrize("unit", [False, True])
def test_staged_hub_and_empty_rows(unit):
    # a hub beyond the staging capacity (scalar path), rows without entries
    # (waves that load nothing), real or +-1 weights
    rng = np.random.default_rng(31)
    rows = [0] * 140 + list(rng.integers(1, 130, 80))
    cols = list(range(1, 141)) + list(rng.integers(1, 130, 80))
    keep = [(r, c) for r, c in zip(rows, cols) if r != c and r < 130 and c < 141]
    vals = rng.normal(size=len(keep))
    if unit:
        vals = np.sign(vals)
    J = sis.Csr.from_coo(150, [r for r, _ in keep], [c for _, c in keep], vals)
    rp, col, val = J.arrays()
    assert np.diff(rp)[0] > 128 and (np.diff(rp)[141:] == 0).all()
    pb = orc.Problem(rp, col, val, S=J.sum)
    A, T, dt = 256, 300, 0.5
    sy = orc.qn_shift(dt, 8)
    for tile in (4 | 8 << 8 | VG, 4 | 16 << 8 | VG | 1 << 16, 4 | 4 << 8 | VG):
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORM_Q8S, tile=tile,
                               xi0=0.2))
        rng2 = np.random.default_rng(32)
        X = rng2.uniform(-0.5, 0.5, (J.n, A)).astype(np.float32)
        Y = rng2.uniform(-0.5, 0.5, (J.n, A)).astype(np.float32)
        s.set_state(X.astype(np.float64), Y.astype(np.float64))
        orc.qn_round(X, Y, sy, 8)
        s.advance(0, 251, T)
        orc.advance_qn(pb, X, Y, 0, 251, T, dt, 0.2, sy, 8, sr=True)
        x, y = s.get_state()
        assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))


def test_autotune_and_full_size_exact_cut():
    J, pb = problem(300, 0.03, sis.W_PM1, 14)
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=256, dt=0.5, form=sis.FORM_Q8S))
    s.init_random(15)
    x0, y0 = s.get_state()
    tile, ms = s.autotune(4)
    assert ms > 0 and s.p.tile == tile
    x, y = s.get_state()
    assert np.array_equal(x, x0) and np.array_equal(y, y0)
    info = s.run(700, 0)
    X, Y = orc.init_state(15, orc.MAXCUT, J.n, 256)
    r = orc.solve_qn(pb, X, Y, 700, 0, P=100, dt=0.5, xi0=s.p.xi0, w=8, sr=True)
    assert info["steps_executed"] == r["steps"]
    assert np.array_equal(s.cuts()[0], orc.cut(pb, r["cur"]))
    # the benchmark graph (N = 2^20, rows in the locality order): the device
    # cut of the returned spins equals an exact integer host evaluation
    J = sis.Csr.gen_regular(1 << 20, 4, sis.W_PM1, 2024)
    rp, col, val = J.arrays()
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=256, dt=0.5, form=sis.FORM_Q8S))
    assert s.p.row_order == 1
    s.init_random(78)
    assert s.run(300, 0)["steps_executed"] == 300
    cuts, _ = s.cuts()
    M = sp.csr_matrix((val, col, rp), shape=(J.n, J.n))
    s64 = s.spins().astype(np.int64)
    H = 0.5 * np.einsum("ia,ia->a", s64, M @ s64)
    assert np.array_equal(cuts, -0.5 * (H - J.sum))


def test_rejected_where_it_does_not_apply():
    J, _ = problem(60)
    bad = [dict(dynamics=sis.BSB), dict(fused=0), dict(dt=0.0), dict(split=8), dict(tile=1 << 20),
           dict(dt=0.015)]
    for extra in bad:
        params = dict(problem=sis.MAXCUT, agents=64, dt=0.5, form=sis.FORM_Q8S)
        params.update(extra)
        with pytest.raises(sis.SisError):
            sis.Solver(J, params)
