"""GPU parity of the row storage order (sis_params.row_order).

The locality order stores the rows of the sign-gather forms (packed, 12-, 10-
and 8-bit codes, the last also with stochastic Y rounding, whose dither is
keyed by node) grouped by graph neighbourhood; states, spins and results
still come and go in node order, each row keeps its entries in CSR order and
the seeded initial state is keyed by node.  Bar: bit-exact -- a solver in
the locality order equals one in node order and the oracle in every state,
spin, lock, step count and cut.
"""
import numpy as np
import pytest

import oracle as orc
import sis

pytestmark = pytest.mark.gpu

WIDTH = {"q12": 12, "q10": 10, "q8": 8, "q8s": 8}


def oracle_solve(pb, form, X, Y, T, uw, dt, xi0):
    if form in WIDTH:
        return orc.solve_qn(pb, X, Y, T, uw, P=100, K=25, dt=dt, xi0=xi0, w=WIDTH[form],
                            sr=form == "q8s")
    return orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, T, uw, P=100, K=25, dt=dt, xi0=xi0)


def solvers(J, form, **kw):
    base = dict(problem=sis.MAXCUT, agents=300, dt=0.5, form=sis.FORMS[form], **kw)
    s1 = sis.Solver(J, dict(base, row_order=1))
    s0 = sis.Solver(J, dict(base, row_order=-1))
    assert s1.p.row_order == 1 and s0.p.row_order == -1
    return s1, s0


def same(s1, s0):
    x1, y1 = s1.get_state()
    x0, y0 = s0.get_state()
    assert np.array_equal(x1, x0) and np.array_equal(y1, y0)
    for which in (sis.SPINS_CURRENT, sis.SPINS_FINAL):
        assert np.array_equal(s1.spins(which), s0.spins(which))
    a1, a0 = s1.agent_state(), s0.agent_state()
    assert all(np.array_equal(u, v) for u, v in zip(a1, a0))
    c1, b1 = s1.cuts()
    c0, b0 = s0.cuts()
    assert np.array_equal(c1, c0) and b1 == b0


@pytest.mark.parametrize("form", ["packed", "q12", "q10", "q8", "q8s"])
@pytest.mark.parametrize("graph", [0, 2])
def test_locality_order_full_run_equals_node_order_and_oracle(form, graph):
    # integer weights in [-3, 3] on a 4-regular graph, 300 agents (ragged),
    # windows with early exit, direct launches or whole-run graph replay
    J = sis.Csr.gen_regular(3000, 4, sis.W_INT3, 8)
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    T, dt = 1700, 0.5
    s1, s0 = solvers(J, form, use_graph=graph)
    for s in (s1, s0):
        s.init_random(31)
    i1, i0 = s1.run(T, 1), s0.run(T, 1)
    assert i1["steps_executed"] == i0["steps_executed"]
    same(s1, s0)
    X, Y = orc.init_state(31, orc.MAXCUT, J.n, 300)
    r = oracle_solve(pb, form, X, Y, T, 1, dt, s1.p.xi0)
    assert i1["steps_executed"] == r["steps"]
    assert np.array_equal(s1.get_state()[0], X.astype(np.float64))
    assert np.array_equal(s1.spins(sis.SPINS_CURRENT), r["cur"])
    assert np.array_equal(s1.spins(sis.SPINS_FINAL), r["fin"])
    assert np.array_equal(s1.cuts()[0], orc.cut(pb, r["cur"]))


@pytest.mark.parametrize("form", ["packed", "q10", "q8", "q8s"])
def test_locality_order_set_state_and_advance(form):
    # explicit states in node order go in and come out in node order; an
    # isolated node and an ER graph with uneven degrees
    J = sis.Csr.gen_er(2500, 0.003, sis.W_PM1, 4)
    rp, col, val = J.arrays()
    assert (np.diff(rp) == 0).any()
    s1, s0 = solvers(J, form, xi0=0.3)
    rng = np.random.default_rng(6)
    X = rng.uniform(-0.9, 0.9, (J.n, 300))
    Y = rng.uniform(-0.9, 0.9, (J.n, 300))
    for s in (s1, s0):
        s.set_state(X, Y)
    assert np.array_equal(s1.get_state()[0], s0.get_state()[0])
    for s in (s1, s0):
        s.prestep()
        s.advance(0, 233, 500)
    x1, y1 = s1.get_state()
    x0, y0 = s0.get_state()
    assert np.array_equal(x1, x0) and np.array_equal(y1, y0)


def test_locality_order_self_loops_duplicates_and_row_split():
    # a COO with self loops and duplicate entries (kept, as the reference's
    # CSR keeps them), integer weights; then a power-law graph whose heavy rows
    # go through the row split: the same sums in either row order
    rng = np.random.default_rng(12)
    n = 1500
    rows = list(rng.integers(0, n, 6000)) + list(range(0, n, 7)) + [3] * 5
    cols = list(rng.integers(0, n, 6000)) + list(range(0, n, 7)) + [9] * 5
    vals = list(rng.choice([-2.0, -1.0, 1.0, 3.0], len(rows)))
    J = sis.Csr.from_coo(n, rows, cols, vals, mirror=False)
    rp, col, val = J.arrays()
    r_of = np.repeat(np.arange(n), np.diff(rp))
    assert (r_of == col).any()  # self loops present
    pb = orc.Problem(rp, col, val, S=J.sum)
    for form in ("packed", "q12", "q8s"):
        s1, s0 = solvers(J, form, xi0=0.2)
        for s in (s1, s0):
            s.init_random(3)
        s1.run(700, 1), s0.run(700, 1)
        same(s1, s0)
        X, Y = orc.init_state(3, orc.MAXCUT, n, 300)
        r = oracle_solve(pb, form, X, Y, 700, 1, 0.5, 0.2)
        assert np.array_equal(s1.spins(sis.SPINS_CURRENT), r["cur"])
    P = sis.Csr.gen_chung_lu(6000, 8.0, 2.1, sis.W_PM1, 4)
    assert np.diff(P.arrays()[0]).max() > 256
    s1, s0 = solvers(P, "packed", split=64)
    for s in (s1, s0):
        s.init_random(8)
        s.advance(0, 250, 1000)
    x1, y1 = s1.get_state()
    x0, y0 = s0.get_state()
    assert np.array_equal(x1, x0) and np.array_equal(y1, y0)


def test_locality_order_default_at_the_benchmark_size():
    # the default (0) chooses the locality order once the sign words exceed
    # 4 MiB (N x agents / 8), and node order below
    J = sis.Csr.gen_regular(1 << 20, 4, sis.W_PM1, 3)
    big = sis.Solver(J, dict(problem=sis.MAXCUT, agents=256, dt=0.5, form=sis.FORM_Q8))
    assert big.p.row_order == 1
    S = sis.Csr.gen_regular(4096, 4, sis.W_PM1, 3)
    small = sis.Solver(S, dict(problem=sis.MAXCUT, agents=256, dt=0.5, form=sis.FORM_Q8))
    assert small.p.row_order == -1


def test_locality_order_declined_where_it_does_not_apply():
    # real weights (the FP64 cut sums would depend on the row order), bSB,
    # the FP32 forms and TSP keep node order even when asked
    real = sis.Csr.gen_er(500, 0.02, sis.W_PM1, 2)
    rp, col, val = real.arrays()
    rows = np.repeat(np.arange(real.n), np.diff(rp))
    R = sis.Csr.from_coo(real.n, rows, col, val * 0.75, mirror=False)
    I = sis.Csr.gen_er(500, 0.02, sis.W_PM1, 2)
    cases = [(R, dict(form=sis.FORM_PACKED)),
             (I, dict(form=sis.FORM_F32)),
             (I, dict(form=sis.FORM_BF16, dynamics=sis.BSB)),
             (I, dict(form=sis.FORM_F32, dynamics=sis.BSB))]
    for J, extra in cases:
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=64, dt=0.5, row_order=1, **extra))
        assert s.p.row_order == -1, extra
    with pytest.raises(sis.SisError):
        sis.Solver(I, dict(problem=sis.MAXCUT, agents=64, form=sis.FORM_PACKED, row_order=2))
    t = sis.Tsp.generate(6, 100, 0, 1)
    Jt, h = t.qubo()
    s = sis.Solver(Jt, dict(problem=sis.TSP, agents=64, ncities=6, dynamics=sis.BSB, row_order=1),
                   field=h)
    assert s.p.row_order == -1
