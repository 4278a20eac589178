"""GPU parity: the HIP solver (through the C ABI) against the CPU oracle.

Bars: FP32 / FP64 / packed / BF16 trajectories are bit-exact with the oracle
evaluated in the same precision and order (explicit FMAs on both sides,
no other contraction); integer cuts, energies and tour lengths are exact.
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
import sis

pytestmark = pytest.mark.gpu


def maxcut_problem(n=200, p=0.05, weights=sis.W_INT3, seed=1):
    J = sis.Csr.gen_er(n, p, weights, seed)
    rp, col, val = J.arrays()
    return J, orc.Problem(rp, col, val, S=J.sum)


def tsp_problem(n=6, seed=3, euc_round=0):
    t = sis.Tsp.generate(n, 100, euc_round, seed)
    J, h = t.qubo()
    rp, col, val = J.arrays()
    return t, J, h, orc.Problem(rp, col, val, field=h, ncities=n)


def seeded_state(n, A, seed, tsp=False):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, A)) if tsp else rng.uniform(-0.1, 0.1, (n, A))
    Y = rng.uniform(-0.1, 0.1, (n, A))
    return X.astype(np.float32).astype(np.float64), Y.astype(np.float32).astype(np.float64)


FORMS = [  # (form, dynamics, fused, oracle precision, oracle problem override)
    ("f32", sis.DSB, 1, "f32"),
    ("f32", sis.BSB, 1, "f32"),
    ("f64", sis.DSB, 1, "f64"),
    ("f64", sis.BSB, 1, "f64"),
    ("packed", sis.DSB, 1, "f32"),
    ("bf16", sis.BSB, 1, "bf16"),
    ("f32", sis.DSB, 0, "f32"),
    ("f64", sis.DSB, 0, "f64"),
]


@pytest.mark.parametrize("seed", [7])
def test_init_random_matches_oracle(seed):
    J, pb = maxcut_problem(50)
    for form, f64 in (("f32", False), ("f64", True)):
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=70, form=sis.FORMS[form]))
        s.init_random(seed)
        x, y = s.get_state()
        X, Y = orc.init_state(seed, orc.MAXCUT, 50, 70, f64=f64)
        assert np.array_equal(x, X) and np.array_equal(y, Y)


@pytest.mark.parametrize("form,dyn,fused,prec", FORMS)
@pytest.mark.parametrize("agents", [96, 256])
def test_maxcut_trajectory_bit_exact(form, dyn, fused, prec, agents):
    J, pb = maxcut_problem()
    X0, Y0 = seeded_state(J.n, agents, 11)
    T, dt = 400, 0.5 if dyn == sis.DSB else 0.1
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=agents, form=sis.FORMS[form], dynamics=dyn,
                           dt=dt, fused=fused))
    xi0 = s.p.xi0
    s.set_state(X0, Y0)
    problem = orc.MAXCUT if fused else orc.MAXCUT_UNFUSED
    np_dt = np.float64 if prec == "f64" else np.float32
    X, Y = X0.astype(np_dt), Y0.astype(np_dt)
    for t0, n in ((0, 150), (150, 300)):  # crosses p = 1 at t = 390
        s.advance(t0, n, T)
        orc.advance(pb, problem, dyn, X, Y, t0, n, T, dt, xi0, 1.0, prec)
        x, y = s.get_state()
        np.testing.assert_array_equal(x, X.astype(np.float64), err_msg=f"X after {t0 + n}")
        np.testing.assert_array_equal(y, Y.astype(np.float64), err_msg=f"Y after {t0 + n}")


@pytest.mark.parametrize("form,dyn,prec", [("f32", sis.BSB, "f32"), ("f64", sis.BSB, "f64"),
                                           ("bf16", sis.BSB, "bf16"), ("packed", sis.DSB, "f32"),
                                           ("f32", sis.DSB, "f32")])
def test_tsp_trajectory_bit_exact(form, dyn, prec):
    t, J, h, pb = tsp_problem(6)
    A = 128
    X0, Y0 = seeded_state(J.n, A, 12, tsp=True)
    T, dt = 1000, 0.1
    s = sis.Solver(J, dict(problem=sis.TSP, agents=A, form=sis.FORMS[form], dynamics=dyn, dt=dt,
                           ncities=6), field=h)
    s.set_state(X0, Y0)
    s.prestep()
    np_dt = np.float64 if prec == "f64" else np.float32
    X, Y = X0.astype(np_dt), Y0.astype(np_dt)
    orc.prestep(X, Y, dt)
    s.advance(0, 600, T)
    orc.advance(pb, orc.TSP, dyn, X, Y, 0, 600, T, dt, 0.002, 1.005, prec)
    x, y = s.get_state()
    np.testing.assert_array_equal(x, X.astype(np.float64))
    np.testing.assert_array_equal(y, Y.astype(np.float64))


@pytest.mark.parametrize("form,dyn,prec", [("f32", sis.DSB, "f32"), ("packed", sis.DSB, "f32"),
                                           ("f64", sis.DSB, "f64"), ("bf16", sis.BSB, "bf16")])
@pytest.mark.parametrize("uw,T", [(1, 6000), (0, 2345)])
@pytest.mark.parametrize("graph", [0, 1, 2])
def test_maxcut_full_run_matches_oracle(form, dyn, prec, uw, T, graph):
    J, pb = maxcut_problem(120, 0.06, sis.W_PM1, 5)
    A = 100  # not a multiple of 64: padded agents must stay invisible
    dt = 0.5 if dyn == sis.DSB else 0.2
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, form=sis.FORMS[form], dynamics=dyn, dt=dt,
                           use_graph=graph))
    X0, Y0 = seeded_state(J.n, A, 21)
    runs = 2 if graph else 1
    for _ in range(runs):  # graph replays across runs must agree with eager
        s.set_state(X0, Y0)
        info = s.run(T, uw)
        np_dt = np.float64 if prec == "f64" else np.float32
        X, Y = X0.astype(np_dt), Y0.astype(np_dt)
        r = orc.solve(pb, orc.MAXCUT, dyn, X, Y, T, uw, P=100, K=25, dt=dt, xi0=s.p.xi0,
                      precision=prec)
        assert info["steps_executed"] == r["steps"]
        x, y = s.get_state()
        np.testing.assert_array_equal(x, X.astype(np.float64))
        np.testing.assert_array_equal(y, Y.astype(np.float64))
        assert np.array_equal(s.spins(sis.SPINS_CURRENT), r["cur"])
        assert np.array_equal(s.spins(sis.SPINS_FINAL), r["fin"])
        st, lk, _ = s.agent_state()
        assert np.array_equal(st, r["stab"]) and np.array_equal(lk, r["bif"])
        cuts, best = s.cuts()
        want = orc.cut(pb, r["cur"])
        assert np.array_equal(cuts, want)
        assert best == int(np.argmax(want))
        assert info["all_bifurcated"] == int(uw == 1 and r["bif"].all())


def test_early_exit_at_window_24_with_dt_zero():
    J, pb = maxcut_problem(64, 0.1, sis.W_PM1, 2)
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=64, dt=0.0))
    X0 = -np.abs(seeded_state(64, 64, 3)[0]) - 0.01
    Y0 = seeded_state(64, 64, 4)[1]
    s.set_state(X0, Y0)
    info = s.run(100000, 1)
    assert info["steps_executed"] == 2400 and info["all_bifurcated"] == 1
    st, lk, _ = s.agent_state()
    assert (st == 24).all() and lk.all()


def test_step_count_edge_cases():
    J, pb = maxcut_problem(40, 0.2, sis.W_PM1, 4)
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=1))
    X0, Y0 = seeded_state(40, 1, 5)
    for T, expect in ((0, 0), (1, 100), (150, 200), (200, 200)):
        s.set_state(X0, Y0)
        assert s.run(T, 0)["steps_executed"] == expect
    cuts, best = s.cuts()
    assert best == 0 and cuts.shape == (1,)


def test_isolated_nodes_and_real_weights(tmp_path):
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate real symmetric\n9 9 5\n2 1 0.5\n4 2 -1.25\n"
                 "5 1 2.0\n9 4 0.75\n7 5 1.5\n")
    J = sis.Csr.read_mtx(p)
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=64, dt=0.3))
    X0, Y0 = seeded_state(9, 64, 6)
    s.set_state(X0, Y0)
    s.run(1000, 0)
    X, Y = X0.astype(np.float32), Y0.astype(np.float32)
    r = orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, 1000, 0, 100, dt=0.3, xi0=s.p.xi0)
    assert np.array_equal(s.get_state()[0], X.astype(np.float64))
    cuts, _ = s.cuts()
    np.testing.assert_allclose(cuts, orc.cut(pb, r["cur"]), rtol=1e-12, atol=1e-12)


@pytest.mark.parametrize("form,dyn,prec", [("f32", sis.BSB, "f32"), ("bf16", sis.BSB, "bf16"),
                                           ("f64", sis.BSB, "f64")])
def test_tsp_full_run_matches_oracle(form, dyn, prec):
    n = 5
    t, J, h, pb = tsp_problem(n, 8)
    A, T, dt = 256, 2000, 0.1
    s = sis.Solver(J, dict(problem=sis.TSP, agents=A, form=sis.FORMS[form], dynamics=dyn, dt=dt,
                           ncities=n), field=h)
    s.init_random(31)
    X, Y = orc.init_state(31, orc.TSP, n * n, A, f64=prec == "f64")
    info = s.run(T, 0)
    r = orc.solve(pb, orc.TSP, dyn, X, Y, T, 0, P=T // 10, K=50, dt=dt, xi0=0.002, coupling=1.005,
                  precision=prec)
    assert info["steps_executed"] == r["steps"] == T
    _, _, vs = s.agent_state()
    assert np.array_equal(vs, r["valid"])
    assert np.array_equal(s.spins(sis.SPINS_FINAL), r["fin"])
    lens, best = s.tsp_lengths(t)
    _, _, d, _ = t.arrays()
    want = np.array([orc.tsp_tour_length(n, d, r["fin"], a)[0] if r["valid"][a] > 0 else -1.0
                     for a in range(A)])
    assert np.array_equal(lens, want)
    if (want >= 0).any():
        assert best == int(np.argmin(np.where(want >= 0, want, np.inf)))
    E = s.energy(sis.SPINS_FINAL)
    assert np.array_equal(E, orc.energy_field(pb, r["fin"]))


def test_runs_are_deterministic():
    J, _ = maxcut_problem(300, 0.03, sis.W_PM1, 9)
    out = []
    for _ in range(2):
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=256, dt=0.5, form=sis.FORM_PACKED))
        s.init_random(5)
        s.run(1500, 1)
        out.append((s.get_state()[0], s.cuts()[0]))
    assert np.array_equal(out[0][0], out[1][0]) and np.array_equal(out[0][1], out[1][1])


PACKED_TILES = [1 | 1 << 8, 2 | 2 << 8 | 1 << 16, 4 | 4 << 8, 8 | 1 << 8 | 1 << 16,
                8 | 4 << 8, 1 | 4 << 8 | 1 << 16]


@pytest.mark.parametrize("tile", PACKED_TILES)
def test_packed_tiles_bit_exact(tile):
    # 203 rows: not a multiple of the rows-per-wavefront tile
    J, pb = maxcut_problem(203, 0.04, sis.W_INT3, 12)
    A, T, dt = 512, 300, 0.5
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORM_PACKED, tile=tile))
    assert s.p.tile == tile
    X0, Y0 = seeded_state(J.n, A, 13)
    s.set_state(X0, Y0)
    s.advance(0, 250, T)
    X, Y = X0.astype(np.float32), Y0.astype(np.float32)
    orc.advance(pb, orc.MAXCUT, orc.DSB, X, Y, 0, 250, T, dt, s.p.xi0, 1.0, "f32")
    x, y = s.get_state()
    assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))


@pytest.mark.parametrize("form,dyn,tile", [("f32", sis.DSB, 4 | 1 << 16), ("f32", sis.BSB, 2 | 1 << 16),
                                           ("f32", sis.DSB, 1 | 1 << 16), ("f64", sis.BSB, 2 | 1 << 16),
                                           ("bf16", sis.BSB, 4 | 1 << 16), ("bf16", sis.BSB, 1)])
def test_plain_and_bf16_tiles_bit_exact(form, dyn, tile):
    J, pb = maxcut_problem(150, 0.05, sis.W_INT3, 16)
    A, T, dt = 256, 300, 0.3
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORMS[form], dynamics=dyn,
                           tile=tile))
    assert s.p.tile == tile
    X0, Y0 = seeded_state(J.n, A, 17)
    s.set_state(X0, Y0)
    s.advance(0, 200, T)
    prec = {"f32": "f32", "f64": "f64", "bf16": "bf16"}[form]
    np_dt = np.float64 if form == "f64" else np.float32
    X, Y = X0.astype(np_dt), Y0.astype(np_dt)
    orc.advance(pb, orc.MAXCUT, dyn, X, Y, 0, 200, T, dt, s.p.xi0, 1.0, prec)
    x, y = s.get_state()
    assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))


RES = 1 << 20  # resident tile: one workgroup per 2^lg agents, state in registers / LDS


@pytest.mark.parametrize("form,dyn,lg", [("f32", sis.DSB, 0), ("f32", sis.DSB, 3), ("f32", sis.BSB, 0),
                                         ("f32", sis.BSB, 2), ("packed", sis.DSB, 1),
                                         ("packed", sis.DSB, 3)])
def test_resident_tiles_bit_exact(form, dyn, lg):
    # +-1 weights (the resident CSR encoding), 203 rows; two launches of odd
    # length so the snapshot parity flips between them
    J, pb = maxcut_problem(203, 0.04, sis.W_PM1, 21)
    A, T, dt = 128, 300, 0.5
    tile = RES | lg << 21
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORMS[form], dynamics=dyn,
                           tile=tile))
    assert s.p.tile == tile
    X0, Y0 = seeded_state(J.n, A, 22)
    s.set_state(X0, Y0)
    s.advance(0, 137, T)
    s.advance(137, 63, T)
    X, Y = X0.astype(np.float32), Y0.astype(np.float32)
    orc.advance(pb, orc.MAXCUT, dyn, X, Y, 0, 200, T, dt, s.p.xi0, 1.0, "f32")
    x, y = s.get_state()
    assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))


@pytest.mark.parametrize("form,dyn,lg", [("f32", sis.DSB, 0), ("f32", sis.BSB, 1), ("packed", sis.DSB, 3)])
def test_resident_hub_and_empty_rows_bit_exact(form, dyn, lg):
    # 150 rows (not a multiple of 64): a hub joined to rows 1..100, a sparse
    # +-1 tail, rows 120..149 without entries.  ELL groups of very different
    # widths, zero-slot padding and empty rows.
    rng = np.random.default_rng(31)
    rows = [0] * 100 + list(rng.integers(1, 120, 60))
    cols = list(range(1, 101)) + list(rng.integers(1, 120, 60))
    keep = [(r, c) for r, c in zip(rows, cols) if r != c]
    vals = rng.choice([-1.0, 1.0], len(keep))
    J = sis.Csr.from_coo(150, [r for r, _ in keep], [c for _, c in keep], vals)
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    A, T, dt = 64, 300, 0.5
    tile = RES | lg << 21
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORMS[form], dynamics=dyn,
                           tile=tile))
    assert s.p.tile == tile
    X0, Y0 = seeded_state(J.n, A, 32)
    s.set_state(X0, Y0)
    s.advance(0, 151, T)
    X, Y = X0.astype(np.float32), Y0.astype(np.float32)
    orc.advance(pb, orc.MAXCUT, dyn, X, Y, 0, 151, T, dt, s.p.xi0, 1.0, "f32")
    x, y = s.get_state()
    assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))


@pytest.mark.parametrize("use_graph", [0, 1, 2])
@pytest.mark.parametrize("form,lg", [("packed", 2), ("f32", 0)])
def test_resident_full_runs_with_windows(use_graph, form, lg):
    # torus 100 x 100 (N = 10^4: ten register slots per thread at G = 1),
    # update window with early exit, graph replay: steps, spins, locks and
    # cuts equal the oracle's
    J = sis.Csr.gen_torus(100, 100, sis.W_PM1, 5) if lg == 0 else sis.Csr.gen_torus(40, 40, sis.W_PM1, 5)
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    A, T, dt = 64, 1500, 0.5
    tile = RES | lg << 21
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORMS[form], tile=tile,
                           use_graph=use_graph, window_threshold=5))
    assert s.p.tile == tile
    s.init_random(7)
    info = s.run(T, 1)
    X, Y = orc.init_state(7, orc.MAXCUT, J.n, A)
    r = orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, T, 1, P=100, K=5, dt=dt, xi0=s.p.xi0)
    assert info["steps_executed"] == r["steps"]
    assert np.array_equal(s.get_state()[0], X.astype(np.float64))
    assert np.array_equal(s.spins(sis.SPINS_FINAL), r["fin"])
    assert np.array_equal(s.cuts()[0], orc.cut(pb, r["cur"]))


RREG = 1 << 23  # resident dSB tile with the ELL in registers (groups at most two chunks wide)


def low_degree_graph(n, seed):
    # +-1 weights, degrees 0..8 (ELL groups of zero, one and two chunks), the
    # last rows empty; mixed enough that the degree-sorted layout applies
    rng = np.random.default_rng(seed)
    deg = np.zeros(n, int)
    rows, cols = [], []
    for _ in range(3 * n):
        a, b = rng.integers(0, n - 9, 2)
        if a != b and deg[a] < 8 and deg[b] < 8 and rng.random() < (0.9 if a % 3 else 0.2):
            rows.append(a)
            cols.append(b)
            deg[a] += 1
            deg[b] += 1
    vals = rng.choice([-1.0, 1.0], len(rows))
    J = sis.Csr.from_coo(n, rows, cols, vals)
    rp, col, val = J.arrays()
    return J, orc.Problem(rp, col, val, S=J.sum)


@pytest.mark.parametrize("graph,form,lg,A", [("torus", "f32", 0, 64), ("torus", "packed", 2, 128),
                                             ("regular", "f32", 1, 128), ("low", "packed", 0, 64),
                                             ("low", "f32", 2, 64)])
def test_resident_register_ell_bit_exact(graph, form, lg, A):
    # ELL entries held in registers, gather slots alone in LDS: the same sums
    # as the LDS tile; two launches of odd length flip the snapshot parity
    if graph == "torus":
        J = sis.Csr.gen_torus(40, 40, sis.W_PM1, 3)
        rp, col, val = J.arrays()
        pb = orc.Problem(rp, col, val, S=J.sum)
    elif graph == "regular":
        J = sis.Csr.gen_regular(1000, 4, sis.W_PM1, 4)
        rp, col, val = J.arrays()
        pb = orc.Problem(rp, col, val, S=J.sum)
    else:
        J, pb = low_degree_graph(700, 9)
        assert np.diff(J.arrays()[0]).max() <= 8 and np.diff(J.arrays()[0])[-9:].max() == 0
    T, dt = 300, 0.5
    tile = RES | lg << 21 | RREG
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORMS[form], tile=tile))
    assert s.p.tile == tile
    X0, Y0 = seeded_state(J.n, A, 23)
    s.set_state(X0, Y0)
    s.advance(0, 137, T)
    s.advance(137, 63, T)
    X, Y = X0.astype(np.float32), Y0.astype(np.float32)
    orc.advance(pb, orc.MAXCUT, orc.DSB, X, Y, 0, 200, T, dt, s.p.xi0, 1.0, "f32")
    x, y = s.get_state()
    assert np.array_equal(x, X.astype(np.float64)) and np.array_equal(y, Y.astype(np.float64))


@pytest.mark.parametrize("use_graph", [0, 1, 2])
@pytest.mark.parametrize("form", ["packed", "f32"])
def test_resident_register_ell_torus_full_runs(use_graph, form):
    # torus 100 x 100 at G = 1: ten register slots per thread, two workgroups
    # per CU; windows with early exit and graph replay against the oracle
    J = sis.Csr.gen_torus(100, 100, sis.W_PM1, 5)
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    A, T, dt = 128, 1500, 0.5
    tile = RES | RREG
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORMS[form], tile=tile,
                           use_graph=use_graph, window_threshold=5))
    assert s.p.tile == tile
    s.init_random(7)
    info = s.run(T, 1)
    X, Y = orc.init_state(7, orc.MAXCUT, J.n, A)
    r = orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, T, 1, P=100, K=5, dt=dt, xi0=s.p.xi0)
    assert info["steps_executed"] == r["steps"]
    assert np.array_equal(s.get_state()[0], X.astype(np.float64))
    assert np.array_equal(s.spins(sis.SPINS_FINAL), r["fin"])
    assert np.array_equal(s.cuts()[0], orc.cut(pb, r["cur"]))


def test_resident_register_ell_rejected_where_it_does_not_apply():
    J, _ = maxcut_problem(203, 0.08, sis.W_PM1, 21)  # rows wider than two chunks
    assert np.diff(J.arrays()[0]).max() > 8
    with pytest.raises(sis.SisError):
        sis.Solver(J, dict(problem=sis.MAXCUT, agents=64, form=sis.FORM_PACKED, tile=RES | RREG))
    L, _ = low_degree_graph(700, 9)  # two-chunk groups beyond 4,096 pairs (G = 8: 5,600)
    with pytest.raises(sis.SisError):
        sis.Solver(L, dict(problem=sis.MAXCUT, agents=64, form=sis.FORM_F32, tile=RES | 3 << 21 | RREG))
    T = sis.Csr.gen_torus(40, 40, sis.W_PM1, 3)
    with pytest.raises(sis.SisError):  # bSB keeps float slots: no register-ELL tile
        sis.Solver(T, dict(problem=sis.MAXCUT, agents=64, form=sis.FORM_F32, dynamics=sis.BSB,
                           tile=RES | RREG))


def test_resident_tile_rejected_where_it_does_not_apply():
    J, _ = maxcut_problem(203, 0.04, sis.W_INT3, 21)  # weights outside +-1
    with pytest.raises(sis.SisError):
        sis.Solver(J, dict(problem=sis.MAXCUT, agents=128, form=sis.FORM_PACKED, tile=RES))
    Jp, _ = maxcut_problem(203, 0.04, sis.W_PM1, 21)
    with pytest.raises(sis.SisError):  # BF16 shadow form
        sis.Solver(Jp, dict(problem=sis.MAXCUT, agents=128, form=sis.FORM_BF16, dynamics=sis.BSB,
                            tile=RES))
    big = sis.Csr.gen_torus(200, 200, sis.W_PM1, 1)  # N = 40,000 > 12,288 pairs
    with pytest.raises(sis.SisError):
        sis.Solver(big, dict(problem=sis.MAXCUT, agents=64, form=sis.FORM_PACKED, tile=RES))


@pytest.mark.parametrize("form", ["packed", "f32", "bf16"])
def test_autotune_keeps_state_and_results(form):
    J, pb = maxcut_problem(300, 0.03, sis.W_PM1, 14)
    A, dt = 256, 0.5
    dyn = sis.BSB if form == "bf16" else sis.DSB
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORMS[form], dynamics=dyn))
    X0, Y0 = seeded_state(J.n, A, 15)
    s.set_state(X0, Y0)
    tile, ms = s.autotune(4)
    assert ms > 0 and s.p.tile == tile
    x, y = s.get_state()
    assert np.array_equal(x, X0) and np.array_equal(y, Y0)
    info = s.run(700, 0)
    X, Y = X0.astype(np.float32), Y0.astype(np.float32)
    r = orc.solve(pb, orc.MAXCUT, dyn, X, Y, 700, 0, P=100, dt=dt, xi0=s.p.xi0,
                  precision="bf16" if form == "bf16" else "f32")
    assert info["steps_executed"] == r["steps"]
    assert np.array_equal(s.get_state()[0], X.astype(np.float64))
    assert np.array_equal(s.cuts()[0], orc.cut(pb, r["cur"]))


def _sparse(J):
    rp, col, val = J.arrays()
    return sp.csr_matrix((val, col, rp), shape=(J.n, J.n))


@pytest.mark.parametrize("n", [1 << 20])
def test_full_size_packed_equals_plain(n):
    # 4-regular N = 2^20, 256 agents: the packed-sign and plain FP32 dSB forms
    # are bit-identical (size-independent property), and the device cut of
    # the returned spins equals an exact integer host evaluation.
    J = sis.Csr.gen_regular(n, 4, sis.W_PM1, 2024)
    res = {}
    for form in ("f32", "packed"):
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=256, dt=0.5, form=sis.FORMS[form]))
        s.init_random(77)
        info = s.run(300, 0)
        assert info["steps_executed"] == 300
        res[form] = (s.get_state(), s.spins(), s.cuts())
        s.close()
    (xa, ya), sa, (ca, ba) = res["f32"]
    (xb, yb), sb, (cb, bb) = res["packed"]
    assert np.array_equal(xa, xb) and np.array_equal(ya, yb) and np.array_equal(sa, sb)
    assert np.array_equal(ca, cb) and ba == bb
    M = _sparse(J)
    s64 = sa.astype(np.int64)
    H = 0.5 * np.einsum("ia,ia->a", s64, (M @ s64))
    assert np.array_equal(ca, -0.5 * (H - J.sum))


@pytest.mark.parametrize("form", ["packed", "f32"])
def test_maxcut_energy_matches_oracle_after_cuts(form):
    """Energy of the current and final spins (s^T J s without a field) right
    after the cuts, which leave other values in the same device buffer."""
    J, _ = maxcut_problem(300, 0.05, sis.W_INT3, 8)
    rp, col, val = J.arrays()
    pb0 = orc.Problem(rp, col, val, field=np.zeros(J.n), S=J.sum)
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=128, dt=0.5, form=sis.FORMS[form]))
    s.init_random(3)
    s.run(3000, 1)
    for which in (sis.SPINS_CURRENT, sis.SPINS_FINAL):
        cuts, _ = s.cuts()
        E = s.energy(which)
        spins = s.spins(which)
        assert np.array_equal(E, orc.energy_field(pb0, spins)), which
    # cut = -(1/2)(H - S) with H = s^T J s / 2, on the current spins
    cuts, _ = s.cuts()
    assert np.array_equal(cuts, (J.sum - s.energy(sis.SPINS_CURRENT) / 2) / 2)


@pytest.mark.parametrize("form", ["f32", "packed", "f64", "bf16"])
def test_graph_without_edges(form):
    """No stored entry at all (every node isolated, Z = 0): the force is zero,
    the step is the free oscillator, the cut is -(1/2)(0 - 0) = 0 and the
    energy 0; the row split and the tuner have nothing to split."""
    n = 70
    J = sis.Csr.from_coo(n, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0), mirror=True)
    assert J.nnz == 0
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    dyn = sis.BSB if form == "bf16" else sis.DSB
    A = 65
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, form=sis.FORMS[form], dynamics=dyn, dt=0.4,
                           xi0=0.5, split=-1))
    X0, Y0 = seeded_state(n, A, 9)
    s.set_state(X0, Y0)
    s.autotune(2)
    assert s.p.split == 0
    s.set_state(X0, Y0)
    info = s.run(300, 0)
    prec = {"f32": "f32", "packed": "f32", "f64": "f64", "bf16": "bf16"}[form]
    np_dt = np.float64 if prec == "f64" else np.float32
    X, Y = X0.astype(np_dt), Y0.astype(np_dt)
    r = orc.solve(pb, orc.MAXCUT, dyn, X, Y, 300, 0, P=100, dt=0.4, xi0=0.5, precision=prec)
    assert info["steps_executed"] == r["steps"]
    x, y = s.get_state()
    np.testing.assert_array_equal(x, X.astype(np.float64))
    np.testing.assert_array_equal(s.spins(), r["cur"])
    cuts, best = s.cuts()
    assert np.all(cuts == 0) and best == 0
    assert np.all(s.energy() == 0)


def test_single_node():
    J = sis.Csr.from_coo(1, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0), mirror=True)
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=3, form=sis.FORMS["packed"], xi0=1.0))
    s.init_random(1)
    info = s.run(250, 1)
    assert info["steps_executed"] == 300
    cuts, best = s.cuts()
    assert cuts.shape == (3,) and np.all(cuts == 0)
