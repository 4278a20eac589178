"""GPU parity at the benchmark configurations, against the CPU oracle.

- ER G(800, 0.06) with +1 weights, 128 agents, dt 0.1, 2000 steps with
  windows: bSB (and dSB) FP32 on the degree-ordered resident tiles (one and
  four (bSB) or eight (dSB) agents per workgroup), the autotuner's pick and a
  row tile, under
  direct launches, per-period graphs and the whole-run graph; with the
  configured coupling 0.5 / (sqrt(mean degree) std J) and the reference's
  0.7 / (std J sqrt N) (src/ising.cu:103-125).
- TSP n = 48 (2304 spins, 435,456 stored entries), 1024 agents, bSB over the
  QUBO's CSR in FP32 and with BF16 gathers: trajectories bit-exact over a few
  hundred steps, and a whole 10,000-step solve whose valid_list, final and
  current spins and X equal the oracle's on a sample of agents; the resident
  structured tile within the reordered-sum bars.
- Random 4-regular N = 2^20, 256 agents: BF16-gather bSB steps bit-exact on a
  sample of agents.

Agents are independent, so the oracle runs a subset of the device's agent
columns (the same initial columns) where the full set would take minutes.
"""
import numpy as np
import pytest

import oracle as orc
import sis

pytestmark = pytest.mark.gpu

RES = 1 << 20  # resident tile: one workgroup per 2^lg agents


def er800():
    J = sis.Csr.gen_er(800, 0.06, sis.W_ONE, 2024)
    rp, col, val = J.arrays()
    return J, orc.Problem(rp, col, val, S=J.sum)


def sorted_ell_applies(J, lg):
    """The solver's rule for degree-ordered resident rows: ELL groups of
    64 / G rows padded to chunks of four entries; the sorted layout when it is
    at least 10% smaller and N G <= 8192."""
    rp, _, _ = J.arrays()
    deg = np.diff(rp)
    G, R = 1 << lg, 64 >> lg
    ng = (J.n + R - 1) // R

    def words(d):
        return sum(int((d[g * R:(g + 1) * R].max() + 3) // 4 * 4 * R) for g in range(ng))

    nat, srt = words(deg), words(deg[np.argsort(-deg, kind="stable")])
    return J.n * G <= 8192 and 10 * srt <= 9 * nat


def c1_xi0(J):
    mean_deg = J.nnz / J.n
    sigma = 0.7 / (J.xi0() * np.sqrt(J.n))  # sample std of all N^2 entries
    return {"configured": 0.5 / (np.sqrt(mean_deg) * sigma), "reference": J.xi0()}


@pytest.mark.parametrize("dyn", [sis.BSB, sis.DSB])
@pytest.mark.parametrize("kind", ["configured", "reference"])
def test_er800_full_runs_every_tile_and_launch_mode(dyn, kind):
    J, pb = er800()
    # bSB gather slots are floats: eight agents per workgroup exceed the LDS,
    # so the larger resident tile is G = 4 there (G = 8 for dSB)
    lg = 2 if dyn == sis.BSB else 3
    assert sorted_ell_applies(J, 0) and sorted_ell_applies(J, lg)
    A, T, dt = 128, 2000, 0.1
    xi0 = c1_xi0(J)[kind]
    X, Y = orc.init_state(11, orc.MAXCUT, J.n, A)
    r = orc.solve(pb, orc.MAXCUT, dyn, X, Y, T, 0, P=100, K=25, dt=dt, xi0=xi0)
    want_cuts = orc.cut(pb, r["cur"])
    tuned = None
    for tile in ("auto", RES | 0 << 21, RES | lg << 21, 2):
        for graph in (0, 1, 2):
            s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, dynamics=dyn,
                                   form=sis.FORM_F32, xi0=xi0, use_graph=graph,
                                   tile=0 if tile == "auto" else tile))
            if tile == "auto":
                s.init_random(12)
                tuned = s.autotune(10)[0]
            else:
                assert s.p.tile == tile
            s.init_random(11)
            info = s.run(T, 0)
            case = (kind, dyn, tile, tuned, graph)
            assert info["steps_executed"] == r["steps"] == T, case
            x, y = s.get_state()
            assert np.array_equal(x, X.astype(np.float64)), case
            assert np.array_equal(y, Y.astype(np.float64)), case
            assert np.array_equal(s.spins(sis.SPINS_CURRENT), r["cur"]), case
            assert np.array_equal(s.spins(sis.SPINS_FINAL), r["fin"]), case
            st, lk, _ = s.agent_state()
            assert np.array_equal(st, r["stab"]) and np.array_equal(lk, r["bif"]), case
            cuts, best = s.cuts()
            assert np.array_equal(cuts, want_cuts), case
            assert best == int(np.argmax(want_cuts)), case
            s.close()


@pytest.mark.parametrize("form", ["packed", "f32"])
@pytest.mark.parametrize("lg", [0, 3])
def test_er800_sorted_resident_dsb_with_early_exit(form, lg):
    """Sorted resident tiles in the dSB forms with the update window's early
    exit on (dt 0.5 so agents lock within the run)."""
    J, pb = er800()
    A, T, dt = 128, 6000, 0.5
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, form=sis.FORMS[form],
                           tile=RES | lg << 21, use_graph=1))
    s.init_random(21)
    info = s.run(T, 1)
    X, Y = orc.init_state(21, orc.MAXCUT, J.n, A)
    r = orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, T, 1, P=100, K=25, dt=dt, xi0=s.p.xi0)
    assert info["steps_executed"] == r["steps"]
    assert np.array_equal(s.get_state()[0], X.astype(np.float64))
    assert np.array_equal(s.spins(sis.SPINS_FINAL), r["fin"])
    assert np.array_equal(s.cuts()[0], orc.cut(pb, r["cur"]))
    assert info["all_bifurcated"] == int(r["bif"].all())


def tsp48():
    t = sis.Tsp.generate(48, 1000, 1, 2024)
    J, h = t.qubo()
    rp, col, val = J.arrays()
    assert J.nnz == 48 * 48 * (4 * 48 - 3)
    return t, J, h, orc.Problem(rp, col, val, field=h, ncities=48)


@pytest.mark.parametrize("form,prec", [("f32", "f32"), ("bf16", "bf16")])
def test_tsp48_csr_trajectory_bit_exact_1024_agents(form, prec):
    t, J, h, pb = tsp48()
    A, T, dt = 1024, 10000, 0.1
    s = sis.Solver(J, dict(problem=sis.TSP, agents=A, form=sis.FORMS[form], dynamics=sis.BSB,
                           dt=dt, ncities=48), field=h)
    s.init_random(5)
    s.prestep()
    X, Y = orc.init_state(5, orc.TSP, 48 * 48, A)
    orc.prestep(X, Y, dt)
    for t0, n in ((0, 150), (150, 150)):
        s.advance(t0, n, T)
        orc.advance(pb, orc.TSP, orc.BSB, X, Y, t0, n, T, dt, 0.002, 1.005, prec)
        x, y = s.get_state()
        assert np.array_equal(x, X.astype(np.float64)), (form, t0 + n)
        assert np.array_equal(y, Y.astype(np.float64)), (form, t0 + n)


@pytest.mark.parametrize("form,prec", [("f32", "f32"), ("bf16", "bf16")])
def test_tsp48_whole_solve_on_agent_sample(form, prec):
    """10,000 steps with the TSP window every 1,000 steps (use_graph 1): X,
    valid_list, final and current spins of every 16th agent equal the
    oracle's solve of those columns; the tour lengths follow."""
    t, J, h, pb = tsp48()
    A, T, dt = 1024, 10000, 0.1
    s = sis.Solver(J, dict(problem=sis.TSP, agents=A, form=sis.FORMS[form], dynamics=sis.BSB,
                           dt=dt, ncities=48, use_graph=1), field=h)
    s.init_random(8)
    info = s.run(T, 0)
    assert info["steps_executed"] == T
    cols = np.arange(0, A, 16)
    X, Y = orc.init_state(8, orc.TSP, 48 * 48, A)
    X, Y = np.ascontiguousarray(X[:, cols]), np.ascontiguousarray(Y[:, cols])
    r = orc.solve(pb, orc.TSP, orc.BSB, X, Y, T, 0, P=T // 10, K=50, dt=dt, xi0=0.002,
                  coupling=1.005, precision=prec)
    x, _ = s.get_state()
    assert np.array_equal(x[:, cols], X.astype(np.float64))
    _, _, vs = s.agent_state()
    assert np.array_equal(vs[cols], r["valid"])
    assert np.array_equal(s.spins(sis.SPINS_FINAL)[:, cols], r["fin"])
    assert np.array_equal(s.spins(sis.SPINS_CURRENT)[:, cols], r["cur"])
    lens, _ = s.tsp_lengths(t)
    _, _, d, _ = t.arrays()
    want = np.array([orc.tsp_tour_length(48, d, r["fin"], k)[0] if r["valid"][k] > 0 else -1.0
                     for k in range(len(cols))])
    assert np.array_equal(lens[cols], want)
    E = s.energy(sis.SPINS_CURRENT)
    assert np.array_equal(E[cols], orc.energy_field(pb, r["cur"]))


def test_tsp16_valid_tours_bit_exact_on_agent_sample():
    """The valid-tour path, where the n = 48 instance ends with none: n = 16 on
    a 1,000-span instance with xi0 = 30 / (n max d), at which nearly every
    agent ends on a valid tour.  bSB FP32 over the CSR, 1024 agents, 10,000
    steps with the window every 1,000: valid_list, final spins and the tour
    lengths of every 16th agent equal the oracle's, most of them valid."""
    n = 16
    t = sis.Tsp.generate(n, 1000, 1, 2024 + n)
    J, h = t.qubo()
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, field=h, ncities=n)
    A, T, dt, xi0 = 1024, 10000, 0.1, 30.0 / (n * t.maxd)
    s = sis.Solver(J, dict(problem=sis.TSP, agents=A, form=sis.FORM_F32, dynamics=sis.BSB,
                           dt=dt, xi0=xi0, ncities=n, use_graph=1), field=h)
    s.init_random(3)
    assert s.run(T, 0)["steps_executed"] == T
    cols = np.arange(0, A, 16)
    X, Y = orc.init_state(3, orc.TSP, n * n, A)
    X, Y = np.ascontiguousarray(X[:, cols]), np.ascontiguousarray(Y[:, cols])
    r = orc.solve(pb, orc.TSP, orc.BSB, X, Y, T, 0, P=T // 10, K=50, dt=dt, xi0=xi0,
                  coupling=1.005, precision="f32")
    assert np.array_equal(s.get_state()[0][:, cols], X.astype(np.float64))
    _, _, vs = s.agent_state()
    assert np.array_equal(vs[cols], r["valid"])
    assert np.array_equal(s.spins(sis.SPINS_FINAL)[:, cols], r["fin"])
    lens, _ = s.tsp_lengths(t)
    _, _, d, _ = t.arrays()
    want = np.array([orc.tsp_tour_length(n, d, r["fin"], k)[0] if r["valid"][k] > 0 else -1.0
                     for k in range(len(cols))])
    assert np.array_equal(lens[cols], want)
    assert (want >= 0).sum() >= len(cols) // 2, (want >= 0).sum()


def test_tsp48_resident_structured_bsb_within_bars():
    """Resident structured tile (bSB FP32, 1024 agents) against the CSR oracle
    on a sample of agents: one step within 1e-6 relative; 1000 steps of the
    10,000-step ramp within 1e-5, or within twice the oracle's own FP32-to-FP64
    distance where the dynamics amplify rounding beyond that."""
    t, J, h, pb = tsp48()
    A, T, dt = 1024, 10000, 0.1
    s = sis.Solver.from_tsp(t, dict(agents=A, form=sis.FORM_F32, dynamics=sis.BSB, dt=dt,
                                    structured=1, tile=3))
    assert s.p.tile == 3
    rng = np.random.default_rng(44)
    X0 = rng.uniform(-0.1, 0.1, (48 * 48, A)).astype(np.float32)
    Y0 = rng.uniform(-0.1, 0.1, (48 * 48, A)).astype(np.float32)
    cols = np.arange(0, A, 32)
    s.set_state(X0.astype(np.float64), Y0.astype(np.float64))
    s.advance(0, 1, T)
    X = np.ascontiguousarray(X0[:, cols])
    Y = np.ascontiguousarray(Y0[:, cols])
    orc.advance(pb, orc.TSP, orc.BSB, X, Y, 0, 1, T, dt, 0.002, 1.005, "f32")
    x = s.get_state()[0][:, cols]
    assert np.abs(x - X).max() / np.abs(X).max() <= 1e-6
    s.advance(1, 999, T)
    orc.advance(pb, orc.TSP, orc.BSB, X, Y, 1, 999, T, dt, 0.002, 1.005, "f32")
    X64 = np.ascontiguousarray(X0[:, cols]).astype(np.float64)
    Y64 = np.ascontiguousarray(Y0[:, cols]).astype(np.float64)
    orc.advance(pb, orc.TSP, orc.BSB, X64, Y64, 0, 1000, T, dt, 0.002, 1.005, "f64")
    x = s.get_state()[0][:, cols]
    scale = np.abs(X64).max()
    err = np.abs(x - X).max() / scale
    own = np.abs(X.astype(np.float64) - X64).max() / scale
    assert err <= max(1e-5, 2 * own), (err, own)


def test_regular_2p20_bf16_bsb_bit_exact_on_agent_sample():
    """Random 4-regular N = 2^20, 256 agents, bSB with BF16 gathers at the
    benchmark's dt 0.5: 16 steps from the seeded state, every 4th agent equal
    to the oracle's BF16 mode bit for bit."""
    J = sis.Csr.gen_regular(1 << 20, 4, sis.W_PM1, 2024)
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    A, T, dt = 256, 20000, 0.5
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=dt, dynamics=sis.BSB,
                           form=sis.FORM_BF16))
    s.init_random(77)
    s.prestep()
    s.advance(0, 16, T)
    cols = np.arange(0, A, 4)
    X, Y = orc.init_state(77, orc.MAXCUT, J.n, A)
    X, Y = np.ascontiguousarray(X[:, cols]), np.ascontiguousarray(Y[:, cols])
    orc.prestep(X, Y, dt)
    orc.advance(pb, orc.MAXCUT, orc.BSB, X, Y, 0, 16, T, dt, s.p.xi0, 1.0, "bf16")
    x, y = s.get_state()
    assert np.array_equal(x[:, cols], X.astype(np.float64))
    assert np.array_equal(y[:, cols], Y.astype(np.float64))
