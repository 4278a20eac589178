"""GPU parity of the row split (k_step_split): rows with more than `split`
stored entries are summed by several wavefronts, chunk sums added in chunk
order in LDS, group sums (rows longer than one workgroup's chunks) added in
group order by the last workgroup to arrive.

Bars:
- dSB with integer weights: every partial sum is an exact integer, so the
  split trajectory is bit-identical to the oracle's sequential sum.
- bSB (value gathers) and real weights: the summation order differs from the
  oracle's CSR order.  Errors are relative, max-norm over all node-agents
  against the max-norm of the oracle state.  Bars:
  - one step from the same state: within a few FP32 roundings (1e-6;
    FP64 1e-15);
  - 1000 steps of a 20000-step ramp: within the FP32 trajectory bound of
    1e-5 (FP64 1e-12), or, where the dynamics amplify rounding beyond that,
    within twice the distance between the oracle in the same precision and
    the FP64 oracle.  Any reordering, e.g. the same code on another CPU,
    moves an FP32 trajectory by that much, and FP64 runs show the same
    amplification at 1e-16 scale.
"""
import numpy as np
import pytest

import oracle as orc
import sis

pytestmark = pytest.mark.gpu


def hub_graph(n=300, seed=4, weights="int"):
    """ER base plus three hubs of degree n-1, n/2 and about n/4: one-group,
    several-group and unsplit rows in one matrix."""
    rng = np.random.default_rng(seed)
    pairs = set()
    for i in range(n):
        for j in range(i + 1, n):
            if rng.random() < 0.03:
                pairs.add((i, j))
    for hub, deg in ((0, n - 1), (1, n // 2), (2, n // 4 + 3)):
        for j in rng.choice(np.arange(3, n), size=min(deg, n - 3), replace=False):
            pairs.add((min(hub, int(j)), max(hub, int(j))))
    pairs.add((0, 1))
    pairs.add((0, 2))
    rows = np.array([p[0] for p in pairs], np.int64)
    cols = np.array([p[1] for p in pairs], np.int64)
    if weights == "int":
        vals = rng.choice([-3, -2, -1, 1, 2, 3], size=len(rows)).astype(np.float64)
    else:
        vals = rng.uniform(-1.5, 1.5, size=len(rows))
    J = sis.Csr.from_coo(n, rows, cols, vals, mirror=True)
    rp, col, val = J.arrays()
    return J, orc.Problem(rp, col, val, S=J.sum)


def seeded_state(n, A, seed, tsp=False):
    rng = np.random.default_rng(seed)
    X = np.zeros((n, A)) if tsp else rng.uniform(-0.1, 0.1, (n, A))
    Y = rng.uniform(-0.1, 0.1, (n, A))
    return X.astype(np.float32).astype(np.float64), Y.astype(np.float32).astype(np.float64)


def max_rel(x, X):
    return float(np.abs(x - X).max() / max(np.abs(X).max(), 1e-30))


EXACT = [  # dSB, integer weights: bit-exact
    ("f32", 1, "f32", 0), ("f64", 1, "f64", 0), ("packed", 1, "f32", 0),
    ("f32", 0, "f32", 0), ("packed", 1, "f32", 1 | 4 << 8), ("packed", 1, "f32", 2 | 2 << 8 | 1 << 16),
]


@pytest.mark.parametrize("form,fused,prec,tile", EXACT)
@pytest.mark.parametrize("tau", [4, 16, 64])
@pytest.mark.parametrize("agents", [96, 200])
def test_split_dsb_integer_weights_bit_exact(form, fused, prec, tile, tau, agents):
    J, pb = hub_graph()
    X0, Y0 = seeded_state(J.n, agents, 31)
    T, dt = 400, 0.5
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=agents, form=sis.FORMS[form], dynamics=sis.DSB,
                           dt=dt, fused=fused, split=tau, tile=tile))
    assert s.p.split == tau
    s.set_state(X0, Y0)
    np_dt = np.float64 if prec == "f64" else np.float32
    X, Y = X0.astype(np_dt), Y0.astype(np_dt)
    problem = orc.MAXCUT if fused else orc.MAXCUT_UNFUSED
    for t0, n in ((0, 150), (150, 250)):
        s.advance(t0, n, T)
        orc.advance(pb, problem, orc.DSB, X, Y, t0, n, T, dt, s.p.xi0, 1.0, prec)
        x, y = s.get_state()
        np.testing.assert_array_equal(x, X.astype(np.float64), err_msg=f"X after {t0 + n}")
        np.testing.assert_array_equal(y, Y.astype(np.float64), err_msg=f"Y after {t0 + n}")


BSB_BARS = [("f32", "f32", 1e-6, 1e-5), ("f64", "f64", 1e-15, 1e-12), ("bf16", "bf16", 1e-6, 1e-5)]


def _bsb_errors(s, pb, problem, X0, Y0, prec, dt, xi0, coupling, T, checkpoints, prestep=False):
    """(step, split vs oracle in the same precision, that oracle vs the FP64
    oracle) at each checkpoint, max over X and Y."""
    s.set_state(X0, Y0)
    np_dt = np.float64 if prec == "f64" else np.float32
    X, Y = X0.astype(np_dt), Y0.astype(np_dt)
    X64, Y64 = X0.copy(), Y0.copy()
    if prestep:
        s.prestep()
        orc.prestep(X, Y, dt)
        orc.prestep(X64, Y64, dt)
    out, t0 = [], 0
    for t1 in checkpoints:
        s.advance(t0, t1 - t0, T)
        orc.advance(pb, problem, orc.BSB, X, Y, t0, t1 - t0, T, dt, xi0, coupling, prec)
        orc.advance(pb, problem, orc.BSB, X64, Y64, t0, t1 - t0, T, dt, xi0, coupling, "f64")
        x, y = s.get_state()
        Xd, Yd = X.astype(np.float64), Y.astype(np.float64)
        out.append((t1, max(max_rel(x, Xd), max_rel(y, Yd)), max(max_rel(Xd, X64), max_rel(Yd, Y64))))
        t0 = t1
    return out


def _check_bsb(e1, e, bar1, bar):
    # one step: the reordered sum is within a few roundings of the sequential one
    assert e1[0][1] <= bar1, e1
    # later steps: the split moves the trajectory no further than the form's
    # own rounding moves it from the FP64 trajectory (both are amplified by
    # the same dynamics), and never beyond `bar` while that is below `bar`
    for t, split_err, prec_err in e:
        assert split_err <= max(bar, 2.0 * prec_err), (t, split_err, prec_err)


@pytest.mark.parametrize("form,prec,bar1,bar", BSB_BARS)
@pytest.mark.parametrize("tau", [4, 32])
def test_split_bsb_maxcut_within_bound(form, prec, bar1, bar, tau):
    """Real weights, value gathers, hubs of degree up to 299, 1000 steps of a
    20000-step ramp."""
    J, pb = hub_graph(weights="real")
    A = 128
    X0, Y0 = seeded_state(J.n, A, 32)
    dt = 0.1
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, form=sis.FORMS[form], dynamics=sis.BSB,
                           dt=dt, split=tau))
    e1 = _bsb_errors(s, pb, orc.MAXCUT, X0, Y0, prec, dt, s.p.xi0, 1.0, 20000, [1])
    e = _bsb_errors(s, pb, orc.MAXCUT, X0, Y0, prec, dt, s.p.xi0, 1.0, 20000, [10, 100, 1000])
    print(f"bSB maxcut {form} tau={tau}: (step, split vs {prec}, {prec} vs f64) {e1 + e}")
    _check_bsb(e1, e, bar1, bar)


@pytest.mark.parametrize("form,prec,bar1,bar", BSB_BARS)
@pytest.mark.parametrize("tau", [4, 8])
def test_split_tsp_within_bound(form, prec, bar1, bar, tau):
    n = 7  # 4n - 3 = 25 entries per row: several groups at tau = 4
    t = sis.Tsp.generate(n, 100, 0, 3)
    J, h = t.qubo()
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, field=h, ncities=n)
    A = 128
    X0, Y0 = seeded_state(J.n, A, 33, tsp=True)
    dt = 0.1
    s = sis.Solver(J, dict(problem=sis.TSP, agents=A, form=sis.FORMS[form], dynamics=sis.BSB, dt=dt,
                           ncities=n, split=tau), field=h)
    e1 = _bsb_errors(s, pb, orc.TSP, X0, Y0, prec, dt, 0.002, 1.005, 20000, [2], prestep=True)
    e = _bsb_errors(s, pb, orc.TSP, X0, Y0, prec, dt, 0.002, 1.005, 20000, [10, 100, 1000],
                    prestep=True)
    print(f"bSB tsp {form} tau={tau}: (step, split vs {prec}, {prec} vs f64) {e1 + e}")
    _check_bsb(e1, e, bar1, bar)


@pytest.mark.parametrize("graph", [0, 1, 2])
def test_split_full_run_matches_oracle(graph):
    J, pb = hub_graph(seed=5)
    A = 100
    dt, T = 0.5, 3000
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, form=sis.FORMS["packed"], dynamics=sis.DSB,
                           dt=dt, use_graph=graph, split=8))
    X0, Y0 = seeded_state(J.n, A, 34)
    for _ in range(2 if graph else 1):
        s.set_state(X0, Y0)
        info = s.run(T, 1)
        X, Y = X0.astype(np.float32), Y0.astype(np.float32)
        r = orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, T, 1, P=100, K=25, dt=dt, xi0=s.p.xi0,
                      precision="f32")
        assert info["steps_executed"] == r["steps"]
        np.testing.assert_array_equal(s.spins(sis.SPINS_CURRENT), r["cur"])
        np.testing.assert_array_equal(s.spins(sis.SPINS_FINAL), r["fin"])
        cuts, best = s.cuts()
        np.testing.assert_array_equal(cuts, orc.cut(pb, r["cur"]))


def test_autotune_chooses_split_and_keeps_results():
    J, pb = hub_graph(n=400, seed=6)
    A = 128
    dt, T = 0.5, 300
    X0, Y0 = seeded_state(J.n, A, 35)
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, form=sis.FORMS["packed"], dynamics=sis.DSB,
                           dt=dt, split=-1))
    s.set_state(X0, Y0)
    tile, ms = s.autotune(5)
    assert ms > 0
    assert s.p.split == 0 or (16 <= s.p.split < 399 and s.p.split & (s.p.split - 1) == 0)
    x, y = s.get_state()  # the tuner restores the state
    np.testing.assert_array_equal(x, X0.astype(np.float32).astype(np.float64))
    s.advance(0, T, T)
    X, Y = X0.astype(np.float32), Y0.astype(np.float32)
    orc.advance(pb, orc.MAXCUT, orc.DSB, X, Y, 0, T, T, dt, s.p.xi0, 1.0, "f32")
    x, y = s.get_state()
    np.testing.assert_array_equal(x, X.astype(np.float64))
    np.testing.assert_array_equal(y, Y.astype(np.float64))


def test_split_threshold_at_or_above_max_degree_is_off():
    J, _ = hub_graph()
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=64, form=sis.FORMS["packed"], split=100000))
    assert s.p.split == 0
