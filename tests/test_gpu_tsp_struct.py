"""GPU parity of the structured TSP product (sis_solver_create_tsp with
structured = 1): J x from row sums, column sums and the distance table instead
of the one-hot QUBO's CSR, against the oracle, which evaluates the CSR.

Bars:
- dSB: every term is a multiple of 1/8 of an integer, so both the CSR and the
  structured sums are exact and the trajectories are bit-identical.
- bSB: the two evaluate the same algebra in a different order.
  - One step from the same state is within 1e-6 relative (FP64: 1e-15).
  - 1000 steps of a 20000-step ramp are within 1e-5, or within twice the
    distance between the oracle in the same precision and the FP64 oracle
    where the dynamics amplify rounding beyond that.
- The energy s^T J s - h.s of the returned spins is exact.
"""
import numpy as np
import pytest

import oracle as orc
import sis

pytestmark = pytest.mark.gpu


def instance(n, seed=3, span=100):
    t = sis.Tsp.generate(n, span, 0, seed)
    J, h = t.qubo()
    rp, col, val = J.arrays()
    return t, J, h, orc.Problem(rp, col, val, field=h, ncities=n)


def seeded_state(n, A, seed):
    rng = np.random.default_rng(seed)
    X = rng.uniform(-0.1, 0.1, (n, A))  # X != 0 so the first step's gathers differ
    Y = rng.uniform(-0.1, 0.1, (n, A))
    return X.astype(np.float32).astype(np.float64), Y.astype(np.float32).astype(np.float64)


def max_rel(x, X):
    return float(np.abs(x - X).max() / max(np.abs(X).max(), 1e-30))


# tile 1: distance product on VALU; 2: on the matrix cores (FP32, default);
# 3: resident workgroups (FP32, n <= 96: state in registers, g in LDS)
FORMS_T = [("f32", "f32", 1), ("f32", "f32", 2), ("f32", "f32", 3), ("f64", "f64", 1)]


@pytest.mark.parametrize("n", [3, 7, 12, 33])
@pytest.mark.parametrize("form,prec,tile", FORMS_T)
def test_structured_dsb_bit_exact(n, form, prec, tile):
    t, J, h, pb = instance(n)
    A = 96
    X0, Y0 = seeded_state(n * n, A, 40 + n)
    T, dt = 800, 0.1
    s = sis.Solver.from_tsp(t, dict(agents=A, form=sis.FORMS[form], dynamics=sis.DSB, dt=dt,
                                    structured=1, tile=tile))
    assert s.p.structured == 1 and s.p.ncities == n and s.p.tile == tile
    s.set_state(X0, Y0)
    np_dt = np.float64 if prec == "f64" else np.float32
    X, Y = X0.astype(np_dt), Y0.astype(np_dt)
    for t0, k in ((0, 300), (300, 500)):
        s.advance(t0, k, T)
        orc.advance(pb, orc.TSP, orc.DSB, X, Y, t0, k, T, dt, 0.002, 1.005, prec)
        x, y = s.get_state()
        np.testing.assert_array_equal(x, X.astype(np.float64), err_msg=f"X after {t0 + k}")
        np.testing.assert_array_equal(y, Y.astype(np.float64), err_msg=f"Y after {t0 + k}")


@pytest.mark.parametrize("form,prec,bar1,bar,tile", [("f32", "f32", 1e-6, 1e-5, 1),
                                                     ("f32", "f32", 1e-6, 1e-5, 2),
                                                     ("f32", "f32", 1e-6, 1e-5, 3),
                                                     ("f64", "f64", 1e-15, 1e-12, 1)])
@pytest.mark.parametrize("n", [7, 16, 40])
def test_structured_bsb_within_bound(form, prec, bar1, bar, tile, n):
    t, J, h, pb = instance(n, seed=5)
    A = 128
    X0, Y0 = seeded_state(n * n, A, 50 + n)
    dt, T = 0.1, 20000
    s = sis.Solver.from_tsp(t, dict(agents=A, form=sis.FORMS[form], dt=dt, structured=1, tile=tile))
    s.set_state(X0, Y0)
    np_dt = np.float64 if prec == "f64" else np.float32
    X, Y = X0.astype(np_dt), Y0.astype(np_dt)
    X64, Y64 = X0.copy(), Y0.copy()
    errs, t0 = [], 0
    for t1 in (1, 10, 100, 1000):
        s.advance(t0, t1 - t0, T)
        orc.advance(pb, orc.TSP, orc.BSB, X, Y, t0, t1 - t0, T, dt, 0.002, 1.005, prec)
        orc.advance(pb, orc.TSP, orc.BSB, X64, Y64, t0, t1 - t0, T, dt, 0.002, 1.005, "f64")
        x, y = s.get_state()
        Xd, Yd = X.astype(np.float64), Y.astype(np.float64)
        errs.append((t1, max(max_rel(x, Xd), max_rel(y, Yd)), max(max_rel(Xd, X64), max_rel(Yd, Y64))))
        t0 = t1
    print(f"structured bSB n={n} {form} tile {tile}: (step, vs oracle {prec}, oracle {prec} vs f64) {errs}")
    assert errs[0][1] <= bar1, errs
    for t1, e, pe in errs[1:]:
        assert e <= max(bar, 2.0 * pe), (t1, e, pe)


@pytest.mark.parametrize("dyn", [sis.DSB, sis.BSB])
@pytest.mark.parametrize("tile", [1, 2, 3])
def test_structured_full_run_and_energy(dyn, tile):
    n = 6
    t, J, h, pb = instance(n, seed=11)
    A, T, dt = 256, 3000, 0.1
    s = sis.Solver.from_tsp(t, dict(agents=A, form=sis.FORM_F32, dynamics=dyn, dt=dt, structured=1,
                                    tile=tile))
    s.init_random(8)
    info = s.run(T, 0)
    fin = s.spins(sis.SPINS_FINAL)
    cur = s.spins(sis.SPINS_CURRENT)
    # exact energies of whatever spins came back
    for which, sp in ((sis.SPINS_FINAL, fin), (sis.SPINS_CURRENT, cur)):
        assert np.array_equal(s.energy(which), orc.energy_field(pb, sp))
    lens, best = s.tsp_lengths(t)
    _, _, d, _ = t.arrays()
    st, lk, valid = s.agent_state()
    for a in range(A):
        if valid[a] > 0:
            assert lens[a] == orc.tsp_tour_length(n, d, fin, a)[0]
        else:
            assert lens[a] == -1
    if dyn == sis.DSB:  # bit-identical to the oracle's CSR evaluation
        X, Y = orc.init_state(8, orc.TSP, n * n, A)
        r = orc.solve(pb, orc.TSP, orc.DSB, X, Y, T, 0, P=T // 10, K=50, dt=dt, xi0=0.002,
                      coupling=1.005)
        assert info["steps_executed"] == r["steps"]
        assert np.array_equal(fin, r["fin"]) and np.array_equal(valid, r["valid"])


def test_create_tsp_matrix_mode_equals_create():
    n = 8
    t, J, h, pb = instance(n, seed=2)
    A = 64
    a = sis.Solver.from_tsp(t, dict(agents=A, form=sis.FORM_F32, dt=0.1))
    b = sis.Solver(J, dict(problem=sis.TSP, agents=A, form=sis.FORM_F32, dt=0.1, ncities=n), field=h)
    assert a.p.structured == 0
    for s in (a, b):
        s.init_random(4)
        s.run(1000, 0)
    assert np.array_equal(a.get_state()[0], b.get_state()[0])
    assert np.array_equal(a.energy(sis.SPINS_FINAL), b.energy(sis.SPINS_FINAL))


@pytest.mark.parametrize("n,form,tile", [(300, "f32", 2), (300, "f32", 1), (200, "f64", 1)])
def test_structured_large_n_dsb_equals_csr(n, form, tile):
    """Above 64 KB of staged v (n > 256 in FP32, n > 128 in FP64): dSB through
    the structured product is bit-identical to dSB through the QUBO's CSR
    (both sums are exact multiples of 1/8)."""
    t = sis.Tsp.generate(n, 1000, 1, 7)
    A, T = 64, 1000
    runs = []
    for structured in (1, 0):
        params = dict(agents=A, form=sis.FORMS[form], dynamics=sis.DSB, dt=0.1,
                      structured=structured)
        if structured:
            params["tile"] = tile
        s = sis.Solver.from_tsp(t, params)
        s.init_random(5)
        s.advance(0, 60, T)
        runs.append(s.get_state())
        s.close()
    (xs, ys), (xc, yc) = runs
    np.testing.assert_array_equal(xs, xc)
    np.testing.assert_array_equal(ys, yc)


def test_structured_n512_tiles_agree():
    """The largest structured FP32 instance (n = 512, 128 KB of staged v):
    matrix cores and VALU give the same dSB trajectory."""
    t = sis.Tsp.generate(512, 1000, 1, 8)
    out = []
    for tile in (1, 2):
        s = sis.Solver.from_tsp(t, dict(agents=64, form=sis.FORM_F32, dynamics=sis.DSB, dt=0.1,
                                        structured=1, tile=tile))
        s.init_random(6)
        s.advance(0, 30, 1000)
        out.append(s.get_state()[0])
        s.close()
    np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize("n,agents", [(48, 64), (48, 100), (48, 1024), (64, 256), (80, 128), (96, 64)])
def test_structured_resident_equals_mfma_dsb(n, agents):
    """Tile 3 (resident workgroups of 1, 2 or 4 agents) against tile 2, at the
    10,000-step TSP size (n = 48) and up to the largest resident n: dSB sums
    are exact, so the trajectories are bit-identical, including agent counts
    that leave padding lanes."""
    t = sis.Tsp.generate(n, 1000, 1, 9)
    out = []
    for tile in (2, 3):
        s = sis.Solver.from_tsp(t, dict(agents=agents, form=sis.FORM_F32, dynamics=sis.DSB, dt=0.1,
                                        structured=1, tile=tile))
        assert s.p.tile == tile
        s.init_random(4)
        s.advance(0, 777, 10000)
        out.append(s.get_state())
        s.close()
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_structured_argument_errors():
    t2 = sis.Tsp.generate(2, 50, 0, 1)
    with pytest.raises(sis.SisError):
        sis.Solver.from_tsp(t2, dict(agents=64, structured=1))
    t = sis.Tsp.generate(8, 50, 0, 1)
    with pytest.raises(sis.SisError):
        sis.Solver.from_tsp(t, dict(agents=64, form=sis.FORM_BF16, structured=1))
    with pytest.raises(sis.SisError):
        sis.Solver.from_tsp(t, dict(agents=64, split=16, structured=1))
    big = sis.Tsp.generate(257, 500, 0, 1)
    with pytest.raises(sis.SisError):
        sis.Solver.from_tsp(big, dict(agents=64, form=sis.FORM_F64, structured=1))
    with pytest.raises(sis.SisError):
        sis.Solver.from_tsp(sis.Tsp.generate(513, 500, 0, 1),
                            dict(agents=64, form=sis.FORM_F32, structured=1))
    J, h = t.qubo()
    with pytest.raises(sis.SisError):
        sis.Solver(J, dict(problem=sis.TSP, agents=64, ncities=8, structured=1), field=h)


def test_structured_default_tile_and_autotune():
    t = sis.Tsp.generate(40, 100, 0, 3)
    s = sis.Solver.from_tsp(t, dict(agents=128, form=sis.FORM_F32, structured=1))
    assert s.p.tile == 2  # matrix cores by default in FP32
    s.init_random(1)
    x0 = s.get_state()[0]
    tile, ms = s.autotune(5)
    assert tile in (1, 2, 3) and ms > 0
    assert np.array_equal(s.get_state()[0], x0)
    s64 = sis.Solver.from_tsp(t, dict(agents=128, form=sis.FORM_F64, structured=1))
    assert s64.p.tile == 1
    with pytest.raises(sis.SisError):
        sis.Solver.from_tsp(t, dict(agents=128, form=sis.FORM_F64, structured=1, tile=2))
    with pytest.raises(sis.SisError):  # resident tile: FP32 only
        sis.Solver.from_tsp(t, dict(agents=128, form=sis.FORM_F64, structured=1, tile=3))
    with pytest.raises(sis.SisError):  # resident tile: n <= 96
        sis.Solver.from_tsp(sis.Tsp.generate(97, 100, 0, 3),
                            dict(agents=128, form=sis.FORM_F32, structured=1, tile=3))
