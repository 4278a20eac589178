"""Pins of the CPU oracle (oracle/sis_oracle.c) by closed forms, invariants,
library routines and brute force -- none of them a retyped copy of the step.

Each check targets a plausible mistake: a transposed or mis-indexed gather,
a flipped coupling sign, a wrong update order or association, a wrong window
count, a wrong cut formula, a wrong QUBO coefficient.
"""
import itertools
import math

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
import readers


def random_problem(n, density, seed, symmetric=True, weights="int"):
    rng = np.random.default_rng(seed)
    M = (rng.random((n, n)) < density).astype(float)
    if weights == "int":
        M *= rng.choice([-3, -2, -1, 1, 2, 3], size=(n, n))
    elif weights == "real":
        M *= rng.normal(size=(n, n))
    np.fill_diagonal(M, 0)
    if symmetric:
        M = np.triu(M, 1)
        M = M + M.T
    csr = sp.csr_matrix(M)
    csr.sort_indices()
    return orc.Problem(csr.indptr, csr.indices, csr.data, S=float(np.triu(M, 1).sum()) if symmetric else float(M.sum())), M


@pytest.mark.parametrize("gather", [orc.DSB, orc.BSB])
@pytest.mark.parametrize("precision", ["f64", "f32"])
def test_kick_is_minus_J_times_g(gather, precision):
    # Non-symmetric J so a transposed gather is caught; c = (0, 0, 1, 0) and
    # Y0 = 0 make the new Y exactly the force F.
    pb, M = random_problem(37, 0.2, 1, symmetric=False, weights="int")
    rng = np.random.default_rng(2)
    dt = np.float64 if precision == "f64" else np.float32
    X = (rng.uniform(-0.9, 0.9, (37, 5))).astype(dt)
    Y = np.zeros_like(X)
    coupling = 1.005 if gather == orc.BSB else 1.0
    Xn = orc.step(pb, orc.MAXCUT, gather, X, Y, [0, 0, 1, 0], coupling, precision)
    g = np.where(np.signbit(X), -1.0, 1.0) if gather == orc.DSB else X.astype(np.float64) * coupling
    F = -(sp.csr_matrix(M) @ g)
    tol = 0 if gather == orc.DSB else (1e-12 if precision == "f64" else 2e-5)
    np.testing.assert_allclose(Y, F, rtol=tol, atol=tol)
    assert np.array_equal(Xn, X)


def test_zero_coupling_conserves_symplectic_area():
    # J = 0: x' = x + c0 y, y' = y + c1 x' has determinant 1, so the area
    # x_a y_b - x_b y_a of two agents is invariant; a step with the old x in
    # the y update (or the updates swapped) breaks it.
    n, A, T, dt = 6, 4, 400, 0.05
    pb = orc.Problem(np.zeros(n + 1, np.int64), np.zeros(0, np.int32), np.zeros(0))
    rng = np.random.default_rng(3)
    X = rng.uniform(-0.05, 0.05, (n, A))
    Y = rng.uniform(-0.05, 0.05, (n, A))
    area0 = X[:, 0] * Y[:, 1] - X[:, 1] * Y[:, 0]
    orc.advance(pb, orc.MAXCUT, orc.DSB, X, Y, 0, 300, T, dt, 0.3, 1.0, "f64")
    assert np.abs(X).max() <= 1.0  # no confinement happened
    area = X[:, 0] * Y[:, 1] - X[:, 1] * Y[:, 0]
    np.testing.assert_allclose(area, area0, rtol=1e-10, atol=1e-16)


def test_saturated_ramp_is_linear_then_confines():
    # After p reaches 1 (t >= T/1.025): c1 = 0, c0 = 2 dt, so y is constant and
    # x grows by 2 dt y per step until |x| > 1, then x = sgn x and y = 0
    # forever (src/kernels.cu:236-242, 261-262).
    T, dt = 100, 0.1
    pb = orc.Problem(np.zeros(2, np.int64), np.zeros(0, np.int32), np.zeros(0))
    X = np.array([[0.2]])
    Y = np.array([[0.3]])
    t0 = 100
    x, y = X[0, 0], Y[0, 0]
    k_conf = next(k for k in range(1, 100) if abs(x + k * 2 * dt * y) > 1)
    for k in range(1, k_conf + 1):
        orc.advance(pb, orc.MAXCUT, orc.DSB, X, Y, t0 + k - 1, 1, T, dt, 0.5, 1.0, "f64")
        assert Y[0, 0] == y
        assert X[0, 0] == pytest.approx(x + k * 2 * dt * y, abs=1e-14)
    orc.advance(pb, orc.MAXCUT, orc.DSB, X, Y, t0 + k_conf, 5, T, dt, 0.5, 1.0, "f64")
    assert X[0, 0] == 1.0 and Y[0, 0] == 0.0


def test_ramp_closed_forms():
    T, dt, xi0 = 1000, 0.2, 0.4
    for t in (0, 1, 500, 975, 976, 2000):
        p = min(1.025 * t / T, 1.0)
        c = orc.sched(orc.MAXCUT, t, T, dt, xi0)
        assert c[0] == pytest.approx(dt * (1 + p)) and c[1] == pytest.approx(dt * (p - 1))
        assert c[2] == pytest.approx(dt * xi0)
        a = min(2.08 * t / T, 2.0)
        b = min(0.5 + 0.5 * (t / T) ** 1.9, 1.0)
        c = orc.sched(orc.TSP, t, T, dt, xi0)
        assert c[1] == pytest.approx(dt * (a - 1)) and c[3] == pytest.approx(b)
    # the unfused path's ramp is not clamped (src/kernels.cu:15)
    assert orc.sched(2, 2000, 1000, 0.1, 1.0)[0] == pytest.approx(0.1 * (1 + 2.05))


def test_dt_zero_is_identity_and_window_locks_at_24_or_25():
    pb, _ = random_problem(20, 0.3, 4)
    rng = np.random.default_rng(5)
    for all_negative, expect in ((True, 2400), (False, 2500)):
        X = rng.uniform(-0.5, 0.5, (20, 8)).astype(np.float32)
        if all_negative:
            X = -np.abs(X) - 0.01
        X0 = X.copy()
        Y = rng.uniform(-0.5, 0.5, (20, 8)).astype(np.float32)
        Y0 = Y.copy()
        r = orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, T=10000, uw=1, P=100, K=25, dt=0.0, xi0=0.3)
        assert np.array_equal(X, X0) and np.array_equal(Y, Y0)
        assert r["steps"] == expect
        assert r["bif"].all() and (r["stab"] == 24).all()
        assert np.array_equal(r["fin"], np.where(np.signbit(X0), -1, 1))


def test_xi0_against_numpy_sample_std():
    for seed, (n, dens) in enumerate(((30, 0.2), (200, 0.03), (64, 0.5))):
        pb, M = random_problem(n, dens, 10 + seed, weights="real")
        want = 0.7 / (np.std(M, ddof=1) * math.sqrt(n))
        assert orc.xi0_maxcut(n, pb.val64) == pytest.approx(want, rel=1e-12)


def _edges(M):
    iu = np.transpose(np.nonzero(np.triu(M, 1)))
    return [(i, j, M[i, j]) for i, j in iu]


def test_cut_formula_brute_force_symmetric():
    pb, M = random_problem(12, 0.4, 6)
    spins = np.array(list(itertools.product([-1, 1], repeat=12)), np.int32).T.copy()
    cuts = orc.cut(pb, spins)
    direct = np.zeros(spins.shape[1])
    for i, j, w in _edges(M):
        direct += w * (spins[i] != spins[j])
    assert np.array_equal(cuts, direct)


def test_cut_general_file_counts_both_triangles(tmp_path):
    # src/csr.cpp keeps both triangles of a "general" file and sums both, so
    # the reported value is W_cut + W/2 (documented reference behaviour).
    p = tmp_path / "g.mtx"
    p.write_text("%%MatrixMarket matrix coordinate integer general\n4 4 6\n"
                 "1 2 2\n2 1 2\n2 3 -1\n3 2 -1\n3 4 4\n4 3 4\n")
    m = readers.read_mtx(str(p))
    pb = orc.Problem(m["row_ptr"], m["col"], m["val"], S=m["sum"])
    spins = np.array(list(itertools.product([-1, 1], repeat=4)), np.int32).T.copy()
    W = 2 - 1 + 4
    direct = sum(w * (spins[i] != spins[j]) for i, j, w in ((0, 1, 2), (1, 2, -1), (2, 3, 4)))
    assert np.array_equal(orc.cut(pb, spins), direct + W / 2)


def test_cut_stats_against_numpy():
    rng = np.random.default_rng(8)
    c = rng.integers(0, 20, 97).astype(float)
    st = orc.cut_stats(c)
    assert st[0] == c.min() and st[1] == c.max()
    assert st[2] == pytest.approx(c.mean()) and st[3] == pytest.approx(np.std(c, ddof=1))
    assert st[4] == np.mean(c == c.max())


def test_dsb_finds_bipartite_max_cut():
    # A +1-weighted even cycle is bipartite: the maximum cut takes every edge.
    # With the coupling sign flipped the dynamics would align spins (cut 0).
    n = 10
    rows = np.arange(n)
    M = np.zeros((n, n))
    M[rows, (rows + 1) % n] = 1
    M = M + M.T
    csr = sp.csr_matrix(M)
    pb = orc.Problem(csr.indptr, csr.indices, csr.data, S=float(n))
    X, Y = orc.init_state(99, orc.MAXCUT, n, 32)
    xi0 = orc.xi0_maxcut(n, pb.val64)
    r = orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, T=2000, uw=0, P=100, dt=0.1, xi0=xi0)
    cuts = orc.cut(pb, r["cur"])
    assert cuts.max() == n


def test_tsp_distances_closed_forms():
    rng = np.random.default_rng(9)
    x = rng.integers(0, 1000, 12).astype(float)
    y = rng.integers(0, 1000, 12).astype(float)
    for rnd in (0, 1):
        d = orc.tsp_distances(x, y, 0, rnd)
        for a in range(12):
            for b in range(a + 1, 12):
                r = math.hypot(x[a] - x[b], y[a] - y[b])
                assert d[orc.tri(a, b, 12)] == (math.floor(r) if rnd == 0 else math.floor(r + 0.5))
    # GEO is the great-circle distance on a 6378.388 km sphere, +1 and
    # truncated (TSPLIB): within 1 km of the haversine distance.
    lat = np.radians(rng.uniform(-60, 60, 8))
    lon = np.radians(rng.uniform(-170, 170, 8))
    d = orc.tsp_distances(lat, lon, 1, 0)
    for a in range(8):
        for b in range(a + 1, 8):
            h = 2 * math.asin(math.sqrt(math.sin((lat[a] - lat[b]) / 2) ** 2 + math.cos(lat[a]) * math.cos(lat[b]) * math.sin((lon[a] - lon[b]) / 2) ** 2))
            assert 0 <= d[orc.tri(a, b, 8)] - 6378.388 * h <= 1.0 + 1e-6


def _tsp_instance(n, seed):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 100, n).astype(float)
    y = rng.integers(0, 100, n).astype(float)
    d = orc.tsp_distances(x, y, 0, 0)
    D, mx = orc.tsp_totals(n, d)
    return d, D, mx


def _tour_spins(perm):
    n = len(perm)
    s = -np.ones(n * n, np.int32)
    for pos, city in enumerate(perm):
        s[pos * n + city] = 1
    return s


@pytest.mark.parametrize("n", [4, 5, 6])
def test_tsp_qubo_energy_identity_and_brute_force(n):
    # For every valid tour, E(s) = s^T J s - h.s equals A * length + const
    # (A = 1); hence the lowest-energy valid state is the optimal tour.
    d, D, mx = _tsp_instance(n, 20 + n)
    J = orc.tsp_qubo_dense(n, d, mx)
    h = orc.tsp_field(n, D, mx)
    csr = sp.csr_matrix(J)
    csr.sort_indices()
    pb = orc.Problem(csr.indptr, csr.indices, csr.data, field=h)
    perms = list(itertools.permutations(range(n)))
    spins = np.stack([_tour_spins(p) for p in perms], axis=1)
    E = orc.energy_field(pb, spins)
    L = np.array([orc.tsp_tour_length(n, d, spins, a)[0] for a in range(len(perms))])
    off = E - L
    assert np.all(off == off[0])
    # nnz of the one-hot QUBO: n^2 (4n - 3) when no distance is zero
    assert csr.nnz == n * n * (4 * n - 3)
    best = min(perms, key=lambda p: sum(d[orc.tri(min(p[i], p[(i + 1) % n]), max(p[i], p[(i + 1) % n]), n)] for i in range(n)))
    assert L[int(np.argmin(E))] == L[perms.index(best)]


def test_tsp_window_validity():
    n = 4
    X = np.full((n * n, 3), -0.5, np.float32)
    for pos, city in enumerate([2, 0, 3, 1]):
        X[pos * n + city, 0] = 0.5          # agent 0: valid tour
    X[0, 1] = X[5, 1] = 0.5                 # agent 1: two cities only
    X[:, 2] = 0.5                           # agent 2: everything on
    cur = np.full((n * n, 3), -1, np.int32)
    fin = cur.copy()
    valid = np.zeros(3, np.int32)
    orc.lib().orc_window_tsp_f32(orc._p(X), orc.C.c_int64(n), orc.C.c_int64(3), orc.C.c_int64(700),
                                 orc._p(cur), orc._p(fin), orc._p(valid))
    assert valid.tolist() == [700, 0, 0]
    v, tour = orc.tsp_tour_length(n, np.arange(1, 7, dtype=float), fin, 0)
    assert tour.tolist() == [2, 0, 3, 1]


def test_splitmix_stream_known_answers():
    # Published splitmix64 outputs for state 0 (Vigna's reference generator):
    # u(0, k) is the k-th output mapped to [0, 1) by its top 53 bits.
    for k, z in enumerate((0xE220A8397B1DCDAF, 0x6E789E6AA1B965F4, 0x06C45D188009454F)):
        assert orc.uniform(0, k) == (z >> 11) * 2.0 ** -53
    X, Y = orc.init_state(5, orc.MAXCUT, 200, 64, f64=True)
    assert X.min() >= -0.1 and X.max() < 0.1 and abs(X.mean()) < 0.005
    assert np.std(Y) == pytest.approx(0.2 / math.sqrt(12), rel=0.03)
    Xt, _ = orc.init_state(5, orc.TSP, 20, 8)
    assert not Xt.any()


def test_prestep_order():
    # y first, then x with the new y (src/kernels.cu:19-20): (0, y) -> (dt y, y);
    # (x, 0) -> (x - dt^2 x, -dt x).
    dt = 0.25
    X = np.array([[0.0, 0.5]])
    Y = np.array([[0.3, 0.0]])
    orc.prestep(X, Y, dt)
    assert X[0, 0] == dt * 0.3 and Y[0, 0] == 0.3
    assert Y[0, 1] == -dt * 0.5 and X[0, 1] == 0.5 - dt * dt * 0.5


def test_tsp_step_structure():
    # TSP step with c = (1, 0, 1, b): y' = 2 F + b h with F = -J (1.005 x)
    # (src/kernels.cu:578-593), x' = x + y'; confined elements skip the update.
    n = 4
    d, D, mx = _tsp_instance(n, 3)
    J = orc.tsp_qubo_dense(n, d, mx)
    h = orc.tsp_field(n, D, mx)
    csr = sp.csr_matrix(J)
    csr.sort_indices()
    pb = orc.Problem(csr.indptr, csr.indices, csr.data, field=h, ncities=n)
    rng = np.random.default_rng(4)
    X = rng.uniform(-0.9, 0.9, (n * n, 3))
    X[5, 1] = 1.5
    Y = np.zeros_like(X)
    Xn = orc.step(pb, orc.TSP, orc.BSB, X, Y, [1.0, 0.0, 1.0, 0.75], 1.005, "f64")
    want_y = 2 * (-(csr @ (X * 1.005))) + 0.75 * h[:, None]
    mask = np.ones_like(X, bool)
    mask[5, 1] = False
    np.testing.assert_allclose(Y[mask], want_y[mask], rtol=1e-12, atol=1e-9)
    np.testing.assert_allclose(Xn[mask], (X + want_y)[mask], rtol=1e-12, atol=1e-9)
    assert Xn[5, 1] == 1.0 and Y[5, 1] == 0.0


def test_unfused_kick_uses_updated_x():
    # Four-kernel path: Update_XY, then mat_mult on the updated X, then confine.
    pb, M = random_problem(29, 0.25, 8, symmetric=False)
    rng = np.random.default_rng(9)
    X = rng.uniform(-0.9, 0.9, (29, 6))
    Y = rng.uniform(-0.9, 0.9, (29, 6))
    c = [1.3, -0.4, 0.05, 0.0]
    Yu = Y + c[1] * X
    Xu = X + c[0] * Yu
    want_y = Yu + c[2] * (-(sp.csr_matrix(M) @ np.where(np.signbit(Xu), -1.0, 1.0)))
    want_x = Xu.copy()
    conf = np.abs(Xu) > 1
    want_x[conf] = np.sign(Xu[conf])
    want_y[conf] = 0.0
    Xn = orc.step(pb, orc.MAXCUT_UNFUSED, orc.DSB, X, Y, c, 1.0, "f64")
    assert conf.any()
    np.testing.assert_allclose(Xn, want_x, rtol=1e-13, atol=1e-14)
    np.testing.assert_allclose(Y, want_y, rtol=1e-12, atol=1e-13)


def test_window_stability_resets_on_change():
    # A sign change at window 3 resets stability; the lock comes 24 unchanged
    # windows after the last change, and final spins are the spins at lock.
    n, A, K = 5, 2, 25
    cur = np.full((n, A), -1, np.int32)
    fin = cur.copy()
    stab = np.zeros(A, np.int32)
    bif = np.zeros(A, np.uint8)
    Xa = np.full((n, A), -0.5, np.float32)
    Xb = Xa.copy()
    Xb[2, 0] = 0.5  # agent 0 differs, agent 1 does not
    seq = [Xa, Xa, Xb] + [Xb] * 30
    lock_at = None
    for w, Xw in enumerate(seq, 1):
        nb = orc.window_maxcut(np.ascontiguousarray(Xw), K, cur, fin, stab, bif)
        if bif[0] and lock_at is None:
            lock_at = w
            assert np.array_equal(fin[:, 0], np.where(np.signbit(Xb[:, 0]), -1, 1))
        if w == 24:
            assert bif[1] == 1 and stab[1] == 24  # never changed: lock at window 24
    assert lock_at == 3 + 24 and nb == 2


def test_executed_steps_closed_form():
    pb, _ = random_problem(10, 0.3, 11)
    for T, P, want in ((0, 100, 0), (1, 100, 100), (250, 100, 300), (300, 100, 300), (17, 5, 20)):
        X, Y = orc.init_state(1, orc.MAXCUT, 10, 3)
        assert orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, T, 0, P, dt=0.1, xi0=0.2)["steps"] == want


def test_tsp_totals_and_tour_length_direct():
    rng = np.random.default_rng(12)
    n = 9
    x = rng.integers(0, 500, n).astype(float)
    y = rng.integers(0, 500, n).astype(float)
    d = orc.tsp_distances(x, y, 0, 1)
    Dm = np.floor(np.hypot(x[:, None] - x[None, :], y[:, None] - y[None, :]) + 0.5)
    D, mx = orc.tsp_totals(n, d)
    assert np.array_equal(D, Dm.sum(1)) and mx == Dm.max()
    perm = rng.permutation(n)
    spins = _tour_spins(perm)[:, None].copy()
    L, tour = orc.tsp_tour_length(n, d, spins, 0)
    assert tour.tolist() == perm.tolist()
    assert L == sum(Dm[perm[i], perm[(i + 1) % n]] for i in range(n))


@pytest.mark.parametrize("n,seed", [(3, 1), (4, 2), (4, 5)])
def test_tsp_field_ground_state_is_optimal_tour(n, seed):
    # Brute force over all 2^(n^2) spin states: the minimum of
    # s^T J s - h.s is a valid one-hot tour of minimum length.  (The
    # valid-tour identity alone cannot see terms of h that are constant over
    # valid tours; a flipped sign, a missing (n-2)(B+C) or a wrong factor on
    # D_k all make the ground state invalid.)
    d, D, mx = _tsp_instance(n, seed)
    J = orc.tsp_qubo_dense(n, d, mx)
    h = orc.tsp_field(n, D, mx)
    S = np.array(list(itertools.product([-1, 1], repeat=n * n)), dtype=float)
    E = np.einsum("si,ij,sj->s", S, J, S) - S @ h
    s = S[int(np.argmin(E))].reshape(n, n)
    assert ((s == 1).sum(0) == 1).all() and ((s == 1).sum(1) == 1).all()
    perm = [int(np.nonzero(s[i] == 1)[0][0]) for i in range(n)]
    L = lambda p: sum(d[orc.tri(min(p[i], p[(i + 1) % n]), max(p[i], p[(i + 1) % n]), n)] for i in range(n))
    assert L(perm) == min(L(p) for p in itertools.permutations(range(n)))


def test_bf16_rounding_matches_torch():
    torch = pytest.importorskip("torch")
    rng = np.random.default_rng(13)
    vals = np.concatenate([rng.normal(size=2000).astype(np.float32),
                           np.array([0.0, -0.0, 1.0, 1.00390625, 1.0078125, 3.0e38, -1e-40,
                                     np.float32(1 + 2 ** -8), np.float32(1 + 3 * 2 ** -9)], np.float32)])
    want = torch.from_numpy(vals).to(torch.bfloat16).to(torch.float32).numpy()
    got = np.array([orc.bf16(float(v)) for v in vals], np.float32)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))
