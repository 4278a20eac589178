"""Batched serving: every instance's result is bit-identical alone and in the
mix (cuts, current spins and executed steps), with admission under capacity
pressure, retirement on lock, uneven sizes and schedules."""
import numpy as np
import pytest

import oracle as orc
import sis

pytestmark = pytest.mark.gpu


def instances(count, seed, nmin=64, nmax=700):
    rng = np.random.default_rng(seed)
    out = []
    for k in range(count):
        n = int(np.exp(rng.uniform(np.log(nmin), np.log(nmax))))
        J = sis.Csr.gen_er(n, min(1.0, 8.0 / (n - 1)), sis.W_INT3, 1000 + k)
        out.append(J)
    return out


def solve_alone(J, agents, dt, T, uw, seed, P=100):
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=agents, dt=dt, form=sis.FORM_PACKED,
                           window_period=P))
    s.init_random(seed)
    info = s.run(T, uw)
    cuts, best = s.cuts()
    spins = s.spins()
    s.close()
    return info["steps_executed"], cuts, spins


@pytest.mark.parametrize("agents_max,agents", [(64, 64), (128, 100)])
def test_batch_results_equal_alone(agents_max, agents):
    Js = instances(24, 5)
    T = [800, 1500, 0, 1200] * 6
    dts = [0.5, 0.3, 0.7] * 8
    # capacity for about a third of the rows: admission happens in waves
    total = sum(J.n for J in Js)
    b = sis.Batch(agents_max=agents_max, max_instances=8, max_rows=max(total // 3, 700),
                  max_nnz=sum(J.nnz for J in Js), chunk_steps=100)
    tickets = [b.submit(J, agents, dts[k], T[k], update_window=k % 2, seed=77 + k)
               for k, J in enumerate(Js)]
    seen_queued = False
    for _ in range(1000):
        running, queued = b.advance(200)
        seen_queued |= queued > 0
        if running == 0 and queued == 0:
            break
    assert seen_queued
    b.advance(0)
    for k, (J, t) in enumerate(zip(Js, tickets)):
        st = b.poll(t)
        assert st["state"] == 2, (k, st)
        steps, cuts, spins = solve_alone(J, agents, dts[k], T[k], k % 2, 77 + k)
        bc, bs = b.result(t, spins=True)
        assert st["steps_executed"] == steps, k
        assert np.array_equal(bc, cuts), k
        assert np.array_equal(bs, spins), k
        assert st["best_cut"] == cuts.max() and st["best_agent"] == int(np.argmax(cuts))
        assert st["submit_s"] <= st["admit_s"] <= st["done_s"]


XCD = 1 << 18  # XCD-affine task order: with 9 uneven instances both placements
# (whole instance on one XCD, and the largest dealt to every XCD) occur
BT = 1 << 19  # bit-transpose integer sums (instances with small integer weights)
PF = 1 << 20  # next-task prefetch (bit-transpose, non-temporal)
LS = 1 << 21  # sign words staged in LDS, segments of 512 rows
PL = 1 << 25  # pipelined bit-transpose step (stores deferred behind the next task's loads)


@pytest.mark.parametrize("agents_max,tile", [(64, 1), (64, 8 | 1 << 16), (64, 2 | 1 << 17),
                                             (64, 8 | 1 << 16 | 1 << 17), (128, 4 | 1 << 17),
                                             (256, 2 | 1 << 16 | 1 << 17), (64, 1 | XCD),
                                             (64, 8 | 1 << 16 | 1 << 17 | XCD),
                                             (256, 2 | 1 << 16 | 1 << 17 | XCD),
                                             (64, 4 | 1 << 16 | BT), (64, 8 | BT | XCD),
                                             (64, 2 | BT), (128, 4 | BT), (256, 2 | 1 << 16 | BT),
                                             (64, 8 | 1 << 16 | BT | PF), (64, 2 | 1 << 16 | BT | PF | XCD),
                                             (256, 4 | 1 << 16 | BT | PF), (64, 4 | LS),
                                             (64, 8 | 1 << 16 | LS | XCD), (128, 2 | LS),
                                             (256, 4 | 1 << 16 | LS), (64, 4 | 1 << 16 | BT | PL),
                                             (64, 8 | BT | PL | XCD), (64, 2 | 1 << 16 | BT | PL),
                                             (128, 4 | BT | PL), (256, 8 | 1 << 16 | BT | PL)])
def test_batch_tiles_bit_exact(agents_max, tile):
    Js = instances(9, 21, 64, 500)
    A = min(agents_max, 100) if agents_max > 64 else 64
    b = sis.Batch(agents_max=agents_max, max_instances=9, max_rows=5000, max_nnz=100000,
                  chunk_steps=100, tile=tile)
    tickets = [b.submit(J, A, 0.5, 600, update_window=0, seed=40 + k) for k, J in enumerate(Js)]
    while b.advance(600)[0]:
        pass
    b.advance(0)
    for k, (J, t) in enumerate(zip(Js, tickets)):
        steps, cuts, spins = solve_alone(J, A, 0.5, 600, 0, 40 + k)
        bc, bs = b.result(t, spins=True)
        assert np.array_equal(bc, cuts) and np.array_equal(bs, spins), (tile, k)


def test_batch_eager_equals_graph():
    Js = instances(6, 33, 64, 300)
    res = []
    for eager in (False, True):
        b = sis.Batch(agents_max=64, max_instances=6, max_rows=3000, max_nnz=50000,
                      chunk_steps=50, eager=eager)
        ts = [b.submit(J, 64, 0.5, 700, update_window=1, seed=3 + k) for k, J in enumerate(Js)]
        while b.advance(200)[0]:
            pass
        b.advance(0)
        res.append([(b.result(t), b.poll(t)["steps_executed"]) for t in ts])
        b.close()
    for (ca, sa), (cb, sb) in zip(*res):
        assert np.array_equal(ca, cb) and sa == sb


def test_batch_matches_oracle_directly():
    J = sis.Csr.gen_er(150, 8 / 149, sis.W_INT3, 3)
    rp, col, val = J.arrays()
    pb = orc.Problem(rp, col, val, S=J.sum)
    b = sis.Batch(agents_max=64, max_instances=4, max_rows=1000, max_nnz=10000, chunk_steps=50)
    t = b.submit(J, 64, 0.5, 1000, update_window=1, seed=9)
    while b.advance(500)[0]:
        pass
    b.advance(0)
    X, Y = orc.init_state(9, orc.MAXCUT, J.n, 64)
    r = orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, 1000, 1, P=100, dt=0.5, xi0=J.xi0())
    cuts, spins = b.result(t, spins=True)
    assert b.poll(t)["steps_executed"] == r["steps"]
    assert np.array_equal(spins, r["cur"]) and np.array_equal(cuts, orc.cut(pb, r["cur"]))


def test_batch_autotune_midrun_keeps_results():
    Js = instances(10, 8, 100, 400)
    b = sis.Batch(agents_max=64, max_instances=16, max_rows=10000, max_nnz=200000, chunk_steps=100)
    tickets = [b.submit(J, 64, 0.5, 1000, update_window=0, seed=5 + k) for k, J in enumerate(Js)]
    b.advance(300)
    tile, ms = b.autotune(10)
    assert ms > 0 and (tile & 0xffff) in (1, 2, 4, 8)
    while b.advance(300)[0]:
        pass
    b.advance(0)
    for k, (J, t) in enumerate(zip(Js, tickets)):
        steps, cuts, spins = solve_alone(J, 64, 0.5, 1000, 0, 5 + k)
        assert np.array_equal(b.result(t), cuts) and b.poll(t)["steps_executed"] == steps


def _weighted(n, p, seed, kind):
    """ER graph with weights of a given kind: 'pm1' +-1 (one bit plane),
    'wide' integers up to 127 and 'big' beyond it (both outside the bit-plane
    range), 'real' non-integers."""
    rng = np.random.default_rng(seed)
    iu, ju = np.triu_indices(n, 1)
    keep = rng.random(len(iu)) < p
    rows, cols = iu[keep].astype(np.int64), ju[keep].astype(np.int64)
    if kind == "pm1":
        vals = rng.choice([-1, 1], size=len(rows))
    elif kind == "wide":
        vals = rng.choice([-127, -100, -64, -37, -5, -1, 1, 3, 37, 64, 100, 127], size=len(rows))
    elif kind == "real":
        vals = rng.uniform(-2.0, 2.0, size=len(rows))
    else:
        vals = rng.choice([-300, 200, 129, -1, 1], size=len(rows))
    return sis.Csr.from_coo(n, rows, cols, vals.astype(np.float64), mirror=True)


@pytest.mark.parametrize("tile", [4 | BT, 8 | 1 << 16 | BT | XCD, 2 | BT, 8 | 1 << 16 | BT | PF,
                                  4 | 1 << 16 | BT | PL, 8 | BT | PL | XCD, 2 | 1 << 16 | BT | PL])
def test_batch_bit_transpose_mixed_weights(tile):
    """Small-integer instances take the bit-plane sums, others the readlane
    path, in the same launch; every result equals the solo solve."""
    Js = instances(4, 17, 64, 400) + [_weighted(150, 0.06, 1, "wide"), _weighted(90, 0.1, 2, "real"),
                                      _weighted(120, 0.08, 3, "big"), _weighted(300, 0.04, 4, "pm1")]
    b = sis.Batch(agents_max=64, max_instances=8, max_rows=6000, max_nnz=200000, chunk_steps=100,
                  tile=tile)
    tickets = [b.submit(J, 64, 0.5, 700, update_window=1, seed=60 + k) for k, J in enumerate(Js)]
    while b.advance(700)[0]:
        pass
    b.advance(0)
    for k, (J, t) in enumerate(zip(Js, tickets)):
        steps, cuts, spins = solve_alone(J, 64, 0.5, 700, 1, 60 + k)
        bc, bs = b.result(t, spins=True)
        assert b.poll(t)["steps_executed"] == steps, k
        assert np.array_equal(bs, spins), (tile, k)
        if k == 5:  # real weights: the two cut kernels add FP64 terms in different orders
            np.testing.assert_allclose(bc, cuts, rtol=1e-12, atol=0)
        else:
            assert np.array_equal(bc, cuts), (tile, k)


@pytest.mark.parametrize("agents_max,tile", [(64, 4 | LS), (64, 2 | 1 << 16 | LS | XCD),
                                             (64, 8 | LS | 1 << 22), (64, 4 | LS | 3 << 22),
                                             (256, 4 | 1 << 16 | LS | 2 << 22)])
def test_batch_lds_staging_segments_and_fallback(agents_max, tile):
    """Instances of one segment, of several (segments of 512 << s rows for
    segment code s; 700 and 1500 rows give partial last segments) and too
    large for the 4096 staged words (5000 rows at 64 agents; above 1024 rows
    at 256 agents), with small integer and real weights, in one launch: each
    equals its solo solve."""
    sizes = [100, 700, 1500, 5000]
    Js = [sis.Csr.gen_er(n, min(1.0, 8.0 / (n - 1)), sis.W_INT3, 70 + k) for k, n in enumerate(sizes)]
    Js.append(_weighted(900, 0.01, 5, "real"))
    A = 64 if agents_max == 64 else 200
    b = sis.Batch(agents_max=agents_max, max_instances=5, max_rows=9000, max_nnz=200000,
                  chunk_steps=100, tile=tile)
    tickets = [b.submit(J, A, 0.5, 500, update_window=1, seed=80 + k) for k, J in enumerate(Js)]
    while b.advance(500)[0]:
        pass
    b.advance(0)
    for k, (J, t) in enumerate(zip(Js, tickets)):
        steps, cuts, spins = solve_alone(J, A, 0.5, 500, 1, 80 + k)
        bc, bs = b.result(t, spins=True)
        assert b.poll(t)["steps_executed"] == steps, (tile, k)
        assert np.array_equal(bs, spins), (tile, k)
        if k == 4:  # real weights: the two cut kernels add FP64 terms in different orders
            np.testing.assert_allclose(bc, cuts, rtol=1e-12, atol=0)
        else:
            assert np.array_equal(bc, cuts), (tile, k)


def test_batch_instances_without_edges():
    """Instances with no stored entry (and a single node) run beside others;
    each equals its solo solve (same xi0 given explicitly: the sample std of
    an all-zero matrix is 0)."""
    empty = lambda n: sis.Csr.from_coo(n, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0))
    Js = instances(3, 8, 64, 200) + [empty(5), empty(1), empty(130)]
    b = sis.Batch(agents_max=64, max_instances=8, max_rows=2000, max_nnz=50000, chunk_steps=50)
    ts = [b.submit(J, 64, 0.5, 400, update_window=1, seed=20 + k, xi0=0.3) for k, J in enumerate(Js)]
    while b.advance(400)[0]:
        pass
    b.advance(0)
    for k, (J, t) in enumerate(zip(Js, ts)):
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=64, dt=0.5, xi0=0.3,
                               form=sis.FORM_PACKED))
        s.init_random(20 + k)
        info = s.run(400, 1)
        cuts, _ = s.cuts()
        bc, bs = b.result(t, spins=True)
        assert b.poll(t)["steps_executed"] == info["steps_executed"], k
        assert np.array_equal(bc, cuts) and np.array_equal(bs, s.spins()), k
        s.close()


RES = 1 << 24  # resident chunks (with the streaming tile as fallback)


@pytest.mark.parametrize("agents_max,tile,chunk,eager", [(64, 8 | 1 << 16 | BT | RES, 100, False),
                                                         (64, 4 | RES, 50, True),
                                                         (128, 2 | 1 << 16 | RES, 100, False),
                                                         (256, 4 | BT | RES, 50, False),
                                                         (64, 8 | BT | RES, 30, False),
                                                         (64, 4 | 1 << 16 | BT | PL | RES, 100, False)])
def test_batch_resident_chunks(agents_max, tile, chunk, eager):
    """Resident chunks against solo solves: instances of 1 to 6000 rows (one
    to 64 agents per workgroup), early exits, admission in waves, and a
    real-weight instance that keeps the batch on the streaming tile while it
    is live (it retires first, then the chunks go resident).  Chunks of 30
    steps do not divide the window period: streaming throughout."""
    empty = sis.Csr.from_coo(1, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0))
    Js = instances(8, 91, 64, 3000) + [sis.Csr.gen_er(6000, 8.0 / 5999, sis.W_INT3, 92),
                                       _weighted(200, 0.05, 7, "real"), empty]
    T = [600, 900, 500, 800, 700, 900, 600, 800, 700, 300, 500]
    A = 64 if agents_max == 64 else 100
    total = sum(J.n for J in Js)
    b = sis.Batch(agents_max=agents_max, max_instances=8, max_rows=max(total // 2, 6000),
                  max_nnz=sum(J.nnz for J in Js), chunk_steps=chunk, tile=tile, eager=eager)
    ts = [b.submit(J, A, 0.5, T[k], update_window=k % 2, seed=50 + k, xi0=0.3)
          for k, J in enumerate(Js)]
    while sum(b.advance(300)):
        pass
    b.advance(0)
    for k, (J, t) in enumerate(zip(Js, ts)):
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=0.5, xi0=0.3, form=sis.FORM_PACKED))
        s.init_random(50 + k)
        info = s.run(T[k], k % 2)
        cuts, _ = s.cuts()
        bc, bs = b.result(t, spins=True)
        assert b.poll(t)["steps_executed"] == info["steps_executed"], (tile, k)
        assert np.array_equal(bs, s.spins()), (tile, k)
        if k == 9:  # real weights: the two cut kernels add FP64 terms in different orders
            np.testing.assert_allclose(bc, cuts, rtol=1e-12, atol=0)
        else:
            assert np.array_equal(bc, cuts), (tile, k)
        s.close()


@pytest.mark.parametrize("seed", [101, 202, 303])
def test_batch_resident_random_mixes(seed):
    """Seeded random mixes on the default tile (resident chunks): sizes from 1
    to 3000 rows, +-1 and small-integer weights, agents below agents_max,
    random step counts and update-window flags, chunks of 25, 50 or 100
    steps; every instance equals its solo solve."""
    rng = np.random.default_rng(seed)
    chunk = int(rng.choice([25, 50, 100]))
    agents_max = int(rng.choice([64, 128]))
    Js, params = [], []
    for k in range(12):
        n = int(np.exp(rng.uniform(0, np.log(3000))))
        if n < 2:
            J = sis.Csr.from_coo(1, np.zeros(0, np.int64), np.zeros(0, np.int64), np.zeros(0))
        else:
            J = sis.Csr.gen_er(n, min(1.0, 6.0 / (n - 1)), sis.W_PM1 if k % 3 == 0 else sis.W_INT3,
                               seed * 100 + k)
        Js.append(J)
        params.append((int(rng.integers(1, agents_max + 1)), int(rng.integers(1, 900)),
                       int(rng.integers(0, 2))))
    b = sis.Batch(agents_max=agents_max, max_instances=6, max_rows=4000, max_nnz=60000,
                  chunk_steps=chunk)
    ts = [b.submit(J, A, 0.5, T, update_window=uw, seed=seed + k, xi0=0.3)
          for k, (J, (A, T, uw)) in enumerate(zip(Js, params))]
    while sum(b.advance(200)):
        pass
    b.advance(0)
    seen_g = set()
    for k, (J, t, (A, T, uw)) in enumerate(zip(Js, ts, params)):
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=0.5, xi0=0.3, form=sis.FORM_PACKED))
        s.init_random(seed + k)
        info = s.run(T, uw)
        cuts, _ = s.cuts()
        bc, bs = b.result(t, spins=True)
        assert b.poll(t)["steps_executed"] == info["steps_executed"], (seed, k)
        assert np.array_equal(bs, s.spins()) and np.array_equal(bc, cuts), (seed, k, J.n, A, T, uw)
        s.close()
        # the first instance of each resident group size G (agents per
        # workgroup: the largest power of two <= 64 with N G <= 8192) also
        # against the oracle directly
        G = 64
        while G > 1 and J.n * G > 8192:
            G //= 2
        if G in seen_g:
            continue
        seen_g.add(G)
        rp, col, val = J.arrays()
        pb = orc.Problem(rp, col, val, S=J.sum)
        X, Y = orc.init_state(seed + k, orc.MAXCUT, J.n, A)
        r = orc.solve(pb, orc.MAXCUT, orc.DSB, X, Y, T, uw, P=100, K=25, dt=0.5, xi0=0.3)
        assert b.poll(t)["steps_executed"] == r["steps"], (seed, k, G)
        assert np.array_equal(bs, r["cur"]) and np.array_equal(bc, orc.cut(pb, r["cur"])), (seed, k, G)
    assert len(seen_g) >= 2
