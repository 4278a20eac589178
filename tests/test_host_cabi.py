"""Host half of the C ABI: exported symbols, the direct-to-CSR TSP QUBO,
xi0, statistics, generators.  Runs without a GPU (no device entry point is
called)."""
import os
import re

import numpy as np
import pytest
import scipy.sparse as sp

import oracle as orc
import readers
import sis

ROOT = os.path.abspath(os.path.join(os.path.dirname(__file__), ".."))


def declared_symbols():
    txt = open(os.path.join(ROOT, "include", "sis.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(sis_[a-z0-9_]+)\s*\(", txt)))


def test_library_exports_every_declared_symbol():
    lib = sis.load()
    names = declared_symbols()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing


def test_binding_covers_header():
    assert set(declared_symbols()) == set(sis._SIGS)


def test_binding_struct_layouts_match_header(tmp_path):
    """Every ctypes structure has the size and field offsets the C compiler
    gives the header's struct."""
    import ctypes
    import subprocess

    structs = {"sis_params": sis.Params, "sis_run_info": sis.RunInfo,
               "sis_batch_params": sis.BatchParams, "sis_batch_status": sis.BatchStatus}
    lines = ['#include <stddef.h>', '#include <stdio.h>', '#include "sis.h"', "int main(void) {"]
    for cname, cls in structs.items():
        lines.append(f'  printf("{cname} size %zu\\n", sizeof({cname}));')
        for f, _ in cls._fields_:
            lines.append(f'  printf("{cname} {f} %zu\\n", offsetof({cname}, {f}));')
    lines += ["  return 0;", "}"]
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines) + "\n")
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c11", "-I", os.path.join(ROOT, "include"), "-o", str(exe), str(src)],
                   check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {tuple(l.split()[:2]): int(l.split()[2]) for l in out if l.strip()}
    for cname, cls in structs.items():
        assert got[(cname, "size")] == ctypes.sizeof(cls), cname
        for f, _ in cls._fields_:
            assert got[(cname, f)] == getattr(cls, f).offset, (cname, f)


@pytest.mark.parametrize("n,seed,euc_round", [(2, 1, 0), (3, 2, 0), (5, 3, 1), (7, 4, 0), (12, 5, 1)])
def test_tsp_qubo_direct_csr_equals_dense(n, seed, euc_round):
    rng = np.random.default_rng(seed)
    x = rng.integers(0, 300, n).astype(float)
    y = rng.integers(0, 300, n).astype(float)
    if n >= 5:
        x[2], y[2] = x[0], y[0]  # coincident cities: zero-distance entries skipped
    t = sis.Tsp.from_coords(x, y, sis.EUC_2D, euc_round)
    J, h = t.qubo()
    rp, col, val = J.arrays()
    d = orc.tsp_distances(x, y, 0, euc_round)
    D, mx = orc.tsp_totals(n, d)
    dense = orc.tsp_qubo_dense(n, d, mx)
    ref = readers.csr_from_dense(dense, n * n)
    assert rp.tolist() == ref["row_ptr"].tolist()
    assert col.tolist() == ref["col"].tolist() and val.tolist() == ref["val"].tolist()
    assert np.array_equal(h, orc.tsp_field(n, D, mx))
    _, _, dist, tot = t.arrays()
    assert np.array_equal(dist, d) and np.array_equal(tot, D) and t.maxd == mx


def test_tsp_qubo_from_reference_dense_fixtures(ref_fixtures):
    for name, fx in ref_fixtures["dense"].items():
        t = sis.Tsp.from_coords(np.array(fx["x"]), np.array(fx["y"]))
        J, _ = t.qubo()
        rp, col, val = J.arrays()
        assert rp.tolist() == fx["csr"]["row_ptr"], name
        assert col.tolist() == fx["csr"]["col"] and val.tolist() == fx["csr"]["val"], name


def test_tsp_qubo_nnz_large_n():
    t = sis.Tsp.generate(48, 1000, 1, 12345)
    J, h = t.qubo()
    _, _, d, _ = t.arrays()
    zeros = int(np.sum(d == 0))
    assert J.n == 2304
    # each zero distance removes 2 * 2 entries per position pair orientation
    assert J.nnz == 48 * 48 * (4 * 48 - 3) - 4 * 48 * zeros
    assert h.shape == (2304,)


def test_xi0_product_equals_oracle():
    for seed in range(3):
        c = sis.Csr.gen_er(300, 0.05, sis.W_INT3, seed)
        _, _, val = c.arrays()
        assert c.xi0() == orc.xi0_maxcut(c.n, val)


def test_cut_stats_product_equals_oracle():
    rng = np.random.default_rng(3)
    for A in (1, 2, 77):
        c = rng.integers(-5, 30, A).astype(float)
        a, b = sis.cut_stats(c), orc.cut_stats(c)
        assert np.array_equal(np.isnan(a), np.isnan(b))
        assert np.array_equal(a[~np.isnan(a)], b[~np.isnan(b)])


def test_generators_shapes_and_determinism():
    er = sis.Csr.gen_er(800, 0.06, sis.W_ONE, 42)
    rp, col, val = er.arrays()
    assert abs(er.nnz / 2 - 0.06 * 800 * 799 / 2) < 5 * np.sqrt(0.06 * 800 * 799 / 2)
    assert np.all(val == 1) and er.sum == er.nnz / 2
    er2 = sis.Csr.gen_er(800, 0.06, sis.W_ONE, 42)
    assert np.array_equal(er2.arrays()[1], col)
    M = sp.csr_matrix((val, col, rp), shape=(800, 800))
    assert (M != M.T).nnz == 0 and M.diagonal().sum() == 0

    tor = sis.Csr.gen_torus(10, 10, sis.W_PM1, 1)
    rp, col, val = tor.arrays()
    assert tor.nnz == 400 and np.all(np.diff(rp) == 4) and set(val) <= {-1.0, 1.0}

    reg = sis.Csr.gen_regular(2000, 4, sis.W_PM1, 2)
    rp, col, val = reg.arrays()
    assert np.all(np.diff(rp) == 4)
    M = sp.csr_matrix((np.ones_like(val), col, rp), shape=(2000, 2000))
    assert M.max() == 1 and M.diagonal().sum() == 0 and (M != M.T).nnz == 0

    c = sis.Csr.gen_er(4096, 8 / 4095, sis.W_INT3, 9)
    _, _, val = c.arrays()
    assert set(val) <= {-3.0, -2.0, -1.0, 1.0, 2.0, 3.0}


def test_chung_lu_power_law():
    n, avg, gamma = 1 << 16, 8.0, 2.1
    g = sis.Csr.gen_chung_lu(n, avg, gamma, sis.W_PM1, 3)
    rp, col, val = g.arrays()
    deg = np.diff(rp)
    M = sp.csr_matrix((np.ones_like(val), col, rp), shape=(n, n))
    assert M.max() == 1 and M.diagonal().sum() == 0 and (M != M.T).nnz == 0  # simple, symmetric
    assert set(val) <= {-1.0, 1.0}
    # drawn pairs: n avg / 2, minus self-loops and repeats (many at the hubs)
    assert 0.6 * n * avg < g.nnz <= n * avg
    # expected degree of rank r is proportional to (r + 1)^(-1/(gamma - 1)):
    # the top rank expects avg n w_0 / sum_r w_r draws; repeated pairs with
    # the other hubs collapse, leaving about 40% of that as distinct neighbours
    w = np.arange(1, n + 1, dtype=np.float64) ** (-1 / (gamma - 1))
    top = avg * n * w[0] / w.sum()
    d = np.sort(deg)[::-1]
    assert 0.3 * top < d[0] < 1.2 * top
    assert d[100] < d[10] < d[0]
    assert np.median(deg) < avg
    g2 = sis.Csr.gen_chung_lu(n, avg, gamma, sis.W_PM1, 3)
    assert np.array_equal(g2.arrays()[1], col)


def test_tour_length_and_random_tour():
    t = sis.Tsp.from_coords([0, 3, 3, 0], [0, 0, 4, 4])
    assert t.tour_length([0, 1, 2, 3]) == 14.0
    assert t.tour_length([0, 2, 1, 3]) == 5 + 4 + 5 + 4
    assert t.random_length(7) in (14.0, 16.0, 18.0) and t.random_length(7) == t.random_length(7)
    with pytest.raises(sis.SisError):
        t.tour_length([0, 0, 1, 2])


@pytest.mark.parametrize("n,seed", [(3, 1), (4, 2), (7, 3), (12, 4)])
def test_tsp_structured_identity(n, seed):
    """The structured product used by sis_solver_create_tsp(structured=1):
    (J g)(i,k) = C/4 sum_j g(j,k) + B/4 sum_l g(i,l) + sum_l d_kl/8 (g(i-1,l) + g(i+1,l))
    equals the direct-to-CSR QUBO times g (n >= 3)."""
    t = sis.Tsp.generate(n, 100, 0, seed)
    J, h = t.qubo()
    rp, col, val = J.arrays()
    M = sp.csr_matrix((val, col, rp), shape=(n * n, n * n))
    _, _, d, _ = t.arrays()
    D = np.zeros((n, n))
    iu = np.triu_indices(n, 1)
    D[iu] = d
    D = D + D.T
    B = C = t.maxd
    rng = np.random.default_rng(seed)
    for _ in range(3):
        g = rng.uniform(-1, 1, n * n)
        G = g.reshape(n, n)  # G[i, k]: position i, city k
        v = np.roll(G, 1, axis=0) + np.roll(G, -1, axis=0)  # g(i-1, .) + g(i+1, .)
        S = C / 4 * G.sum(axis=0)[None, :] + B / 4 * G.sum(axis=1)[:, None] + (v @ D.T) / 8
        np.testing.assert_allclose(S.ravel(), M @ g, rtol=1e-12, atol=1e-9)


@pytest.mark.parametrize("n", [2, 3, 4, 5, 7, 9])
@pytest.mark.parametrize("seed,span,rnd", [(1, 100, 0), (4, 1000, 1), (9, 13, 0)])
def test_tsp_scaled_xi0_against_dense_qubo(n, seed, span, rnd):
    """The reference's alternative coupling (src/ising_tsp.cu:62-83):
    0.5 / (sigma sqrt N), sigma^2 = sum of squared QUBO entries / (N^2 - 1),
    recomputed with numpy over the oracle's dense n^2 x n^2 matrix (and over the
    stored CSR entries), so a wrong entry count (the adjacent positions of
    n = 2 and n >= 3, the diagonal) or a centred moment would fail."""
    import oracle as orc

    t = sis.Tsp.generate(n, span, rnd, seed)
    _, _, d, _ = t.arrays()
    D, maxd = orc.tsp_totals(n, d)
    dense = orc.tsp_qubo_dense(n, d, maxd)
    N = n * n
    want = 0.5 / (np.sqrt(np.sum(dense ** 2) / (N * N - 1)) * np.sqrt(N))
    np.testing.assert_allclose(t.xi0(), want, rtol=1e-15)
    J, _ = t.qubo()
    _, _, val = J.arrays()
    assert np.sum(val ** 2) == np.sum(dense ** 2)  # multiples of 1/64: exact
    # the mean the reference computes but drops would change the value
    cen = 0.5 / (np.sqrt((np.sum((dense - dense.mean()) ** 2)) / (N * N - 1)) * np.sqrt(N))
    assert t.xi0() != cen


def test_tsp_scaled_xi0_rejects_one_city():
    t = sis.Tsp.from_coords(np.array([1.0]), np.array([2.0]))
    with pytest.raises(sis.SisError):
        t.xi0()
