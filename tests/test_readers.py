"""Readers and CSR construction against the reference's own compiled readers.

Fixtures in tests/golden/ref_readers.json come from oracle/_ref/ref_driver,
which links the reference's unmodified src/csr.cpp and src/tsp_cpu.cpp.  Both
the oracle restatement (oracle/readers.py) and the product (libsis host entry
points) are checked against them.
"""
import os

import numpy as np
import pytest

import readers
import sis

INPUTS = os.path.join(os.path.dirname(__file__), "golden", "inputs")


def _ref_rowptr_ok(ref):
    # The reference's row pointers are correct exactly when no row is empty.
    return len(ref["row_ptr"]) == ref["size"] + 1


@pytest.mark.parametrize("name", ["pattern_sym.mtx", "int_general.mtx", "real_sym.mtx",
                                  "int_sym_pm1.mtx", "empty_rows.mtx"])
def test_oracle_mtx_matches_reference(name, ref_fixtures):
    ref = ref_fixtures["mtx"][name]
    o = readers.read_mtx(os.path.join(INPUTS, name), reference_rowptr=True)
    assert o["size"] == ref["size"]
    assert o["nnz"] == ref["nnz"]
    assert o["sum"] == ref["sum"]
    assert o["row_ptr"].tolist() == ref["row_ptr"]
    assert o["col"].tolist() == ref["col"]
    assert o["val"].tolist() == ref["val"]


@pytest.mark.parametrize("name", ["pattern_sym.mtx", "int_general.mtx", "real_sym.mtx",
                                  "int_sym_pm1.mtx", "empty_rows.mtx"])
def test_product_mtx_matches_reference(name, ref_fixtures):
    ref = ref_fixtures["mtx"][name]
    c = sis.Csr.read_mtx(os.path.join(INPUTS, name))
    rp, col, val = c.arrays()
    assert c.n == ref["size"] and c.nnz == ref["nnz"] and c.sum == ref["sum"]
    assert col.tolist() == ref["col"]
    assert val.tolist() == ref["val"]
    if _ref_rowptr_ok(ref):
        assert rp.tolist() == ref["row_ptr"]
    else:
        # Reference defect (src/csr.cpp:171-188): pointers lost across empty
        # rows.  The product builds the correct CSR; the oracle's correct mode
        # agrees with it.
        o = readers.read_mtx(os.path.join(INPUTS, name))
        assert rp.tolist() == o["row_ptr"].tolist()
        assert rp[-1] == ref["nnz"] and len(rp) == c.n + 1
        assert np.all(np.diff(rp) >= 0)


def test_empty_rows_degree_counts():
    c = sis.Csr.read_mtx(os.path.join(INPUTS, "empty_rows.mtx"))
    rp, col, _ = c.arrays()
    # edges 2-1, 4-2, 5-1 (1-based): degrees of nodes 1..7
    assert np.diff(rp).tolist() == [2, 2, 0, 1, 1, 0, 0]


@pytest.mark.parametrize("name", ["euc8.tsp", "geo6.tsp", "euc_float5.tsp"])
def test_tsp_readers_match_reference(name, ref_fixtures):
    ref = ref_fixtures["tsp"][name]
    o = readers.read_tsp(os.path.join(INPUTS, name))
    assert o["n"] == ref["n"] and o["dtype"] == ref["dtype"]
    assert o["x"].tolist() == ref["x"] and o["y"].tolist() == ref["y"]
    t = sis.Tsp.read(os.path.join(INPUTS, name))
    x, y, _, _ = t.arrays()
    assert t.n == ref["n"]
    assert t.kind() == (sis.GEO if ref["dtype"] == "GEO" else sis.EUC_2D)
    assert x.tolist() == ref["x"] and y.tolist() == ref["y"]


def test_tsp_tolerant_header(tmp_path):
    p = tmp_path / "spaced.tsp"
    p.write_text("NAME : spaced\nTYPE : TSP\nDIMENSION : 3\nEDGE_WEIGHT_TYPE : EUC_2D\n"
                 "NODE_COORD_SECTION\n  1 0 0\n  2 3 4\n  3 6 8\nEOF\n")
    t = sis.Tsp.read(p)
    _, _, d, _ = t.arrays()
    assert t.n == 3 and d.tolist() == [5.0, 10.0, 5.0]


def test_mtx_bad_banner_is_an_error(tmp_path):
    p = tmp_path / "bad.mtx"
    p.write_text("%%MatrixMarket matrix coordinate integer\n3 3 1\n1 2 1\n")
    with pytest.raises(sis.SisError):
        sis.Csr.read_mtx(p)
    p.write_text("%%MatrixMarket matrix coordinate integer skew-symmetric\n3 3 1\n2 1 1\n")
    with pytest.raises(sis.SisError):
        sis.Csr.read_mtx(p)


def test_mtx_truncated_and_out_of_range(tmp_path):
    p = tmp_path / "t.mtx"
    p.write_text("%%MatrixMarket matrix coordinate integer symmetric\n3 3 2\n2 1 1\n")
    with pytest.raises(sis.SisError):
        sis.Csr.read_mtx(p)
    p.write_text("%%MatrixMarket matrix coordinate integer symmetric\n3 3 1\n4 1 1\n")
    with pytest.raises(sis.SisError):
        sis.Csr.read_mtx(p)
    with pytest.raises(sis.SisError):
        sis.Csr.read_mtx(tmp_path / "missing.mtx")


def test_tsp_unsupported_type(tmp_path):
    p = tmp_path / "att.tsp"
    p.write_text("NAME: a\nDIMENSION: 2\nEDGE_WEIGHT_TYPE: ATT\nNODE_COORD_SECTION\n1 0 0\n2 1 1\nEOF\n")
    with pytest.raises(sis.SisError):
        sis.Tsp.read(p)


@pytest.mark.parametrize("gen", ["er", "torus", "regular"])
def test_mtx_writer_round_trip(tmp_path, gen):
    if gen == "er":
        c = sis.Csr.gen_er(60, 0.1, sis.W_INT3, 5)
    elif gen == "torus":
        c = sis.Csr.gen_torus(5, 7, sis.W_PM1, 6)
    else:
        c = sis.Csr.gen_regular(50, 4, sis.W_PM1, 7)
    p = tmp_path / "g.mtx"
    c.write_mtx(p)
    r = sis.Csr.read_mtx(p)
    a, b = c.arrays(), r.arrays()
    assert r.n == c.n and r.sum == c.sum
    for u, v in zip(a, b):
        assert u.tolist() == v.tolist()
    # and the oracle reader (reference semantics) agrees with the writer's file
    o = readers.read_mtx(str(p))
    assert o["col"].tolist() == b[1].tolist() and o["sum"] == r.sum


def test_coo_and_dense_builders(ref_fixtures):
    for name, fx in ref_fixtures["dense"].items():
        import oracle as orc

        n = fx["n"]
        d = orc.tsp_distances(np.array(fx["x"]), np.array(fx["y"]), 0, 0)
        D, mx = orc.tsp_totals(n, d)
        dense = orc.tsp_qubo_dense(n, d, mx)
        c = sis.Csr.from_dense(dense)
        rp, col, val = c.arrays()
        ref = fx["csr"]
        assert rp.tolist() == ref["row_ptr"], name
        assert col.tolist() == ref["col"] and val.tolist() == ref["val"], name
