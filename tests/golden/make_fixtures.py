"""Writes the golden reader/CSR fixtures.

Inputs are crafted here (written to tests/golden/inputs/); expected outputs
come from oracle/_ref/ref_driver, which links the reference's own, unmodified
src/csr.cpp and src/tsp_cpu.cpp (built by `make -C oracle ref`).  The outputs
are stored as JSON next to this script, so tests need neither the reference
tree nor the fixture writer.

    python tests/golden/make_fixtures.py
"""
import json
import os
import subprocess
import sys

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.abspath(os.path.join(HERE, "..", ".."))
INP = os.path.join(HERE, "inputs")
REF = os.path.join(ROOT, "oracle", "_ref", "ref_driver")
sys.path.insert(0, os.path.join(ROOT, "oracle"))

MTX = {
    # pattern + symmetric, comment lines, one diagonal entry (dropped)
    "pattern_sym.mtx": """%%MatrixMarket matrix coordinate pattern symmetric
% a comment line
% another one
6 6 8
2 1
3 1
3 2
4 4
5 3
6 5
6 1
4 2
""",
    # integer + general: both triangles given, negative weights, duplicates
    "int_general.mtx": """%%MatrixMarket matrix coordinate integer general
5 5 10
1 2 3
2 1 3
1 3 -2
3 1 -2
2 3 1
3 2 1
4 5 7
5 4 7
4 5 -1
5 4 -1
""",
    # real + symmetric with non-integral weights and a duplicated pair
    "real_sym.mtx": """%%MatrixMarket matrix coordinate real symmetric
%
4 4 6
2 1 0.5
3 1 -1.25
3 2 3e-1
4 3 2.0
4 1 0.125
4 3 -0.75
""",
    # integer symmetric, Gset-like: weights +-1, larger
    "int_sym_pm1.mtx": None,  # generated below
    # empty row 3 (0-based 2) and empty last rows: the reference's row
    # pointers lose entries here (src/csr.cpp:171-188)
    "empty_rows.mtx": """%%MatrixMarket matrix coordinate integer symmetric
7 7 3
2 1 1
4 2 1
5 1 2
""",
}

TSP = {
    "euc8.tsp": """NAME: euc8
TYPE: TSP
COMMENT: crafted
DIMENSION: 8
EDGE_WEIGHT_TYPE: EUC_2D
NODE_COORD_SECTION
1 565.0 575.0
2 25.0 185.0
3 345.0 750.0
4 945.0 685.0
5 845.0 655.0
6 880.0 660.0
7 25.0 230.0
8 525.0 1000.0
EOF
""",
    "geo6.tsp": """NAME: geo6
TYPE: TSP
COMMENT: crafted GEO instance (DDD.MM coordinates)
DIMENSION: 6
EDGE_WEIGHT_TYPE: GEO
EDGE_WEIGHT_FORMAT: FUNCTION
DISPLAY_DATA_TYPE: COORD_DISPLAY
NODE_COORD_SECTION
1 16.47 96.10
2 16.47 94.44
3 20.09 92.54
4 22.39 93.37
5 25.23 97.24
6 -22.00 -96.05
EOF
""",
    "euc_float5.tsp": """NAME: f5
TYPE: TSP
DIMENSION: 5
EDGE_WEIGHT_TYPE: EUC_2D
NODE_COORD_SECTION
1 0.5 0.25
2 3.75 4.0
3 10.0 0.0
4 6.1 8.2
5 1e2 2.5e1
EOF
""",
}


def gen_pm1():
    rng = np.random.default_rng(1234)
    n = 40
    lines = []
    for i in range(1, n):
        for j in range(0, i):
            if rng.random() < 0.15:
                lines.append(f"{i + 1} {j + 1} {1 if rng.random() < 0.5 else -1}")
    return "%%MatrixMarket matrix coordinate integer symmetric\n" + f"{n} {n} {len(lines)}\n" + "\n".join(lines) + "\n"


def run_ref(mode, path):
    out = subprocess.run([REF, mode, path], check=True, capture_output=True, text=True).stdout
    return out[out.index("BEGIN\n") + len("BEGIN\n"):]


def parse_csr(txt):
    lines = txt.split("\n")
    size, s, nnz, nrp = lines[0].split()
    rp = [int(v) for v in lines[1].split()]
    col = [int(v) for v in lines[2].split()]
    val = [float(v) for v in lines[3].split()]
    return dict(size=int(size), sum=float(s), nnz=int(nnz), row_ptr=rp, col=col, val=val)


def main():
    os.makedirs(INP, exist_ok=True)
    MTX["int_sym_pm1.mtx"] = gen_pm1()
    out = {"mtx": {}, "tsp": {}, "dense": {}}
    for name, text in MTX.items():
        p = os.path.join(INP, name)
        with open(p, "w") as f:
            f.write(text)
        out["mtx"][name] = parse_csr(run_ref("mtx", p))
    for name, text in TSP.items():
        p = os.path.join(INP, name)
        with open(p, "w") as f:
            f.write(text)
        txt = run_ref("tsp", p).split("\n")
        n, dtype = txt[0].split()
        xy = [[float(v) for v in ln.split()] for ln in txt[1:1 + int(n)]]
        out["tsp"][name] = dict(n=int(n), dtype=dtype, x=[a for a, _ in xy], y=[b for _, b in xy])
    # Dense one-hot QUBOs (oracle construction) through the reference's
    # CSR(vector<double>, dim) (src/csr.cpp:137-155).
    import oracle as orc
    for name, n, seed, dup in (("tsp4", 4, 7, False), ("tsp5dup", 5, 11, True), ("tsp2", 2, 3, False)):
        rng = np.random.default_rng(seed)
        x = rng.integers(0, 100, n).astype(float)
        y = rng.integers(0, 100, n).astype(float)
        if dup:  # two coincident cities: zero distance entries are skipped
            x[1], y[1] = x[0], y[0]
        d = orc.tsp_distances(x, y, 0, 0)
        D, mx = orc.tsp_totals(n, d)
        J = orc.tsp_qubo_dense(n, d, mx)
        p = os.path.join(INP, name + "_dense.txt")
        with open(p, "w") as f:
            f.write(f"{n * n}\n" + " ".join(repr(float(v)) for v in J.ravel()) + "\n")
        csr = parse_csr(run_ref("dense", p))
        out["dense"][name] = dict(n=n, x=x.tolist(), y=y.tolist(), csr=csr)
        os.remove(p)
    with open(os.path.join(HERE, "ref_readers.json"), "w") as f:
        json.dump(out, f, indent=1, sort_keys=True)
    print("wrote", os.path.join(HERE, "ref_readers.json"))


if __name__ == "__main__":
    main()
