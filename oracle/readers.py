"""Restatement of the reference's input readers.  TEST INFRASTRUCTURE ONLY.

Only tests/ (and bench.py's cpu_baseline leg) import this module, as the
checker.  It follows the reference's stream semantics token by token so that
fixtures written by the reference's own compiled readers (oracle/_ref) pin it:

* ``read_mtx``  -- src/csr.cpp:15-109 (banner scan, comment skip, diagonal
  drop, ``sum``, symmetric mirror) and src/csr.cpp:157-192 (lexicographic
  tuple sort, row pointers).  ``reference_rowptr=True`` reproduces the
  reference's row-pointer construction, which loses pointers across empty
  rows (src/csr.cpp:171-188); the default builds the correct CSR.
* ``csr_from_dense`` -- src/csr.cpp:137-155 (row-major scan, zeros skipped,
  diagonal kept).
* ``read_tsp``  -- src/tsp_cpu.cpp:9-45 (line-start peeks for 'D', 'E', '1',
  GEO degrees.minutes to radians).
"""
import numpy as np


class _Stream:
    """Minimal std::istream emulation over a text buffer."""

    def __init__(self, text):
        self.t = text
        self.p = 0

    def peek(self):
        return self.t[self.p] if self.p < len(self.t) else None

    def ignore_line(self, limit=2048):
        n = 0
        while self.p < len(self.t) and n < limit:
            c = self.t[self.p]
            self.p += 1
            n += 1
            if c == "\n":
                return

    def token(self):
        while self.p < len(self.t) and self.t[self.p].isspace():
            self.p += 1
        if self.p >= len(self.t):
            raise EOFError("unexpected end of input")
        s = self.p
        while self.p < len(self.t) and not self.t[self.p].isspace():
            self.p += 1
        return self.t[s:self.p]


def _build_csr(rows, cols, vals, n, reference_rowptr):
    order = sorted(range(len(vals)), key=lambda e: (rows[e], cols[e], vals[e]))
    r = [rows[e] for e in order]
    c = np.array([cols[e] for e in order], dtype=np.int32)
    v = np.array([vals[e] for e in order], dtype=np.float64)
    if reference_rowptr:
        rp = [0]
        prev = 0
        for e in range(len(r)):
            if r[e] != prev:
                rp.append(e)
                prev = r[e]
        rp.append(len(r))
        rp = np.array(rp, dtype=np.int64)
    else:
        rp = np.zeros(n + 1, dtype=np.int64)
        for x in r:
            rp[x + 1] += 1
        rp = np.cumsum(rp).astype(np.int64)
    return rp, c, v


def read_mtx(path, reference_rowptr=False):
    """Returns dict(size, sum, nnz, row_ptr, col, val, symmetric, entries)."""
    with open(path, "r") as f:
        s = _Stream(f.read())
    symmetric, pattern = True, False
    while True:
        w = s.token()
        if w == "symmetric":
            symmetric = True
            break
        if w == "pattern":
            pattern = True
        elif w == "integer":
            pattern = False
        elif w == "general":
            symmetric = False
            break
    s.ignore_line()
    while s.peek() == "%":
        s.ignore_line()
    M, _N, L = int(s.token()), int(s.token()), int(s.token())
    rows, cols, vals = [], [], []
    total = 0.0
    for _ in range(L):
        r, c = int(s.token()), int(s.token())
        d = 1.0 if pattern else float(s.token())
        if r != c:
            rows.append(r - 1)
            cols.append(c - 1)
            vals.append(d)
            total += d
    if symmetric:
        for e in range(len(vals)):
            if cols[e] != rows[e]:
                rows.append(cols[e])
                cols.append(rows[e])
                vals.append(vals[e])
    rp, c, v = _build_csr(rows, cols, vals, M, reference_rowptr)
    return dict(size=M, sum=total, nnz=len(v), row_ptr=rp, col=c, val=v, symmetric=symmetric)


def csr_from_dense(dense, dim):
    dense = np.asarray(dense, dtype=np.float64).reshape(dim, dim)
    rows, cols, vals = [], [], []
    for i in range(dim):
        for j in range(dim):
            if dense[i, j] != 0:
                rows.append(i)
                cols.append(j)
                vals.append(float(dense[i, j]))
    rp, c, v = _build_csr(rows, cols, vals, dim, reference_rowptr=True)
    return dict(size=dim, sum=0.0, nnz=len(v), row_ptr=rp, col=c, val=v)


def read_tsp(path):
    """Returns dict(n, dtype, x, y) with GEO coordinates converted to radians."""
    with open(path, "r") as f:
        s = _Stream(f.read())
    while s.peek() != "D":
        if s.peek() is None:
            raise EOFError("no DIMENSION line")
        s.ignore_line()
    s.token()
    n = int(s.token())
    while s.peek() != "E":
        if s.peek() is None:
            raise EOFError("no EDGE_WEIGHT_TYPE line")
        s.ignore_line()
    s.token()
    dtype = s.token()
    while s.peek() != "1":
        if s.peek() is None:
            raise EOFError("no node section")
        s.ignore_line()
    xs, ys = [], []
    for _ in range(n):
        _idx, a, b = float(s.token()), float(s.token()), float(s.token())
        if dtype == "GEO":
            PI = 3.14159265358979323846
            ad, bd = int(a), int(b)
            am, bm = a - ad, b - bd
            a = PI * (ad + 5.0 * am / 3.0) / 180.0
            b = PI * (bd + 5.0 * bm / 3.0) / 180.0
        xs.append(a)
        ys.append(b)
    return dict(n=n, dtype=dtype, x=np.array(xs), y=np.array(ys))


def dense_from_csr(rp, col, val, n):
    J = np.zeros((n, n))
    for i in range(n):
        for k in range(rp[i], rp[i + 1]):
            J[i, col[k]] += val[k]
    return J

