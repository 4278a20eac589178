"""ctypes front end of liboracle.so (sis_oracle.c).  TEST INFRASTRUCTURE ONLY.

Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg load this
module, and only as the checker (or, for cpu_baseline, as the timed CPU
port).  See sis_oracle.c for the semantics and the reference lines followed.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
MAXCUT, TSP = 0, 1
DSB, BSB = 0, 1

_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = C.CDLL(LIB_PATH)
        _lib.orc_uniform.restype = C.c_double
        _lib.orc_uniform.argtypes = [C.c_uint64, C.c_uint64]
        _lib.orc_xi0_maxcut.restype = C.c_double
        _lib.orc_xi0_maxcut.argtypes = [C.c_int64, C.c_int64, C.c_void_p]
        _lib.orc_bf16.restype = C.c_float
        _lib.orc_bf16.argtypes = [C.c_float]
        _lib.orc_solve_f32.restype = C.c_int64
        _lib.orc_solve_f64.restype = C.c_int64
        _lib.orc_window_maxcut_f32.restype = C.c_int64
        _lib.orc_tsp_tour_length.restype = C.c_double
    return _lib


def set_threads(n):
    """Threads of the oracle's OpenMP loops."""
    lib().orc_set_threads(int(n))


def max_threads():
    return int(lib().orc_max_threads())


def _p(a):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def xi0_maxcut(n, val):
    val = np.ascontiguousarray(val, np.float64)
    return lib().orc_xi0_maxcut(C.c_int64(n), C.c_int64(len(val)), _p(val))


def uniform(seed, k):
    return lib().orc_uniform(C.c_uint64(seed), C.c_uint64(k))


def bf16(x):
    return lib().orc_bf16(C.c_float(x))


def init_state(seed, problem, n, agents, f64=False):
    dt = np.float64 if f64 else np.float32
    X = np.zeros((n, agents), dt)
    Y = np.zeros((n, agents), dt)
    fn = lib().orc_init_state_f64 if f64 else lib().orc_init_state_f32
    fn(C.c_uint64(seed), C.c_int(problem), C.c_int64(n), C.c_int64(agents), _p(X), _p(Y))
    return X, Y


def sched(problem, t, T, dt, xi0):
    c = np.zeros(4)
    lib().orc_sched(C.c_int(problem), C.c_int64(t), C.c_int64(T), C.c_double(dt), C.c_double(xi0), _p(c))
    return c


class Problem:
    """CSR (+ field) in the dtypes the oracle entry points expect."""

    def __init__(self, rp, col, val, field=None, S=0.0, ncities=0):
        self.rp = np.ascontiguousarray(rp, np.int64)
        self.col = np.ascontiguousarray(col, np.int32)
        self.val64 = np.ascontiguousarray(val, np.float64)
        self.val32 = self.val64.astype(np.float32)
        self.n = len(self.rp) - 1
        self.h64 = None if field is None else np.ascontiguousarray(field, np.float64)
        self.h32 = None if field is None else self.h64.astype(np.float32)
        self.S = S
        self.ncities = ncities


def advance(pb, problem, gather, X, Y, t0, nsteps, T, dt, xi0, coupling, precision="f32"):
    """Steps t0..t0+nsteps-1 in place on X, Y (no pre-step, no windows)."""
    n, A = X.shape
    if precision == "f64":
        lib().orc_advance_f64(C.c_int(problem), C.c_int(gather), _p(pb.rp), _p(pb.col), _p(pb.val64),
                              _p(pb.h64), C.c_int64(n), C.c_int64(A), C.c_int64(t0),
                              C.c_int64(nsteps), C.c_int64(T), C.c_double(dt), C.c_double(xi0),
                              C.c_double(coupling), _p(X), _p(Y))
    else:
        prec = 1 if precision == "bf16" else 0
        lib().orc_advance_f32(C.c_int(problem), C.c_int(gather), C.c_int(prec), _p(pb.rp),
                              _p(pb.col), _p(pb.val32), _p(pb.h32), C.c_int64(n), C.c_int64(A),
                              C.c_int64(t0), C.c_int64(nsteps), C.c_int64(T), C.c_double(dt),
                              C.c_double(xi0), C.c_double(coupling), _p(X), _p(Y))


def step(pb, problem, gather, Xs, Y, c, coupling=1.0, precision="f32"):
    """One step with explicit coefficients c[0..3]; returns the new X (Y in place)."""
    n, A = Xs.shape
    Xn = np.zeros_like(Xs)
    c = np.ascontiguousarray(c, np.float64)
    if precision == "f64":
        lib().orc_step_f64(C.c_int(problem), C.c_int(gather), _p(pb.rp), _p(pb.col), _p(pb.val64),
                           _p(pb.h64), C.c_int64(n), C.c_int64(A), _p(Xs), _p(Xn), _p(Y), _p(c),
                           C.c_double(coupling))
    else:
        lib().orc_step_f32(C.c_int(problem), C.c_int(gather), C.c_int(precision == "bf16"),
                           _p(pb.rp), _p(pb.col), _p(pb.val32), _p(pb.h32), C.c_int64(n),
                           C.c_int64(A), _p(Xs), _p(Xn), _p(Y), _p(c), C.c_double(coupling))
    return Xn


def prestep(X, Y, dt):
    fn = lib().orc_prestep_f64 if X.dtype == np.float64 else lib().orc_prestep_f32
    fn(C.c_int64(X.size), C.c_double(dt), _p(X), _p(Y))


MAXCUT_UNFUSED = 2  # problem code of the four-kernel MaxCut path


def solve(pb, problem, gather, X, Y, T, uw, P, K=25, dt=0.1, xi0=0.0, coupling=1.0,
          precision="f32"):
    """Whole solve in place on X, Y.  Returns dict(steps, cur, fin, stab, bif, valid)."""
    n, A = X.shape
    cur = np.full((n, A), -1, np.int32)
    fin = np.full((n, A), -1, np.int32)
    stab = np.zeros(A, np.int32)
    bif = np.zeros(A, np.uint8)
    valid = np.zeros(A, np.int32)
    if precision == "f64":
        steps = lib().orc_solve_f64(C.c_int(problem), C.c_int(gather), _p(pb.rp), _p(pb.col),
                                    _p(pb.val64), _p(pb.h64), C.c_int64(n), C.c_int64(A),
                                    C.c_int64(T), C.c_int(uw), C.c_int64(P), C.c_int(K),
                                    C.c_int64(pb.ncities), C.c_double(dt), C.c_double(xi0),
                                    C.c_double(coupling), _p(X), _p(Y), _p(cur), _p(fin),
                                    _p(stab), _p(bif), _p(valid))
    else:
        prec = 1 if precision == "bf16" else 0
        steps = lib().orc_solve_f32(C.c_int(problem), C.c_int(gather), C.c_int(prec), _p(pb.rp),
                                    _p(pb.col), _p(pb.val32), _p(pb.h32), C.c_int64(n),
                                    C.c_int64(A), C.c_int64(T), C.c_int(uw), C.c_int64(P),
                                    C.c_int(K), C.c_int64(pb.ncities), C.c_double(dt),
                                    C.c_double(xi0), C.c_double(coupling), _p(X), _p(Y), _p(cur),
                                    _p(fin), _p(stab), _p(bif), _p(valid))
    return dict(steps=steps, cur=cur, fin=fin, stab=stab, bif=bif.astype(np.int32), valid=valid)


def cut(pb, spins):
    spins = np.ascontiguousarray(spins, np.int32)
    n, A = spins.shape
    out = np.zeros(A)
    lib().orc_cut(_p(pb.rp), _p(pb.col), _p(pb.val64), C.c_int64(n), C.c_int64(A), _p(spins),
                  C.c_double(pb.S), _p(out))
    return out


def energy_field(pb, spins):
    spins = np.ascontiguousarray(spins, np.int32)
    n, A = spins.shape
    out = np.zeros(A)
    lib().orc_energy_field(_p(pb.rp), _p(pb.col), _p(pb.val64), _p(pb.h64), C.c_int64(n),
                           C.c_int64(A), _p(spins), _p(out))
    return out


def cut_stats(cuts):
    cuts = np.ascontiguousarray(cuts, np.float64)
    out = np.zeros(5)
    lib().orc_cut_stats(_p(cuts), C.c_int64(len(cuts)), _p(out))
    return out


def window_maxcut(X, K, cur, fin, stab, bif):
    n, A = X.shape
    return lib().orc_window_maxcut_f32(_p(X), C.c_int64(n), C.c_int64(A), C.c_int(K), _p(cur),
                                       _p(fin), _p(stab), _p(bif))


def tsp_distances(x, y, kind=0, euc_round=0):
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    n = len(x)
    d = np.zeros(max(n * (n - 1) // 2, 1))
    lib().orc_tsp_distances(C.c_int64(n), _p(x), _p(y), C.c_int(kind), C.c_int(euc_round), _p(d))
    return d[: n * (n - 1) // 2]


def tsp_totals(n, d):
    D = np.zeros(n)
    mx = C.c_double()
    lib().orc_tsp_totals(C.c_int64(n), _p(np.ascontiguousarray(d, np.float64)), _p(D), C.byref(mx))
    return D, mx.value


def tsp_qubo_dense(n, d, maxd):
    J = np.zeros((n * n, n * n))
    lib().orc_tsp_qubo_dense(C.c_int64(n), _p(np.ascontiguousarray(d, np.float64)),
                             C.c_double(maxd), _p(J))
    return J


def tsp_field(n, D, maxd):
    h = np.zeros(n * n)
    lib().orc_tsp_field(C.c_int64(n), _p(np.ascontiguousarray(D, np.float64)), C.c_double(maxd), _p(h))
    return h


def tsp_tour_length(n, d, spins, agent):
    spins = np.ascontiguousarray(spins, np.int32)
    tour = np.zeros(n, np.int32)
    v = lib().orc_tsp_tour_length(C.c_int64(n), _p(np.ascontiguousarray(d, np.float64)),
                                  _p(spins), C.c_int64(spins.shape[1]), C.c_int64(agent), _p(tour))
    return v, tour


def tri(c1, c2, n):
    return c2 + c1 * n - ((c1 + 2) * (c1 + 1)) // 2


# ---------------------------------------------------------------- narrow-state dSB forms
# Code width w: 12 (X codes +-1024 and the wall code +-1025, Y codes +-2047),
# 10 (X codes +-256, wall +-257, Y codes +-511) or 8 (X codes +-64, wall +-65,
# Y codes +-127).  The 8-bit form also comes with stochastic Y rounding
# (sr=True, SIS_FORM_Q8S): Y rounds to floor(y 2^sy + u), u a hashed dither
# of (step, node, agent).

def qn_shift(dt, w):
    """Y code shift sy of the w-bit form for time step dt (-1 if dt <= 0)."""
    lib().orc_qn_shift.restype = C.c_int
    return int(lib().orc_qn_shift(C.c_int(w), C.c_double(dt)))


def qn_round(X, Y, sy, w):
    """Rounds float32 X, Y in place to the w-bit codes (decoded values)."""
    assert X.dtype == np.float32 and Y.dtype == np.float32
    lib().orc_qn_round_state(C.c_int(w), C.c_int64(X.size), _p(X), _p(Y), C.c_int(sy))


def step_qn(pb, Xs, Y, c, sy, w, sr=False, t=0):
    """One narrow-state step from the snapshot Xs: returns the new X, updates Y.
    sr (8-bit only): Y rounded stochastically with the dither of step index t."""
    n, A = Xs.shape
    Xn = np.zeros_like(Xs)
    cc = np.ascontiguousarray(c, np.float64)
    if sr:
        assert w == 8
        lib().orc_step_q8s(_p(pb.rp), _p(pb.col), _p(pb.val32), C.c_int64(n), C.c_int64(A),
                           _p(Xs), _p(Xn), _p(Y), _p(cc), C.c_int(sy), C.c_int64(t))
    else:
        lib().orc_step_qn(C.c_int(w), _p(pb.rp), _p(pb.col), _p(pb.val32), C.c_int64(n),
                          C.c_int64(A), _p(Xs), _p(Xn), _p(Y), _p(cc), C.c_int(sy))
    return Xn


def advance_qn(pb, X, Y, t0, nsteps, T, dt, xi0, sy, w, sr=False):
    """Narrow-state dSB MaxCut steps t0..t0+nsteps-1 in place (decoded codes);
    sr: the 8-bit form with stochastic Y rounding."""
    n, A = X.shape
    args = (_p(pb.rp), _p(pb.col), _p(pb.val32), C.c_int64(n), C.c_int64(A), C.c_int64(t0),
            C.c_int64(nsteps), C.c_int64(T), C.c_double(dt), C.c_double(xi0), C.c_int(sy),
            _p(X), _p(Y))
    if sr:
        assert w == 8
        lib().orc_advance_q8s(*args)
    else:
        lib().orc_advance_qn(C.c_int(w), *args)


def prestep_qn(X, Y, dt, sy, w):
    lib().orc_prestep_qn(C.c_int(w), C.c_int64(X.size), C.c_double(dt), C.c_int(sy), _p(X), _p(Y))


def solve_qn(pb, X, Y, T, uw, P, K=25, dt=0.5, xi0=0.0, sy=None, w=12, sr=False):
    """Whole narrow-state dSB MaxCut solve in place on float32 X, Y (rounded to
    the codes first).  sr: the 8-bit form with stochastic Y rounding.
    Returns dict(steps, cur, fin, stab, bif)."""
    n, A = X.shape
    sy = qn_shift(dt, w) if sy is None else sy
    cur = np.full((n, A), -1, np.int32)
    fin = np.full((n, A), -1, np.int32)
    stab = np.zeros(A, np.int32)
    bif = np.zeros(A, np.uint8)
    args = (_p(pb.rp), _p(pb.col), _p(pb.val32), C.c_int64(n), C.c_int64(A), C.c_int64(T),
            C.c_int(uw), C.c_int64(P), C.c_int(K), C.c_double(dt), C.c_double(xi0), C.c_int(sy),
            _p(X), _p(Y), _p(cur), _p(fin), _p(stab), _p(bif))
    if sr:
        assert w == 8
        lib().orc_solve_q8s.restype = C.c_int64
        steps = lib().orc_solve_q8s(*args)
    else:
        lib().orc_solve_qn.restype = C.c_int64
        steps = lib().orc_solve_qn(C.c_int(w), *args)
    return dict(steps=steps, cur=cur, fin=fin, stab=stab, bif=bif.astype(np.int32))


def q12_shift(dt):
    return qn_shift(dt, 12)


def q12_round(X, Y, sy):
    qn_round(X, Y, sy, 12)


def advance_q12(pb, X, Y, t0, nsteps, T, dt, xi0, sy):
    advance_qn(pb, X, Y, t0, nsteps, T, dt, xi0, sy, 12)


def prestep_q12(X, Y, dt, sy):
    prestep_qn(X, Y, dt, sy, 12)


def solve_q12(pb, X, Y, T, uw, P, K=25, dt=0.5, xi0=0.0, sy=None):
    return solve_qn(pb, X, Y, T, uw, P, K, dt, xi0, sy, 12)
