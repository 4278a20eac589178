"""ctypes binding of libsis.so (include/sis.h) for tests and benchmarks.

The library is loaded from ``SIS_LIB`` or from the in-tree build
(``sparse-ising-solver_amd/build/libsis.so``).  There is no fallback: if the
library is missing or a call fails, an exception is raised.
"""
import ctypes as C
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.environ.get("SIS_LIB", os.path.join(_HERE, "..", "build", "libsis.so"))

MAXCUT, TSP = 0, 1
DSB, BSB = 0, 1
FORM_F32, FORM_F64, FORM_PACKED, FORM_BF16, FORM_Q12, FORM_Q8, FORM_Q10, FORM_Q8S = range(8)
W_ONE, W_PM1, W_INT3 = 0, 1, 2
EUC_2D, GEO = 0, 1
SPINS_CURRENT, SPINS_FINAL = 0, 1
FORMS = {"f32": FORM_F32, "f64": FORM_F64, "packed": FORM_PACKED, "bf16": FORM_BF16,
         "q12": FORM_Q12, "q10": FORM_Q10, "q8": FORM_Q8, "q8s": FORM_Q8S}


class SisError(RuntimeError):
    pass


class Params(C.Structure):
    _fields_ = [
        ("problem", C.c_int32), ("dynamics", C.c_int32), ("form", C.c_int32),
        ("agents", C.c_int32), ("dt", C.c_double), ("xi0", C.c_double),
        ("coupling", C.c_double), ("window_period", C.c_int32),
        ("window_threshold", C.c_int32), ("ncities", C.c_int32),
        ("use_graph", C.c_int32), ("tile", C.c_int32), ("fused", C.c_int32),
        ("split", C.c_int32), ("structured", C.c_int32), ("row_order", C.c_int32),
    ]


class RunInfo(C.Structure):
    _fields_ = [
        ("steps_executed", C.c_int64), ("windows", C.c_int64),
        ("all_bifurcated", C.c_int32), ("reserved", C.c_int32),
        ("elapsed_us", C.c_double), ("device_ms", C.c_double),
        ("step_kernel_ms", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_ if k != "reserved"}


class BatchParams(C.Structure):
    _fields_ = [
        ("agents_max", C.c_int32), ("max_instances", C.c_int32), ("max_rows", C.c_int64),
        ("max_nnz", C.c_int64), ("chunk_steps", C.c_int32), ("tile", C.c_int32),
        ("eager", C.c_int32),
    ]


class BatchStatus(C.Structure):
    _fields_ = [
        ("state", C.c_int32), ("agents", C.c_int32), ("n", C.c_int64),
        ("steps_executed", C.c_int64), ("best_agent", C.c_int64), ("best_cut", C.c_double),
        ("submit_s", C.c_double), ("admit_s", C.c_double), ("done_s", C.c_double),
    ]

    def as_dict(self):
        return {k: getattr(self, k) for k, _ in self._fields_}


_P = C.c_void_p
_I64P = C.POINTER(C.c_int64)
_I32P = C.POINTER(C.c_int32)
_DP = C.POINTER(C.c_double)

# name -> (restype, argtypes); restype None means int status.
_SIGS = {
    "sis_last_error": (C.c_char_p, []),
    "sis_csr_read_mtx": (None, [C.c_char_p, C.POINTER(_P)]),
    "sis_csr_from_coo": (None, [C.c_int64, C.c_int64, _I32P, _I32P, _DP, C.c_int, C.POINTER(_P)]),
    "sis_csr_from_dense": (None, [C.c_int64, _DP, C.POINTER(_P)]),
    "sis_csr_shape": (None, [_P, _I64P, _I64P, _DP]),
    "sis_csr_export": (None, [_P, _I64P, _I32P, _DP]),
    "sis_csr_write_mtx": (None, [_P, C.c_char_p]),
    "sis_csr_free": ("void", [_P]),
    "sis_maxcut_xi0": (None, [_P, _DP]),
    "sis_maxcut_random_cut": (None, [_P, C.c_uint64, _DP]),
    "sis_cut_stats": (None, [_DP, C.c_int64, _DP]),
    "sis_tsp_read": (None, [C.c_char_p, C.c_int, C.POINTER(_P)]),
    "sis_tsp_from_coords": (None, [C.c_int64, _DP, _DP, C.c_int, C.c_int, C.POINTER(_P)]),
    "sis_tsp_info": (None, [_P, _I64P, _DP, _DP]),
    "sis_tsp_kind_of": (None, [_P, C.POINTER(C.c_int)]),
    "sis_tsp_export": (None, [_P, _DP, _DP, _DP, _DP]),
    "sis_tsp_qubo": (None, [_P, C.POINTER(_P), _DP]),
    "sis_tsp_tour_length": (None, [_P, _I32P, _DP]),
    "sis_tsp_xi0": (None, [_P, _DP]),
    "sis_tsp_random_length": (None, [_P, C.c_uint64, _DP]),
    "sis_tsp_write": (None, [_P, C.c_char_p]),
    "sis_tsp_free": ("void", [_P]),
    "sis_gen_er": (None, [C.c_int64, C.c_double, C.c_int, C.c_uint64, C.POINTER(_P)]),
    "sis_gen_torus": (None, [C.c_int64, C.c_int64, C.c_int, C.c_uint64, C.POINTER(_P)]),
    "sis_gen_regular": (None, [C.c_int64, C.c_int, C.c_int, C.c_uint64, C.POINTER(_P)]),
    "sis_gen_chung_lu": (None, [C.c_int64, C.c_double, C.c_double, C.c_int, C.c_uint64,
                                C.POINTER(_P)]),
    "sis_gen_tsp": (None, [C.c_int64, C.c_int64, C.c_int, C.c_uint64, C.POINTER(_P)]),
    "sis_params_default": (None, [C.c_int, C.POINTER(Params)]),
    "sis_solver_create": (None, [_P, _DP, C.POINTER(Params), C.POINTER(_P)]),
    "sis_solver_create_tsp": (None, [_P, C.POINTER(Params), C.POINTER(_P)]),
    "sis_solver_destroy": (None, [_P]),
    "sis_solver_params": (None, [_P, C.POINTER(Params)]),
    "sis_solver_init_random": (None, [_P, C.c_uint64]),
    "sis_solver_set_state": (None, [_P, _DP, _DP]),
    "sis_solver_get_state": (None, [_P, _DP, _DP]),
    "sis_solver_run": (None, [_P, C.c_int64, C.c_int, C.POINTER(RunInfo)]),
    "sis_solver_prestep": (None, [_P]),
    "sis_solver_autotune": (None, [_P, C.c_int64, _I32P, _DP]),
    "sis_solver_advance": (None, [_P, C.c_int64, C.c_int64, C.c_int64]),
    "sis_solver_get_spins": (None, [_P, C.c_int, _I32P]),
    "sis_solver_maxcut_cuts": (None, [_P, _DP, _I64P]),
    "sis_maxcut_cuts_noalloc": (None, [_P, _P, _P, C.c_int64, _P, C.c_int64, C.c_double, _P, _P,
                                       _P]),
    "sis_solver_energy": (None, [_P, C.c_int, _DP]),
    "sis_solver_agent_state": (None, [_P, _I32P, _I32P, _I32P]),
    "sis_solver_tsp_lengths": (None, [_P, _P, _DP, _I64P]),
    "sis_batch_create": (None, [C.POINTER(BatchParams), C.POINTER(_P)]),
    "sis_batch_destroy": (None, [_P]),
    "sis_batch_submit": (None, [_P, _P, C.POINTER(Params), C.c_int64, C.c_int, C.c_uint64, _I64P]),
    "sis_batch_advance": (None, [_P, C.c_int64, _I32P, _I32P]),
    "sis_batch_poll": (None, [_P, C.c_int64, C.POINTER(BatchStatus)]),
    "sis_batch_result": (None, [_P, C.c_int64, _DP, _I32P]),
    "sis_batch_release": (None, [_P, C.c_int64]),
    "sis_batch_autotune": (None, [_P, C.c_int64, _I32P, _DP]),
}


def load(path=None):
    lib = C.CDLL(path or LIB_PATH)
    for name, (res, args) in _SIGS.items():
        fn = getattr(lib, name)
        fn.argtypes = args
        fn.restype = None if res == "void" else (C.c_int if res is None else res)
    return lib


_lib = None


def lib():
    global _lib
    if _lib is None:
        _lib = load()
    return _lib


def _ck(status, what):
    if status != 0:
        msg = lib().sis_last_error().decode(errors="replace")
        raise SisError(f"{what} failed with status {status}: {msg}")


def _ptr(a, ctype):
    return a.ctypes.data_as(C.POINTER(ctype)) if a is not None else None


class Csr:
    def __init__(self, handle):
        self.h = C.c_void_p(handle) if not isinstance(handle, C.c_void_p) else handle
        n, nnz, s = C.c_int64(), C.c_int64(), C.c_double()
        _ck(lib().sis_csr_shape(self.h, C.byref(n), C.byref(nnz), C.byref(s)), "sis_csr_shape")
        self.n, self.nnz, self.sum = n.value, nnz.value, s.value

    @classmethod
    def read_mtx(cls, path):
        h = C.c_void_p()
        _ck(lib().sis_csr_read_mtx(str(path).encode(), C.byref(h)), f"reading {path}")
        return cls(h)

    @classmethod
    def from_coo(cls, n, rows, cols, vals, mirror=True):
        rows = np.ascontiguousarray(rows, dtype=np.int32)
        cols = np.ascontiguousarray(cols, dtype=np.int32)
        vals = np.ascontiguousarray(vals, dtype=np.float64)
        h = C.c_void_p()
        _ck(lib().sis_csr_from_coo(n, len(vals), _ptr(rows, C.c_int32), _ptr(cols, C.c_int32),
                                   _ptr(vals, C.c_double), int(mirror), C.byref(h)), "sis_csr_from_coo")
        return cls(h)

    @classmethod
    def from_dense(cls, dense):
        dense = np.ascontiguousarray(dense, dtype=np.float64)
        h = C.c_void_p()
        _ck(lib().sis_csr_from_dense(dense.shape[0], _ptr(dense, C.c_double), C.byref(h)),
            "sis_csr_from_dense")
        return cls(h)

    @classmethod
    def gen_er(cls, n, p, weights, seed):
        h = C.c_void_p()
        _ck(lib().sis_gen_er(n, p, weights, seed, C.byref(h)), "sis_gen_er")
        return cls(h)

    @classmethod
    def gen_torus(cls, rows, cols, weights, seed):
        h = C.c_void_p()
        _ck(lib().sis_gen_torus(rows, cols, weights, seed, C.byref(h)), "sis_gen_torus")
        return cls(h)

    @classmethod
    def gen_regular(cls, n, d, weights, seed):
        h = C.c_void_p()
        _ck(lib().sis_gen_regular(n, d, weights, seed, C.byref(h)), "sis_gen_regular")
        return cls(h)

    @classmethod
    def gen_chung_lu(cls, n, avg_degree, exponent, weights, seed):
        h = C.c_void_p()
        _ck(lib().sis_gen_chung_lu(n, avg_degree, exponent, weights, seed, C.byref(h)),
            "sis_gen_chung_lu")
        return cls(h)

    def arrays(self):
        rp = np.zeros(self.n + 1, np.int64)
        col = np.zeros(max(self.nnz, 1), np.int32)
        val = np.zeros(max(self.nnz, 1), np.float64)
        _ck(lib().sis_csr_export(self.h, _ptr(rp, C.c_int64), _ptr(col, C.c_int32),
                                 _ptr(val, C.c_double)), "sis_csr_export")
        return rp, col[: self.nnz], val[: self.nnz]

    def write_mtx(self, path):
        _ck(lib().sis_csr_write_mtx(self.h, str(path).encode()), "sis_csr_write_mtx")

    def xi0(self):
        v = C.c_double()
        _ck(lib().sis_maxcut_xi0(self.h, C.byref(v)), "sis_maxcut_xi0")
        return v.value

    def random_cut(self, seed):
        v = C.c_double()
        _ck(lib().sis_maxcut_random_cut(self.h, seed, C.byref(v)), "sis_maxcut_random_cut")
        return v.value

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.sis_csr_free(self.h)
            self.h = None


class Tsp:
    def __init__(self, handle):
        self.h = handle
        n, mx, sm = C.c_int64(), C.c_double(), C.c_double()
        _ck(lib().sis_tsp_info(self.h, C.byref(n), C.byref(mx), C.byref(sm)), "sis_tsp_info")
        self.n, self.maxd, self.sumd = n.value, mx.value, sm.value

    @classmethod
    def read(cls, path, euc_round=0):
        h = C.c_void_p()
        _ck(lib().sis_tsp_read(str(path).encode(), euc_round, C.byref(h)), f"reading {path}")
        return cls(h)

    @classmethod
    def from_coords(cls, x, y, kind=EUC_2D, euc_round=0):
        x = np.ascontiguousarray(x, dtype=np.float64)
        y = np.ascontiguousarray(y, dtype=np.float64)
        h = C.c_void_p()
        _ck(lib().sis_tsp_from_coords(len(x), _ptr(x, C.c_double), _ptr(y, C.c_double), kind,
                                      euc_round, C.byref(h)), "sis_tsp_from_coords")
        return cls(h)

    @classmethod
    def generate(cls, n, span, euc_round, seed):
        h = C.c_void_p()
        _ck(lib().sis_gen_tsp(n, span, euc_round, seed, C.byref(h)), "sis_gen_tsp")
        return cls(h)

    def kind(self):
        k = C.c_int()
        _ck(lib().sis_tsp_kind_of(self.h, C.byref(k)), "sis_tsp_kind_of")
        return k.value

    def arrays(self):
        n = self.n
        x, y, t = np.zeros(n), np.zeros(n), np.zeros(n)
        d = np.zeros(max(n * (n - 1) // 2, 1))
        _ck(lib().sis_tsp_export(self.h, _ptr(x, C.c_double), _ptr(y, C.c_double),
                                 _ptr(d, C.c_double), _ptr(t, C.c_double)), "sis_tsp_export")
        return x, y, d[: n * (n - 1) // 2], t

    def qubo(self):
        h = C.c_void_p()
        field = np.zeros(self.n * self.n)
        _ck(lib().sis_tsp_qubo(self.h, C.byref(h), _ptr(field, C.c_double)), "sis_tsp_qubo")
        return Csr(h), field

    def tour_length(self, tour):
        tour = np.ascontiguousarray(tour, dtype=np.int32)
        v = C.c_double()
        _ck(lib().sis_tsp_tour_length(self.h, _ptr(tour, C.c_int32), C.byref(v)),
            "sis_tsp_tour_length")
        return v.value

    def xi0(self):
        """The reference's alternative coupling scale 0.5 / (sigma sqrt N)."""
        v = C.c_double()
        _ck(lib().sis_tsp_xi0(self.h, C.byref(v)), "sis_tsp_xi0")
        return v.value

    def random_length(self, seed):
        v = C.c_double()
        _ck(lib().sis_tsp_random_length(self.h, seed, C.byref(v)), "sis_tsp_random_length")
        return v.value

    def write(self, path):
        _ck(lib().sis_tsp_write(self.h, str(path).encode()), "sis_tsp_write")

    def __del__(self):
        if getattr(self, "h", None) and _lib is not None:
            _lib.sis_tsp_free(self.h)
            self.h = None


def cut_stats(cut):
    cut = np.ascontiguousarray(cut, dtype=np.float64)
    out = np.zeros(5)
    _ck(lib().sis_cut_stats(_ptr(cut, C.c_double), len(cut), _ptr(out, C.c_double)), "sis_cut_stats")
    return out


def default_params(problem):
    p = Params()
    _ck(lib().sis_params_default(problem, C.byref(p)), "sis_params_default")
    return p


def _params(params):
    if isinstance(params, dict):
        p = default_params(params.get("problem", MAXCUT))
        for k, v in params.items():
            setattr(p, k, v)
        return p
    return params


class Solver:
    """A device solver.  ``params`` is a Params or a dict of Params fields."""

    def __init__(self, J, params, field=None):
        params = _params(params)
        self.J = J
        self.field = None if field is None else np.ascontiguousarray(field, dtype=np.float64)
        h = C.c_void_p()
        _ck(lib().sis_solver_create(J.h, _ptr(self.field, C.c_double), C.byref(params),
                                    C.byref(h)), "sis_solver_create")
        self._attach(h, J.n)

    @classmethod
    def from_tsp(cls, tsp, params):
        """Solver for a TSP instance (sis_solver_create_tsp); params may set
        structured=1 to take J x from the QUBO's structure (no matrix)."""
        params = _params(dict(params, problem=TSP) if isinstance(params, dict) else params)
        h = C.c_void_p()
        _ck(lib().sis_solver_create_tsp(tsp.h, C.byref(params), C.byref(h)), "sis_solver_create_tsp")
        self = cls.__new__(cls)
        self.J, self.field, self.tsp = None, None, tsp
        self._attach(h, tsp.n * tsp.n)
        return self

    def _attach(self, h, N):
        self.h = h
        self.p = Params()
        _ck(lib().sis_solver_params(self.h, C.byref(self.p)), "sis_solver_params")
        self.N, self.A = N, self.p.agents

    def init_random(self, seed):
        _ck(lib().sis_solver_init_random(self.h, seed), "sis_solver_init_random")

    def set_state(self, x, y):
        x = np.ascontiguousarray(x, dtype=np.float64).reshape(self.N, self.A)
        y = np.ascontiguousarray(y, dtype=np.float64).reshape(self.N, self.A)
        _ck(lib().sis_solver_set_state(self.h, _ptr(x, C.c_double), _ptr(y, C.c_double)),
            "sis_solver_set_state")

    def get_state(self):
        x = np.zeros((self.N, self.A))
        y = np.zeros((self.N, self.A))
        _ck(lib().sis_solver_get_state(self.h, _ptr(x, C.c_double), _ptr(y, C.c_double)),
            "sis_solver_get_state")
        return x, y

    def prestep(self):
        _ck(lib().sis_solver_prestep(self.h), "sis_solver_prestep")

    def autotune(self, trial_steps=10):
        tile, ms = C.c_int32(), C.c_double()
        _ck(lib().sis_solver_autotune(self.h, trial_steps, C.byref(tile), C.byref(ms)),
            "sis_solver_autotune")
        _ck(lib().sis_solver_params(self.h, C.byref(self.p)), "sis_solver_params")
        return tile.value, ms.value

    def advance(self, t0, nsteps, T):
        _ck(lib().sis_solver_advance(self.h, t0, nsteps, T), "sis_solver_advance")

    def run(self, max_steps, update_window=0):
        info = RunInfo()
        _ck(lib().sis_solver_run(self.h, max_steps, update_window, C.byref(info)), "sis_solver_run")
        return info.as_dict()

    def spins(self, which=SPINS_CURRENT):
        out = np.zeros((self.N, self.A), np.int32)
        _ck(lib().sis_solver_get_spins(self.h, which, _ptr(out, C.c_int32)), "sis_solver_get_spins")
        return out

    def cuts(self):
        out = np.zeros(self.A)
        best = C.c_int64()
        _ck(lib().sis_solver_maxcut_cuts(self.h, _ptr(out, C.c_double), C.byref(best)),
            "sis_solver_maxcut_cuts")
        return out, best.value

    def energy(self, which=SPINS_CURRENT):
        out = np.zeros(self.A)
        _ck(lib().sis_solver_energy(self.h, which, _ptr(out, C.c_double)), "sis_solver_energy")
        return out

    def agent_state(self):
        st = np.zeros(self.A, np.int32)
        lk = np.zeros(self.A, np.int32)
        vs = np.zeros(self.A, np.int32)
        _ck(lib().sis_solver_agent_state(self.h, _ptr(st, C.c_int32), _ptr(lk, C.c_int32),
                                         _ptr(vs, C.c_int32)), "sis_solver_agent_state")
        return st, lk, vs

    def tsp_lengths(self, tsp):
        out = np.zeros(self.A)
        best = C.c_int64()
        _ck(lib().sis_solver_tsp_lengths(self.h, tsp.h, _ptr(out, C.c_double), C.byref(best)),
            "sis_solver_tsp_lengths")
        return out, best.value

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.sis_solver_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()


class Batch:
    """Continuous-batching engine for many small MaxCut dSB solves."""

    def __init__(self, agents_max=64, max_instances=1024, max_rows=1 << 20, max_nnz=1 << 23,
                 chunk_steps=100, tile=0, eager=False):
        bp = BatchParams(agents_max, max_instances, max_rows, max_nnz, chunk_steps, tile, int(eager))
        h = C.c_void_p()
        _ck(lib().sis_batch_create(C.byref(bp), C.byref(h)), "sis_batch_create")
        self.h = h
        self.agents = {}
        self.sizes = {}

    def submit(self, J, agents, dt, max_steps, update_window=1, seed=1, xi0=0.0, window_period=100):
        p = default_params(MAXCUT)
        p.agents, p.dt, p.xi0, p.window_period, p.form = agents, dt, xi0, window_period, FORM_PACKED
        t = C.c_int64()
        _ck(lib().sis_batch_submit(self.h, J.h, C.byref(p), max_steps, update_window, seed,
                                   C.byref(t)), "sis_batch_submit")
        self.agents[t.value] = agents
        self.sizes[t.value] = J.n
        return t.value

    def advance(self, steps):
        r, q = C.c_int32(), C.c_int32()
        _ck(lib().sis_batch_advance(self.h, steps, C.byref(r), C.byref(q)), "sis_batch_advance")
        return r.value, q.value

    def poll(self, ticket):
        s = BatchStatus()
        _ck(lib().sis_batch_poll(self.h, ticket, C.byref(s)), "sis_batch_poll")
        return s.as_dict()

    def result(self, ticket, spins=False):
        A, n = self.agents[ticket], self.sizes[ticket]
        cuts = np.zeros(A)
        sp = np.zeros((n, A), np.int32) if spins else None
        _ck(lib().sis_batch_result(self.h, ticket, _ptr(cuts, C.c_double), _ptr(sp, C.c_int32)),
            "sis_batch_result")
        return (cuts, sp) if spins else cuts

    def release(self, ticket):
        _ck(lib().sis_batch_release(self.h, ticket), "sis_batch_release")

    def autotune(self, trial_steps=20):
        tile, ms = C.c_int32(), C.c_double()
        _ck(lib().sis_batch_autotune(self.h, trial_steps, C.byref(tile), C.byref(ms)),
            "sis_batch_autotune")
        return tile.value, ms.value

    def close(self):
        if getattr(self, "h", None) is not None and _lib is not None:
            _lib.sis_batch_destroy(self.h)
            self.h = None

    def __del__(self):
        self.close()
