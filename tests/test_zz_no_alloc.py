"""The allocation-free cut entry point (sis_maxcut_cuts_noalloc) against the
allocating one (sis_solver_maxcut_cuts) on the same spins: identical cuts and
best agent, bit for bit, on two calls in a row (the second on a smaller
problem, so the shared scratch array holds the first call's partials), plus
the size limit.  Device buffers come from the HIP runtime the library itself
links (hipMalloc / hipMemcpy through ctypes)."""
import ctypes as C

import numpy as np
import pytest

import sis

pytestmark = pytest.mark.gpu

_hip = None


def hip():
    global _hip
    if _hip is None:
        sis.lib()  # loads libsis.so and with it its HIP runtime
        _hip = C.CDLL("libamdhip64.so.7")
        _hip.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
        _hip.hipFree.argtypes = [C.c_void_p]
        _hip.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        _hip.hipDeviceSynchronize.argtypes = []
    return _hip


class DevArray:
    """A device copy of a numpy array (hipMemcpy kinds: 1 host->device, 2 device->host)."""

    def __init__(self, a):
        self.a = np.ascontiguousarray(a)
        self.p = C.c_void_p()
        assert hip().hipMalloc(C.byref(self.p), max(self.a.nbytes, 8)) == 0
        if self.a.nbytes:
            assert hip().hipMemcpy(self.p, self.a.ctypes.data, self.a.nbytes, 1) == 0

    def get(self):
        out = np.empty_like(self.a)
        assert hip().hipDeviceSynchronize() == 0
        assert hip().hipMemcpy(out.ctypes.data, self.p, out.nbytes, 2) == 0
        return out

    def __del__(self):
        if self.p:
            hip().hipFree(self.p)


def pack_spins(spins):
    """N x A int32 +-1 -> N x W uint64 words, bit a % 64 of word a / 64 set for -1."""
    n, A = spins.shape
    W = (A + 63) // 64
    words = np.zeros((n, W), np.uint64)
    neg = spins < 0
    for a in range(A):
        words[:, a // 64] |= neg[:, a].astype(np.uint64) << np.uint64(a % 64)
    return words


def noalloc_cuts(J, spins):
    rp, col, val = J.arrays()
    A = spins.shape[1]
    d_rp, d_col, d_val = DevArray(rp.astype(np.int64)), DevArray(col.astype(np.int32)), \
        DevArray(val.astype(np.float64))
    d_spins = DevArray(pack_spins(spins))
    d_cut, d_best = DevArray(np.full(A, np.nan)), DevArray(np.full(1, -7, np.int64))
    st = sis.lib().sis_maxcut_cuts_noalloc(d_rp.p, d_col.p, d_val.p, J.n, d_spins.p, A, J.sum,
                                           d_cut.p, d_best.p, None)
    assert st == 0, sis.lib().sis_last_error()
    return d_cut.get(), int(d_best.get()[0])


@pytest.mark.parametrize("form", ["packed", "f32"])
def test_noalloc_cuts_equal_solver_cuts_twice(form):
    cases = [(sis.Csr.gen_er(900, 0.02, sis.W_INT3, 3), 200, 2000),   # larger first
             (sis.Csr.gen_torus(12, 12, sis.W_PM1, 4), 70, 500)]       # then smaller
    for J, A, T in cases:
        s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=A, dt=0.5, form=sis.FORMS[form]))
        s.init_random(11)
        s.run(T, 0)
        want, best = s.cuts()
        got, gbest = noalloc_cuts(J, s.spins(sis.SPINS_CURRENT))
        assert np.array_equal(got, want), (J.n, A)
        assert gbest == best
        s.close()


def test_noalloc_real_weights_and_size_limit():
    # real weights: identical summation order, so identical doubles
    rng = np.random.default_rng(5)
    rows = rng.integers(0, 300, 1500)
    cols = rng.integers(0, 300, 1500)
    keep = rows != cols
    J = sis.Csr.from_coo(300, rows[keep], cols[keep], rng.normal(size=int(keep.sum())))
    s = sis.Solver(J, dict(problem=sis.MAXCUT, agents=130, dt=0.3, form=sis.FORM_F32))
    s.init_random(2)
    s.run(700, 0)
    want, best = s.cuts()
    got, gbest = noalloc_cuts(J, s.spins(sis.SPINS_CURRENT))
    assert np.array_equal(got, want) and gbest == best
    # partial sums beyond the scratch array: rejected before any launch
    lib = sis.lib()
    one = DevArray(np.zeros(8, np.int64))
    st = lib.sis_maxcut_cuts_noalloc(one.p, one.p, one.p, 1 << 26, one.p, 4096, 0.0, one.p,
                                     None, None)
    assert st == 1  # SIS_ERR_ARG
    assert b"scratch" in lib.sis_last_error()
