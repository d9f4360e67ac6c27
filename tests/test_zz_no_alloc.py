"""elfel_hip_sparse_noalloc (the allocation-free COO -> CSR/CSC build) against
elfel_hip_sparse, the entry point it is a variant of: the same pattern and
values bit for bit (tolerance 0) on two calls in a row, the second smaller
than the first, through the C ABI.  The device arrays are allocated with the
HIP runtime the library itself links (libamdhip64 through ctypes), so the
test process holds one HIP runtime."""
import ctypes

import numpy as np
import pytest

from elfel_hip import _lib as L

pytestmark = pytest.mark.gpu

H2D, D2H = 1, 2


class Hip:
    """hipMalloc / hipMemcpy / hipFree of the runtime libelfel_hip.so uses."""

    def __init__(self):
        self.rt = ctypes.CDLL("libamdhip64.so.7", mode=ctypes.RTLD_GLOBAL)
        self.rt.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
        self.rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
        self.rt.hipFree.argtypes = [ctypes.c_void_p]
        self.live = []

    def alloc(self, nbytes):
        p = ctypes.c_void_p()
        assert self.rt.hipMalloc(ctypes.byref(p), max(nbytes, 16)) == 0
        self.live.append(p.value)
        return p.value

    def put(self, a):
        a = np.ascontiguousarray(a)
        d = self.alloc(a.nbytes)
        assert self.rt.hipMemcpy(d, a.ctypes.data, a.nbytes, H2D) == 0
        return d

    def get(self, d, n, dtype):
        out = np.empty(n, dtype)
        if n:
            assert self.rt.hipMemcpy(out.ctypes.data, d, out.nbytes, D2H) == 0
        return out

    def free(self):
        for p in self.live:
            self.rt.hipFree(p)
        self.live = []


@pytest.fixture
def hip():
    h = Hip()
    yield h
    h.free()


def _coo(n, nrow, ncol, seed):
    rng = np.random.default_rng(seed)
    # few distinct entries, so that many duplicates are summed; explicit zeros
    I = rng.integers(1, nrow + 1, n)
    J = rng.integers(1, ncol + 1, n)
    I[: n // 3] = I[n // 3: 2 * (n // 3)]
    J[: n // 3] = J[n // 3: 2 * (n // 3)]
    V = rng.standard_normal(n)
    V[::17] = 0.0
    return I.astype(np.int64), J.astype(np.int64), V


def _noalloc(ctx, hip, I, J, V, nrow, ncol, layout):
    n = I.size
    nmaj = ncol if layout == L.CSC else nrow
    dI, dJ, dV = hip.put(I), hip.put(J), hip.put(V)
    ptr = hip.put(np.full(nmaj + 1, -7, np.int32))
    ind = hip.put(np.full(max(n, 1), -7, np.int32))
    out = hip.put(np.full(max(n, 1), np.nan))
    nnz = ctx.sparse_noalloc(nrow, ncol, n, dI, dJ, dV, layout, ptr, ind, out)
    return (nnz, hip.get(ptr, nmaj + 1, np.int32).astype(np.int64), hip.get(ind, nnz, np.int32).astype(np.int64),
            hip.get(out, nnz, np.float64))


def _reference(ctx, I, J, V, nrow, ncol, layout):
    nnz = ctx.sparse(I, J, V, nrow, ncol, layout=layout)
    ptr, ind, val = ctx.copy_sparse(base=0)[:3]
    return nnz, ptr, ind, val


@pytest.mark.parametrize("layout", [L.CSC, L.CSR])
def test_noalloc_equals_sparse_twice_second_smaller(layout, hip):
    ctx = L.Context(0)
    for n, nrow, ncol, seed in ((900_000, 3001, 2503, 1), (40_000, 700, 901, 2)):
        I, J, V = _coo(n, nrow, ncol, seed)
        got = _noalloc(ctx, hip, I, J, V, nrow, ncol, layout)
        ref = _reference(ctx, I, J, V, nrow, ncol, layout)
        assert got[0] == ref[0]
        for a, b, what in zip(got[1:], ref[1:], ("pointers", "indices", "values")):
            assert a.shape == b.shape and np.array_equal(a, b), (n, what)
    ctx.close()


def test_noalloc_limits_and_errors(hip):
    ctx = L.Context(0)
    # empty stream: every pointer 0
    ptr = hip.put(np.full(6, -1, np.int32))
    assert ctx.sparse_noalloc(5, 5, 0, 0, 0, 0, L.CSR, ptr, 0, 0) == 0
    assert not hip.get(ptr, 6, np.int32).any()
    # beyond the scratch limit: refused on the host before any launch
    with pytest.raises(L.ElfelHipError) as e:
        ctx.sparse_noalloc(5, 5, L.NOALLOC_MAX_COO + 1, ptr, ptr, ptr, L.CSR, ptr, ptr, ptr)
    assert e.value.code == L.EINVAL
    # an index out of range (the key kernel clamps it and raises the flag)
    I = np.array([1, 2, 6], np.int64)
    J = np.array([1, 2, 3], np.int64)
    V = np.ones(3)
    with pytest.raises(L.ElfelRangeError):
        _noalloc(ctx, hip, I, J, V, 5, 5, L.CSR)
    # the context keeps working afterwards
    got = _noalloc(ctx, hip, I[:2], J[:2], V[:2], 5, 5, L.CSR)
    assert got[0] == 2 and np.array_equal(got[2], [0, 1])
    ctx.close()
