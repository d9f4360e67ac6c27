"""Global assemblers backed by the device (drop-in for Elfel's Assemblers module).

SysmatAssemblerHIP mirrors SysmatAssemblerSparse (src/Assemblers.jl:19-123):
start(a, nrow, ncol), assemble(a, r, c, v), assemble(a, lma),
assemble(a, lma.T) (transpose(lma)), finish(a) -> CSC sparse matrix.  The COO
stream collected on the host is handed to the device, which sorts it stably and
sums duplicates in stream order (the semantics of SparseArrays.sparse).  The
device-integrated weak forms of elfel_hip.integrate leave their matrix on the
device instead; finish(a) then copies it out.

SysvecAssemblerHIP mirrors SysvecAssembler (:170-230).

Errors mirror the reference: an index outside the matrix raises
ElfelRangeError (an IndexError), as sparse() raises ArgumentError and the
vector assembler BoundsError.
"""
import numpy as np

from . import _lib
from .localassemblers import LocalMatrixAssembler, LocalVectorAssembler, TransposedLMA

_default_ctx = None


def default_context():
    """The process-wide device context (device 0)."""
    global _default_ctx
    if _default_ctx is None:
        _default_ctx = _lib.Context(0)
    return _default_ctx


class _Chunks:
    def __init__(self, dtype):
        self.dtype = dtype
        self.parts = []
        self.scalars = []

    def add(self, a):
        if self.scalars:
            self.parts.append(np.asarray(self.scalars, self.dtype))
            self.scalars = []
        # copy: the caller re-initialises its local assembler for the next element
        self.parts.append(np.array(a, dtype=self.dtype, copy=True).ravel())

    def add1(self, v):
        self.scalars.append(v)

    def array(self):
        if self.scalars:
            self.parts.append(np.asarray(self.scalars, self.dtype))
            self.scalars = []
        if not self.parts:
            return np.zeros(0, self.dtype)
        return np.concatenate(self.parts)


class SysmatAssemblerHIP:
    """System-matrix assembler whose sparse build runs on the GPU."""

    def __init__(self, zero=0.0, ctx=None):
        self.zero = zero
        self._ctx = ctx
        self.nrow = self.ncol = 0
        self._reset()

    def _reset(self):
        self._I, self._J, self._V = _Chunks(np.int64), _Chunks(np.int64), _Chunks(np.float64)
        self._device_result = False

    def context(self):
        if self._ctx is None:
            self._ctx = _lib.Context(0)
        return self._ctx

    # ---- Elfel protocol -------------------------------------------------------
    def start(self, nrow, ncol):
        self.nrow, self.ncol = int(nrow), int(ncol)
        self._reset()
        return self

    def assemble(self, *args):
        if len(args) == 3:
            r, c, v = args
            self._I.add1(int(r)); self._J.add1(int(c)); self._V.add1(float(v))
        elif len(args) == 1 and isinstance(args[0], LocalMatrixAssembler):
            lma = args[0]
            self._I.add(lma.row.ravel(order="F"))
            self._J.add(lma.col.ravel(order="F"))
            self._V.add(lma.M.ravel(order="F"))
        elif len(args) == 1 and isinstance(args[0], TransposedLMA):
            p = args[0].parent
            self._I.add(p.col.ravel(order="C"))
            self._J.add(p.row.ravel(order="C"))
            self._V.add(p.M.ravel(order="C"))
        else:
            raise TypeError("assemble(a, r, c, v), assemble(a, lma) or assemble(a, lma.T)")
        return self

    def coo(self):
        return self._I.array(), self._J.array(), self._V.array()

    def finish(self, as_scipy=True):
        ctx = self.context()
        if not self._device_result:
            I, J, V = self.coo()
            ctx.sparse(I, J, V, self.nrow, self.ncol, _lib.CSC)
        ptr, ind, val, shape, lay = ctx.copy_sparse(base=0)
        if not as_scipy:
            return ptr, ind, val, shape, lay
        import scipy.sparse as sp
        if lay == _lib.CSC:
            return sp.csc_matrix((val, ind, ptr), shape=shape)
        return sp.csr_matrix((val, ind, ptr), shape=shape)


class SysvecAssemblerHIP:
    """System-vector assembler whose scatter-add runs on the GPU."""

    def __init__(self, zero=0.0, ctx=None):
        self.zero = zero
        self._ctx = ctx
        self.ndofs = 0
        self._reset()

    def _reset(self):
        self._I, self._V = _Chunks(np.int64), _Chunks(np.float64)
        self._device_src = None

    def context(self):
        if self._ctx is None:
            self._ctx = _lib.Context(0)
        return self._ctx

    def start(self, ndofs):
        self.ndofs = int(ndofs)
        self._reset()
        return self

    def assemble(self, *args):
        if len(args) == 2:
            self._I.add1(int(args[0])); self._V.add1(float(args[1]))
        elif len(args) == 1 and isinstance(args[0], LocalVectorAssembler):
            self._I.add(args[0].row)
            self._V.add(args[0].V)
        else:
            raise TypeError("assemble(v, i, val) or assemble(v, lva)")
        return self

    def finish(self):
        if self._device_src is not None:
            ctx, n = self._device_src
            return ctx.copy_vector(n)
        return self.context().vector(self._I.array(), self._V.array(), self.ndofs)


# ---- the reference's generic-function protocol ------------------------------
def start(a, *dims):
    return a.start(*dims)


def assemble(a, *args):
    return a.assemble(*args)


def finish(a):
    return a.finish()


# Aliases of start / finish under the names assemble_init, makematrix and
# makevector; the same functions.
assemble_init = start
makematrix = finish
makevector = finish
