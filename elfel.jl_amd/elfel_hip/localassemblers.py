"""Element-level (local) assemblers (mirror of src/LocalAssemblers.jl).

LocalMatrixAssembler keeps a dense element block with the global row and
column dof numbers of every entry; init! fills the index matrices column-major
(k runs i-fastest, src/LocalAssemblers.jl:68-83) and zeroes the block.  Python
indexing is 0-based: ke[i, j] is Elfel's ke[i+1, j+1].
"""
import numpy as np


class LocalMatrixAssembler:
    def __init__(self, nrow, ncol, z=0.0):
        self.row = np.zeros((nrow, ncol), np.int64)
        self.col = np.zeros((nrow, ncol), np.int64)
        self.M = np.full((nrow, ncol), z, dtype=np.float64)

    @property
    def shape(self):
        return self.M.shape

    def __getitem__(self, ij):
        return self.M[ij]

    def __setitem__(self, ij, v):
        self.M[ij] = v

    def init(self, rdofs, cdofs):
        rdofs = np.asarray(rdofs, np.int64).ravel()
        cdofs = np.asarray(cdofs, np.int64).ravel()
        self.row[:, :] = rdofs[:, None]
        self.col[:, :] = cdofs[None, :]
        self.M[:, :] = 0.0
        return self

    @property
    def T(self):
        """transpose(lma): assembled row-major from the parent (src/Assemblers.jl:109-114)."""
        return TransposedLMA(self)


class TransposedLMA:
    def __init__(self, parent):
        self.parent = parent


class LocalVectorAssembler:
    def __init__(self, nrow, z=0.0):
        self.row = np.zeros(nrow, np.int64)
        self.V = np.full(nrow, z, dtype=np.float64)

    def __getitem__(self, i):
        return self.V[i]

    def __setitem__(self, i, v):
        self.V[i] = v

    def init(self, rdofs):
        self.row[:] = np.asarray(rdofs, np.int64).ravel()
        self.V[:] = 0.0
        return self


def init(a, *dofs):
    """init!(ke, rdofs, cdofs) / init!(fe, rdofs)."""
    return a.init(*dofs)
