"""Finite element spaces and dof numbering (mirror of Elfel's FEFields/FESpaces).

Host-side: the numbering runs once per problem and defines the sparsity
pattern the device assembles.  Semantics follow the reference:

* dofs live on the entities the element names (FEData, src/FElements.jl):
  vertices for the nodal elements, vertices and the cell for
  FEH1_T3_BUBBLE, the cell only for FEL2_Q4;
* numberfreedofs!: free dofs numbered term-major, component-minor from
  `firstnum`, data dofs zeroed (src/FEFields.jl:137-154), field by field
  (src/FESpaces.jl:141-155);
* numberdatadofs!: data dofs numbered after, same order (:164-177);
* numberdofs!(spaces): free dofs of each space in turn, each starting one past
  the highest free number of the previous space, then the data dofs likewise
  (src/FESpaces.jl:250-259);
* element dof order: the nodes' dofs, then the cell's, each entity's
  components in order (src/FESpaces.jl:87-105, src/FEIterators.jl:161-171);
* gathersysvec!/scattersysvec! (src/FEFields.jl:247-279).

The reference visits a space's fields in the key order of a Dict keyed by
the entity dimension (FESpace._irsfields, src/FESpaces.jl:76-86), which the
reference tree does not fix; this mirror numbers the vertex field before the
cell field.  That order decides only which numbers the dofs receive: the
assembled entry of every pair of dofs is the same either way.

Entity ids and component numbers passed by callers are 1-based, as in Elfel.
"""
import numpy as np


class _Field:
    """One FEField: dof numbers, data flags and values of nterms entities."""

    def __init__(self, nterms, ncomp, T):
        self.dofnums = np.zeros((nterms, ncomp), np.int64)
        self.isdatum = np.zeros((nterms, ncomp), bool)
        self.dofvals = np.zeros((nterms, ncomp), T)


class FESpace:
    def __init__(self, mesh, fe, nfecopies=1, T=np.float64):
        self.mesh = mesh
        self.fe = fe
        self.nfecopies = int(nfecopies)
        self.T = T
        nc = self.nfecopies
        self._node = _Field(mesh.nnodes if fe.ndofperfeat(0) > 0 else 0, nc, T)
        self._cell = _Field(mesh.nelements if fe.ndofperfeat(fe.mdim) > 0 else 0, nc, T)

    # node-field arrays (the only field of the nodal elements)
    @property
    def dofnums(self):
        return self._node.dofnums

    @dofnums.setter
    def dofnums(self, v):
        self._node.dofnums = v

    @property
    def isdatum(self):
        return self._node.isdatum

    @property
    def dofvals(self):
        return self._node.dofvals

    @dofvals.setter
    def dofvals(self, v):
        self._node.dofvals = v

    # cell-field arrays (FEH1_T3_BUBBLE, FEL2_Q4)
    @property
    def cell_dofnums(self):
        return self._cell.dofnums

    @property
    def cell_dofvals(self):
        return self._cell.dofvals

    def _fields(self):
        return [f for f in (self._node, self._cell) if f.dofnums.shape[0] > 0]

    def _field(self, m):
        if m == 0 and self.fe.ndofperfeat(0) > 0:
            return self._node
        if m == self.fe.mdim and self.fe.ndofperfeat(m) > 0:
            return self._cell
        raise ValueError(f"{self.fe!r} carries no dofs on entities of dimension {m}")

    # ---- element dof metadata (src/FESpaces.jl:87-105) -------------------
    def ndofsperel(self):
        return self.fe.ndofsperel() * self.nfecopies

    def edofmdim(self):
        dims = [0] * self.fe.ngeo * (self.fe.ndofperfeat(0) > 0) + [self.fe.mdim] * (self.fe.ndofperfeat(self.fe.mdim) > 0)
        return np.repeat(np.array(dims, np.int64), self.nfecopies)

    def edofbfnum(self):
        return np.repeat(np.arange(1, self.fe.nbf + 1), self.nfecopies)

    def edofcompnt(self):
        return np.tile(np.arange(1, self.nfecopies + 1), self.fe.nbf)

    def doftype(self):
        return self.T

    # ---- boundary conditions -------------------------------------------------
    def setebc(self, m, eid, comp, val):
        """setebc!(fesp, m, eid, comp, val): eid and comp 1-based; eid may be an array."""
        f = self._field(m)
        eid = np.asarray(eid, np.int64) - 1
        f.isdatum[eid, comp - 1] = True
        f.dofvals[eid, comp - 1] = val
        return self

    def setdofval(self, m, eid, comp, val):
        f = self._field(m)
        eid = np.asarray(eid, np.int64) - 1
        f.dofvals[eid, comp - 1] = val
        return self

    # ---- numbering -----------------------------------------------------------
    def numberfreedofs(self, firstnum=1):
        for f in self._fields():
            free = ~f.isdatum.ravel()
            d = np.zeros(free.size, np.int64)
            d[free] = firstnum + np.arange(int(free.sum()), dtype=np.int64)
            f.dofnums = d.reshape(f.isdatum.shape)
            firstnum += int(free.sum())
        return self

    def numberdatadofs(self, firstnum=0):
        if firstnum == 0:
            firstnum = self.nunknowns() + 1
        for f in self._fields():
            data = f.isdatum.ravel()
            d = f.dofnums.ravel().copy()
            d[data] = firstnum + np.arange(int(data.sum()), dtype=np.int64)
            f.dofnums = d.reshape(f.isdatum.shape)
            firstnum += int(data.sum())
        return self

    def ndofs(self):
        return sum(f.dofnums.size for f in self._fields())

    def nunknowns(self):
        return sum(int((~f.isdatum).sum()) for f in self._fields())

    def highestfreedofnum(self):
        return max([int(f.dofnums[~f.isdatum].max()) for f in self._fields() if (~f.isdatum).any()] or [0])

    def highestdatadofnum(self):
        return max([int(f.dofnums[f.isdatum].max()) for f in self._fields() if f.isdatum.any()] or [0])

    def dofnum(self, m, eid, comp):
        return int(self._field(m).dofnums[eid - 1, comp - 1])

    # ---- system vectors ------------------------------------------------------
    def gathersysvec(self, v):
        for f in self._fields():
            v[f.dofnums.ravel() - 1] = f.dofvals.ravel()
        return v

    def scattersysvec(self, v):
        for f in self._fields():
            d = f.dofnums.ravel()
            ok = d <= len(v)
            vals = f.dofvals.ravel().copy()
            vals[ok] = v[d[ok] - 1]
            f.dofvals = vals.reshape(f.dofvals.shape)
        return self

    # ---- element-level views used by the device path ------------------------
    def element_dofs(self):
        """Element dofs (E x nedof, 1-based): the nodes' dofs, then the cell's."""
        parts = []
        if self._node.dofnums.shape[0]:
            parts.append(self._node.dofnums[self.mesh.conn - 1].reshape(self.mesh.nelements, -1))
        if self._cell.dofnums.shape[0]:
            parts.append(self._cell.dofnums)
        return np.concatenate(parts, 1)


def numberdofs(spaces):
    """numberdofs!(fesp) for one space or a list of spaces (src/FESpaces.jl:250-273)."""
    if isinstance(spaces, FESpace):
        spaces = [spaces]
    spaces[0].numberfreedofs(1)
    for i in range(1, len(spaces)):
        spaces[i].numberfreedofs(spaces[i - 1].highestfreedofnum() + 1)
    spaces[0].numberdatadofs(spaces[-1].highestfreedofnum() + 1)
    for i in range(1, len(spaces)):
        spaces[i].numberdatadofs(spaces[i - 1].highestdatadofnum() + 1)


def gathersysvec(v, spaces):
    for s in (spaces if isinstance(spaces, (list, tuple)) else [spaces]):
        s.gathersysvec(v)
    return v


def scattersysvec(spaces, v):
    for s in (spaces if isinstance(spaces, (list, tuple)) else [spaces]):
        s.scattersysvec(v)
    return spaces


def ndofs(spaces):
    return sum(s.ndofs() for s in (spaces if isinstance(spaces, (list, tuple)) else [spaces]))


def nunknowns(spaces):
    return sum(s.nunknowns() for s in (spaces if isinstance(spaces, (list, tuple)) else [spaces]))
