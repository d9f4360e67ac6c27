"""Finite element spaces and dof numbering (mirror of Elfel's FEFields/FESpaces).

Host-side: the numbering runs once per problem and defines the sparsity
pattern the device assembles.  Semantics follow the reference exactly:

* dofs live on vertices only for T3/T6/Q4 (ndofperfeat = [1, 0, 0, 0]);
* numberfreedofs!: free dofs numbered term-major, component-minor from
  `firstnum`, data dofs zeroed (src/FEFields.jl:137-154);
* numberdatadofs!: data dofs numbered after, same order (:164-177);
* numberdofs!(spaces): free dofs of each space in turn, each starting one past
  the highest free number of the previous space, then the data dofs likewise
  (src/FESpaces.jl:250-259);
* element dof order: entity-major, component-minor (src/FESpaces.jl:87-105,
  src/FEIterators.jl:161-171);
* gathersysvec!/scattersysvec! (src/FEFields.jl:247-279).

Entity ids and component numbers passed by callers are 1-based, as in Elfel.
"""
import numpy as np


class FESpace:
    def __init__(self, mesh, fe, nfecopies=1, T=np.float64):
        self.mesh = mesh
        self.fe = fe
        self.nfecopies = int(nfecopies)
        self.T = T
        n = mesh.nnodes
        self.dofnums = np.zeros((n, self.nfecopies), np.int64)
        self.isdatum = np.zeros((n, self.nfecopies), bool)
        self.dofvals = np.zeros((n, self.nfecopies), T)

    # ---- element dof metadata (src/FESpaces.jl:87-105) -------------------
    def ndofsperel(self):
        return self.fe.ndofsperel() * self.nfecopies

    def edofmdim(self):
        return np.zeros(self.ndofsperel(), np.int64)

    def edofbfnum(self):
        return np.repeat(np.arange(1, self.fe.nbf + 1), self.nfecopies)

    def edofcompnt(self):
        return np.tile(np.arange(1, self.nfecopies + 1), self.fe.nbf)

    def doftype(self):
        return self.T

    # ---- boundary conditions -------------------------------------------------
    def setebc(self, m, eid, comp, val):
        """setebc!(fesp, m, eid, comp, val): eid and comp 1-based; eid may be an array."""
        if m != 0:
            raise ValueError("T3/T6/Q4 spaces carry dofs on vertices only (m = 0)")
        eid = np.asarray(eid, np.int64) - 1
        self.isdatum[eid, comp - 1] = True
        self.dofvals[eid, comp - 1] = val
        return self

    def setdofval(self, m, eid, comp, val):
        eid = np.asarray(eid, np.int64) - 1
        self.dofvals[eid, comp - 1] = val
        return self

    # ---- numbering -----------------------------------------------------------
    def numberfreedofs(self, firstnum=1):
        free = ~self.isdatum.ravel()
        d = np.zeros(free.size, np.int64)
        d[free] = firstnum + np.arange(int(free.sum()), dtype=np.int64)
        self.dofnums = d.reshape(self.isdatum.shape)
        return self

    def numberdatadofs(self, firstnum=0):
        if firstnum == 0:
            firstnum = self.nunknowns() + 1
        data = self.isdatum.ravel()
        d = self.dofnums.ravel().copy()
        d[data] = firstnum + np.arange(int(data.sum()), dtype=np.int64)
        self.dofnums = d.reshape(self.isdatum.shape)
        return self

    def ndofs(self):
        return self.dofnums.size

    def nunknowns(self):
        return int((~self.isdatum).sum())

    def highestfreedofnum(self):
        m = ~self.isdatum
        return int(self.dofnums[m].max()) if m.any() else 0

    def highestdatadofnum(self):
        m = self.isdatum
        return int(self.dofnums[m].max()) if m.any() else 0

    def dofnum(self, m, eid, comp):
        return int(self.dofnums[eid - 1, comp - 1])

    # ---- system vectors ------------------------------------------------------
    def gathersysvec(self, v):
        v[self.dofnums.ravel() - 1] = self.dofvals.ravel()
        return v

    def scattersysvec(self, v):
        d = self.dofnums.ravel()
        ok = d <= len(v)
        vals = self.dofvals.ravel().copy()
        vals[ok] = v[d[ok] - 1]
        self.dofvals = vals.reshape(self.dofvals.shape)
        return self

    # ---- element-level views used by the device path ------------------------
    def element_dofs(self):
        """Element dofs (E x nedof, 1-based), entity-major, component-minor."""
        return self.dofnums[self.mesh.conn - 1].reshape(self.mesh.nelements, -1)


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
