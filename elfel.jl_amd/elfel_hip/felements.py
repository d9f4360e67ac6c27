"""Scalar finite elements (mirror of Elfel's FElements module).

Basis functions and parametric gradients follow src/FElements.jl:239-246
(T3), :264-288 (T6: vertices, then midsides of 1-2, 2-3, 3-1), :306-320
(Q4), :203-225 (L2), :322-354 (T3 with a cubic bubble on the cell), :356-384
(T4) and :386-413 (FEL2_Q4: one constant function on the cell).  The nodal
elements are their own geometry carriers (:196); FEL2_Q4's is the nodal Q4
(:393).  ndofperfeat(m) gives the dofs on each entity of dimension m (the
FEData vectors of the constructors).
"""
import numpy as np

from . import _lib


class FE:
    code = None
    name = None
    refshape = None
    nbf = 0
    _perfeat = (1, 0, 0, 0)

    @property
    def ngeo(self):
        """Nodes of the geometry carrier (the element's own nodes)."""
        return self._nfeat[0]

    @property
    def mdim(self):
        return {"interval": 1, "triangle": 2, "square": 2, "tetrahedron": 3}[self.refshape]

    def ndofsperel(self):
        return self.nbf

    def nfeatofdim(self, m):
        return self._nfeat[m]

    def ndofperfeat(self, m):
        return self._perfeat[m]

    def __repr__(self):
        return f"FEH1_{self.name}()"


class FEH1_T3(FE):
    code, name, refshape, nbf = _lib.T3, "T3", "triangle", 3
    _nfeat = (3, 3, 1, 0)

    def bfun(self, pc):
        r, s = pc
        return np.array([(1 - r - s), r, s])

    def bfungradpar(self, pc):
        return np.array([[-1., -1.], [+1., 0.], [0., +1.]])


class FEH1_T6(FE):
    code, name, refshape, nbf = _lib.T6, "T6", "triangle", 6
    _nfeat = (6, 3, 1, 0)

    def bfun(self, pc):
        r, s = pc
        t = 1. - r - s
        return np.array([t * (t + t - 1), r * (r + r - 1), s * (s + s - 1), 4 * r * t, 4 * r * s, 4 * s * t])

    def bfungradpar(self, pc):
        r, s = pc
        return np.array([[-3 + 4 * r + 4 * s, -3 + 4 * r + 4 * s], [4 * r - 1, 0.0], [0.0, 4 * s - 1],
                         [4 - 8 * r - 4 * s, -4 * r], [4 * s, 4 * r], [-4 * s, 4 - 4 * r - 8 * s]])


class FEH1_Q4(FE):
    code, name, refshape, nbf = _lib.Q4, "Q4", "square", 4
    _nfeat = (4, 4, 1, 0)

    def bfun(self, pc):
        r, s = pc
        return np.array([0.25 * (1. - r) * (1. - s), 0.25 * (1. + r) * (1. - s),
                         0.25 * (1. + r) * (1. + s), 0.25 * (1. - r) * (1. + s)])

    def bfungradpar(self, pc):
        r, s = pc
        return np.array([[-(1. - s) * 0.25, -(1. - r) * 0.25], [(1. - s) * 0.25, -(1. + r) * 0.25],
                         [(1. + s) * 0.25, (1. + r) * 0.25], [-(1. + s) * 0.25, (1. - r) * 0.25]])


class FEH1_L2(FE):
    code, name, refshape, nbf = _lib.L2, "L2", "interval", 2
    _nfeat = (2, 1, 0, 0)

    def bfun(self, pc):
        r = pc[0]
        return np.array([(1. - r), (1. + r)]) / 2.0

    def bfungradpar(self, pc):
        return np.array([[-1.0], [+1.0]]) / 2.0


class FEH1_T3_BUBBLE(FE):
    code, name, refshape, nbf = _lib.T3B, "T3_BUBBLE", "triangle", 4
    _nfeat = (3, 3, 1, 0)
    _perfeat = (1, 0, 1, 0)

    def bfun(self, pc):
        r, s = pc
        return np.array([(1 - r - s), r, s, (1 - r - s) * r * s])

    def bfungradpar(self, pc):
        r, s = pc
        return np.array([[-1., -1.], [+1., 0.], [0., +1.],
                         [(-r * s + (1 - r - s) * s), (-r * s + (1 - r - s) * r)]])


class FEH1_T4(FE):
    code, name, refshape, nbf = _lib.T4, "T4", "tetrahedron", 4
    _nfeat = (4, 6, 4, 1)

    def bfun(self, pc):
        r, s, t = pc
        return np.array([(1 - r - s - t), r, s, t])

    def bfungradpar(self, pc):
        return np.array([[-1.0, -1.0, -1.0], [+1.0, 0.0, 0.0], [0.0, +1.0, 0.0], [0.0, 0.0, +1.0]])


class FEL2_Q4(FE):
    code, name, refshape, nbf = _lib.Q4L2, "Q4", "square", 1
    _nfeat = (4, 4, 1, 0)
    _perfeat = (0, 0, 1, 0)

    def bfun(self, pc):
        return np.array([1.0])

    def bfungradpar(self, pc):
        return np.array([[0.0, 0.0]])

    def __repr__(self):
        return "FEL2_Q4()"
