"""Scalar finite elements (mirror of Elfel's FElements module for T3, T6, Q4).

Basis functions and parametric gradients follow src/FElements.jl:239-246
(T3), :264-288 (T6: vertices, then midsides of 1-2, 2-3, 3-1) and :306-320
(Q4).  The geometry carrier of these nodal elements is the element itself
(:196).
"""
import numpy as np

from . import _lib


class FE:
    code = None
    name = None
    refshape = None
    nbf = 0

    def ndofsperel(self):
        return self.nbf

    def nfeatofdim(self, m):
        return self._nfeat[m]

    def ndofperfeat(self, m):
        return 1 if m == 0 else 0

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
