"""Reference shapes and quadrature rules (mirror of Elfel's RefShapes module).

Rules and literal constants follow src/RefShapes.jl: Gauss rules `_gauss1`
(:85-110), triangle rules `_triangle` (:113-230), the square tensor rule with
i outer / j inner (:333-366), tetrahedron rules `_tetrahedron` (:232-258).
Settings mirror the reference's NamedTuples:
quadrature("triangle", kind="default", npts=3), quadrature("square",
kind="Gauss", order=2).
"""
import numpy as np

_GAUSS = {
    1: ([0.0], [2.0]),
    2: ([-0.577350269189626, 0.577350269189626], [1.0, 1.0]),
    3: ([-0.774596669241483, 0.0, 0.774596669241483], [0.5555555555555556, 0.8888888888888889, 0.5555555555555556]),
    4: ([-0.86113631159405, -0.33998104358486, 0.33998104358486, 0.86113631159405],
        [0.34785484513745, 0.65214515486255, 0.65214515486255, 0.34785484513745]),
    5: ([-0.906179845938664, -0.538469310105683, 0.000000000000000, 0.538469310105683, 0.906179845938664],
        [0.236926885056189, 0.478628670499367, 0.568888888888889, 0.478628670499367, 0.236926885056189]),
}


class IntegRule:
    def __init__(self, param_coords, weights):
        self.param_coords = np.asarray(param_coords, np.float64)
        self.weights = np.asarray(weights, np.float64)
        self.npts = self.weights.size


def _gauss1(order):
    if order in _GAUSS:
        pc, w = _GAUSS[order]
        return np.array(pc), np.array(w)
    # Golub-Welsch, as the reference's `opengauss` (src/RefShapes.jl:86-89)
    n = np.arange(1, order)
    off = n / np.sqrt(4.0 * n ** 2 - 1)
    vals, vecs = np.linalg.eigh(np.diag(off, 1) + np.diag(off, -1))
    return vals, 2 * vecs[0, :] ** 2


def _triangle(npts):
    if npts == 1:
        return np.array([[1.0 / 3., 1.0 / 3.]]), np.array([1.0 / 2.0])
    if npts == 3:
        return (np.array([[2.0 / 3, 1.0 / 6], [1.0 / 6, 2.0 / 3], [1.0 / 6, 1.0 / 6]]),
                np.array([1.0 / 3, 1.0 / 3, 1.0 / 3]) / 2)
    if npts == 6:
        pc = [[0.816847572980459, 0.091576213509771], [0.091576213509771, 0.816847572980459],
              [0.091576213509771, 0.091576213509771], [0.108103018168070, 0.445948490915965],
              [0.445948490915965, 0.108103018168070], [0.445948490915965, 0.445948490915965]]
        w = np.array([0.109951743655322] * 3 + [0.223381589678011] * 3) / 2
        return np.array(pc), w
    raise ValueError(f"Unknown number of integration points {npts}")


def _tetrahedron(npts):
    if npts == 1:
        return np.array([[0.25, 0.25, 0.25]]), np.array([1.0 / 6.0])
    if npts == 4:
        a, b = 0.13819660, 0.58541020
        return np.array([[a, a, a], [b, a, a], [a, b, a], [a, a, b]]), np.full(4, 0.041666666666666666667)
    if npts == 5:
        a, b, c, d, e = 1.0 / 6.0, 0.25, 0.5, -0.8, 0.45
        return np.array([[b, b, b], [c, a, a], [a, c, a], [a, a, c], [a, a, a]]), np.array([d, e, e, e, e]) / 6
    raise ValueError(f"Unknown number of integration points {npts}")


def quadrature(refshape, kind="default", npts=None, order=None):
    if refshape == "interval":
        if kind not in ("default", "Gauss"):
            raise ValueError(f"Integration rule {kind} not available")
        pc, w = _gauss1(order or 1)
        return IntegRule(pc.reshape(-1, 1), w)
    if refshape == "triangle":
        if kind != "default":
            raise ValueError(f"Integration rule {kind} not available")
        pc, w = _triangle(npts or 1)
        return IntegRule(pc, w)
    if refshape == "square":
        if kind not in ("default", "Gauss"):
            raise ValueError(f"Integration rule {kind} not available")
        o = order or 1
        p1, w1 = _gauss1(o)
        pc = np.array([[p1[i], p1[j]] for i in range(o) for j in range(o)])
        w = np.array([w1[i] * w1[j] for i in range(o) for j in range(o)])
        return IntegRule(pc, w)
    if refshape == "tetrahedron":
        if kind != "default":
            raise ValueError(f"Integration rule {kind} not available")
        pc, w = _tetrahedron(npts or 1)
        return IntegRule(pc, w)
    raise ValueError(f"unknown reference shape {refshape}")


def rule_id(refshape, settings):
    """Device rule id: triangle, tetrahedron -> number of points; square,
    interval -> Gauss order."""
    if refshape in ("triangle", "tetrahedron"):
        return int(settings.get("npts", 1))
    return int(settings.get("order", 1))
