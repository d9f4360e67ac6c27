"""Pins of the oracle's element families (runs without a GPU).

FEH1_L2, FEH1_T4, FEL2_Q4 and FEH1_T3_BUBBLE (src/FElements.jl:203-225,
322-413) against the reference's own tests (test/test_felements.jl,
test/test_fespaces.jl:117-129, test/test_refshapes.jl:96-115), and the weak
forms that use them against closed forms and invariants:
- L2 heat / mass: the textbook 1D element matrices;
- T4 heat: ke = kappa V G G' from the barycentric gradients of an independent
  4x4 solve; the example's exact nodal temperature on a T4 block
  (examples/heat/poisson/t4.jl:82-90);
- the Reddy Stokes form (mu |grad u|^2 + mu (div u)^2, with -p div u) for
  P1b/P1 and Q1/Q0: element blocks against finite-difference gradients and
  numpy's solve, the energy of a linear velocity field, the divergence of a
  linear field against each pressure function, symmetry; the T6/T3 case
  reproduces the pinned three-space Reddy oracle bit for bit.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from helpers import pins

P = pins()
F = P["felements_families"]


# ---- reference pins --------------------------------------------------------------
def test_family_basis_matches_reference_tests():
    assert np.allclose(O.fam_bfun(O.L2, [1 / 3]), F["L2_bfun_third"])
    assert np.allclose(O.fam_bfungradpar(O.L2, [1 / 3]), F["L2_grad"])
    assert np.allclose(O.fam_bfun(O.T4, [0.25, 0.25, 0.25]), F["T4_bfun_quarter"])
    assert np.allclose(O.fam_bfungradpar(O.T4, [0.25, 0.25, 0.25]), F["T4_grad"])
    assert np.allclose(O.fam_bfun(O.Q4L2, [0.25, 0.25]), F["Q4L2_bfun_quarter"])
    assert np.allclose(O.fam_bfungradpar(O.Q4L2, [0.25, 0.25]), F["Q4L2_grad"])
    assert np.allclose(O.fam_bfun(O.T3B, [1 / 3, 1 / 3]), F["T3B_bfun_centroid"])
    assert O.fam_nbf(O.T3B) == F["T3B_ndofsperel"]


def test_tetrahedron_and_interval_rules_match_reference_tests():
    r = P["refshapes_tetrahedron"]
    pc, w = O.fam_quad(O.T4, 1)
    assert np.allclose(pc, r["npts1_pc"]) and np.allclose(w, r["npts1_w"])
    pc, w = O.fam_quad(O.T4, 4)
    assert np.allclose(pc, r["npts4_pc"]) and np.allclose(w, r["npts4_w"])
    pc, w = O.fam_quad(O.L2, 3)
    assert np.allclose(pc[:, 0], P["refshapes"]["gauss3_pc"]) and np.allclose(w, P["refshapes"]["gauss3_w"])
    # the 4-point tetrahedron rule integrates quadratics: int x^2 = 1/60, int xy = 1/120
    pc, w = O.fam_quad(O.T4, 4)
    assert abs(np.sum(w * pc[:, 0] ** 2) - 1 / 60) < 1e-8 and abs(np.sum(w * pc[:, 0] * pc[:, 1]) - 1 / 120) < 1e-8


def _fd_grad(fn, x, h=1e-6):
    x = np.asarray(x, float)
    g = []
    for d in range(x.size):
        e = np.zeros_like(x)
        e[d] = h
        g.append((fn(x + e) - fn(x - e)) / (2 * h))
    return np.stack(g, axis=-1)


@pytest.mark.parametrize("fe,nodes", [
    (O.L2, [[-1.0], [1.0]]),
    (O.T4, [[0, 0, 0], [1, 0, 0], [0, 1, 0], [0, 0, 1]]),
    (O.T3B, [[0, 0], [1, 0], [0, 1]]),
])
def test_family_kronecker_and_gradients(fe, nodes):
    """Nodal functions are Kronecker deltas at the nodes; the bubble vanishes on
    the nodes and edges; every gradient equals central differences of bfun."""
    for a, x in enumerate(nodes):
        N = O.fam_bfun(fe, x)
        assert np.allclose(N[: len(nodes)], np.eye(len(nodes))[a], atol=1e-15)
        if fe == O.T3B:
            assert abs(N[3]) < 1e-15
    rng = np.random.default_rng(3)
    for _ in range(6):
        x = rng.uniform(0.05, 0.3, O.fam_mdim(fe))
        g = O.fam_bfungradpar(fe, x)
        assert np.allclose(g, _fd_grad(lambda y: O.fam_bfun(fe, y), x), atol=1e-8)
    if fe == O.T3B:
        for t in np.linspace(0, 1, 5):  # bubble is zero on the three edges
            for x in ([t, 0], [0, t], [t, 1 - t]):
                assert abs(O.fam_bfun(fe, x)[3]) < 1e-15


# ---- scalar forms -----------------------------------------------------------------
def test_l2_heat_and_mass_closed_forms():
    x = np.array([[0.3], [0.35], [0.9], [1.7]])  # a nonuniform 1D mesh
    conn = np.array([[1, 2], [2, 3], [3, 4]])
    dn = np.arange(1, 5)
    I, J, V, FI, FV = O.fam_scalar_coo(O.HEAT, O.L2, 2, conn, x, dn, kappa=1.5, Q=2.0)
    _, _, M, _, _ = O.fam_scalar_coo(O.MASS, O.L2, 2, conn, x, dn, kappa=10.0)
    for e in range(3):
        h = x[e + 1, 0] - x[e, 0]
        ke = V[4 * e: 4 * e + 4].reshape(2, 2).T
        me = M[4 * e: 4 * e + 4].reshape(2, 2).T
        assert np.allclose(ke, 1.5 / h * np.array([[1, -1], [-1, 1]]), rtol=1e-13)
        assert np.allclose(me, 10.0 * h / 6 * np.array([[2, 1], [1, 2]]), rtol=1e-13)
        assert np.allclose(FV[2 * e: 2 * e + 2], 2.0 * h / 2, rtol=1e-14)
    assert np.array_equal(I[:4], [1, 2, 1, 2]) and np.array_equal(J[:4], [1, 1, 2, 2])


def _tet_closed_form(X):
    M = np.hstack([np.ones((4, 1)), X])
    G = np.linalg.inv(M)[1:, :].T  # row i: gradient of the barycentric coordinate i
    return abs(np.linalg.det(M)) / 6, G


def test_t4_heat_closed_form():
    X = np.array([[0.1, 0.0, 0.2], [1.2, 0.1, 0.0], [0.3, 0.9, 0.1], [0.2, 0.3, 1.1]])
    conn = np.arange(1, 5).reshape(1, 4)
    I, J, V, FI, FV = O.fam_scalar_coo(O.HEAT, O.T4, 1, conn, X, np.arange(1, 5), kappa=2.5, Q=-6.0)
    vol, G = _tet_closed_form(X)
    assert np.allclose(V.reshape(4, 4).T, 2.5 * vol * G @ G.T, rtol=1e-12)
    assert np.allclose(FV, -6.0 * vol / 4, rtol=1e-13)
    _, _, M, _, _ = O.fam_scalar_coo(O.MASS, O.T4, 4, conn, X, np.arange(1, 5), kappa=1.0)
    assert np.allclose(M.reshape(4, 4).T, vol / 20 * (np.ones((4, 4)) + np.eye(4)), rtol=1e-6)


def test_t4block_is_a_conforming_positive_mesh():
    nA, nB, nC = 3, 2, 4
    conn, xyz = O.t4block(1.0, 2.0, 0.5, nA, nB, nC)
    vols = [_tet_closed_form(xyz[c - 1])[0] * np.sign(np.linalg.det(np.hstack([np.ones((4, 1)), xyz[c - 1]])))
            for c in conn]
    assert min(vols) > 0 and abs(sum(vols) - 1.0) < 1e-12
    faces = {}
    for c in conn:
        for drop in range(4):
            f = tuple(sorted(np.delete(c, drop)))
            faces[f] = faces.get(f, 0) + 1
    assert max(faces.values()) == 2
    nbnd = sum(1 for v in faces.values() if v == 1)
    assert nbnd == 4 * (nA * nB + nB * nC + nA * nC)


def _fam_heat_solve(fe, rule, conn, x, bnd, tempf, kappa=1.0, Q=-6.0):
    Nn = x.shape[0]
    free = np.flatnonzero(~bnd)
    dn = np.zeros(Nn, np.int64)
    dn[free] = np.arange(1, free.size + 1)
    dn[bnd] = np.arange(free.size + 1, Nn + 1)
    I, J, V, FI, FV = O.fam_scalar_coo(O.HEAT, fe, rule, conn, x, dn, kappa=kappa, Q=Q)
    K = O.sparse_scipy(I, J, V, Nn, Nn)
    Fv = O.vec_assemble(FI, FV, Nn)
    T = np.zeros(Nn)
    T[dn[bnd] - 1] = tempf(x[bnd])
    nu = free.size
    KT = K @ T
    T[:nu] = spla.spsolve(K[:nu, :nu].tocsc(), Fv[:nu] - KT[:nu])
    return T[dn - 1]


def test_t4_heat_nodal_exactness():
    """examples/heat/poisson/t4.jl: Q = -6, T = 1 + x^2 + 2 y^2 on the boundary,
    the mean nodal error of the solution <= 1e-9."""
    n = 5
    conn, xyz = O.t4block(1.0, 1.0, 1.0, n, n, n)
    bnd = np.any((xyz < 1e-12) | (xyz > 1 - 1e-12), axis=1)
    tempf = lambda X: 1.0 + X[:, 0] ** 2 + 2.0 * X[:, 1] ** 2
    T = _fam_heat_solve(O.T4, 1, conn, xyz, bnd, tempf)
    assert np.mean(np.abs(T - tempf(xyz))) <= P["heat_t4_nodal_mean_error_max"]["value"]


def test_l2_heat_exact_for_quadratic():
    """1D Poisson -T'' = 2 with T = x^2 at both ends: linear elements are nodally exact."""
    conn, x = O.l2block(2.0, 9)
    bnd = np.zeros(x.shape[0], bool)
    bnd[[0, -1]] = True
    T = _fam_heat_solve(O.L2, 2, conn, x, bnd, lambda X: X[:, 0] ** 2, Q=-2.0)
    assert np.allclose(T, x[:, 0] ** 2, atol=1e-12)


@pytest.mark.parametrize("fe,rule", [(O.T3, 1), (O.T3, 3), (O.Q4, 2), (O.T6, 3)])
def test_scalar_heat_equals_pinned_2d_heat(fe, rule):
    """On the 2D nodal elements the family kernel is the pinned heat oracle, bit for bit."""
    gen = {O.T3: O.t3block, O.Q4: O.q4block, O.T6: O.t6block}[fe]
    conn, xy = gen(1.0, 1.3, 4, 3)
    rng = np.random.default_rng(5)
    xy = xy + rng.uniform(-0.03, 0.03, xy.shape)
    dn = np.arange(1, xy.shape[0] + 1).reshape(-1, 1)
    p = O.make_problem(O.HEAT, fe, rule, conn, xy, dn, kappa=1.7, Q=0.3)
    a = O.coo(p)
    b = O.fam_scalar_coo(O.HEAT, fe, rule, conn, xy, dn, kappa=1.7, Q=0.3)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


# ---- Reddy Stokes ------------------------------------------------------------------
def _geo_map(gfe, X):
    return lambda xi: O.fam_bfun(gfe, xi)[: X.shape[0]] @ X


def _reddy_element_fd(vfe, pfe, gfe, Xg, rule, mu):
    """Element blocks from finite-difference parametric gradients, numpy's
    determinant and solve: independent of the oracle's tables and formulas."""
    pc, w = O.fam_quad(vfe, rule)
    nu, np_ = O.fam_nbf(vfe), O.fam_nbf(pfe)
    kxx, kyy, kxy = np.zeros((nu, nu)), np.zeros((nu, nu)), np.zeros((nu, nu))
    kxp, kyp = np.zeros((nu, np_)), np.zeros((nu, np_))
    xmap = _geo_map(gfe, Xg)
    for q in range(len(w)):
        Jm = _fd_grad(xmap, pc[q])                 # J[s, d] = dx_s / dxi_d
        gxi = _fd_grad(lambda y: O.fam_bfun(vfe, y), pc[q])  # [k, d]
        g = np.linalg.solve(Jm.T, gxi.T).T         # grad_x N_k
        JxW = np.linalg.det(Jm) * w[q]
        Np = O.fam_bfun(pfe, pc[q])
        gx, gy = g[:, 0], g[:, 1]
        kxx += mu * JxW * (2 * np.outer(gx, gx) + np.outer(gy, gy))
        kyy += mu * JxW * (np.outer(gx, gx) + 2 * np.outer(gy, gy))
        kxy += mu * JxW * np.outer(gx, gy)
        kxp += -JxW * np.outer(gx, Np)
        kyp += -JxW * np.outer(gy, Np)
    return kxx, kyy, kxy, kxp, kyp


def _one_element_reddy(vfe, pfe, Xg, mu):
    nu, npp = O.fam_nbf(vfe), O.fam_nbf(pfe)
    nv = Xg.shape[0]
    conn = np.arange(1, nv + 1).reshape(1, nv)
    dn_ux, dn_uy = np.arange(1, nv + 1), np.arange(nu + 1, nu + nv + 1)
    kw = {}
    if vfe == O.T3B:
        kw.update(de_ux=[nu], de_uy=[2 * nu])
    if pfe == O.T3:
        kw.update(conn_p=conn, xy_p=Xg, dn_p=np.arange(2 * nu + 1, 2 * nu + 4))
    else:
        kw.update(de_p=[2 * nu + 1])
    rule = 3 if vfe != O.Q4 else 2
    I, J, V = O.fam_reddy_coo(vfe, pfe, rule, conn, Xg, dn_ux, dn_uy, mu=mu, **kw)
    n = 2 * nu + npp
    return O.sparse_scipy(I, J, V, n, n).toarray(), rule


@pytest.mark.parametrize("vfe,pfe", [(O.T3B, O.T3), (O.Q4, O.Q4L2)])
def test_reddy_element_blocks_against_finite_differences(vfe, pfe):
    Xg = {O.T3B: np.array([[0.1, 0.2], [1.3, 0.4], [0.5, 1.7]]),
          O.Q4: np.array([[0.0, 0.1], [1.1, 0.0], [1.3, 0.9], [-0.1, 1.2]])}[vfe]
    mu = 0.7
    K, rule = _one_element_reddy(vfe, pfe, Xg, mu)
    gfe = O.T3 if pfe == O.T3 else vfe
    kxx, kyy, kxy, kxp, kyp = _reddy_element_fd(vfe, pfe, gfe, Xg, rule, mu)
    nu = O.fam_nbf(vfe)
    ux, uy, p = slice(0, nu), slice(nu, 2 * nu), slice(2 * nu, None)
    scale = np.abs(K).max()
    for got, want in ((K[ux, ux], kxx), (K[uy, uy], kyy), (K[ux, uy], kxy), (K[uy, ux], kxy.T),
                      (K[ux, p], kxp), (K[p, ux], kxp.T), (K[uy, p], kyp), (K[p, uy], kyp.T)):
        assert np.abs(got - want).max() <= 1e-7 * scale
    assert np.abs(K[p, p]).max() == 0


def _jittered(gen, N, frac, seed):
    conn, xy = gen(2.0, 2.0, N, N)
    xy = xy - 1.0
    inner = np.all(np.abs(xy) < 1 - 1e-12, axis=1)
    rng = np.random.default_rng(seed)
    xy[inner] += rng.uniform(-frac, frac, (int(inner.sum()), 2)) * (2.0 / N)
    return conn, xy


def _reddy_global(vfe, pfe, N, mu=1.3):
    gen = O.t3block if vfe == O.T3B else O.q4block
    conn, xy = _jittered(gen, N, 0.2, 11)
    E, Nn = conn.shape[0], xy.shape[0]
    dn_ux, dn_uy = np.arange(1, Nn + 1), np.arange(Nn + 1, 2 * Nn + 1)
    o = 2 * Nn
    kw = {}
    if vfe == O.T3B:
        kw.update(de_ux=np.arange(o + 1, o + E + 1), de_uy=np.arange(o + E + 1, o + 2 * E + 1))
        o += 2 * E
    if pfe == O.T3:
        kw.update(conn_p=conn, xy_p=xy, dn_p=np.arange(o + 1, o + Nn + 1))
        npd = Nn
    else:
        kw.update(de_p=np.arange(o + 1, o + E + 1))
        npd = E
    rule = 3 if vfe == O.T3B else 2
    I, J, V = O.fam_reddy_coo(vfe, pfe, rule, conn, xy, dn_ux, dn_uy, mu=mu, **kw)
    n = o + npd
    return O.sparse_scipy(I, J, V, n, n, csc=False), conn, xy, o, mu


@pytest.mark.parametrize("vfe,pfe", [(O.T3B, O.T3), (O.Q4, O.Q4L2)])
def test_reddy_linear_field_energy_divergence_symmetry(vfe, pfe):
    K, conn, xy, o, mu = _reddy_global(vfe, pfe, 6)
    Nn = xy.shape[0]
    a, b, c, d = 0.3, -1.1, 0.7, 0.45
    u = np.zeros(K.shape[0])
    u[:Nn] = a * xy[:, 0] + b * xy[:, 1]
    u[Nn:2 * Nn] = c * xy[:, 0] + d * xy[:, 1]  # bubble coefficients stay 0
    area = 4.0
    Kuu = K[:o, :o]
    # mu |grad u|^2 + mu (div u)^2: the Reddy velocity form
    assert abs(u[:o] @ (Kuu @ u[:o]) - mu * area * (a * a + b * b + c * c + d * d + (a + d) ** 2)) < 1e-11
    # the transposed form 2 mu eps:eps would give a different number for b != c
    assert abs(mu * area * (2 * a * a + 2 * d * d + (b + c) ** 2) - mu * area * (a * a + b * b + c * c + d * d + (a + d) ** 2)) > 0.1
    # constants are in the kernel
    one = np.zeros(o)
    one[:Nn] = 1.0
    assert np.abs(Kuu @ one).max() < 1e-12
    # divergence against each pressure function: -(a + d) int Np
    div = (K[o:, :o] @ u[:o])
    if pfe == O.T3:
        w = np.zeros(Nn)
        for e in conn:
            X = xy[e - 1]
            ar = 0.5 * abs((X[1, 0] - X[0, 0]) * (X[2, 1] - X[0, 1]) - (X[2, 0] - X[0, 0]) * (X[1, 1] - X[0, 1]))
            w[e - 1] += ar / 3
    else:
        def quad_area(e):
            p, q = xy[e[2] - 1] - xy[e[0] - 1], xy[e[3] - 1] - xy[e[1] - 1]
            return 0.5 * abs(p[0] * q[1] - p[1] * q[0])
        w = np.array([quad_area(e) for e in conn])
    assert np.allclose(div, -(a + d) * w, atol=1e-12)
    assert abs(K - K.T).max() == 0


def test_reddy_t6t3_equals_pinned_three_space_oracle():
    """The family kernel's T6/T3 case is the pinned Reddy oracle
    (test/test_stokes.jl:373-431, golden errors at :551-553), bit for bit."""
    c6, xy6 = O.t6block(2.0, 2.0, 4, 4)
    xy6 = xy6 - 1.0
    c3, xy3 = O.t6tot3(c6, xy6)
    Nn6, Nn3 = xy6.shape[0], xy3.shape[0]
    dx, dy = np.arange(1, Nn6 + 1), np.arange(Nn6 + 1, 2 * Nn6 + 1)
    dp = np.arange(2 * Nn6 + 1, 2 * Nn6 + Nn3 + 1)
    a = O.stokes_reddy_coo(c6, xy6, c3, xy3, dx, dy, dp, 0.8)
    b = O.fam_reddy_coo(O.T6, O.T3, 3, c6, xy6, dx, dy, mu=0.8, conn_p=c3, xy_p=xy3, dn_p=dp)
    for u, v in zip(a, b):
        assert np.array_equal(u, v)


def test_reddy_rejects_bad_pairs():
    conn, xy = O.t3block(1.0, 1.0, 2, 2)
    dn = np.arange(1, xy.shape[0] + 1)
    with pytest.raises(ValueError):
        O.fam_reddy_coo(O.T3, O.T3, 3, conn, xy, dn, dn, conn_p=conn, xy_p=xy, dn_p=dn)
    with pytest.raises(ValueError):  # the bubble needs its cell dofs
        O.fam_reddy_coo(O.T3B, O.T3, 3, conn, xy, dn, dn, conn_p=conn, xy_p=xy, dn_p=dn)


# ---- the colliding-flow drivers converge (p1b_p1.jl, q1_q0.jl) ----------------------
def _flow_errors(pair, N):
    import flow_helpers as FH
    mesh, Ux, Uy, P = FH.spaces(pair, N)
    I, J, V = FH.oracle_coo(O, mesh, Ux, Uy, P)
    n = Ux.ndofs() + Uy.ndofs() + P.ndofs()
    K = O.sparse_scipy(I, J, V, n, n, csc=False)
    U = FH.solve(K, Ux, Uy, P)
    return (FH.errors_p1b if pair == "p1b" else FH.errors_q1q0)(mesh, Ux, Uy, P, U)


def test_p1b_p1_colliding_flow_converges():
    """The P1b/P1 discretisation of the exact colliding flow: velocity L2 error
    O(h^2), pressure L2 error faster than O(h^1.5) (measured ratios 4.0 and 3.4
    per halving).  A dropped or transposed term breaks the convergence."""
    e = [_flow_errors("p1b", N) for N in (8, 16, 32)]
    for a, b in zip(e, e[1:]):
        assert a[1] / b[1] > 3.7 and a[0] / b[0] > 2.9
    assert e[-1][1] < 0.06 and e[-1][0] < 1.7


def test_q1_q0_velocity_converges():
    """Q1/Q0 is not inf-sup stable (its pressure shows checkerboard modes and
    does not converge here), but the velocity L2 error is O(h^2)."""
    e = [_flow_errors("q1q0", N) for N in (8, 16, 32)]
    for a, b in zip(e, e[1:]):
        assert a[1] / b[1] > 3.7
    assert e[-1][1] < 0.05
