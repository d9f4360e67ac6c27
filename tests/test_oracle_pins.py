"""Pins of the CPU oracle (runs without a GPU).

Each oracle function is checked against the reference's own known answers
(tests/golden/reference_pins.json, transcribed from Elfel.jl's test files)
and, where those leave it open, against closed forms, invariants or brute
force chosen so that a dropped term, a wrong sign or index, or a transposed
operand fails.  A mutated copy of the oracle's arithmetic is run through the
same checks to show they bite (test_pins_catch_planted_mistakes).
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import oracle as O
from helpers import GOLDEN, pins

P = pins()


# ---- quadrature (src/RefShapes.jl) ----------------------------------------------
def test_quadrature_matches_reference_tests():
    r = P["refshapes"]
    pc, w = O.gauss1(3)
    assert np.allclose(pc, r["gauss3_pc"]) and np.allclose(w, r["gauss3_w"])
    pc, w = O.quad_triangle(3)
    assert np.allclose(pc, r["tri3_pc"]) and np.allclose(w, r["tri3_w"])
    assert all(x == 0.16666666666666666 for x in w)  # exact equality, test_qpiterators.jl:60
    pc, w = O.quad_square(2)
    assert pc.shape == (4, 2) and np.allclose(pc[2], r["square2_pc3"])
    res = [np.sum(np.exp(O.gauss1(k)[0]) * O.gauss1(k)[1]) for k in range(1, 6)]
    assert np.allclose(res, r["gauss_exp_integrals"])
    with pytest.raises(ValueError):
        O.quad_triangle(5)


def test_quadrature_exactness():
    # triangle 3-pt integrates quadratics exactly: int x^2 = 1/12, int xy = 1/24
    pc, w = O.quad_triangle(3)
    assert abs(np.sum(w * pc[:, 0] ** 2) - 1 / 12) < 1e-15
    assert abs(np.sum(w * pc[:, 0] * pc[:, 1]) - 1 / 24) < 1e-15
    pc, w = O.quad_square(2)  # 2x2 Gauss integrates cubics: int x^2 y^2 over [-1,1]^2 = 4/9
    assert abs(np.sum(w * pc[:, 0] ** 2 * pc[:, 1] ** 2) - 4 / 9) < 1e-13


# ---- basis functions (src/FElements.jl) -----------------------------------------
NODES = {O.T3: [(0, 0), (1, 0), (0, 1)],
         O.T6: [(0, 0), (1, 0), (0, 1), (0.5, 0), (0.5, 0.5), (0, 0.5)],
         O.Q4: [(-1, -1), (1, -1), (1, 1), (-1, 1)]}


@pytest.mark.parametrize("fe", [O.T3, O.T6, O.Q4])
def test_basis_kronecker_partition_and_gradients(fe):
    nodes = NODES[fe]
    for a, x in enumerate(nodes):
        assert np.allclose(O.bfun(fe, x), np.eye(len(nodes))[a], atol=1e-15)
    rng = np.random.default_rng(1)
    for _ in range(5):
        x = rng.uniform(0, 0.5, 2) if fe != O.Q4 else rng.uniform(-1, 1, 2)
        assert abs(O.bfun(fe, x).sum() - 1) < 1e-14
        g = O.bfungradpar(fe, x)
        assert np.allclose(g.sum(0), 0, atol=1e-14)
        for d in range(2):  # central differences of bfun
            h = 1e-6
            e = np.zeros(2); e[d] = h
            fd = (O.bfun(fe, x + e) - O.bfun(fe, x - e)) / (2 * h)
            assert np.allclose(fd, g[:, d], atol=1e-8)


def test_t3_bfun_golden():
    assert np.allclose(O.bfun(O.T3, [1 / 3, 1 / 3]), P["felements"]["T3_bfun_centroid"])
    q = P["qpiterator"]
    pc, _ = O.quad_triangle(1)
    assert np.allclose(O.bfun(O.T3, pc[0]), q["N_1pt"][0])
    pc, _ = O.quad_triangle(3)
    # the reference iterates qps in the order of its table: (2/3,1/6), (1/6,2/3), (1/6,1/6)
    assert np.allclose([O.bfun(O.T3, x) for x in pc], q["N_3pt"])
    assert np.allclose(O.bfungradpar(O.T3, pc[0]), q["gradN_T3"])


# ---- meshes and numbering ---------------------------------------------------------
def test_t3block_matches_qmesh_fixture():
    conn, xy = O.t3block(1.0, 1.0, 2, 3)
    assert np.array_equal(conn, np.loadtxt(f"{GOLDEN}/qmesh-conn.dat", dtype=np.int64))
    assert np.allclose(xy, np.loadtxt(f"{GOLDEN}/qmesh-xyz.dat"), atol=1e-15)
    assert np.array_equal(conn, np.array(P["feiterator_elements"]["value"]))


def test_numbering_feiterator_golden():
    conn = np.loadtxt(f"{GOLDEN}/qmesh-conn.dat", dtype=np.int64)
    isd = np.zeros((12, 1), bool)
    (dn,) = O.numberdofs([isd])
    assert dn[:, 0].tolist() == list(range(1, 13))
    isd[[0, 3, 6, 9]] = True
    (dn,) = O.numberdofs([isd])
    assert dn[:, 0].tolist() == P["feiterator_dofs_ebc"]["value"]
    # three copies, EBC on component 1 of nodes 1, 4, 7, 10 (test_feiterators.jl:145-167)
    isd3 = np.zeros((12, 3), bool)
    isd3[[0, 3, 6, 9], 0] = True
    (dn3,) = O.numberdofs([isd3])
    eld = dn3[conn - 1].reshape(12, -1)
    assert eld.tolist() == P["feiterator_dofs_3comp"]["value"]


def test_numbering_fefield_golden():
    xy = np.loadtxt(f"{GOLDEN}/mt3gen3-xyz.dat")
    isd = np.zeros((xy.shape[0], 2), bool)
    (dn,) = O.numberdofs([isd])
    assert dn.max() == P["fefield_mt3gen3"]["highest_free_no_data"]
    isd[2, 1] = True
    isd[4, 0] = True
    (dn,) = O.numberdofs([isd])
    assert dn[~isd].max() == P["fefield_mt3gen3"]["highest_free"]
    assert dn[isd].max() == P["fefield_mt3gen3"]["highest_data"]
    assert dn[2, 1] == 41  # v[41] == 3.0 in test_fefields.jl:52 (the (3,2) dof)


def test_stokes_counts_golden():
    """(tndof, tnunk) == (91003, 89402), test/test_stokes.jl:130."""
    N = 100
    c6, xy6 = O.t6block(1.0, 1.0, N, N)
    c3, xy3 = O.t6tot3(c6, xy6)
    bnd = (np.isclose(xy6[:, 0], 0) | np.isclose(xy6[:, 0], 1) | np.isclose(xy6[:, 1], 0) | np.isclose(xy6[:, 1], 1))
    pd = np.zeros((xy3.shape[0], 1), bool)
    pd[0] = True
    dx, dy, dp = O.numberdofs([bnd.reshape(-1, 1), bnd.reshape(-1, 1), pd])
    tndof = dx.size + dy.size + dp.size
    tnunk = 2 * int((~bnd).sum()) + xy3.shape[0] - 1
    assert (tndof, tnunk) == tuple(P["stokes_driven_N100_counts"]["value"])
    assert max(dx.max(), dy.max(), dp.max()) == tndof


# ---- sparse() (src/Assemblers.jl:121-123) ---------------------------------------------
def _assembler_coo():
    pz = P["assembler_blocks"]
    I, J, V = [], [], []
    for m, gi, gj in ((pz["m1"], pz["gi1"], pz["gj1"]), (pz["m2"], pz["gi2"], pz["gj2"])):
        m = np.array(m)
        for j in range(m.shape[1]):
            for i in range(m.shape[0]):
                I.append(gi[i]); J.append(gj[j]); V.append(m[i, j])
    return np.array(I), np.array(J), np.array(V)


@pytest.mark.parametrize("csc", [True, False])
def test_sparse_assembler_golden(csc):
    I, J, V = _assembler_coo()
    A = O.sparse_scipy(I, J, V, 7, 7, csc=csc).toarray()
    B = np.array(P["assembler_B"]["value"])
    assert np.linalg.norm(A - B) / np.linalg.norm(B) < P["assembler_B"]["tol"]


def test_sparse_brute_force_and_structural_zeros():
    rng = np.random.default_rng(11)
    I = rng.integers(1, 40, 3000)
    J = rng.integers(1, 30, 3000)
    V = rng.standard_normal(3000)
    V[:50] = 0.0
    ptr, ind, val = O.sparse(I, J, V, 40, 30, csc=True)
    dense = np.zeros((40, 30))
    seen = set()
    for i, j, v in zip(I, J, V):
        dense[i - 1, j - 1] += v
        seen.add((i - 1, j - 1))
    got = np.zeros((40, 30))
    for c in range(30):
        rows = ind[ptr[c]:ptr[c + 1]]
        assert np.all(np.diff(rows) > 0)
        got[rows, c] = val[ptr[c]:ptr[c + 1]]
    assert np.allclose(got, dense, atol=1e-13)
    assert ptr[-1] == len(seen)  # every (i, j) present, zeros included
    # left-to-right sums: duplicates of one key added in stream order
    k = (I[0], J[0])
    order = [v for i, j, v in zip(I, J, V) if (i, j) == k]
    s = order[0]
    for v in order[1:]:
        s = s + v
    r = ind[ptr[k[1] - 1]:ptr[k[1]]].tolist().index(k[0] - 1)
    assert val[ptr[k[1] - 1] + r] == s
    with pytest.raises(ValueError):
        O.sparse([1, 41], [1, 1], [1.0, 1.0], 40, 30)


def test_vec_assemble_golden_order():
    F = O.vec_assemble([1, 3, 1], [1.0, 2.0, 1e-17], 3)
    assert F.tolist() == [1.0 + 1e-17, 0.0, 2.0]
    with pytest.raises(IndexError):
        O.vec_assemble([4], [1.0], 3)


# ---- weak forms -----------------------------------------------------------------------
def _heat_problem(N, fe=O.T3, rule=1, Q=-6.0):
    gen = {O.T3: O.t3block, O.Q4: O.q4block, O.T6: O.t6block}[fe]
    conn, xy = gen(1.0, 1.0, N, N)
    bnd = (np.isclose(xy[:, 0], 0) | np.isclose(xy[:, 0], 1) | np.isclose(xy[:, 1], 0) | np.isclose(xy[:, 1], 1))
    (dn,) = O.numberdofs([bnd.reshape(-1, 1)])
    return conn, xy, bnd, dn, O.make_problem(O.HEAT, fe, rule, conn, xy, dn, kappa=1.0, Q=Q)


def _heat_solve(N, fe, rule):
    conn, xy, bnd, dn, p = _heat_problem(N, fe, rule)
    nd = xy.shape[0]
    ptr, ind, val, F = O.assemble(p, nd, csc=True)
    import scipy.sparse as sp
    K = sp.csc_matrix((val, ind, ptr), shape=(nd, nd))
    T = np.zeros(nd)
    tempf = lambda x, y: 1.0 + x ** 2 + 2.0 * y ** 2  # noqa: E731
    T[dn[bnd, 0] - 1] = tempf(xy[bnd, 0], xy[bnd, 1])
    nu = int((~bnd).sum())
    KT = K @ T
    T[:nu] = spla.spsolve(K[:nu, :nu].tocsc(), F[:nu] - KT[:nu])
    return T, xy, dn, nu


def test_heat_t3_golden_temperature():
    T, xy, dn, nu = _heat_solve(4, O.T3, 1)
    assert np.allclose(T, P["heat_t3_N4_T"]["value"], rtol=1.5e-8)


def test_heat_q4_N100_nodal_exactness():
    """test/test_heat.jl:203-224: nunknowns == 9801 and mean nodal error <= 1e-9."""
    T, xy, dn, nu = _heat_solve(100, O.Q4, 2)
    assert nu == P["heat_q4_N100_nunknowns"]["value"]
    exact = 1.0 + xy[:, 0] ** 2 + 2.0 * xy[:, 1] ** 2
    assert np.mean(np.abs(T[dn[:, 0] - 1] - exact)) <= P["heat_nodal_mean_error_max"]["value"]


def _single_element(fe, X):
    conn = np.arange(1, len(X) + 1, dtype=np.int64).reshape(1, -1)
    return conn, np.asarray(X, float)


def test_t3_element_closed_form():
    """ke = (b_i b_j + c_i c_j) / (4A) for a linear triangle, fe = Q A / 3."""
    X = np.array([[0.1, 0.2], [1.3, 0.4], [0.5, 1.7]])
    conn, xy = _single_element(O.T3, X)
    dn = np.arange(1, 4).reshape(-1, 1)
    p = O.make_problem(O.HEAT, O.T3, 1, conn, xy, dn, kappa=2.0, Q=3.0)
    I, J, V, FI, FV = O.coo(p)
    ke = V.reshape(3, 3).T
    x, y = X[:, 0], X[:, 1]
    A = 0.5 * ((x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]))
    b = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]])
    c = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]])
    assert np.allclose(ke, 2.0 * (np.outer(b, b) + np.outer(c, c)) / (4 * A), rtol=1e-13)
    assert np.allclose(FV, 3.0 * A / 3, rtol=1e-14)
    assert np.array_equal(I, np.tile(np.arange(1, 4), 3)) and np.array_equal(J, np.repeat(np.arange(1, 4), 3))


def test_q4_rectangle_closed_form():
    a, b = 0.7, 0.3  # rectangle sides
    X = np.array([[0, 0], [a, 0], [a, b], [0, b]], float)
    conn, xy = _single_element(O.Q4, X)
    p = O.make_problem(O.HEAT, O.Q4, 2, conn, xy, np.arange(1, 5).reshape(-1, 1), kappa=1.0, Q=0.0)
    ke = O.coo(p)[2].reshape(4, 4).T
    r = b / a
    Kx = (r / 6) * np.array([[2, -2, -1, 1], [-2, 2, 1, -1], [-1, 1, 2, -2], [1, -1, -2, 2]])
    Ky = (1 / (6 * r)) * np.array([[2, 1, -1, -2], [1, 2, -2, -1], [-1, -2, 2, 1], [-2, -1, 1, 2]])
    assert np.allclose(ke, Kx + Ky, rtol=1e-12)


def test_elasticity_cst_rigid_modes_symmetry():
    X = np.array([[0.1, 0.2], [1.3, 0.4], [0.5, 1.7]])
    conn, xy = _single_element(O.T3, X)
    dn = np.arange(1, 7).reshape(3, 2)
    E, nu = 1.0, 0.3
    D = E / ((1 + nu) * (1 - 2 * nu)) * np.array([[1 - nu, nu, 0], [nu, 1 - nu, 0], [0, 0, (1 - 2 * nu) / 2]])
    p = O.make_problem(O.ELAST, O.T3, 1, conn, xy, dn, D=D)
    ke = O.coo(p)[2].reshape(6, 6).T
    x, y = X[:, 0], X[:, 1]
    A = 0.5 * ((x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]))
    b = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]]) / (2 * A)
    c = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]]) / (2 * A)
    B = np.zeros((3, 6))
    B[0, 0::2], B[1, 1::2], B[2, 0::2], B[2, 1::2] = b, c, c, b
    assert np.allclose(ke, A * B.T @ D @ B, rtol=1e-12)
    assert np.allclose(ke, ke.T, rtol=1e-13)
    modes = [np.tile([1, 0], 3), np.tile([0, 1], 3), np.column_stack([-y, x]).ravel()]
    for m in modes:
        assert np.abs(ke @ m).max() < 1e-12 * np.abs(ke).max()


def test_t6_curved_elasticity_rigid_modes_and_brute_force():
    rng = np.random.default_rng(5)
    conn, xy = O.t6block(1.0, 1.0, 3, 3)
    xy = xy + rng.uniform(-0.03, 0.03, xy.shape)
    E = conn.shape[0]
    dn = np.arange(1, 2 * xy.shape[0] + 1).reshape(-1, 2)
    D = np.array([[2.0, 0.6, 0.0], [0.6, 2.0, 0.0], [0.0, 0.0, 0.7]])
    p = O.make_problem(O.ELAST, O.T6, 3, conn, xy, dn, D=D)
    I, J, V, _, _ = O.coo(p)
    K = O.sparse_scipy(I, J, V, dn.size, dn.size).toarray()
    dense = np.zeros_like(K)
    for k in range(I.size):
        dense[I[k] - 1, J[k] - 1] += V[k]
    assert np.allclose(K, dense, atol=1e-14)
    assert np.allclose(K, K.T, atol=1e-12)
    x, y = xy[:, 0], xy[:, 1]
    for m in (np.tile([1, 0], xy.shape[0]), np.tile([0, 1], xy.shape[0]), np.column_stack([-y, x]).ravel()):
        assert np.abs(K @ m).max() < 1e-11 * np.abs(K).max()
    assert E == 18


def test_heat_row_sums_and_load_total():
    conn, xy, bnd, dn, p = _heat_problem(7, O.T6, 3, Q=2.5)
    nd = xy.shape[0]
    ptr, ind, val, F = O.assemble(p, nd, csc=False)
    import scipy.sparse as sp
    K = sp.csr_matrix((val, ind, ptr), shape=(nd, nd))
    assert np.abs(K @ np.ones(nd)).max() < 1e-12
    assert abs(F.sum() - 2.5 * 1.0) < 1e-13


# ---- Stokes (golden error norms) --------------------------------------------------------
def _colliding_mesh(N):
    c6, xy6 = O.t6block(2.0, 2.0, N, N)
    xy6 = xy6 - 1.0
    c3, xy3 = O.t6tot3(c6, xy6)
    bnd = np.isclose(np.abs(xy6[:, 0]), 1) | np.isclose(np.abs(xy6[:, 1]), 1)
    pd = np.zeros((xy3.shape[0], 1), bool)
    pd[np.argmin(np.hypot(xy3[:, 0], xy3[:, 1]))] = True
    return c6, xy6, c3, xy3, bnd, pd


def _trueu(x, y):
    return 20 * x * y ** 3, 5 * x ** 4 - 5 * y ** 4


def _solve(K, U, nu):
    KT = K @ U
    U[:nu] = spla.spsolve(K[:nu, :nu].tocsc(), -KT[:nu])
    return U


@pytest.mark.parametrize("k", [0, 1, 2])
def test_stokes_reddy_golden(k):
    N = P["stokes_reddy_errors"]["N"][k]
    c6, xy6, c3, xy3, bnd, pd = _colliding_mesh(N)
    dx, dy, dp = O.numberdofs([bnd.reshape(-1, 1), bnd.reshape(-1, 1), pd])
    nd = dx.size + dy.size + dp.size
    nu = 2 * int((~bnd).sum()) + xy3.shape[0] - 1
    I, J, V = O.stokes_reddy_coo(c6, xy6, c3, xy3, dx[:, 0], dy[:, 0], dp[:, 0], 1.0)
    K = O.sparse_scipy(I, J, V, nd, nd)
    U = np.zeros(nd)
    ux, uy = _trueu(xy6[bnd, 0], xy6[bnd, 1])
    U[dx[bnd, 0] - 1], U[dy[bnd, 0] - 1] = ux, uy
    U = _solve(K, U, nu)
    ep = O.colliding_p_error(c3, xy3, dp[:, 0], U)
    ev = O.colliding_v_error_scalar(c6, xy6, dx[:, 0], dy[:, 0], U)
    assert np.allclose([ep, ev], P["stokes_reddy_errors"]["value"][k], rtol=1.5e-8)


@pytest.mark.parametrize("kind,key", [(O.STOKES_VECLAP, "veclap"), (O.STOKES_GEN, "gen")])
def test_stokes_vector_space_golden(kind, key):
    """veclap: test/test_stokes.jl:779.  gen: differs from the Reddy form only by a
    null-Lagrangian term, so its solution carries the Reddy N=4 errors."""
    c6, xy6, c3, xy3, bnd, pd = _colliding_mesh(4)
    du, dp = O.numberdofs([np.stack([bnd, bnd], 1), pd])
    nd = du.size + dp.size
    nu = 2 * int((~bnd).sum()) + xy3.shape[0] - 1
    p = O.make_problem(kind, O.T6, 3, c6, xy6, du, conn_p=c3, dofnums_p=dp, mu=1.0, D=np.diag([2.0, 2.0, 1.0]))
    I, J, V, _, _ = O.coo(p)
    K = O.sparse_scipy(I, J, V, nd, nd)
    U = np.zeros(nd)
    ux, uy = _trueu(xy6[bnd, 0], xy6[bnd, 1])
    U[du[bnd, 0] - 1], U[du[bnd, 1] - 1] = ux, uy
    U = _solve(K, U, nu)
    e = [O.colliding_p_error(c3, xy3, dp[:, 0], U), O.colliding_v_error_vector(c6, xy6, du, U)]
    gold = P["stokes_veclap_N4_errors"]["value"] if key == "veclap" else P["stokes_reddy_errors"]["value"][0]
    assert np.allclose(e, gold, rtol=1.5e-8)


# ---- solver -------------------------------------------------------------------------------
def test_pcg_matches_direct_solve():
    conn, xy, bnd, dn, p = _heat_problem(30, O.Q4, 2, Q=1.0)
    nd = xy.shape[0]
    ptr, ind, val, F = O.assemble(p, nd, csc=False)
    nu = int((~bnd).sum())
    x, it, rr = O.pcg(ptr, ind, val, F, nu, tol=1e-12)
    import scipy.sparse as sp
    K = sp.csr_matrix((val, ind, ptr), shape=(nd, nd))
    xd = spla.spsolve(K[:nu, :nu].tocsc(), F[:nu])
    assert rr <= 1e-12 and np.allclose(x, xd, rtol=1e-9, atol=1e-12)


# ---- the pins bite: planted mistakes in a copy of the oracle arithmetic fail them ---------
def test_pins_catch_planted_mistakes():
    """Recompute a T3 heat element with the oracle's Jacobian/gradient formulas
    and three plausible mistakes; only the correct one passes the closed form."""
    X = np.array([[0.1, 0.2], [1.3, 0.4], [0.5, 1.7]])
    g = O.bfungradpar(O.T3, [1 / 3, 1 / 3])

    def element(mistake):
        j11, j21 = X[:, 0] @ g[:, 0], X[:, 1] @ g[:, 0]
        j12, j22 = X[:, 0] @ g[:, 1], X[:, 1] @ g[:, 1]
        if mistake == "transposed J":
            j12, j21 = j21, j12
        det = j11 * j22 - j21 * j12
        if mistake == "sign":
            det = j11 * j22 + j21 * j12
        gx = (j22 * g[:, 0] - j21 * g[:, 1]) / det
        gy = (j11 * g[:, 1] - j12 * g[:, 0]) / det
        if mistake == "dropped term":
            gy = (j11 * g[:, 1]) / det
        return (np.outer(gx, gx) + np.outer(gy, gy)) * det * 0.5

    x, y = X[:, 0], X[:, 1]
    A = 0.5 * ((x[1] - x[0]) * (y[2] - y[0]) - (x[2] - x[0]) * (y[1] - y[0]))
    b = np.array([y[1] - y[2], y[2] - y[0], y[0] - y[1]])
    c = np.array([x[2] - x[1], x[0] - x[2], x[1] - x[0]])
    exact = (np.outer(b, b) + np.outer(c, c)) / (4 * A)
    conn, xy = _single_element(O.T3, X)
    p = O.make_problem(O.HEAT, O.T3, 1, conn, xy, np.arange(1, 4).reshape(-1, 1), kappa=1.0)
    assert np.allclose(O.coo(p)[2].reshape(3, 3).T, element(None), rtol=1e-14)
    assert np.allclose(element(None), exact, rtol=1e-13)
    for m in ("transposed J", "sign", "dropped term"):
        assert not np.allclose(element(m), exact, rtol=1e-6), m
