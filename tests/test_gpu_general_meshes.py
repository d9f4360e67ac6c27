"""Meshes and dof maps that the REUSE tiles do not hold, and the context's
state checks.

The reference's sparse() assembles any COO stream (src/Assemblers.jl:97-123),
so the library must too.  The REUSE method's tiles need every node in at most
16 elements, rows of at most 128 entries and each dof owned by one (node,
component); on a mesh that breaks one of these, REUSE caches the plain
method's sort order instead and sums the runs of the element stream in FP64
(elfel_hip_assembly_info says so), and the result equals the oracle's bit for
bit (tolerance 0).
"""
import numpy as np
import pytest

import oracle as O
from elfel_hip import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture
def gctx():
    c = L.Context(0)
    yield c
    c.close()


def fan_mesh(K, rings=2):
    """T3 fan: a centre node in K triangles, then `rings` rings of K nodes each
    joined by quadrilateral strips split into two triangles.  The centre has
    valence K, ring nodes 6."""
    xy = [(0.0, 0.0)]
    for r in range(1, rings + 1):
        for k in range(K):
            t = 2 * np.pi * (k + 0.3 * (r % 2)) / K
            xy.append((r * np.cos(t), r * np.sin(t)))
    xy = np.array(xy)
    node = lambda r, k: 1 + (r - 1) * K + (k % K)  # noqa: E731  (0-based)
    conn = []
    for k in range(K):
        conn.append((0, node(1, k), node(1, k + 1)))
    for r in range(1, rings):
        for k in range(K):
            a, b, c, d = node(r, k), node(r, k + 1), node(r + 1, k + 1), node(r + 1, k)
            conn.append((a, d, c))
            conn.append((a, c, b))
    return np.array(conn, np.int64) + 1, xy


def _oracle_heat(conn, xy, dn, n, layout, kappa=1.3, Q=-2.0):
    p = O.make_problem(O.HEAT, O.T3, 1, conn, xy, dn.reshape(-1, 1), kappa=kappa, Q=Q)
    return O.assemble(p, n, csc=(layout == L.CSC))


def _same(dev, ref):
    for a, b, what in zip(dev[:3], ref[:3], ("pointers", "indices", "values")):
        assert a.shape == b.shape and np.array_equal(a, b), what


@pytest.mark.parametrize("layout", [L.CSR, L.CSC])
@pytest.mark.parametrize("precision", [L.FP64, L.MIXED])
def test_fan_mesh_valence_32(gctx, layout, precision):
    """A node in 32 triangles (valence 32 > 16): the REUSE tiles do not hold
    it, so REUSE sums the runs of the plain method's stream in its cached sort
    order, in FP64, equal to the oracle bit for bit, and says so."""
    conn, xy = fan_mesh(32, rings=3)
    Nn = xy.shape[0]
    dn = np.random.default_rng(5).permutation(Nn) + 1
    gctx.set_mesh(0, conn, xy)
    gctx.set_dofs(0, dn, Nn)
    ref = _oracle_heat(conn, xy, dn, Nn, layout)
    for _ in range(3):  # later calls reuse the cached decision and sort order
        gctx.assemble(L.HEAT, L.T3, 1, precision=precision, method=L.REUSE, layout=layout, kappa=1.3, Q=-2.0)
        _same(gctx.copy_sparse(base=0), ref)
        assert np.array_equal(gctx.copy_vector(Nn), ref[3])
        m, p, note = gctx.assembly_info()
        assert (m, p) == (L.REUSE, L.FP64) and "too highly connected" in note, note


def test_fan_mesh_elasticity(gctx):
    """Vector problem on the fan: the centre's two rows have 2 * 33 entries."""
    conn, xy = fan_mesh(24, rings=2)
    Nn = xy.shape[0]
    n = 2 * Nn
    dn = (np.random.default_rng(9).permutation(n) + 1).reshape(Nn, 2)
    D = np.array([[1.35, 0.58, 0.0], [0.58, 1.35, 0.0], [0.0, 0.0, 0.38]])
    gctx.set_mesh(0, conn, xy)
    gctx.set_dofs(0, dn, n)
    p = O.make_problem(O.ELAST, O.T3, 1, conn, xy, dn, D=D)
    ref = O.assemble(p, n, csc=True)
    for _ in range(2):
        gctx.assemble(L.ELAST, L.T3, 1, precision=L.MIXED, method=L.REUSE, layout=L.CSC, D=D)
        _same(gctx.copy_sparse(base=0), ref)
        assert gctx.assembly_info()[:2] == (L.REUSE, L.FP64)


def test_shared_dof_two_nodes(gctx):
    """Periodic tie: the nodes of the right edge of a T3 block get the dof
    numbers of the left edge's nodes, so one dof belongs to two nodes.  The
    oracle (sparse() of the COO stream) sums both nodes' contributions into
    one row; REUSE sums the cached sort order's runs and gives the same bits."""
    N = 9
    conn, xy = O.t3block(1.0, 1.0, N, N)
    Nn = xy.shape[0]
    node = np.arange(Nn).reshape(N + 1, N + 1)  # x fastest: node[j, i]
    owner = np.arange(Nn)
    owner[node[:, N]] = node[:, 0]
    uniq, inv = np.unique(owner, return_inverse=True)
    dn = inv + 1
    n = uniq.size
    gctx.set_mesh(0, conn, xy)
    gctx.set_dofs(0, dn, n)
    for layout in (L.CSR, L.CSC):
        ref = _oracle_heat(conn, xy, dn, n, layout)
        gctx.assemble(L.HEAT, L.T3, 1, precision=L.MIXED, method=L.REUSE, layout=layout, kappa=1.3, Q=-2.0)
        _same(gctx.copy_sparse(base=0), ref)
        assert np.array_equal(gctx.copy_vector(n), ref[3])
        m, p, note = gctx.assembly_info()
        assert (m, p) == (L.REUSE, L.FP64) and "shared" in note, note
    # an ordinary map again: the REUSE method runs
    gctx.set_dofs(0, np.arange(1, Nn + 1), Nn)
    gctx.assemble(L.HEAT, L.T3, 1, precision=L.MIXED, method=L.REUSE, kappa=1.3, Q=-2.0)
    assert gctx.assembly_info()[:2] == (L.REUSE, L.MIXED)


@pytest.mark.parametrize("N", [8, 600])
def test_q4_shared_dof_two_nodes(gctx, N):
    """Heat Q4 (2x2 Gauss, per-element coefficients) with the right edge tied
    to the left edge: the row-owner tiles do not hold a dof of two nodes; the
    cached sort order's FP64 runs equal the oracle's matrix and load vector
    bit for bit (600^2: past the device builder's 2^18-row threshold)."""
    conn, xy = O.q4block(1.0, 1.0, N, N)
    Nn, E = xy.shape[0], conn.shape[0]
    node = np.arange(Nn).reshape(N + 1, N + 1)
    owner = np.arange(Nn)
    owner[node[:, N]] = node[:, 0]
    uniq, inv = np.unique(owner, return_inverse=True)
    dn = inv + 1
    n = uniq.size
    rng = np.random.default_rng(8)
    kap, Qe = np.exp(rng.normal(size=E)), rng.uniform(-1.0, 1.0, E)
    gctx.set_mesh(0, conn, xy)
    gctx.set_dofs(0, dn, n)
    gctx.set_coefficients(E, kap, Qe, None)
    ref = O.assemble(O.make_problem(O.HEAT, O.Q4, 2, conn, xy, dn.reshape(-1, 1), kappa_e=kap, Q_e=Qe), n,
                     csc=False)
    gctx.assemble(L.HEAT, L.Q4, 2, precision=L.MIXED, method=L.REUSE, layout=L.CSR)
    _same(gctx.copy_sparse(base=0), ref)
    assert np.array_equal(gctx.copy_vector(n), ref[3])
    m, p, note = gctx.assembly_info()
    assert (m, p) == (L.REUSE, L.FP64) and "shared" in note, note


@pytest.mark.parametrize("form", [L.STOKES_GEN, L.STOKES_VECLAP])
def test_stokes_shared_velocity_dof(gctx, form):
    """Taylor-Hood T6/T3 with two velocity nodes tied to one pair of dofs:
    the two-field problem's runs (kuu, kup and kup^T in one stream) summed in
    the cached sort order equal the oracle's sparse() bit for bit; then the
    pressure map changes and the cached order is rebuilt."""
    from elfel_hip import workloads as W
    cfg = W.stokes_t6t3(7)
    u, p = cfg["ufesp"], cfg["pfesp"]
    du, dp = u.dofnums.copy(), p.dofnums.copy()
    du[9] = du[8]  # node 10 takes node 9's two dofs
    used = np.unique(np.concatenate([du.ravel(), dp.ravel()]))
    remap = np.zeros(used.max() + 1, np.int64)
    remap[used] = np.arange(1, used.size + 1)
    du, dp, n = remap[du], remap[dp], used.size
    okind = O.STOKES_GEN if form == L.STOKES_GEN else O.STOKES_VECLAP
    gctx.set_mesh(0, u.mesh.conn, u.mesh.xy)
    gctx.set_dofs(0, du, n)
    gctx.set_mesh(1, p.mesh.conn, p.mesh.xy)
    for perm in (False, True):
        dpp = dp if not perm else dp[np.random.default_rng(4).permutation(dp.shape[0])]
        gctx.set_dofs(1, dpp, n)
        ref = O.assemble(O.make_problem(okind, O.T6, 3, u.mesh.conn, u.mesh.xy, du, conn_p=p.mesh.conn,
                                        dofnums_p=dpp, D=cfg["D"], mu=1.0), n, csc=True)
        gctx.assemble(form, L.T6, 3, precision=L.MIXED, method=L.REUSE, layout=L.CSC, D=cfg["D"], mu=1.0)
        _same(gctx.copy_sparse(base=0), ref)
        m, prec, note = gctx.assembly_info()
        assert (m, prec) == (L.REUSE, L.FP64) and "shared" in note, note


def test_new_mesh_invalidates_dof_map(gctx):
    """A dof map set for one mesh does not describe a larger mesh: assembling
    after set_mesh without set_dofs is a state error (raised on the host,
    before any launch), not an out-of-bounds read."""
    c1, x1 = O.t3block(1.0, 1.0, 4, 4)
    c2, x2 = O.t3block(1.0, 1.0, 8, 8)
    gctx.set_mesh(0, c1, x1)
    gctx.set_dofs(0, np.arange(1, x1.shape[0] + 1), x1.shape[0])
    gctx.assemble(L.HEAT, L.T3, 1, kappa=1.0, Q=1.0)
    gctx.set_mesh(0, c2, x2)
    for meth in (L.SORT, L.REUSE):
        with pytest.raises(L.ElfelHipError) as e:
            gctx.assemble(L.HEAT, L.T3, 1, method=meth, kappa=1.0, Q=1.0)
        assert e.value.code == L.ESTATE
    # the same node count keeps the map (new coordinates of the same mesh)
    gctx.set_dofs(0, np.arange(1, x2.shape[0] + 1), x2.shape[0])
    gctx.set_mesh(0, c2, x2 * 2.0)
    gctx.assemble(L.HEAT, L.T3, 1, kappa=1.0, Q=1.0)


def test_device_mesh_invalidates_dof_map(gctx):
    """The device generators replace a field's mesh the same way."""
    gctx.block_mesh(0, L.Q4, 1.0, 1.0, 3, 3)
    gctx.set_dofs(0, np.arange(1, 17), 16)
    gctx.assemble(L.HEAT, L.Q4, 2, kappa=1.0, Q=1.0)
    gctx.block_mesh(0, L.Q4, 1.0, 1.0, 5, 5)
    with pytest.raises(L.ElfelHipError) as e:
        gctx.assemble(L.HEAT, L.Q4, 2, kappa=1.0, Q=1.0)
    assert e.value.code == L.ESTATE


def test_check_geom_flag(gctx):
    """ELFEL_CHECK_GEOM returns EGEOM before assembling when an element has
    det J <= 0 at a quadrature point; without the flag the inverted element is
    assembled as the reference assembles it (jacjac does not check), equal to
    the oracle bit for bit."""
    N = 6
    conn, xy = O.t3block(1.0, 1.0, N, N)
    Nn = xy.shape[0]
    dn = np.arange(1, Nn + 1)
    gctx.set_mesh(0, conn, xy)
    gctx.set_dofs(0, dn, Nn)
    gctx.assemble(L.HEAT, L.T3, 1, method=L.REUSE, precision=L.MIXED, kappa=1.0, Q=1.0, flags=L.CHECK_GEOM)
    bad = xy.copy()
    k = (N + 1) * 3 + 3  # an interior node moved across its neighbours
    bad[k] += np.array([1.6, 0.9]) / N
    gctx.set_mesh(0, conn, bad)
    for meth, prec in ((L.SORT, L.FP64), (L.REUSE, L.MIXED)):
        with pytest.raises(L.ElfelHipError) as e:
            gctx.assemble(L.HEAT, L.T3, 1, method=meth, precision=prec, kappa=1.0, Q=1.0, flags=L.CHECK_GEOM)
        assert e.value.code == L.EGEOM
    mindet, _ = (None, None)
    gctx.assemble(L.HEAT, L.T3, 1, method=L.REUSE, precision=L.FP64, layout=L.CSR, kappa=1.0, Q=1.0)
    mindet, nbad = gctx.geometry_stats()
    assert mindet <= 0 and nbad >= 1
    _same(gctx.copy_sparse(base=0), _oracle_heat(conn, bad, dn, Nn, L.CSR, kappa=1.0, Q=1.0))
    with pytest.raises(L.ElfelHipError) as e:
        gctx.assemble(L.HEAT, L.T3, 1, kappa=1.0, Q=1.0, flags=4)
    assert e.value.code == L.EINVAL


def test_reddy_q1q0_ignores_larger_node_map(gctx):
    """(Q4, FEL2_Q4) Reddy: the system is sized by the maps the pair reads.
    A larger node map left on field 1 by an earlier problem does not inflate
    the matrix."""
    N = 6
    conn, xy = O.q4block(2.0, 2.0, N, N)
    xy = xy - 1.0
    E, Nn = conn.shape[0], xy.shape[0]
    total = 2 * Nn + E
    perm = np.random.default_rng(3).permutation(total) + 1
    dux, duy, dp = perm[:Nn], perm[Nn:2 * Nn], perm[2 * Nn:]
    # an earlier problem's pressure T3 mesh and node map on field 1, numbered up to 5000
    c3, x3 = O.t3block(1.0, 1.0, 20, 20)
    gctx.set_mesh(1, c3, x3)
    gctx.set_dofs(1, np.arange(5000 - x3.shape[0] + 1, 5001), 5000)
    gctx.set_mesh(0, conn, xy)
    gctx.set_dofs(0, np.stack([dux, duy], 1), total)
    gctx.set_element_dofs(1, dp, total)
    nnz = gctx.assemble(L.STOKES_REDDY, L.Q4, 2, mu=1.0, fe_p=L.Q4L2)
    I, J, V = O.fam_reddy_coo(L.Q4, L.Q4L2, 2, conn, xy, dux, duy, mu=1.0, de_p=dp)
    ref = O.sparse(I, J, V, total, total, csc=False)
    assert nnz == ref[1].size
    assert gctx.result_info()[:2] == (total, total)
    _same(gctx.copy_sparse(base=0), ref)
