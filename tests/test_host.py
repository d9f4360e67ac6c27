"""Host-side logic of the elfel_hip package (runs without a GPU): meshes,
numbering, COO stream order of the host assemblers, workload generators."""
import os

import numpy as np
import pytest

import elfel_hip as eh
import oracle as O
from elfel_hip import workloads as W
from helpers import GOLDEN, pins

P = pins()


def test_load_and_t3block_match_fixture():
    m = eh.load(os.path.join(GOLDEN, "qmesh.mesh"))
    b = eh.T3block(1.0, 1.0, 2, 3)
    assert m.shape == "T3" and np.array_equal(m.conn, b.conn) and np.allclose(m.xy, b.xy, atol=1e-15)


@pytest.mark.parametrize("N", [1, 3, 8])
def test_generators_match_oracle(N):
    for gen, og in ((eh.T3block, O.t3block), (eh.Q4block, O.q4block), (eh.T6block, O.t6block)):
        m = gen(1.0, 0.5, N, N + 1)
        c, xy = og(1.0, 0.5, N, N + 1)
        assert np.array_equal(m.conn, c)
        assert np.array_equal(m.xy, xy)  # bitwise: same formulas
    m6 = eh.T6block(2.0, 2.0, N, N)
    p = eh.T6toT3(m6)
    c3, xy3 = O.t6tot3(m6.conn, m6.xy)
    assert np.array_equal(p.conn, c3) and np.array_equal(p.xy, xy3)


def test_t6_midside_nodes_are_edge_midpoints():
    m = eh.T6block(1.0, 1.0, 5, 4)
    for a, b, mid in ((0, 1, 3), (1, 2, 4), (2, 0, 5)):
        assert np.allclose((m.xy[m.conn[:, a] - 1] + m.xy[m.conn[:, b] - 1]) / 2, m.xy[m.conn[:, mid] - 1])
    assert m.nnodes == 11 * 9
    # each interior edge shared by two elements gets one node
    assert np.unique(m.conn[:, 3:]).size == m.nnodes - 6 * 5


def test_boundary_nodes():
    m = eh.T6block(1.0, 1.0, 6, 6)
    b = m.boundary_nodes()
    xy = m.xy[b - 1]
    on = np.isclose(xy[:, 0], 0) | np.isclose(xy[:, 0], 1) | np.isclose(xy[:, 1], 0) | np.isclose(xy[:, 1], 1)
    assert on.all() and b.size == 4 * 12


def test_numbering_matches_reference_pins():
    m = eh.load(os.path.join(GOLDEN, "qmesh.mesh"))
    f = eh.FESpace(m, eh.FEH1_T3())
    eh.numberdofs(f)
    assert f.ndofs() == 12 and f.nunknowns() == 12
    f.setebc(0, [1, 4, 7, 10], 1, 0.0)
    eh.numberdofs(f)
    assert f.nunknowns() == 8
    assert f.dofnums[:, 0].tolist() == P["feiterator_dofs_ebc"]["value"]
    f3 = eh.FESpace(m, eh.FEH1_T3(), 3)
    f3.setebc(0, [1, 4, 7, 10], 1, 0.0)
    f3.numberfreedofs()
    f3.numberdatadofs()
    assert f3.edofcompnt().tolist() == P["feiterator_dofs_3comp"]["eldofcomps"]
    assert f3.element_dofs().tolist() == P["feiterator_dofs_3comp"]["value"]


def test_fefield_mt3gen3_pins():
    m = eh.load(os.path.join(GOLDEN, "mt3gen3.mesh"))
    f = eh.FESpace(m, eh.FEH1_T3(), 2)
    f.numberfreedofs()
    f.numberdatadofs(f.highestfreedofnum() + 1)
    assert f.highestfreedofnum() == 42 and f.highestdatadofnum() == 0
    f.setebc(0, 3, 2, 3.0)
    f.setebc(0, 5, 1, 5.0)
    f.numberfreedofs(1)
    f.numberdatadofs(f.highestfreedofnum() + 1)
    assert (f.highestfreedofnum(), f.highestdatadofnum()) == (40, 42)
    v = np.zeros(f.ndofs())
    f.gathersysvec(v)
    assert v[41 - 1] == 3.0
    f.scattersysvec(np.arange(1.0, 43.0))
    assert f.dofvals[2, 1] == 41.0


def test_numbering_multi_space_matches_oracle():
    cfg = W.stokes_t6t3(8)
    u, p = cfg["ufesp"], cfg["pfesp"]
    ou, op = O.numberdofs([u.isdatum, p.isdatum])
    assert np.array_equal(u.dofnums, ou) and np.array_equal(p.dofnums, op)
    assert cfg["ndofs"] == u.dofnums.size + p.dofnums.size


def test_stokes_counts_pin():
    cfg = W.stokes_t6t3(100)
    assert (cfg["ndofs"], cfg["nfree"]) == tuple(P["stokes_driven_N100_counts"]["value"])


def test_heat_q4_nunknowns_pin():
    cfg = W.heat_q4(100)
    assert cfg["nfree"] == P["heat_q4_N100_nunknowns"]["value"]


def test_host_coo_stream_order_matches_oracle():
    """assemble!(a, lma) appends column-major, assemble!(a, transpose(lma))
    row-major of the parent (src/Assemblers.jl:97-114): a host loop over the
    oracle's element blocks reproduces the oracle's own COO stream."""
    cfg = W.stokes_t6t3(3)
    u, p = cfg["ufesp"], cfg["pfesp"]
    n = cfg["ndofs"]
    op = O.make_problem(O.STOKES_GEN, O.T6, 3, u.mesh.conn, u.mesh.xy, u.dofnums, conn_p=p.mesh.conn,
                        dofnums_p=p.dofnums, D=cfg["D"])
    I, J, V, _, _ = O.coo(op)
    a = eh.SysmatAssemblerHIP()
    eh.start(a, n, n)
    kuu = eh.LocalMatrixAssembler(12, 12)
    kup = eh.LocalMatrixAssembler(12, 3)
    ud, pd = u.element_dofs(), p.element_dofs()
    blk = 144 + 36 + 36
    for e in range(u.mesh.nelements):
        eh.init(kuu, ud[e], ud[e])
        eh.init(kup, ud[e], pd[e])
        s = V[e * blk:(e + 1) * blk]
        kuu.M[:, :] = s[:144].reshape(12, 12, order="F")
        kup.M[:, :] = s[144:180].reshape(12, 3, order="F")
        eh.assemble(a, kuu)
        eh.assemble(a, kup)
        eh.assemble(a, kup.T)
    hI, hJ, hV = a.coo()
    assert np.array_equal(hI, I) and np.array_equal(hJ, J) and np.array_equal(hV, V)


def test_quadrature_mirror():
    q = eh.quadrature("triangle", npts=3)
    assert np.allclose(q.param_coords, P["refshapes"]["tri3_pc"]) and (q.weights == 0.16666666666666666).all()
    q = eh.quadrature("square", kind="Gauss", order=2)
    assert np.allclose(q.param_coords[2], P["refshapes"]["square2_pc3"])
    q = eh.quadrature("interval", order=7)
    assert abs(np.sum(np.exp(q.param_coords[:, 0]) * q.weights) - (np.e - 1 / np.e)) < 1e-14


def test_felements_mirror_oracle():
    rng = np.random.default_rng(2)
    for fe, code in ((eh.FEH1_T3(), O.T3), (eh.FEH1_T6(), O.T6), (eh.FEH1_Q4(), O.Q4)):
        x = rng.uniform(0, 0.4, 2)
        assert np.array_equal(fe.bfun(x), O.bfun(code, x))
        assert np.array_equal(fe.bfungradpar(x), O.bfungradpar(code, x))


def test_workloads_deterministic_and_in_range():
    a = W.elast_t6(20, frac=0.15)
    b = W.elast_t6(20, frac=0.15)
    assert np.array_equal(a["fesp"].mesh.xy, b["fesp"].mesh.xy)
    base = eh.T6block(1.0, 1.0, 20, 20)
    d = np.abs(a["fesp"].mesh.xy - base.xy)
    assert d.max() <= 0.15 / 20 + 1e-15 and d.max() > 0.1 / 20
    x, y = base.xy[:, 0], base.xy[:, 1]
    onx = np.isclose(x, 0) | np.isclose(x, 1)
    assert np.all(d[onx, 0] == 0)
    u = W.uniform(W.SEED, 1, np.arange(100000, dtype=np.uint64))
    assert 0 <= u.min() and u.max() < 1 and abs(u.mean() - 0.5) < 0.01
    z = W.normal(W.SEED, 2, 100000)
    assert abs(z.mean()) < 0.02 and abs(z.std() - 1) < 0.02
    specs = W.batch_specs()
    E = np.array([2 * n * n for _, n in specs])
    assert len(specs) == 256 and E.min() >= 800 and E.max() <= 2.1e5


# ---- element families in the mirror ---------------------------------------------------
def test_family_generators_match_oracle():
    for args in ((1.0, 2.0, 0.5, 3, 2, 4), (1.0, 1.0, 1.0, 5, 5, 5)):
        m = eh.T4block(*args)
        c, x = O.t4block(*args)
        assert np.array_equal(m.conn, c) and np.array_equal(m.xy, x)
    m = eh.L2block(2.0, 9)
    c, x = O.l2block(2.0, 9)
    assert np.array_equal(m.conn, c) and np.array_equal(m.xy, x)
    # boundary facets: the cube's surface nodes, the segment's end points
    b = eh.T4block(1.0, 1.0, 1.0, 4, 3, 2).boundary_nodes()
    assert b.size == 5 * 4 * 3 - 3 * 2 * 1
    assert np.array_equal(eh.L2block(1.0, 7).boundary_nodes(), [1, 8])


def test_family_basis_mirror_matches_oracle():
    rng = np.random.default_rng(4)
    for fe, code in ((eh.FEH1_T3_BUBBLE(), O.T3B), (eh.FEH1_T4(), O.T4), (eh.FEH1_L2(), O.L2), (eh.FEL2_Q4(), O.Q4L2)):
        for _ in range(4):
            pc = rng.uniform(0.05, 0.3, fe.mdim)
            assert np.array_equal(fe.bfun(pc), O.fam_bfun(code, pc))
            assert np.array_equal(fe.bfungradpar(pc), O.fam_bfungradpar(code, pc))
    for m in range(4):  # FEData of the constructors (src/FElements.jl:339,372,400)
        assert eh.FEH1_T3_BUBBLE().ndofperfeat(m) == (1, 0, 1, 0)[m]
        assert eh.FEL2_Q4().ndofperfeat(m) == (0, 0, 1, 0)[m]
        assert eh.FEH1_T4().nfeatofdim(m) == (4, 6, 4, 1)[m]


def test_cell_dof_numbering():
    """numberdofs!([Ux, Uy, P]) with bubble and cell-constant fields: free dofs
    of each space (vertices, then cells) from one past the previous space's
    highest free number, then the data dofs likewise (src/FESpaces.jl:141-173,
    250-259); element dofs are the nodes' then the cell's."""
    mesh = eh.T3block(1.0, 1.0, 3, 2)
    ux, uy = eh.FESpace(mesh, eh.FEH1_T3_BUBBLE()), eh.FESpace(mesh, eh.FEH1_T3_BUBBLE())
    p = eh.FESpace(mesh, eh.FEH1_T3())
    ux.setebc(0, [1, 2], 1, 0.5)
    uy.setebc(2, [3], 1, -1.0)
    p.setebc(0, 1, 1, 0.0)
    eh.numberdofs([ux, uy, p])
    Nn, E = mesh.nnodes, mesh.nelements
    nf_ux, nf_uy, nf_p = Nn - 2 + E, Nn + E - 1, Nn - 1
    assert ux.nunknowns() == nf_ux and uy.nunknowns() == nf_uy and p.nunknowns() == nf_p
    # free numbers: ux nodes, ux cells, uy nodes, uy cells, p nodes; then the data dofs
    assert ux.dofnums[2, 0] == 1 and ux.cell_dofnums[0, 0] == Nn - 2 + 1
    assert uy.dofnums[0, 0] == nf_ux + 1 and uy.cell_dofnums[0, 0] == nf_ux + Nn + 1
    assert p.dofnums[1, 0] == nf_ux + nf_uy + 1
    nu = nf_ux + nf_uy + nf_p
    assert ux.dofnums[0, 0] == nu + 1 and ux.dofnums[1, 0] == nu + 2
    assert uy.cell_dofnums[2, 0] == nu + 3 and p.dofnums[0, 0] == nu + 4
    assert sorted(np.concatenate([s.element_dofs().ravel() for s in (ux, uy, p)]).tolist())[-1] == eh.ndofs([ux, uy, p])
    ed = ux.element_dofs()
    assert ed.shape == (E, 4) and np.array_equal(ed[:, :3], ux.dofnums[mesh.conn - 1, 0])
    assert np.array_equal(ed[:, 3], ux.cell_dofnums[:, 0])
    U = np.zeros(eh.ndofs([ux, uy, p]))
    eh.gathersysvec(U, [ux, uy, p])
    assert U[nu] == 0.5 and U[nu + 2] == -1.0
    q = eh.FESpace(eh.Q4block(1.0, 1.0, 2, 2), eh.FEL2_Q4())
    q.setebc(2, 4, 1, 0.0)
    eh.numberdofs(q)
    assert q.ndofs() == 4 and q.dofnums.shape[0] == 0 and q.cell_dofnums[3, 0] == 4
    with pytest.raises(ValueError):
        q.setebc(0, 1, 1, 0.0)
