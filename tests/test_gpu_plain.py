"""GPU parity of the plain FP64 method and of the generic drop-in assemblers.

The device result is compared with the oracle on the same mesh arrays: the
sparsity pattern must match exactly and the values within 1e-12 relative
(the FP64 kernels repeat the reference's operation order, so the observed
difference is expected to be 0).
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import elfel_hip as eh
import oracle as O
from elfel_hip import workloads as W
from helpers import compare_sparse, oracle_problem, pins

pytestmark = pytest.mark.gpu


def _device_csx(am):
    return am.finish(as_scipy=False)


def test_heat_t3_golden_temperature():
    """test/test_heat.jl:89-111 end to end with the device assembly."""
    N = 4
    mesh = eh.T3block(1.0, 1.0, N, N)
    fesp = eh.FESpace(mesh, eh.FEH1_T3())
    b = mesh.boundary_nodes()
    xy = mesh.xy[b - 1]
    fesp.setebc(0, b, 1, 1.0 + xy[:, 0] ** 2 + 2.0 * xy[:, 1] ** 2)
    eh.numberdofs(fesp)
    n = fesp.ndofs()
    am = eh.start(eh.SysmatAssemblerHIP(0.0), n, n)
    av = eh.start(eh.SysvecAssemblerHIP(0.0), n)
    eh.integrate_heat(am, av, fesp, dict(kind="default"), kappa=1.0, Q=-6.0)
    K, F = eh.finish(am), eh.finish(av)
    T = np.zeros(n)
    eh.gathersysvec(T, fesp)
    nu = fesp.nunknowns()
    KT = K @ T
    T[:nu] = spla.spsolve(K[:nu, :nu].tocsc(), F[:nu] - KT[:nu])
    gold = np.array(pins()["heat_t3_N4_T"]["value"])
    assert np.allclose(T, gold, rtol=1.5e-8, atol=0)


@pytest.mark.parametrize("layout", ["csc", "csr"])
def test_plain_heat_t3_c1_bitexact(layout):
    cfg = W.heat_t3(100)
    fesp = cfg["fesp"]
    n = cfg["ndofs"]
    am = eh.start(eh.SysmatAssemblerHIP(), n, n)
    av = eh.start(eh.SysvecAssemblerHIP(), n)
    eh.integrate_heat(am, av, fesp, cfg["qsettings"], kappa=1.0, Q=-6.0, layout=layout)
    dev = _device_csx(am)
    F = eh.finish(av)
    p = oracle_problem(O, O.HEAT, fesp, cfg["qsettings"], kappa=1.0, Q=-6.0)
    ref = O.assemble(p, n, csc=(layout == "csc"))
    err = compare_sparse(dev, ref, rtol=1e-12)
    assert dev[0][-1] == 70601  # nnz = 7N^2 + 6N + 1
    assert np.array_equal(F, ref[3]), "load vector differs"
    print(f"heat-T3 C1 {layout}: nnz={dev[0][-1]} max rel err={err:.2e}")


def test_plain_elast_t6_jittered():
    cfg = W.elast_t6(24, frac=0.15)
    fesp = cfg["fesp"]
    n = cfg["ndofs"]
    am = eh.start(eh.SysmatAssemblerHIP(), n, n)
    eh.integrate_elasticity(am, fesp, cfg["qsettings"], cfg["D"])
    dev = _device_csx(am)
    p = oracle_problem(O, O.ELAST, fesp, cfg["qsettings"], D=cfg["D"])
    ref = O.assemble(p, n, csc=True)
    err = compare_sparse(dev, ref, rtol=1e-12)
    assert dev[0][-1] == 46 * 24 * 24 * 4 + 16 * 24 * 4 + 4  # 4 * (46N^2 + 16N + 1)
    print(f"elasticity-T6 N=24: max rel err={err:.2e}")


@pytest.mark.parametrize("form", ["gen", "veclap"])
def test_plain_stokes(form):
    cfg = W.stokes_t6t3(10)
    u, pf = cfg["ufesp"], cfg["pfesp"]
    n = cfg["ndofs"]
    am = eh.start(eh.SysmatAssemblerHIP(), n, n)
    eh.integrate_stokes(am, u, pf, cfg["qsettings"], form=form, D=cfg["D"], mu=1.0)
    dev = _device_csx(am)
    kind = O.STOKES_GEN if form == "gen" else O.STOKES_VECLAP
    p = oracle_problem(O, kind, u, cfg["qsettings"], pfesp=pf, D=cfg["D"], mu=1.0)
    ref = O.assemble(p, n, csc=True)
    compare_sparse(dev, ref, rtol=1e-12)


def test_plain_heat_q4_lognormal():
    cfg = W.heat_q4(40, sigma=1.0)
    fesp = cfg["fesp"]
    n = cfg["ndofs"]
    am = eh.start(eh.SysmatAssemblerHIP(), n, n)
    av = eh.start(eh.SysvecAssemblerHIP(), n)
    eh.integrate_heat(am, av, fesp, cfg["qsettings"], kappa_e=cfg["kappa_e"], Q_e=cfg["Q_e"])
    dev = _device_csx(am)
    F = eh.finish(av)
    p = oracle_problem(O, O.HEAT, fesp, cfg["qsettings"], kappa_e=cfg["kappa_e"], Q_e=cfg["Q_e"])
    ref = O.assemble(p, n, csc=True)
    compare_sparse(dev, ref, rtol=1e-12)
    assert dev[0][-1] == (3 * 40 + 1) ** 2
    assert np.array_equal(F, ref[3])


# ---- generic drop-in (host COO -> device sparse) ------------------------------
def _blocks():
    pz = pins()["assembler_blocks"]
    return [(np.array(pz["m1"]), pz["gi1"], pz["gj1"]), (np.array(pz["m2"]), pz["gi2"], pz["gj2"])]


def test_assembler_entries_B():
    """test/test_assemblers.jl:5-36: scalar assemble! with duplicates."""
    a = eh.start(eh.SysmatAssemblerHIP(0.0), 7, 7)
    for m, gi, gj in _blocks():
        for j in range(m.shape[1]):
            for i in range(m.shape[0]):
                eh.assemble(a, gi[i], gj[j], m[i, j])
    A = eh.finish(a).toarray()
    B = np.array(pins()["assembler_B"]["value"])
    assert np.linalg.norm(A - B) / np.linalg.norm(B) < 1e-5


def test_assembler_local_and_transposed_B():
    """test/test_assemblers.jl:41-136: LocalMatrixAssembler and transpose."""
    B = np.array(pins()["assembler_B"]["value"])
    a = eh.start(eh.SysmatAssemblerHIP(0.0), 7, 7)
    for m, gi, gj in _blocks():
        la = eh.LocalMatrixAssembler(m.shape[0], m.shape[1], 0.0)
        eh.init(la, gi, gj)
        la.M += m
        eh.assemble(a, la)
    assert np.linalg.norm(eh.finish(a).toarray() - B) / np.linalg.norm(B) < 1e-5
    a = eh.start(eh.SysmatAssemblerHIP(0.0), 7, 7)
    for m, gi, gj in _blocks():
        la = eh.LocalMatrixAssembler(m.shape[1], m.shape[0], 0.0)
        eh.init(la, gj, gi)
        la.M += m.T
        eh.assemble(a, la.T)
    assert np.linalg.norm(eh.finish(a).toarray() - B) / np.linalg.norm(B) < 1e-5


@pytest.mark.parametrize("layout", ["csr", "csc"])
def test_random_coo_matches_oracle_sparse(ctx, layout):
    """Stable sort + in-order duplicate sums: bit-identical to the oracle's sparse()."""
    rng = np.random.default_rng(7)
    nrow, ncol, n = 3000, 2500, 400000
    I = rng.integers(1, nrow + 1, n)
    J = rng.integers(1, ncol + 1, n)
    V = rng.standard_normal(n)
    V[::97] = 0.0  # explicit zeros stay structural
    csc = layout == "csc"
    ctx.sparse(I, J, V, nrow, ncol, eh._lib.CSC if csc else eh._lib.CSR)
    dev = ctx.copy_sparse(base=0)
    ref = O.sparse(I, J, V, nrow, ncol, csc=csc)
    assert np.array_equal(dev[0], ref[0]) and np.array_equal(dev[1], ref[1])
    assert np.array_equal(dev[2], ref[2]), "duplicate sums are not in stream order"


def test_copy_one_based(ctx):
    ctx.sparse([1, 2, 2], [1, 1, 1], [1.0, 2.0, 3.0], 2, 2, eh._lib.CSC)
    ptr, ind, val, shape, lay = ctx.copy_sparse(base=1)
    assert ptr.tolist() == [1, 3, 3] and ind.tolist() == [1, 2] and val.tolist() == [1.0, 5.0]


def test_vector_assembler_stream_order(ctx):
    rng = np.random.default_rng(3)
    idx = rng.integers(1, 501, 20000)
    val = rng.standard_normal(20000)
    out = ctx.vector(idx, val, 600)
    assert np.array_equal(out, O.vec_assemble(idx, val, 600))


def test_feiterator_brute_force_dense_equals_sparse():
    """test/test_feiterators.jl:88-120 on the qmesh fixture."""
    import os
    from helpers import GOLDEN
    mesh = eh.load(os.path.join(GOLDEN, "qmesh.mesh"))
    fesp = eh.FESpace(mesh, eh.FEH1_T3())
    fesp.setebc(0, [1, 4, 7, 10], 1, 0.0)
    eh.numberdofs(fesp)
    A = np.array(pins()["feiterator_A"]["value"])
    ass = eh.start(eh.SysmatAssemblerHIP(0.0), 12, 12)
    ke = eh.LocalMatrixAssembler(3, 3, 0.0)
    D = np.zeros((12, 12))
    for d in fesp.element_dofs():
        eh.init(ke, d, d)
        ke.M += A
        eh.assemble(ass, ke)
        D[np.ix_(d - 1, d - 1)] += A
    S = eh.finish(ass).toarray()
    assert np.allclose(D, S, rtol=0, atol=1e-15)


# ---- error behaviour ----------------------------------------------------------
def test_out_of_range_raises(ctx):
    with pytest.raises(IndexError):
        ctx.sparse([1, 8], [1, 1], [1.0, 1.0], 7, 7, eh._lib.CSC)
    with pytest.raises(IndexError):
        ctx.sparse([0], [1], [1.0], 7, 7, eh._lib.CSR)
    with pytest.raises(IndexError):
        ctx.vector([1, 11], [1.0, 2.0], 10)
    # the context stays usable after an error
    ctx.sparse([1], [1], [2.5], 1, 1, eh._lib.CSR)
    assert ctx.copy_sparse()[2].tolist() == [2.5]


def test_unnumbered_dof_raises():
    mesh = eh.T3block(1.0, 1.0, 2, 2)
    fesp = eh.FESpace(mesh, eh.FEH1_T3())  # never numbered: all dof numbers are 0
    am = eh.start(eh.SysmatAssemblerHIP(), 9, 9)
    with pytest.raises(IndexError):
        eh.integrate_heat(am, None, fesp, dict(kind="default"))


def test_empty_stream(ctx):
    ctx.sparse([], [], [], 5, 4, eh._lib.CSC)
    ptr, ind, val, shape, lay = ctx.copy_sparse()
    assert ptr.tolist() == [0] * 5 and ind.size == 0 and shape == (5, 4)
