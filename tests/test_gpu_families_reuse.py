"""The element families on the REUSE method: the sort of the COO stream done
once per mesh + dof map + weak form and kept as a permutation; each assembly
runs the element kernel (values only) and sums every entry's run through it.

FP64: the plain method's result bit for bit, hence the oracle's (pattern,
values, load vector), on the first call (symbolic phase) and on cached calls.
MIXED (FP64 geometry, FP32 gradients, products and element values, FP64
sums): the same pattern, values and load vector within 1e-6 of the oracle's in
the Frobenius norm relative to the oracle's (a few FP32 roundings per value,
each 6e-8 relative).
"""
import numpy as np
import pytest

import oracle as O
from elfel_hip import _lib as L
from helpers import frob_rel

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6


@pytest.fixture
def fctx():
    c = L.Context(0)
    yield c
    c.close()


def _ref(I, J, V, n, layout):
    return O.sparse(I, J, V, n, n, csc=(layout == L.CSC))


def _check(ctx, ref, F_ref, prec, what):
    ptr, ind, val = ctx.copy_sparse(base=0)[:3]
    assert np.array_equal(ptr, ref[0]) and np.array_equal(ind, ref[1]), what
    if prec == L.FP64:
        assert np.array_equal(val, ref[2]), what
    else:
        assert frob_rel(val, ref[2]) <= MIXED_TOL, what
    if F_ref is not None:
        F = ctx.copy_vector(F_ref.size)
        if prec == L.FP64:
            assert np.array_equal(F, F_ref), what
        else:
            assert frob_rel(F, F_ref) <= MIXED_TOL, what


def _t4_mesh(n=7, seed=7):
    conn, xyz = O.t4block(1.0, 1.2, 0.8, n, n + 1, n - 1)
    bnd = np.any((xyz < 1e-12) | (xyz > np.array([1.0, 1.2, 0.8]) - 1e-12), axis=1)
    rng = np.random.default_rng(seed)
    xyz = xyz.copy()
    xyz[~bnd] += rng.uniform(-0.15, 0.15, (int((~bnd).sum()), 3)) / n
    return conn, xyz


@pytest.mark.parametrize("kind,rule", [(L.HEAT, 1), (L.HEAT, 4), (L.MASS, 1), (L.MASS, 4)])
@pytest.mark.parametrize("layout", [L.CSR, L.CSC])
def test_t4_reuse(fctx, kind, rule, layout):
    conn, xyz = _t4_mesh()
    Nn = xyz.shape[0]
    dn = np.random.default_rng(3).permutation(Nn) + 1
    fctx.set_mesh(0, conn, xyz)
    fctx.set_dofs(0, dn, Nn)
    I, J, V, FI, FV = O.fam_scalar_coo(O.HEAT if kind == L.HEAT else O.MASS, O.T4, rule, conn, xyz, dn,
                                       kappa=1.7, Q=-6.0)
    ref = _ref(I, J, V, Nn, layout)
    F_ref = O.vec_assemble(FI, FV, Nn) if kind == L.HEAT else None
    for prec in (L.FP64, L.MIXED, L.FP64):  # symbolic phase, then cached calls
        nnz = fctx.assemble(kind, L.T4, rule, precision=prec, method=L.REUSE, layout=layout, kappa=1.7, Q=-6.0)
        assert nnz == ref[1].size
        assert fctx.assembly_info()[:2] == (L.REUSE, prec)
        _check(fctx, ref, F_ref, prec, (kind, rule, prec))


@pytest.mark.parametrize("kind", [L.HEAT, L.MASS])
def test_l2_reuse(fctx, kind):
    conn, x = O.l2block(2.0, 150)
    Nn = x.shape[0]
    dn = np.random.default_rng(9).permutation(Nn) + 1
    fctx.set_mesh(0, conn, x)
    fctx.set_dofs(0, dn, Nn)
    I, J, V, FI, FV = O.fam_scalar_coo(O.HEAT if kind == L.HEAT else O.MASS, O.L2, 2, conn, x, dn, kappa=1.3, Q=2.0)
    ref = _ref(I, J, V, Nn, L.CSR)
    F_ref = O.vec_assemble(FI, FV, Nn) if kind == L.HEAT else None
    for prec in (L.FP64, L.MIXED):
        fctx.assemble(kind, L.L2, 2, precision=prec, method=L.REUSE, kappa=1.3, Q=2.0)
        _check(fctx, ref, F_ref, prec, (kind, prec))


@pytest.mark.parametrize("fe,gen,rule", [(L.T3, O.t3block, 3), (L.Q4, O.q4block, 2), (L.T6, O.t6block, 3)])
def test_mass_2d_reuse(fctx, fe, gen, rule):
    conn, xy = gen(1.0, 1.0, 13, 11)
    Nn = xy.shape[0]
    dn = np.random.default_rng(2).permutation(Nn) + 1
    fctx.set_mesh(0, conn, xy)
    fctx.set_dofs(0, dn, Nn)
    I, J, V, _, _ = O.fam_scalar_coo(O.MASS, fe, rule, conn, xy, dn, kappa=0.5)
    ref = _ref(I, J, V, Nn, L.CSR)
    for prec in (L.FP64, L.MIXED):
        fctx.assemble(L.MASS, fe, rule, precision=prec, method=L.REUSE, kappa=0.5)
        _check(fctx, ref, None, prec, (fe, prec))


@pytest.mark.parametrize("vfe,rule", [(L.T3B, 3), (L.Q4, 2), (L.T6, 3)])
@pytest.mark.parametrize("layout", [L.CSR, L.CSC])
def test_reddy_reuse(fctx, vfe, rule, layout):
    from test_gpu_families import _reddy_setup
    conn, xy, dux, duy, pfe, total, kw = _reddy_setup(fctx, vfe, 12)
    I, J, V = O.fam_reddy_coo(vfe, pfe, rule, conn, xy, dux, duy, mu=0.9, **kw)
    ref = _ref(I, J, V, total, layout)
    for prec in (L.FP64, L.MIXED, L.FP64):
        nnz = fctx.assemble(L.STOKES_REDDY, vfe, rule, precision=prec, method=L.REUSE, layout=layout, mu=0.9,
                            fe_p=pfe)
        assert nnz == ref[1].size
        _check(fctx, ref, None, prec, (vfe, prec))


def test_reuse_cache_follows_mesh_dofs_and_form(fctx):
    """A new dof map, another weak form on the same mesh, new cell dofs and a
    plain assembly in between: every REUSE call equals the oracle's result for
    what the context holds at that call."""
    conn, xyz = _t4_mesh(5, seed=4)
    Nn = xyz.shape[0]
    fctx.set_mesh(0, conn, xyz)
    for seed in (1, 2):
        dn = np.random.default_rng(seed).permutation(Nn) + 1
        fctx.set_dofs(0, dn, Nn)
        for kind in (L.HEAT, L.MASS, L.HEAT):
            ok = O.HEAT if kind == L.HEAT else O.MASS
            I, J, V, FI, FV = O.fam_scalar_coo(ok, O.T4, 1, conn, xyz, dn, kappa=1.0, Q=1.0)
            fctx.assemble(kind, L.T4, 1, method=L.REUSE, kappa=1.0, Q=1.0)
            _check(fctx, _ref(I, J, V, Nn, L.CSR), O.vec_assemble(FI, FV, Nn) if kind == L.HEAT else None, L.FP64,
                   (seed, kind))
            fctx.assemble(kind, L.T4, 1, method=L.SORT, kappa=1.0, Q=1.0)  # overwrites the plain buffers
    # the bubble's cell dofs renumbered: the Reddy pattern is rebuilt
    from test_gpu_families import _reddy_setup
    conn, xy, dux, duy, pfe, total, kw = _reddy_setup(fctx, L.T3B, 6)
    fctx.assemble(L.STOKES_REDDY, L.T3B, 3, method=L.REUSE, fe_p=pfe)
    E = conn.shape[0]
    eux, euy = kw["de_uy"].copy(), kw["de_ux"].copy()  # swap the bubble's two components
    fctx.set_element_dofs(0, np.stack([eux, euy], 1), total)
    kw.update(de_ux=eux, de_uy=euy)
    I, J, V = O.fam_reddy_coo(L.T3B, pfe, 3, conn, xy, dux, duy, mu=1.0, **kw)
    fctx.assemble(L.STOKES_REDDY, L.T3B, 3, method=L.REUSE, fe_p=pfe)
    _check(fctx, _ref(I, J, V, total, L.CSR), None, L.FP64, "cell dofs")
    assert E > 0


def test_t4_heat_reuse_at_size(fctx):
    """T4 heat on a 40^3 block (384,000 tetrahedra)."""
    conn, xyz = O.t4block(1.0, 1.0, 1.0, 40, 40, 40)
    Nn = xyz.shape[0]
    dn = np.arange(1, Nn + 1)
    fctx.set_mesh(0, conn, xyz)
    fctx.set_dofs(0, dn, Nn)
    I, J, V, FI, FV = O.fam_scalar_coo(O.HEAT, O.T4, 1, conn, xyz, dn, kappa=1.0, Q=-6.0)
    ref, F_ref = _ref(I, J, V, Nn, L.CSR), O.vec_assemble(FI, FV, Nn)
    for prec in (L.FP64, L.MIXED):
        fctx.assemble(L.HEAT, L.T4, 1, precision=prec, method=L.REUSE, kappa=1.0, Q=-6.0)
        _check(fctx, ref, F_ref, prec, prec)


def test_mass_reuse_solves_like_plain(fctx):
    """PCG on the REUSE method's matrix (its pattern lives in the family
    cache): the same iterate bits as on the plain method's matrix."""
    conn, xyz = _t4_mesh(6, seed=5)
    Nn = xyz.shape[0]
    fctx.set_mesh(0, conn, xyz)
    fctx.set_dofs(0, np.arange(1, Nn + 1), Nn)
    b = np.sin(np.arange(Nn, dtype=np.float64))
    out = []
    for m in (L.SORT, L.REUSE):
        fctx.assemble(L.MASS, L.T4, 4, method=m, kappa=1.0)
        out.append(fctx.pcg(Nn, b, np.zeros(Nn), tol=1e-12, maxit=2000))
    assert np.array_equal(out[0][0], out[1][0]) and out[0][1] == out[1][1] and out[1][3] <= 1e-12


def test_empty_mesh_reuse(fctx):
    fctx.set_mesh(0, np.zeros((0, 2), np.int64), np.linspace(0, 1, 5).reshape(-1, 1))
    fctx.set_dofs(0, np.arange(1, 6), 5)
    for prec in (L.FP64, L.MIXED):
        nnz = fctx.assemble(L.HEAT, L.L2, 2, precision=prec, method=L.REUSE, kappa=1.0, Q=3.0)
        ptr, ind, val = fctx.copy_sparse(base=0)[:3]
        assert nnz == 0 and ind.size == 0 and not ptr.any()
        assert np.array_equal(fctx.copy_vector(5), np.zeros(5))


def test_reuse_any_coo_stream(fctx):
    """The families' REUSE method keeps the sort of whatever stream the mesh
    and dof map give: two nodes sharing a dof (a periodic tie), here also two
    nodes of one element, so that an element block holds the same (row,
    column) pair more than once, still equal the oracle's sparse() bit for bit."""
    conn, xyz = _t4_mesh(4, seed=6)
    Nn = xyz.shape[0]
    dn = np.arange(1, Nn + 1)
    dn[7] = dn[3]  # node 8 takes node 4's dof
    dn[conn[0, 1] - 1] = dn[conn[0, 0] - 1]  # two nodes of element 1 share a dof
    used = np.unique(dn)
    remap = np.zeros(used.max() + 1, np.int64)
    remap[used] = np.arange(1, used.size + 1)
    dn = remap[dn]
    n = used.size
    fctx.set_mesh(0, conn, xyz)
    fctx.set_dofs(0, dn, n)
    for kind in (L.HEAT, L.MASS):
        ok = O.HEAT if kind == L.HEAT else O.MASS
        I, J, V, FI, FV = O.fam_scalar_coo(ok, O.T4, 1, conn, xyz, dn, kappa=1.0, Q=2.0)
        ref = _ref(I, J, V, n, L.CSR)
        F_ref = O.vec_assemble(FI, FV, n) if kind == L.HEAT else None
        for prec in (L.FP64, L.MIXED):
            fctx.assemble(kind, L.T4, 1, precision=prec, method=L.REUSE, kappa=1.0, Q=2.0)
            assert fctx.assembly_info()[:2] == (L.REUSE, prec)
            _check(fctx, ref, F_ref, prec, (kind, prec))
