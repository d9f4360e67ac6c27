"""GPU parity of the element families (FEH1_T4, FEH1_L2, FEH1_T3_BUBBLE,
FEL2_Q4) and the Reddy / mass forms on the plain FP64 method (the REUSE
method: test_gpu_families_reuse.py).

The device result is compared with the oracle's family kernels
(oracle/elfel_oracle_fam.c, pinned in test_oracle_families.py) on the same
mesh arrays: pattern, values and load vector bit for bit (the FP64 kernels
repeat the oracle's operation order; the tolerance written here is 0).
"""
import numpy as np
import pytest

import oracle as O
from elfel_hip import _lib as L

pytestmark = pytest.mark.gpu


@pytest.fixture
def fctx():
    """A context of its own: the families refuse per-element coefficients
    that other tests leave on the shared one."""
    c = L.Context(0)
    yield c
    c.close()


def _ref_csx(I, J, V, n, layout):
    return O.sparse(I, J, V, n, n, csc=(layout == L.CSC))


def _same(dev, ref):
    for a, b, what in zip(dev[:3], ref[:3], ("pointers", "indices", "values")):
        assert a.shape == b.shape and np.array_equal(a, b), what


def _jitter(x, h, frac, seed, fixed):
    rng = np.random.default_rng(seed)
    x = x.copy()
    move = ~fixed
    x[move] += rng.uniform(-frac, frac, (int(move.sum()), x.shape[1])) * h
    return x


def _shuffled_dofs(n, seed):
    return np.random.default_rng(seed).permutation(n) + 1


# ---- heat and mass on the nodal families --------------------------------------
T4_CASES = [(L.HEAT, 1), (L.HEAT, 4), (L.MASS, 1), (L.MASS, 4)]


@pytest.mark.parametrize("kind,rule", T4_CASES)
@pytest.mark.parametrize("layout", [L.CSR, L.CSC])
def test_t4_bitexact(fctx, kind, rule, layout):
    n = 7
    conn, xyz = O.t4block(1.0, 1.2, 0.8, n, n + 1, n - 1)
    bnd = np.any((xyz < 1e-12) | (xyz > np.array([1.0, 1.2, 0.8]) - 1e-12), axis=1)
    xyz = _jitter(xyz, 1.0 / n, 0.15, 7, bnd)
    Nn = xyz.shape[0]
    dn = _shuffled_dofs(Nn, 3)
    fctx.set_mesh(0, conn, xyz)
    fctx.set_dofs(0, dn, Nn)
    nnz = fctx.assemble(kind, L.T4, rule, layout=layout, kappa=1.7, Q=-6.0)
    I, J, V, FI, FV = O.fam_scalar_coo(O.HEAT if kind == L.HEAT else O.MASS, O.T4, rule, conn, xyz, dn,
                                       kappa=1.7, Q=-6.0)
    ref = _ref_csx(I, J, V, Nn, layout)
    dev = fctx.copy_sparse(base=0)
    assert nnz == ref[1].size
    _same(dev, ref)
    if kind == L.HEAT:
        assert np.array_equal(fctx.copy_vector(Nn), O.vec_assemble(FI, FV, Nn))


@pytest.mark.parametrize("kind", [L.HEAT, L.MASS])
@pytest.mark.parametrize("rule", [1, 2, 3])
def test_l2_bitexact(fctx, kind, rule):
    conn, x = O.l2block(2.0, 150)
    x = _jitter(x, 2.0 / 150, 0.3, 5, np.isin(np.arange(x.shape[0]), [0, x.shape[0] - 1]))
    Nn = x.shape[0]
    dn = _shuffled_dofs(Nn, 9)
    fctx.set_mesh(0, conn, x)
    fctx.set_dofs(0, dn, Nn)
    fctx.assemble(kind, L.L2, rule, kappa=10.0, Q=2.0)
    I, J, V, FI, FV = O.fam_scalar_coo(O.HEAT if kind == L.HEAT else O.MASS, O.L2, rule, conn, x, dn,
                                       kappa=10.0, Q=2.0)
    _same(fctx.copy_sparse(base=0), _ref_csx(I, J, V, Nn, L.CSR))
    if kind == L.HEAT:
        assert np.array_equal(fctx.copy_vector(Nn), O.vec_assemble(FI, FV, Nn))


@pytest.mark.parametrize("fe,gen,rule", [(L.T3, O.t3block, 3), (L.Q4, O.q4block, 2), (L.T6, O.t6block, 3),
                                         (L.T3, O.t3block, 1)])
def test_mass_2d_bitexact(fctx, fe, gen, rule):
    conn, xy = gen(1.0, 1.0, 13, 11)
    Nn = xy.shape[0]
    dn = _shuffled_dofs(Nn, 2)
    fctx.set_mesh(0, conn, xy)
    fctx.set_dofs(0, dn, Nn)
    fctx.assemble(L.MASS, fe, rule, kappa=0.5)
    I, J, V, _, _ = O.fam_scalar_coo(O.MASS, fe, rule, conn, xy, dn, kappa=0.5)
    _same(fctx.copy_sparse(base=0), _ref_csx(I, J, V, Nn, L.CSR))


# ---- Reddy-form Stokes ----------------------------------------------------------------
def _reddy_setup(fctx, vfe, N, seed=11):
    """Colliding-flow square [-1, 1]^2 with jittered interior nodes; dof numbers
    of the three spaces shuffled over one system.  Returns what the oracle needs."""
    if vfe == L.T6:
        conn, xy = O.t6block(2.0, 2.0, N, N)
    elif vfe == L.T3B:
        conn, xy = O.t3block(2.0, 2.0, N, N)
    else:
        conn, xy = O.q4block(2.0, 2.0, N, N)
    xy = xy - 1.0
    if vfe != L.T6:  # straight-sided T6 only: its pressure Jacobian comes from the corners
        xy = _jitter(xy, 2.0 / N, 0.2, seed, np.any(np.abs(xy) > 1 - 1e-12, axis=1))
    E, Nn = conn.shape[0], xy.shape[0]
    kw = {}
    if vfe == L.T6:
        c3, xy3 = O.t6tot3(conn, xy)
        npd, pfe = xy3.shape[0], L.T3
    elif vfe == L.T3B:
        c3, xy3, npd, pfe = conn, xy, Nn, L.T3
    else:
        npd, pfe = E, L.Q4L2
    nb = E if vfe == L.T3B else 0
    total = 2 * Nn + 2 * nb + npd
    perm = _shuffled_dofs(total, seed + 1)
    dux, duy = perm[:Nn], perm[Nn:2 * Nn]
    o = 2 * Nn
    fctx.set_mesh(0, conn, xy)
    fctx.set_dofs(0, np.stack([dux, duy], 1), total)
    if vfe == L.T3B:
        eux, euy = perm[o:o + E], perm[o + E:o + 2 * E]
        o += 2 * E
        fctx.set_element_dofs(0, np.stack([eux, euy], 1), total)
        kw.update(de_ux=eux, de_uy=euy)
    if pfe == L.T3:
        dp = perm[o:o + npd]
        fctx.set_mesh(1, c3, xy3)
        fctx.set_dofs(1, dp, total)
        kw.update(conn_p=c3, xy_p=xy3, dn_p=dp)
    else:
        dp = perm[o:o + E]
        fctx.set_element_dofs(1, dp, total)
        kw.update(de_p=dp)
    return conn, xy, dux, duy, pfe, total, kw


@pytest.mark.parametrize("vfe,rule", [(L.T3B, 3), (L.Q4, 2), (L.T6, 3), (L.T3B, 1), (L.Q4, 1)])
@pytest.mark.parametrize("layout", [L.CSR, L.CSC])
def test_reddy_bitexact(fctx, vfe, rule, layout):
    conn, xy, dux, duy, pfe, total, kw = _reddy_setup(fctx, vfe, 12)
    nnz = fctx.assemble(L.STOKES_REDDY, vfe, rule, layout=layout, mu=0.9, fe_p=pfe)
    I, J, V = O.fam_reddy_coo(vfe, pfe, rule, conn, xy, dux, duy, mu=0.9, **kw)
    ref = _ref_csx(I, J, V, total, layout)
    assert nnz == ref[1].size
    _same(fctx.copy_sparse(base=0), ref)


def test_reddy_p1b_at_size(fctx):
    """P1b/P1 on a 200 x 200 block (80,000 elements), the matrix of the
    example's finest run and beyond (examples/stokes/colliding_flow/p1b_p1.jl:
    N = 4 ... 64)."""
    conn, xy, dux, duy, pfe, total, kw = _reddy_setup(fctx, L.T3B, 200)
    fctx.assemble(L.STOKES_REDDY, L.T3B, 3, mu=1.0, fe_p=pfe)
    I, J, V = O.fam_reddy_coo(L.T3B, pfe, 3, conn, xy, dux, duy, mu=1.0, **kw)
    _same(fctx.copy_sparse(base=0), _ref_csx(I, J, V, total, L.CSR))


def test_t4_heat_at_size(fctx):
    """T4 heat on a 40^3 block (384,000 tetrahedra), twice the example's N = 20."""
    conn, xyz = O.t4block(1.0, 1.0, 1.0, 40, 40, 40)
    Nn = xyz.shape[0]
    dn = np.arange(1, Nn + 1)
    fctx.set_mesh(0, conn, xyz)
    fctx.set_dofs(0, dn, Nn)
    fctx.assemble(L.HEAT, L.T4, 1, kappa=1.0, Q=-6.0)
    I, J, V, FI, FV = O.fam_scalar_coo(O.HEAT, O.T4, 1, conn, xyz, dn, kappa=1.0, Q=-6.0)
    _same(fctx.copy_sparse(base=0), _ref_csx(I, J, V, Nn, L.CSR))
    assert np.array_equal(fctx.copy_vector(Nn), O.vec_assemble(FI, FV, Nn))


# ---- error behaviour ---------------------------------------------------------------
def test_family_errors(fctx):
    conn, xyz = O.t4block(1.0, 1.0, 1.0, 2, 2, 2)
    Nn = xyz.shape[0]
    fctx.set_mesh(0, conn, xyz)
    fctx.set_dofs(0, np.arange(1, Nn + 1), Nn)
    with pytest.raises(L.ElfelHipError) as e:  # the sort method is the plain FP64 method
        fctx.assemble(L.HEAT, L.T4, 1, precision=L.MIXED)
    assert e.value.code == L.ENOTSUP
    with pytest.raises(L.ElfelHipError) as e:
        fctx.assemble(L.HEAT, L.T4, 1, method=7)
    assert e.value.code == L.EINVAL
    with pytest.raises(L.ElfelHipError) as e:  # no 3-point tetrahedron rule
        fctx.assemble(L.HEAT, L.T4, 3)
    assert e.value.code == L.ENOTSUP
    with pytest.raises(L.ElfelHipError) as e:  # a 3D mesh is not a Q4 mesh (nen matches, sdim does not)
        fctx.assemble(L.HEAT, L.Q4, 2)
    assert e.value.code == L.EINVAL
    with pytest.raises(L.ElfelHipError) as e:  # 2D weak forms refuse a 3D mesh
        fctx.assemble(L.ELAST, L.Q4, 2, D=np.eye(3))
    assert e.value.code == L.EINVAL
    fctx.set_coefficients(conn.shape[0], np.ones(conn.shape[0]), None, None)
    with pytest.raises(L.ElfelHipError) as e:
        fctx.assemble(L.HEAT, L.T4, 1)
    assert e.value.code == L.ENOTSUP
    fctx.set_coefficients(conn.shape[0], None, None, None)
    fctx.assemble(L.HEAT, L.T4, 1)
    # the bubble needs its cell dofs; a cell dof out of range is a range error
    c3, xy = O.t3block(1.0, 1.0, 3, 3)
    fctx.set_mesh(0, c3, xy)
    fctx.set_dofs(0, np.stack([np.arange(1, 17), np.arange(17, 33)], 1), 100)
    fctx.set_mesh(1, c3, xy)
    fctx.set_dofs(1, np.arange(33, 49), 100)
    with pytest.raises(L.ElfelHipError) as e:
        fctx.assemble(L.STOKES_REDDY, L.T3B, 3, fe_p=L.T3)
    assert e.value.code == L.ESTATE
    with pytest.raises(L.ElfelRangeError):
        fctx.set_element_dofs(0, np.zeros((c3.shape[0], 2), np.int64), 100)
    fctx.set_element_dofs(0, np.arange(49, 49 + 2 * c3.shape[0]).reshape(-1, 2), 100)
    with pytest.raises(L.ElfelHipError) as e:  # (T3B, Q4L2) is not a pair
        fctx.assemble(L.STOKES_REDDY, L.T3B, 3, fe_p=L.Q4L2)
    assert e.value.code == L.ENOTSUP
    fctx.assemble(L.STOKES_REDDY, L.T3B, 3, fe_p=L.T3)


def test_empty_meshes_give_empty_matrix_and_zero_load(fctx):
    """A mesh without elements: no entries, every pointer 0, and the load
    vector of a heat assembly all zeros (start! zeros it, src/Assemblers.jl:194-198)."""
    for fe, conn, x in ((L.L2, np.zeros((0, 2), np.int64), np.linspace(0, 1, 5).reshape(-1, 1)),
                        (L.Q4, np.zeros((0, 4), np.int64), np.zeros((5, 2)))):
        fctx.set_mesh(0, conn, x)
        fctx.set_dofs(0, np.arange(1, 6), 5)
        nnz = fctx.assemble(L.HEAT, fe, 2, kappa=1.0, Q=3.0)
        ptr, ind, val = fctx.copy_sparse(base=0)[:3]
        assert nnz == 0 and ind.size == 0 and not ptr.any()
        assert np.array_equal(fctx.copy_vector(5), np.zeros(5))
