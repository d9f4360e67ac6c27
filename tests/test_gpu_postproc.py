"""Colliding-flow post-processing on the device (the L2 error norms of
test/test_stokes.jl:438-496 and :675-731) against the reference's own golden
error values (test/test_stokes.jl:779, transcribed into
tests/golden/reference_pins.json) and against the oracle's norms.

The whole chain runs through the C ABI except the sparse direct solve: dof
numbering, FP64 assembly of the vector-Laplacian Taylor-Hood form, then the
device error integrals of the solved system vector.  The device sums the
element terms in a fixed tree order, the oracle left to right; the norms
agree to 1e-12 relative.
"""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import oracle as O
from elfel_hip import _lib as L
from helpers import pins

pytestmark = pytest.mark.gpu

P = pins()
SUM_ORDER_RTOL = 1e-12


def _colliding(ctx, N, kind):
    c6, xy6 = O.t6block(2.0, 2.0, N, N)
    xy6 = xy6 - 1.0
    c3, xy3 = O.t6tot3(c6, xy6)
    bnd = np.isclose(np.abs(xy6[:, 0]), 1) | np.isclose(np.abs(xy6[:, 1]), 1)
    pd = np.zeros((xy3.shape[0], 1), bool)
    pd[np.argmin(np.hypot(xy3[:, 0], xy3[:, 1]))] = True
    du, dp = ctx.number_dofs([np.stack([bnd, bnd], 1), pd])
    nd = du.size + dp.size
    nu = 2 * int((~bnd).sum()) + xy3.shape[0] - 1
    ctx.set_mesh(0, c6, xy6)
    ctx.set_dofs(0, du, nd)
    ctx.set_mesh(1, c3, xy3)
    ctx.set_dofs(1, dp, nd)
    ctx.assemble(kind, L.T6, 3, precision=L.FP64, method=L.SORT, layout=L.CSC, mu=1.0, D=np.diag([2.0, 2.0, 1.0]))
    ptr, ind, val, shape, lay = ctx.copy_sparse(base=0)
    K = sp.csc_matrix((val, ind, ptr), shape=(nd, nd))
    U = np.zeros(nd)
    U[du[bnd, 0] - 1] = 20 * xy6[bnd, 0] * xy6[bnd, 1] ** 3
    U[du[bnd, 1] - 1] = 5 * xy6[bnd, 0] ** 4 - 5 * xy6[bnd, 1] ** 4
    KT = K @ U
    U[:nu] = spla.spsolve(K[:nu, :nu].tocsc(), -KT[:nu])
    return c6, xy6, c3, xy3, du, dp, U


@pytest.mark.parametrize("kind,key", [(L.STOKES_VECLAP, "veclap"), (L.STOKES_GEN, "gen")])
def test_colliding_errors_golden_n4(ctx, kind, key):
    c6, xy6, c3, xy3, du, dp, U = _colliding(ctx, 4, kind)
    ep, ev = ctx.colliding_errors(U)
    gold = P["stokes_veclap_N4_errors"]["value"] if key == "veclap" else P["stokes_reddy_errors"]["value"][0]
    assert np.allclose([ep, ev], gold, rtol=1.5e-8)
    op = O.colliding_p_error(c3, xy3, dp[:, 0], U)
    ov = O.colliding_v_error_vector(c6, xy6, du, U)
    assert abs(ep - op) <= SUM_ORDER_RTOL * op and abs(ev - ov) <= SUM_ORDER_RTOL * ov


@pytest.mark.parametrize("N", [16, 64])
def test_colliding_errors_match_oracle(ctx, N):
    c6, xy6, c3, xy3, du, dp, U = _colliding(ctx, N, L.STOKES_VECLAP)
    ep, ev = ctx.colliding_errors(U)
    op = O.colliding_p_error(c3, xy3, dp[:, 0], U)
    ov = O.colliding_v_error_vector(c6, xy6, du, U)
    assert abs(ep - op) <= SUM_ORDER_RTOL * op and abs(ev - ov) <= SUM_ORDER_RTOL * ov
    # Taylor-Hood P2/P1: velocity error O(h^3), so halving h cuts it about 8x
    if N == 64:
        c6b, xy6b, c3b, xy3b, dub, dpb, Ub = _colliding(ctx, 32, L.STOKES_VECLAP)
        ev32 = ctx.colliding_errors(Ub)[1]
        assert 6.0 < ev32 / ev < 10.0


def test_colliding_errors_zero_vector_and_short_vector(ctx):
    """For U = 0 the norms are the quadrature norms of the exact fields
    themselves (no cancellation in the integrand); a vector shorter than the
    fields' dof count is refused."""
    c6, xy6, c3, xy3, du, dp, U = _colliding(ctx, 8, L.STOKES_VECLAP)
    Z = np.zeros_like(U)
    ep, ev = ctx.colliding_errors(Z)
    assert abs(ep - O.colliding_p_error(c3, xy3, dp[:, 0], Z)) <= SUM_ORDER_RTOL * ep
    assert abs(ev - O.colliding_v_error_vector(c6, xy6, du, Z)) <= SUM_ORDER_RTOL * ev
    with pytest.raises(L.ElfelHipError):
        ctx.colliding_errors(U[:-5])
