"""The reference's element-family drivers through the host mirror and the
device assembly (GPU):

* examples/heat/poisson/t4.jl: T4 block N = 20, Q = -6, the exact nodal
  temperature 1 + x^2 + 2 y^2 to a mean error <= 1e-9 (the example's check);
* examples/stokes/colliding_flow/p1b_p1.jl and q1_q0.jl: the device-assembled
  matrix equals the oracle's bit for bit and the solution carries the same
  errors;
* examples/klein_gordon/klein_gordon.jl: M, K and D = m^2 M on the L2 mesh
  against the 1D closed forms.
"""
import numpy as np
import pytest
import scipy.sparse.linalg as spla

import elfel_hip as eh
import flow_helpers as FH
import oracle as O
from helpers import pins

pytestmark = pytest.mark.gpu


def test_t4_example_nodal_exactness():
    N = 20
    mesh = eh.T4block(1.0, 1.0, 1.0, N, N, N)
    fesp = eh.FESpace(mesh, eh.FEH1_T4())
    vl = mesh.boundary_nodes()
    X = mesh.xy[vl - 1]
    tempf = lambda x, y: 1.0 + x ** 2 + 2.0 * y ** 2
    fesp.setebc(0, vl, 1, tempf(X[:, 0], X[:, 1]))
    fesp.numberfreedofs()
    fesp.numberdatadofs()
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
    eh.scattersysvec(fesp, T)
    Tn = fesp.dofvals[:, 0]
    exact = tempf(mesh.xy[:, 0], mesh.xy[:, 1])
    assert np.mean(np.abs(Tn - exact)) <= pins()["heat_t4_nodal_mean_error_max"]["value"]


@pytest.mark.parametrize("pair", ["p1b", "q1q0"])
def test_colliding_flow_drivers_match_oracle(pair):
    N = 32
    mesh, Ux, Uy, P = FH.spaces(pair, N)
    n = eh.ndofs([Ux, Uy, P])
    am = eh.start(eh.SysmatAssemblerHIP(0.0), n, n)
    qs = dict(kind="default", npts=3) if pair == "p1b" else dict(kind="default", order=2)
    eh.integrate_stokes_reddy(am, Ux, Uy, P, qs, mu=1.0, layout="csr")
    K = eh.finish(am).tocsr()
    I, J, V = FH.oracle_coo(O, mesh, Ux, Uy, P)
    Kr = O.sparse_scipy(I, J, V, n, n, csc=False)
    K.sort_indices()
    assert np.array_equal(K.indptr, Kr.indptr) and np.array_equal(K.indices, Kr.indices)
    assert np.array_equal(K.data, Kr.data)
    U = FH.solve(K, Ux, Uy, P)
    e = (FH.errors_p1b if pair == "p1b" else FH.errors_q1q0)(mesh, Ux, Uy, P, U)
    Ur = FH.solve(Kr, *FH.spaces(pair, N)[1:])
    er = (FH.errors_p1b if pair == "p1b" else FH.errors_q1q0)(mesh, Ux, Uy, P, Ur)
    assert np.array_equal(np.array(e), np.array(er))
    assert e[1] < (0.06 if pair == "p1b" else 0.05)


def test_klein_gordon_matrices():
    L, n, m2 = 2.0, 150, 10.0
    mesh = eh.L2block(L, n)
    psi = eh.FESpace(mesh, eh.FEH1_L2())
    eh.numberdofs(psi)
    nd = psi.ndofs()
    qs = dict(kind="Gauss", order=2)
    mats = []
    for form in ("M", "K", "D"):
        am = eh.start(eh.SysmatAssemblerHIP(0.0), nd, nd)
        if form == "K":
            eh.integrate_heat(am, None, psi, qs, kappa=1.0, Q=0.0)
        else:
            eh.integrate_mass(am, psi, qs, coef=1.0 if form == "M" else m2)
        mats.append(eh.finish(am).toarray())
    M, K, D = mats
    h = L / n
    tri = lambda d, o: np.diag(np.full(n + 1, d)) + np.diag(np.full(n, o), 1) + np.diag(np.full(n, o), -1)
    Mx = tri(4 * h / 6, h / 6)
    Mx[0, 0] = Mx[-1, -1] = 2 * h / 6
    Kx = tri(2 / h, -1 / h)
    Kx[0, 0] = Kx[-1, -1] = 1 / h
    assert np.allclose(M, Mx, rtol=1e-12, atol=1e-15)
    assert np.allclose(K, Kx, rtol=1e-12, atol=1e-12)
    assert np.allclose(D, m2 * Mx, rtol=1e-12, atol=1e-14)
