"""Host side of the colliding-flow drivers with the P1b/P1 and Q1/Q0 pairs
(examples/stokes/colliding_flow/p1b_p1.jl:36-240, q1_q0.jl:35-210): mesh,
essential conditions, numbering, solve and the error integrals, around an
assembly passed in (the oracle's or the device's).  Test code only."""
import numpy as np
import scipy.sparse.linalg as spla

import elfel_hip as eh

A = 1.0


def trueux(x, y):
    return 20 * x * y ** 3


def trueuy(x, y):
    return 5 * x ** 4 - 5 * y ** 4


def truep(x, y):
    return 60 * x ** 2 * y - 20 * y ** 3


def spaces(pair, N):
    """p1b_p1.jl:178-197 / q1_q0.jl:130-152: the spaces with their essential
    conditions, numbered by numberdofs!([Ux, Uy, P])."""
    gen = eh.T3block if pair == "p1b" else eh.Q4block
    mesh = eh.meshes.transform(gen(2 * A, 2 * A, N, N), lambda x: x - A)
    ufe = eh.FEH1_T3_BUBBLE if pair == "p1b" else eh.FEH1_Q4
    Ux, Uy = eh.FESpace(mesh, ufe()), eh.FESpace(mesh, ufe())
    inflate = A / N / 100
    for box in ([-A, A, -A, -A], [-A, -A, -A, A], [A, A, -A, A], [-A, A, A, A]):
        vl = mesh.vselect(box, inflate)
        xy = mesh.xy[vl - 1]
        Ux.setebc(0, vl, 1, trueux(xy[:, 0], xy[:, 1]))
        Uy.setebc(0, vl, 1, trueuy(xy[:, 0], xy[:, 1]))
    if pair == "p1b":
        P = eh.FESpace(mesh, eh.FEH1_T3())
        P.setebc(0, mesh.nearestto([0.0, 0.0]), 1, 0.0)
    else:
        P = eh.FESpace(mesh, eh.FEL2_Q4())
        # q1_q0.jl:150-151 pins the cell whose number is that of the node nearest the centre
        P.setebc(2, mesh.nearestto([0.0, 0.0]), 1, 0.0)
    eh.numberdofs([Ux, Uy, P])
    return mesh, Ux, Uy, P


def oracle_coo(O, mesh, Ux, Uy, P, mu=1.0):
    pair_p1b = isinstance(Ux.fe, eh.FEH1_T3_BUBBLE)
    kw = {}
    if pair_p1b:
        kw.update(de_ux=Ux.cell_dofnums[:, 0], de_uy=Uy.cell_dofnums[:, 0], conn_p=mesh.conn, xy_p=mesh.xy,
                  dn_p=P.dofnums[:, 0])
        return O.fam_reddy_coo(O.T3B, O.T3, 3, mesh.conn, mesh.xy, Ux.dofnums[:, 0], Uy.dofnums[:, 0], mu=mu, **kw)
    kw.update(de_p=P.cell_dofnums[:, 0])
    return O.fam_reddy_coo(O.Q4, O.Q4L2, 2, mesh.conn, mesh.xy, Ux.dofnums[:, 0], Uy.dofnums[:, 0], mu=mu, **kw)


def solve(K, Ux, Uy, P):
    """solve! of the drivers: U[1:nu] = K[1:nu, 1:nu] \\ (F - K U)[1:nu], F = 0."""
    sp = [Ux, Uy, P]
    n = eh.ndofs(sp)
    nu = eh.nunknowns(sp)
    U = np.zeros(n)
    eh.gathersysvec(U, sp)
    KU = K @ U
    U[:nu] = spla.spsolve(K[:nu, :nu].tocsc(), -KU[:nu])
    eh.scattersysvec(sp, U)
    return U


def _tri_rule():
    return np.array([[2.0 / 3, 1.0 / 6], [1.0 / 6, 2.0 / 3], [1.0 / 6, 1.0 / 6]]), np.full(3, 1.0 / 6)


def errors_p1b(mesh, Ux, Uy, P, U):
    """evaluate_pressure_error / evaluate_velocity_error of p1b_p1.jl:110-172:
    3-point rule, the velocity with its bubble."""
    pc, w = _tri_rule()
    X = mesh.xy[mesh.conn - 1]  # E x 3 x 2
    ux = U[Ux.element_dofs() - 1]
    uy = U[Uy.element_dofs() - 1]
    pv = U[P.element_dofs() - 1]
    ep = ev = 0.0
    fe = eh.FEH1_T3_BUBBLE()
    for q in range(3):
        N = fe.bfun(pc[q])
        x = np.einsum("k,ekd->ed", N[:3], X)
        J = X[:, 1] - X[:, 0], X[:, 2] - X[:, 0]
        det = J[0][:, 0] * J[1][:, 1] - J[0][:, 1] * J[1][:, 0]
        JxW = det * w[q]
        ph = pv @ N[:3]
        ep += np.sum(JxW * (ph - truep(x[:, 0], x[:, 1])) ** 2)
        ev += np.sum(JxW * ((ux @ N - trueux(x[:, 0], x[:, 1])) ** 2 + (uy @ N - trueuy(x[:, 0], x[:, 1])) ** 2))
    return np.sqrt(ep), np.sqrt(ev)


def errors_q1q0(mesh, Ux, Uy, P, U):
    """evaluate_error of q1_q0.jl:119-128 (relative pressure error at the cell
    centroids) and the velocity L2 error with the 2 x 2 Gauss rule."""
    X = mesh.xy[mesh.conn - 1]  # E x 4 x 2
    cen = X.mean(axis=1)
    p = U[P.cell_dofnums[:, 0] - 1]
    pt = truep(cen[:, 0], cen[:, 1])
    e_p = np.linalg.norm(p - pt) / np.linalg.norm(pt)
    fe = eh.FEH1_Q4()
    g = 0.577350269189626
    ux, uy = U[Ux.element_dofs() - 1], U[Uy.element_dofs() - 1]
    ev = 0.0
    for r in (-g, g):
        for s in (-g, g):
            N, G = fe.bfun([r, s]), fe.bfungradpar([r, s])
            x = np.einsum("k,ekd->ed", N, X)
            J = np.einsum("ekd,km->edm", X, G)
            det = J[:, 0, 0] * J[:, 1, 1] - J[:, 1, 0] * J[:, 0, 1]
            ev += np.sum(det * ((ux @ N - trueux(x[:, 0], x[:, 1])) ** 2 + (uy @ N - trueuy(x[:, 0], x[:, 1])) ** 2))
    return e_p, np.sqrt(ev)
