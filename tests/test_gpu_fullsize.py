"""Full-size workloads on the GPU, checked through properties that do not
need the CPU oracle at that size:

* the pattern's entry count against the closed-form counts of the meshes
  (T6 node graph 46N^2+16N+1 per component pair, T6 x T3 coupling
  19N^2+10N+1, Q4 (3N+1)^2);
* the FP64 reuse method equals the FP64 plain method bit for bit;
* the mixed method stays within 1e-6 of FP64 (relative Frobenius norm);
* elasticity: the three rigid-body modes lie in the kernel of K, K is
  symmetric; heat: K annihilates constants and sum(F) = sum_e f_e |e|.
The small-size parity against the oracle is in test_gpu_plain.py and
test_gpu_reuse.py.
"""
import numpy as np
import pytest

from elfel_hip import _lib as L
from elfel_hip import workloads as W
from helpers import frob_rel

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6


def _assemble(ctx, kind, fe, rule, precision, method, layout, **kw):
    ctx.assemble(kind, fe, rule, precision=precision, method=method, layout=layout, **kw)
    ptr, ind, val, shape, lay = ctx.copy_sparse(base=0)
    return ptr, ind, val


def _csr_matvec(ptr, ind, val, x):
    rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    return np.bincount(rows, weights=val * x[ind], minlength=len(ptr) - 1)


def test_c2_elasticity_t6_400():
    N = 400
    cfg = W.elast_t6(N, frac=0.15)
    f = cfg["fesp"]
    ctx = L.Context(0)
    ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
    ctx.set_dofs(0, f.dofnums, cfg["ndofs"])
    D = cfg["D"]
    plain = _assemble(ctx, L.ELAST, L.T6, 3, L.FP64, L.SORT, L.CSR, D=D)
    reuse = _assemble(ctx, L.ELAST, L.T6, 3, L.FP64, L.REUSE, L.CSR, D=D)
    mixed = _assemble(ctx, L.ELAST, L.T6, 3, L.MIXED, L.REUSE, L.CSR, D=D)
    ctx.close()
    assert len(plain[1]) == 4 * (46 * N * N + 16 * N + 1)
    for r in (reuse, mixed):
        assert np.array_equal(r[0], plain[0]) and np.array_equal(r[1], plain[1])
    assert np.array_equal(reuse[2], plain[2])
    err = frob_rel(mixed[2], plain[2])
    print(f"C2 mixed vs FP64: {err:.2e}")
    assert err <= MIXED_TOL
    # rigid-body modes: dof d of node k holds component d % 2 (numbering is
    # per node, components minor), so build the modes through dofnums
    ptr, ind, val = plain
    n = cfg["ndofs"]
    dn = f.dofnums - 1
    xy = f.mesh.xy
    scale = np.abs(val).max()
    for mode in ("tx", "ty", "rot"):
        u = np.zeros(n)
        if mode == "tx":
            u[dn[:, 0]] = 1.0
        elif mode == "ty":
            u[dn[:, 1]] = 1.0
        else:
            u[dn[:, 0]] = -xy[:, 1]
            u[dn[:, 1]] = xy[:, 0]
        r = _csr_matvec(ptr, ind, val, u)
        assert np.abs(r).max() <= 1e-10 * scale * np.abs(u).max(), mode
    # symmetry: K[i, j] == K[j, i] (the CSR transpose via CSC of the same data)
    import scipy.sparse as sp
    K = sp.csr_matrix((val, ind, ptr), shape=(n, n))
    d = (K - K.T).tocsr()
    assert np.abs(d.data).max(initial=0.0) <= 1e-12 * scale


def test_c3_stokes_t6t3_500():
    N = 500
    cfg = W.stokes_t6t3(N)
    u, p = cfg["ufesp"], cfg["pfesp"]
    n = cfg["ndofs"]
    ctx = L.Context(0)
    ctx.set_mesh(0, u.mesh.conn, u.mesh.xy)
    ctx.set_dofs(0, u.dofnums, n)
    ctx.set_mesh(1, p.mesh.conn, p.mesh.xy)
    ctx.set_dofs(1, p.dofnums, n)
    kw = dict(D=cfg["D"], mu=1.0)
    plain = _assemble(ctx, L.STOKES_GEN, L.T6, 3, L.FP64, L.SORT, L.CSC, **kw)
    reuse = _assemble(ctx, L.STOKES_GEN, L.T6, 3, L.FP64, L.REUSE, L.CSC, **kw)
    mixed = _assemble(ctx, L.STOKES_GEN, L.T6, 3, L.MIXED, L.REUSE, L.CSC, **kw)
    ctx.close()
    nuu = 4 * (46 * N * N + 16 * N + 1)
    nup = 2 * (19 * N * N + 10 * N + 1)
    assert len(plain[1]) == nuu + 2 * nup
    for r in (reuse, mixed):
        assert np.array_equal(r[0], plain[0]) and np.array_equal(r[1], plain[1])
    assert np.array_equal(reuse[2], plain[2])
    err = frob_rel(mixed[2], plain[2])
    print(f"C3 mixed vs FP64: {err:.2e}")
    assert err <= MIXED_TOL


def test_c5_heat_q4_1024_properties():
    N = 1024
    cfg = W.heat_q4(N, sigma=1.0)
    f = cfg["fesp"]
    n = cfg["ndofs"]
    ctx = L.Context(0)
    ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
    ctx.set_dofs(0, f.dofnums, n)
    ctx.set_coefficients(f.mesh.nelements, cfg["kappa_e"], cfg["Q_e"], None)
    res = {}
    for key, prec, meth in (("plain", L.FP64, L.SORT), ("reuse", L.FP64, L.REUSE), ("mixed", L.MIXED, L.REUSE)):
        ptr, ind, val = _assemble(ctx, L.HEAT, L.Q4, 2, prec, meth, L.CSR)
        res[key] = (ptr, ind, val, ctx.copy_vector(n))
    ctx.close()
    plain, reuse, mixed = res["plain"], res["reuse"], res["mixed"]
    assert len(plain[1]) == (3 * N + 1) ** 2
    for r in (reuse, mixed):
        assert np.array_equal(r[0], plain[0]) and np.array_equal(r[1], plain[1])
    assert np.array_equal(reuse[2], plain[2]) and np.array_equal(reuse[3], plain[3])
    assert frob_rel(mixed[2], plain[2]) <= MIXED_TOL
    assert frob_rel(mixed[3], plain[3]) <= MIXED_TOL
    ptr, ind, val, F = plain
    rs = _csr_matvec(ptr, ind, val, np.ones(n))
    assert np.abs(rs).max() <= 1e-12 * np.abs(val).max()
    h = 1.0 / N
    assert abs(F.sum() - cfg["Q_e"].sum() * h * h) <= 1e-12 * np.abs(cfg["Q_e"]).sum() * h * h


def test_c4_full_batch_against_plain():
    specs = W.batch_specs()
    probs = [W.batch_problem(k, kind, nn) for k, (kind, nn) in enumerate(specs)]
    ctx = L.Context(0)
    ctx.batch_begin()
    for p in probs:
        f = p["fesp"]
        if p["kind"] == "heat":
            ctx.batch_add(L.HEAT, L.T3, 1, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(), coef_e=p["kappa_e"],
                          kappa=1.0, Q=p["Q"])
        else:
            ctx.batch_add(L.ELAST, L.T6, 3, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(), coef_e=p["scale_e"],
                          D=p["D"])
    ctx.batch_build(autotune=False)
    ctx.batch_assemble()
    st = ctx.batch_stats()
    assert st["nprob"] == 256
    c2 = L.Context(0)
    for k in range(0, 256, 17):
        p = probs[k]
        f = p["fesp"]
        c2.set_mesh(0, f.mesh.conn, f.mesh.xy)
        c2.set_dofs(0, f.dofnums, f.ndofs())
        if p["kind"] == "heat":
            c2.set_coefficients(f.mesh.nelements, p["kappa_e"], None, None)
            ref = _assemble(c2, L.HEAT, L.T3, 1, L.FP64, L.SORT, L.CSC, kappa=1.0, Q=p["Q"])
            Fref = c2.copy_vector(f.ndofs())
        else:
            c2.set_coefficients(f.mesh.nelements, None, None, p["scale_e"])
            ref = _assemble(c2, L.ELAST, L.T6, 3, L.FP64, L.SORT, L.CSC, D=p["D"])
        got = ctx.batch_copy(k)
        assert np.array_equal(got[0], ref[0]) and np.array_equal(got[1], ref[1]), k
        assert frob_rel(got[2], ref[2]) <= MIXED_TOL, k
        if p["kind"] == "heat":
            assert frob_rel(got[3], Fref) <= MIXED_TOL, k
    c2.close()
    ctx.close()
