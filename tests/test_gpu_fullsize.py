"""The benchmark workloads at full size on the GPU, against the oracle.

* C2 elasticity T6 400^2, C3 Stokes T6/T3 500^2, C5 heat Q4 4096^2: the oracle
  (element loop + sparse(), one thread) assembles the same mesh arrays; the
  FP64 plain and REUSE methods equal it bit for bit (pattern, values, load
  vector; tolerance 0), the mixed method is within 1e-6 relative in the
  Frobenius norm (the mixed path's accuracy gate).
* C4: all 256 batched problems (one persistent launch) against the oracle.
* C5: both Jacobi-PCG variants to ||b - A x|| / ||b|| <= 1e-10, the residual
  recomputed on the host in FP64 from the copied matrix.
Size-independent properties are checked as well: the pattern's closed-form
entry count, rigid-body modes and symmetry (elasticity), K 1 = 0 and
sum F = sum_e f_e |e| (heat).
"""
import numpy as np
import pytest
import scipy.sparse as sp

import oracle as O
from elfel_hip import _lib as L
from elfel_hip import workloads as W
from helpers import frob_rel, oracle_problem

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6


def _assemble(ctx, kind, fe, rule, precision, method, layout, **kw):
    ctx.assemble(kind, fe, rule, precision=precision, method=method, layout=layout, **kw)
    ptr, ind, val, shape, lay = ctx.copy_sparse(base=0)
    return ptr, ind, val


def _same_pattern(a, ref):
    assert np.array_equal(a[0], ref[0]), "pointers differ from the oracle"
    assert np.array_equal(a[1], ref[1]), "indices differ from the oracle"


def _csr_matvec(ptr, ind, val, x):
    rows = np.repeat(np.arange(len(ptr) - 1), np.diff(ptr))
    return np.bincount(rows, weights=val * x[ind], minlength=len(ptr) - 1)


@pytest.mark.timeout(600)
def test_c2_elasticity_t6_400_vs_oracle():
    N = 400
    cfg = W.elast_t6(N, frac=0.15)
    f, n, D = cfg["fesp"], cfg["ndofs"], cfg["D"]
    ref = O.assemble(oracle_problem(O, O.ELAST, f, cfg["qsettings"], D=D), n, csc=True)
    assert ref[1].size == 4 * (46 * N * N + 16 * N + 1)
    ctx = L.Context(0)
    ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
    ctx.set_dofs(0, f.dofnums, n)
    for prec, meth in ((L.FP64, L.SORT), (L.FP64, L.REUSE), (L.MIXED, L.REUSE)):
        got = _assemble(ctx, L.ELAST, L.T6, 3, prec, meth, L.CSC, D=D)
        assert ctx.assembly_info()[:2] == (meth, prec)
        _same_pattern(got, ref)
        if prec == L.FP64:
            assert np.array_equal(got[2], ref[2]), f"FP64 method {meth} differs from the oracle"
        else:
            err = frob_rel(got[2], ref[2])
            print(f"C2 mixed vs oracle: {err:.2e}")
            assert err <= MIXED_TOL
    ctx.close()
    # rigid-body modes (dof d of node k holds component d % 2) and symmetry
    ptr, ind, val = ref[:3]
    K = sp.csc_matrix((val, ind, ptr), shape=(n, n)).tocsr()
    dn, xy = f.dofnums - 1, f.mesh.xy
    scale = np.abs(val).max()
    for mode in ("tx", "ty", "rot"):
        u = np.zeros(n)
        if mode == "tx":
            u[dn[:, 0]] = 1.0
        elif mode == "ty":
            u[dn[:, 1]] = 1.0
        else:
            u[dn[:, 0]], u[dn[:, 1]] = -xy[:, 1], xy[:, 0]
        assert np.abs(K @ u).max() <= 1e-10 * scale * np.abs(u).max(), mode
    assert np.abs((K - K.T).tocsr().data).max(initial=0.0) <= 1e-12 * scale


@pytest.mark.timeout(600)
def test_c3_stokes_t6t3_500_vs_oracle():
    N = 500
    cfg = W.stokes_t6t3(N)
    u, p, n = cfg["ufesp"], cfg["pfesp"], cfg["ndofs"]
    kw = dict(D=cfg["D"], mu=1.0)
    ref = O.assemble(oracle_problem(O, O.STOKES_GEN, u, cfg["qsettings"], pfesp=p, **kw), n, csc=True)
    assert ref[1].size == 4 * (46 * N * N + 16 * N + 1) + 4 * (19 * N * N + 10 * N + 1)
    ctx = L.Context(0)
    ctx.set_mesh(0, u.mesh.conn, u.mesh.xy)
    ctx.set_dofs(0, u.dofnums, n)
    ctx.set_mesh(1, p.mesh.conn, p.mesh.xy)
    ctx.set_dofs(1, p.dofnums, n)
    for prec, meth in ((L.FP64, L.SORT), (L.FP64, L.REUSE), (L.MIXED, L.REUSE)):
        got = _assemble(ctx, L.STOKES_GEN, L.T6, 3, prec, meth, L.CSC, **kw)
        assert ctx.assembly_info()[:2] == (meth, prec)
        _same_pattern(got, ref)
        if prec == L.FP64:
            assert np.array_equal(got[2], ref[2]), f"FP64 method {meth} differs from the oracle"
        else:
            err = frob_rel(got[2], ref[2])
            print(f"C3 mixed vs oracle: {err:.2e}")
            assert err <= MIXED_TOL
    ctx.close()


@pytest.mark.timeout(900)
def test_c4_all_256_problems_vs_oracle():
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
    assert ctx.batch_stats()["nprob"] == 256
    worst = 0.0
    for k, p in enumerate(probs):
        f = p["fesp"]
        if p["kind"] == "heat":
            op = oracle_problem(O, O.HEAT, f, p["qsettings"], kappa_e=p["kappa_e"], Q=p["Q"])
        else:
            op = oracle_problem(O, O.ELAST, f, p["qsettings"], D=p["D"], scale_e=p["scale_e"])
        ref = O.assemble(op, f.ndofs(), csc=True)
        got = ctx.batch_copy(k)
        _same_pattern(got, ref)
        e = frob_rel(got[2], ref[2])
        assert e <= MIXED_TOL, (k, e)
        if p["kind"] == "heat":
            e = max(e, frob_rel(got[3], ref[3]))
            assert e <= MIXED_TOL, (k, e)
        worst = max(worst, e)
    print(f"C4: 256 problems, worst mixed error vs oracle {worst:.2e}")
    ctx.close()


@pytest.fixture(scope="module")
def c5():
    cfg = W.heat_q4(4096, sigma=1.0)
    yield cfg


@pytest.mark.timeout(900)
def test_c5_heat_q4_4096_vs_oracle(c5):
    N = 4096
    f, n = c5["fesp"], c5["ndofs"]
    op = oracle_problem(O, O.HEAT, f, c5["qsettings"], kappa_e=c5["kappa_e"], Q_e=c5["Q_e"])
    ref = O.assemble(op, n, csc=False)
    del op
    assert ref[1].size == (3 * N + 1) ** 2
    ctx = L.Context(0)
    ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
    ctx.set_dofs(0, f.dofnums, n)
    ctx.set_coefficients(f.mesh.nelements, c5["kappa_e"], c5["Q_e"], None)
    for prec, meth in ((L.FP64, L.SORT), (L.FP64, L.REUSE), (L.MIXED, L.REUSE)):
        got = _assemble(ctx, L.HEAT, L.Q4, 2, prec, meth, L.CSR)
        F = ctx.copy_vector(n)
        assert ctx.assembly_info()[:2] == (meth, prec)
        _same_pattern(got, ref)
        if prec == L.FP64:
            assert np.array_equal(got[2], ref[2]), f"FP64 method {meth} differs from the oracle"
            assert np.array_equal(F, ref[3]), f"FP64 method {meth}: load vector differs from the oracle"
        else:
            ev, eF = frob_rel(got[2], ref[2]), frob_rel(F, ref[3])
            print(f"C5 mixed vs oracle: matrix {ev:.2e}, load vector {eF:.2e}")
            assert ev <= MIXED_TOL and eF <= MIXED_TOL
        del got
    ctx.close()
    ptr, ind, val, F = ref
    assert np.abs(_csr_matvec(ptr, ind, val, np.ones(n))).max() <= 1e-12 * np.abs(val).max()
    h = 1.0 / N
    assert abs(F.sum() - c5["Q_e"].sum() * h * h) <= 1e-12 * np.abs(c5["Q_e"]).sum() * h * h


@pytest.mark.timeout(900)
def test_c5_pcg_both_precisions_to_1e10(c5):
    """Both PCG variants on the FP64 REUSE matrix of the full 4096^2 problem
    (16,769,025 unknowns) to relative residual 1e-10; the residual is
    recomputed on the host in FP64 from the copied CSR matrix (data dofs are
    0, so b = F on the free rows)."""
    f, n, nu = c5["fesp"], c5["ndofs"], c5["nfree"]
    ctx = L.Context(0)
    ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
    ctx.set_dofs(0, f.dofnums, n)
    ctx.set_coefficients(f.mesh.nelements, c5["kappa_e"], c5["Q_e"], None)
    ctx.assemble(L.HEAT, L.Q4, 2, precision=L.FP64, method=L.REUSE, layout=L.CSR)
    F = ctx.copy_vector(n)
    ptr, ind, val = ctx.copy_sparse(base=0)[:3]
    Kf = sp.csr_matrix((val[:ptr[nu]], ind[:ptr[nu]], ptr[:nu + 1]), shape=(nu, n))
    del ind, val
    b = F[:nu]
    sols = {}
    for prec in (L.MIXED, L.FP64):
        x, it, outer, rr = ctx.pcg(nu, F, np.zeros(n), precision=prec, tol=1e-10, maxit=60000)
        host = float(np.linalg.norm(b - Kf @ x) / np.linalg.norm(b))
        print(f"C5 PCG {'mixed' if prec == L.MIXED else 'fp64'}: {it} iterations, {outer} FP64 residual "
              f"updates, device relres {rr:.3e}, host relres {host:.3e}")
        assert rr <= 1e-10
        assert host <= 1.05e-10
        sols[prec] = x[:nu]
    ctx.close()
    d = np.linalg.norm(sols[L.MIXED] - sols[L.FP64]) / np.linalg.norm(sols[L.FP64])
    print(f"C5 PCG: mixed vs FP64 solution {d:.2e}")
    assert d <= 1e-7
