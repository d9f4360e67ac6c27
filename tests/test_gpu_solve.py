"""GPU Jacobi-PCG (FP64, and FP32 inner iterations inside FP64 iterative
refinement) on assembled matrices, checked against direct solves."""
import numpy as np
import pytest
import scipy.sparse as sp
import scipy.sparse.linalg as spla

import elfel_hip as eh
from elfel_hip import workloads as W
from helpers import pins

pytestmark = pytest.mark.gpu


def _assemble_heat(cfg, **kw):
    fesp, n = cfg["fesp"], cfg["ndofs"]
    am = eh.start(eh.SysmatAssemblerHIP(), n, n)
    av = eh.start(eh.SysvecAssemblerHIP(), n)
    eh.integrate_heat(am, av, fesp, cfg["qsettings"], method="reuse", layout="csr", **kw)
    return am, av


@pytest.mark.parametrize("precision", [eh._lib.FP64, eh._lib.MIXED])
def test_pcg_heat_t3_golden_temperature(precision):
    cfg = W.heat_t3(4)
    am, av = _assemble_heat(cfg, kappa=1.0, Q=-6.0)
    F = eh.finish(av)
    n, nu = cfg["ndofs"], cfg["nfree"]
    T = np.zeros(n)
    eh.gathersysvec(T, cfg["fesp"])
    x, it, outer, rr = am.context().pcg(nu, F, T, precision=precision, tol=1e-12, maxit=1000)
    T[:nu] = x[:nu]
    assert rr <= 1e-12
    assert np.allclose(T, pins()["heat_t3_N4_T"]["value"], rtol=1.5e-8)


@pytest.mark.parametrize("precision", [eh._lib.FP64, eh._lib.MIXED])
def test_pcg_heat_q4_lognormal(precision):
    cfg = W.heat_q4(96, sigma=1.0)
    am, av = _assemble_heat(cfg, kappa_e=cfg["kappa_e"], Q_e=cfg["Q_e"])
    F = eh.finish(av)
    n, nu = cfg["ndofs"], cfg["nfree"]
    x0 = np.zeros(n)
    x, it, outer, rr = am.context().pcg(nu, F, x0, precision=precision, tol=1e-10, maxit=20000)
    K = am.finish()
    A = K[:nu, :nu].tocsc()
    xd = spla.spsolve(A, F[:nu])
    res = np.linalg.norm(F[:nu] - A @ x[:nu]) / np.linalg.norm(F[:nu])
    print(f"PCG precision={precision}: {it} iterations, {outer} refinements, relres {rr:.2e} (host {res:.2e})")
    assert rr <= 1e-10 and res <= 2e-10
    assert np.linalg.norm(x[:nu] - xd) / np.linalg.norm(xd) <= 1e-8


@pytest.mark.parametrize("precision", [eh._lib.FP64, eh._lib.MIXED])
def test_pcg_elasticity_t6_with_data_dofs(precision):
    cfg = W.elast_t6(16, frac=0.1)
    fesp, n, nu = cfg["fesp"], cfg["ndofs"], cfg["nfree"]
    am = eh.start(eh.SysmatAssemblerHIP(), n, n)
    eh.integrate_elasticity(am, fesp, cfg["qsettings"], cfg["D"], method="reuse", layout="csr")
    U = np.zeros(n)
    eh.gathersysvec(U, fesp)
    F = np.zeros(n)
    x, it, outer, rr = am.context().pcg(nu, F, U, precision=precision, tol=1e-10, maxit=50000)
    K = am.finish()
    ref = U.copy()
    KT = K @ ref
    ref[:nu] = spla.spsolve(K[:nu, :nu].tocsc(), (F - KT)[:nu])
    assert rr <= 1e-10
    assert np.linalg.norm(x[:nu] - ref[:nu]) / np.linalg.norm(ref[:nu]) <= 1e-7
