"""GPU parity of the REUSE method (cached pattern + fused owner-computes kernel).

FP64 reuse repeats the reference's arithmetic and summation order: its result
must equal the oracle's bit for bit (checked with tolerance 0).  The mixed
method (FP32 element arithmetic, FP64 sums) must stay within 1e-6 relative in
the Frobenius norm (the accuracy gate of the mixed path), and its largest entrywise error is
reported.
"""
import numpy as np
import pytest

import elfel_hip as eh
import oracle as O
from elfel_hip import workloads as W
from helpers import compare_sparse, frob_rel, oracle_problem

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6


def _heat(cfg, precision, layout, am=None, av=None, **coef):
    fesp, n = cfg["fesp"], cfg["ndofs"]
    am = am or eh.start(eh.SysmatAssemblerHIP(), n, n)
    av = av or eh.start(eh.SysvecAssemblerHIP(), n)
    eh.integrate_heat(am, av, fesp, cfg["qsettings"], precision=precision, method="reuse", layout=layout, **coef)
    return am.finish(as_scipy=False), eh.finish(av), am, av


@pytest.mark.parametrize("layout", ["csc", "csr"])
def test_reuse_heat_t3_c1(layout):
    cfg = W.heat_t3(100)
    p = oracle_problem(O, O.HEAT, cfg["fesp"], cfg["qsettings"], kappa=1.0, Q=-6.0)
    ref = O.assemble(p, cfg["ndofs"], csc=(layout == "csc"))
    dev, F, _, _ = _heat(cfg, "fp64", layout, kappa=1.0, Q=-6.0)
    compare_sparse(dev, ref, rtol=0.0)
    assert np.array_equal(F, ref[3])
    dev, F, _, _ = _heat(cfg, "mixed", layout, kappa=1.0, Q=-6.0)
    compare_sparse(dev, ref, rtol=1e-5)
    assert frob_rel(dev[2], ref[2]) <= MIXED_TOL
    assert np.abs(F - ref[3]).max() <= 1e-6 * np.abs(ref[3]).max()


@pytest.mark.parametrize("layout", ["csc", "csr"])
def test_reuse_elast_t6_jittered(layout):
    cfg = W.elast_t6(30, frac=0.15)
    fesp, n = cfg["fesp"], cfg["ndofs"]
    p = oracle_problem(O, O.ELAST, fesp, cfg["qsettings"], D=cfg["D"])
    ref = O.assemble(p, n, csc=(layout == "csc"))
    for prec, tol in (("fp64", 0.0), ("mixed", None)):
        am = eh.start(eh.SysmatAssemblerHIP(), n, n)
        eh.integrate_elasticity(am, fesp, cfg["qsettings"], cfg["D"], precision=prec, method="reuse", layout=layout)
        dev = am.finish(as_scipy=False)
        if tol is not None:
            compare_sparse(dev, ref, rtol=tol)
        else:
            compare_sparse(dev, ref, rtol=1e-5)
            e = frob_rel(dev[2], ref[2])
            print(f"elasticity-T6 mixed {layout}: frobenius rel err {e:.2e}")
            assert e <= MIXED_TOL


@pytest.mark.parametrize("form", ["gen", "veclap"])
@pytest.mark.parametrize("layout", ["csc", "csr"])
def test_reuse_stokes(form, layout):
    cfg = W.stokes_t6t3(12)
    u, pf, n = cfg["ufesp"], cfg["pfesp"], cfg["ndofs"]
    kind = O.STOKES_GEN if form == "gen" else O.STOKES_VECLAP
    p = oracle_problem(O, kind, u, cfg["qsettings"], pfesp=pf, D=cfg["D"], mu=1.0)
    ref = O.assemble(p, n, csc=(layout == "csc"))
    for prec in ("fp64", "mixed"):
        am = eh.start(eh.SysmatAssemblerHIP(), n, n)
        eh.integrate_stokes(am, u, pf, cfg["qsettings"], form=form, D=cfg["D"], mu=1.0, precision=prec,
                            method="reuse", layout=layout)
        dev = am.finish(as_scipy=False)
        if prec == "fp64":
            compare_sparse(dev, ref, rtol=0.0)
        else:
            compare_sparse(dev, ref, rtol=1e-5)
            assert frob_rel(dev[2], ref[2]) <= MIXED_TOL


def test_reuse_heat_q4_coefficients_change_between_calls():
    """The pattern is built once; new coefficients only rerun the numeric kernel."""
    cfg = W.heat_q4(48, sigma=1.0)
    fesp, n = cfg["fesp"], cfg["ndofs"]
    am = eh.start(eh.SysmatAssemblerHIP(), n, n)
    av = eh.start(eh.SysvecAssemblerHIP(), n)
    for seed_shift in (0, 1, 2):
        ke = cfg["kappa_e"] * (1.0 + 0.5 * seed_shift)
        qe = cfg["Q_e"] - 0.1 * seed_shift
        dev, F, am, av = _heat(cfg, "fp64", "csc", am=am, av=av, kappa_e=ke, Q_e=qe)
        p = oracle_problem(O, O.HEAT, fesp, cfg["qsettings"], kappa_e=ke, Q_e=qe)
        ref = O.assemble(p, n, csc=True)
        compare_sparse(dev, ref, rtol=0.0)
        assert np.array_equal(F, ref[3])


def test_reuse_rebuilds_after_mesh_change():
    n = None
    am = None
    for N in (10, 14):
        cfg = W.elast_t6(N, frac=0.1)
        fesp, n = cfg["fesp"], cfg["ndofs"]
        am = eh.start(am or eh.SysmatAssemblerHIP(), n, n)
        eh.integrate_elasticity(am, fesp, cfg["qsettings"], cfg["D"], precision="fp64", method="reuse")
        dev = am.finish(as_scipy=False)
        p = oracle_problem(O, O.ELAST, fesp, cfg["qsettings"], D=cfg["D"])
        compare_sparse(dev, O.assemble(p, n, csc=True), rtol=0.0)


def test_reuse_equals_plain_bitwise():
    cfg = W.elast_t6(20, frac=0.15)
    fesp, n = cfg["fesp"], cfg["ndofs"]
    res = []
    for meth in ("sort", "reuse"):
        am = eh.start(eh.SysmatAssemblerHIP(), n, n)
        eh.integrate_elasticity(am, fesp, cfg["qsettings"], cfg["D"], precision="fp64", method=meth)
        res.append(am.finish(as_scipy=False))
    for a, b in zip(res[0][:3], res[1][:3]):
        assert np.array_equal(a, b)


def test_mixed_accuracy_on_sliver_elements():
    """Slivers of aspect ratio 1e4 stay within the mixed tolerance (the mixed
    kernels form J, det J and 1/det J in FP64) and are counted by the
    geometry statistics."""
    cfg = W.elast_t6(16, frac=0.0)
    fesp, n = cfg["fesp"], cfg["ndofs"]
    xy = fesp.mesh.xy
    # the upper half of the square becomes slivers of aspect ratio 1e4 (no inversion)
    up = xy[:, 1] > 0.5
    xy[up, 1] = 0.5 + (xy[up, 1] - 0.5) * 1e-4
    p = oracle_problem(O, O.ELAST, fesp, cfg["qsettings"], D=cfg["D"])
    ref = O.assemble(p, n, csc=True)
    am = eh.start(eh.SysmatAssemblerHIP(), n, n)
    eh.integrate_elasticity(am, fesp, cfg["qsettings"], cfg["D"], precision="mixed", method="reuse")
    dev = am.finish(as_scipy=False)
    mindet, nbad = am.context().geometry_stats()
    compare_sparse(dev, ref, rtol=1e-4)
    assert frob_rel(dev[2], ref[2]) <= MIXED_TOL
    assert nbad > 0 and mindet > 0


@pytest.mark.parametrize("layout", ["csc", "csr"])
@pytest.mark.parametrize("numbering", ["node order", "shuffled"])
def test_reuse_heat_band_tiles(layout, numbering):
    """Scalar meshes of 2^18 rows or more are packed into band tiles (rows
    taken in a locality order of their nodes, laid out in dof order, written
    back run by run).  On a jittered T3 mesh of 271,441 nodes, in node order
    (long runs) and with shuffled dof numbers (runs of one row), FP64 reuse
    equals the oracle bit for bit and the mixed method stays within 1e-6."""
    N = 520
    mesh = eh.T3block(1.0, 1.0, N, N)
    xy = mesh.xy.copy()
    inner = np.all((xy > 1e-12) & (xy < 1 - 1e-12), axis=1)
    rng = np.random.default_rng(17)
    xy[inner] += rng.uniform(-0.2, 0.2, (int(inner.sum()), 2)) / N
    mesh = eh.Mesh(mesh.conn, xy, "T3")
    fesp = eh.FESpace(mesh, eh.FEH1_T3())
    eh.numberdofs(fesp)
    n = fesp.ndofs()
    assert n >= 1 << 18
    if numbering == "shuffled":
        fesp.dofnums = (rng.permutation(n) + 1).reshape(-1, 1)
    cfg = dict(fesp=fesp, ndofs=n, qsettings=dict(kind="default", npts=1))
    p = oracle_problem(O, O.HEAT, fesp, cfg["qsettings"], kappa=1.3, Q=-2.0)
    ref = O.assemble(p, n, csc=(layout == "csc"))
    dev, F, _, _ = _heat(cfg, "fp64", layout, kappa=1.3, Q=-2.0)
    compare_sparse(dev, ref, rtol=0.0)
    assert np.array_equal(F, ref[3])
    dev, F, _, _ = _heat(cfg, "mixed", layout, kappa=1.3, Q=-2.0)
    assert frob_rel(dev[2], ref[2]) <= MIXED_TOL
    assert frob_rel(F, ref[3]) <= MIXED_TOL
