"""Batched uneven problems in one persistent fused launch replayed from a
hipGraph: every problem's matrix and load vector against the oracle, and graph
replays (with eager work and synchronisations between them) against the first
result."""
import numpy as np
import pytest

import elfel_hip as eh
import oracle as O
from elfel_hip import workloads as W
from helpers import frob_rel, oracle_problem

pytestmark = pytest.mark.gpu


def _heat_q4_problem(k, n, sigma=0.5, frac=0.2):
    """A jittered Q4 heat problem with lognormal conductivity, built like the
    batch's T3 heat problems (examples/heat/poisson/q4.jl)."""
    mesh = eh.Q4block(1.0, 1.0, n, n)
    W.jitter(mesh, 1.0 / n, frac, W.SEED, stream=500 + 3 * k)
    fesp = eh.FESpace(mesh, eh.FEH1_Q4())
    fesp.setebc(0, mesh.boundary_nodes(), 1, 0.0)
    eh.numberdofs(fesp)
    coef = np.exp(sigma * W.normal(W.SEED, 501 + 3 * k, mesh.nelements))
    return dict(kind="heatq4", mesh=mesh, fesp=fesp, kappa_e=coef, Q=-0.5, qsettings=dict(kind="Gauss", order=2),
                ndofs=fesp.ndofs())


def _small_batch(ctx, specs, precision=eh._lib.MIXED, layout=eh._lib.CSC):
    probs = []
    ctx.batch_begin()
    for k, (kind, n) in enumerate(specs):
        p = _heat_q4_problem(k, n) if kind == "heatq4" else W.batch_problem(k, kind, n)
        f = p["fesp"]
        if kind == "heat":
            ctx.batch_add(eh._lib.HEAT, eh._lib.T3, 1, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(),
                          coef_e=p["kappa_e"], kappa=1.0, Q=p["Q"], precision=precision, layout=layout)
        elif kind == "heatq4":
            ctx.batch_add(eh._lib.HEAT, eh._lib.Q4, 2, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(),
                          coef_e=p["kappa_e"], kappa=1.0, Q=p["Q"], precision=precision, layout=layout)
        else:
            ctx.batch_add(eh._lib.ELAST, eh._lib.T6, 3, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(),
                          coef_e=p["scale_e"], D=p["D"], precision=precision, layout=layout)
        probs.append(p)
    return probs


def _oracle(p, csc=True):
    f = p["fesp"]
    if p["kind"] in ("heat", "heatq4"):
        op = oracle_problem(O, O.HEAT, f, p["qsettings"], kappa_e=p["kappa_e"], Q=p["Q"])
    else:
        op = oracle_problem(O, O.ELAST, f, p["qsettings"], D=p["D"], scale_e=p["scale_e"])
    return O.assemble(op, f.ndofs(), csc=csc)


@pytest.mark.parametrize("autotune", [False, True])
def test_batch_matches_oracle_and_replays(ctx, autotune):
    specs = [("heat", 7), ("elast", 5), ("heat", 20), ("elast", 12), ("elast", 3), ("heat", 33)]
    probs = _small_batch(ctx, specs)
    ctx.batch_build(autotune=autotune)
    ctx.batch_assemble()
    first = []
    for k, p in enumerate(probs):
        ptr, ind, val, F = ctx.batch_copy(k)
        rptr, rind, rval, rF = _oracle(p)
        assert np.array_equal(ptr, rptr) and np.array_equal(ind, rind)
        assert frob_rel(val, rval) <= 1e-6
        if p["kind"] != "elast":
            assert np.abs(F - rF).max() <= 1e-6 * np.abs(rF).max()
        first.append(val.copy())
    for _ in range(3):  # replays with eager host work and syncs in between
        ctx.synchronize()
        ctx.batch_assemble()
        for k in range(len(probs)):
            assert np.array_equal(ctx.batch_copy(k)[2], first[k])
    st = ctx.batch_stats()
    assert st["nprob"] == len(specs)
    info = ctx.batch_info(0)
    assert info["block"] in (128, 256)


MIXED_SPECS = [("heat", 7), ("heatq4", 9), ("elast", 5), ("heat", 20), ("heatq4", 31), ("elast", 12),
               ("elast", 3), ("heat", 33), ("heatq4", 2)]


@pytest.mark.parametrize("layout", [eh._lib.CSC, eh._lib.CSR])
def test_batch_fp64_with_q4_equals_oracle_bitwise(layout):
    """An FP64 batch of heat T3, heat Q4 and elasticity T6 problems: every
    matrix and load vector equals the oracle's bit for bit (tolerance 0), as
    the single-problem FP64 method does."""
    ctx = eh._lib.Context(0)
    probs = _small_batch(ctx, MIXED_SPECS, precision=eh._lib.FP64, layout=layout)
    ctx.batch_build(autotune=False)
    ctx.batch_assemble()
    for k, p in enumerate(probs):
        ptr, ind, val, F = ctx.batch_copy(k)
        rptr, rind, rval, rF = _oracle(p, csc=(layout == eh._lib.CSC))
        assert np.array_equal(ptr, rptr) and np.array_equal(ind, rind), k
        assert np.array_equal(val, rval), (k, p["kind"])
        if p["kind"] != "elast":
            assert np.array_equal(F, rF), k
    ctx.close()


def test_batch_mixed_with_q4_and_replays():
    """A mixed batch with heat Q4 problems: within 1e-6 of the oracle, and
    graph replays give the first result's bits."""
    ctx = eh._lib.Context(0)
    probs = _small_batch(ctx, MIXED_SPECS)
    ctx.batch_build(autotune=True)
    ctx.batch_assemble()
    first = []
    for k, p in enumerate(probs):
        ptr, ind, val, F = ctx.batch_copy(k)
        rptr, rind, rval, rF = _oracle(p)
        assert np.array_equal(ptr, rptr) and np.array_equal(ind, rind), k
        assert frob_rel(val, rval) <= 1e-6, k
        if p["kind"] != "elast":
            assert frob_rel(F, rF) <= 1e-6, k
        first.append(val.copy())
    for _ in range(2):
        ctx.synchronize()
        ctx.batch_assemble()
        for k in range(len(probs)):
            assert np.array_equal(ctx.batch_copy(k)[2], first[k])
    ctx.close()


def test_batch_refuses_mixed_precisions_and_flags():
    ctx = eh._lib.Context(0)
    p = W.batch_problem(0, "heat", 4)
    f = p["fesp"]
    ctx.batch_begin()
    ctx.batch_add(eh._lib.HEAT, eh._lib.T3, 1, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(), kappa=1.0, Q=1.0)
    with pytest.raises(eh._lib.ElfelHipError) as e:
        ctx.batch_add(eh._lib.HEAT, eh._lib.T3, 1, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(), kappa=1.0, Q=1.0,
                      precision=eh._lib.FP64)
    assert e.value.code == eh._lib.ENOTSUP
    with pytest.raises(eh._lib.ElfelHipError) as e:  # Q4 with a 1-point rule is not a batch kind
        q = _heat_q4_problem(1, 3)
        g = q["fesp"]
        ctx.batch_add(eh._lib.HEAT, eh._lib.Q4, 1, g.mesh.conn, g.mesh.xy, g.dofnums, g.ndofs(), kappa=1.0, Q=1.0)
    assert e.value.code == eh._lib.ENOTSUP
    ctx.close()
