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


@pytest.mark.parametrize("block", [128, 256])
def test_batch_fp64_each_tile_form_bitwise(monkeypatch, block):
    """The batch at each workgroup size, no timing: heat tiles are row-owner
    tiles at 256 and task tiles at 128; both equal the oracle bit for bit."""
    monkeypatch.setenv("ELFEL_HIP_BATCH_BLOCK", str(block))
    ctx = eh._lib.Context(0)
    probs = _small_batch(ctx, MIXED_SPECS, precision=eh._lib.FP64)
    ctx.batch_build(autotune=False)
    ctx.batch_assemble()
    assert ctx.batch_info(0)["block"] == block
    for k, p in enumerate(probs):
        ptr, ind, val, F = ctx.batch_copy(k)
        rptr, rind, rval, rF = _oracle(p)
        assert np.array_equal(ptr, rptr) and np.array_equal(ind, rind), k
        assert np.array_equal(val, rval), (k, p["kind"])
        if p["kind"] != "elast":
            assert np.array_equal(F, rF), k
    ctx.close()


def _fan(K):
    """T3 fan: a centre node in K triangles inside two rings of K nodes."""
    xy = [(0.0, 0.0)]
    for r in (1, 2):
        for k in range(K):
            t = 2 * np.pi * (k + 0.3 * (r % 2)) / K
            xy.append((r * np.cos(t), r * np.sin(t)))
    node = lambda r, k: 1 + (r - 1) * K + (k % K)  # noqa: E731
    conn = [(0, node(1, k), node(1, k + 1)) for k in range(K)]
    for k in range(K):
        a, b, c, d = node(1, k), node(1, k + 1), node(2, k + 1), node(2, k)
        conn += [(a, d, c), (a, c, b)]
    return np.array(conn, np.int64) + 1, np.array(xy)


@pytest.mark.parametrize("precision", [eh._lib.FP64, eh._lib.MIXED])
def test_batch_heat_kind_falls_back_to_task_tiles(monkeypatch, precision):
    """A heat T3 problem with a node in 12 elements (more than a row record
    holds, fewer than the task tiles' 16) puts every heat T3 problem of the
    batch on task tiles; all still match the oracle (FP64: bit for bit)."""
    monkeypatch.setenv("ELFEL_HIP_BATCH_BLOCK", "256")
    ctx = eh._lib.Context(0)
    probs = _small_batch(ctx, [("heat", 9), ("heatq4", 7), ("elast", 4)], precision=precision)
    conn, xy = _fan(12)
    dn = np.arange(1, xy.shape[0] + 1, dtype=np.int64).reshape(-1, 1)
    Q = probs[0]["Q"]  # problems of one kind share their uniform coefficients
    ctx.batch_add(eh._lib.HEAT, eh._lib.T3, 1, conn, xy, dn, xy.shape[0], kappa=1.0, Q=Q, precision=precision,
                  layout=eh._lib.CSC)
    ctx.batch_build(autotune=False)
    ctx.batch_assemble()
    refs = [_oracle(p) for p in probs]
    refs.append(O.assemble(O.make_problem(O.HEAT, O.T3, 1, conn, xy, dn, kappa=1.0, Q=Q), xy.shape[0], csc=True))
    for k, ref in enumerate(refs):
        ptr, ind, val, F = ctx.batch_copy(k)
        assert np.array_equal(ptr, ref[0]) and np.array_equal(ind, ref[1]), k
        if precision == eh._lib.FP64:
            assert np.array_equal(val, ref[2]), k
        else:
            assert frob_rel(val, ref[2]) <= 1e-6, k
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
