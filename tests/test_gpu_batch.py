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


def _small_batch(ctx, specs):
    probs = []
    ctx.batch_begin()
    for k, (kind, n) in enumerate(specs):
        p = W.batch_problem(k, kind, n)
        f = p["fesp"]
        if kind == "heat":
            ctx.batch_add(eh._lib.HEAT, eh._lib.T3, 1, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(),
                          coef_e=p["kappa_e"], kappa=1.0, Q=p["Q"])
        else:
            ctx.batch_add(eh._lib.ELAST, eh._lib.T6, 3, f.mesh.conn, f.mesh.xy, f.dofnums, f.ndofs(),
                          coef_e=p["scale_e"], D=p["D"])
        probs.append(p)
    return probs


def _oracle(p):
    f = p["fesp"]
    if p["kind"] == "heat":
        op = oracle_problem(O, O.HEAT, f, p["qsettings"], kappa_e=p["kappa_e"], Q=p["Q"])
    else:
        op = oracle_problem(O, O.ELAST, f, p["qsettings"], D=p["D"], scale_e=p["scale_e"])
    return O.assemble(op, f.ndofs(), csc=True)


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
        if p["kind"] == "heat":
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
