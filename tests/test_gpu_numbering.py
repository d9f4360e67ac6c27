"""Device dof numbering and system-vector gather/scatter against the oracle's
numberdofs (a restatement of src/FESpaces.jl:250-273, pinned by the
reference's numbering tests in test_oracle_pins.py) and the host FESpace."""
import time

import numpy as np
import pytest

import oracle as O
from elfel_hip import _lib as L
from elfel_hip import workloads as W
from elfel_hip.fespaces import gathersysvec, scattersysvec

pytestmark = pytest.mark.gpu


def _rng_flags(rng, n, nc, p):
    return rng.random((n, nc)) < p


@pytest.mark.parametrize("seed", [0, 1, 2])
def test_number_dofs_matches_oracle(ctx, seed):
    rng = np.random.default_rng(seed)
    cases = [
        [_rng_flags(rng, 1000, 2, 0.1), _rng_flags(rng, 300, 1, 0.05)],           # velocity + pressure
        [_rng_flags(rng, 57, 1, 0.5)],                                           # one space
        [_rng_flags(rng, 40, 2, 0.2), np.ones((10, 1), bool), _rng_flags(rng, 25, 1, 0.3)],  # a space without free dofs
        [np.zeros((0, 2), bool), _rng_flags(rng, 33, 2, 0.4)],                   # an empty space
        [np.zeros((20, 1), bool), np.zeros((5, 3), bool)],                        # no data dofs at all
    ]
    for flags in cases:
        dev = ctx.number_dofs(flags)
        ref = O.numberdofs(flags)
        for d, r in zip(dev, ref):
            assert np.array_equal(d, r)


def test_number_dofs_matches_host_spaces(ctx):
    cfg = W.stokes_t6t3(30)
    u, p = cfg["ufesp"], cfg["pfesp"]
    dev = ctx.number_dofs([u.isdatum, p.isdatum])
    assert np.array_equal(dev[0], u.dofnums) and np.array_equal(dev[1], p.dofnums)


def test_gather_scatter_sysvec(ctx):
    cfg = W.heat_t3(20)
    f = cfg["fesp"]
    n = cfg["ndofs"]
    ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
    ctx.set_dofs(0, f.dofnums, n)
    v_host = gathersysvec(np.full(n, -7.0), f)
    v_dev = ctx.gather_sysvec(0, f.dofvals, np.full(n, -7.0))
    assert np.array_equal(v_dev, v_host)
    rng = np.random.default_rng(5)
    w = rng.standard_normal(n)
    before = f.dofvals.copy()
    scattersysvec(f, w)
    got = ctx.scatter_sysvec(0, w, before.ravel().copy())
    assert np.array_equal(got, f.dofvals.ravel())
    # a shorter vector: dofs past its end keep their values (the free part only)
    nu = cfg["nfree"]
    got2 = ctx.scatter_sysvec(0, w[:nu], before.ravel().copy())
    exp2 = before.ravel().copy()
    d = f.dofnums.ravel()
    exp2[d <= nu] = w[d[d <= nu] - 1]
    assert np.array_equal(got2, exp2)
    # gather into a vector too short for the data dofs: BoundsError in the reference
    with pytest.raises(L.ElfelRangeError):
        ctx.gather_sysvec(0, f.dofvals, np.zeros(nu))


def test_number_dofs_c5_size(ctx):
    cfg = W.heat_q4(2048, sigma=1.0)
    f = cfg["fesp"]
    t0 = time.perf_counter()
    dev = ctx.number_dofs([f.isdatum])[0]
    dt = time.perf_counter() - t0
    assert np.array_equal(dev, f.dofnums)
    print(f"number_dofs {f.isdatum.size} dofs: {dt * 1e3:.1f} ms (host round trip included)")
