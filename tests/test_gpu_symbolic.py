"""The REUSE method's symbolic phase for vector problems (elasticity, Stokes)
built on the device (symbolic.hip) against the oracle: the pattern and the
FP64 values bit for bit, the mixed values within 1e-6, on dof numberings that
split a node's two rows (one free and one data dof per node), scatter them
(random permutation), and on the host builder's own numbering; the fallbacks
(a shared dof, a node in more than 16 elements) and the first-call time."""
import numpy as np
import pytest

import elfel_hip as eh
import oracle as O
from elfel_hip import _lib as L
from elfel_hip import workloads as W
from helpers import frob_rel

pytestmark = pytest.mark.gpu

MIXED_TOL = 1e-6


def _elast_mesh(N, numbering, seed=3):
    mesh = eh.T6block(1.0, 1.0, N, N)
    W.jitter(mesh, 1.0 / N, 0.15, W.SEED)
    fesp = eh.FESpace(mesh, eh.FEH1_T6(), 2)
    x, y = mesh.xy[:, 0], mesh.xy[:, 1]
    if numbering in ("split", "perm"):
        # one component fixed on the bottom edge, the other on the left edge:
        # those nodes have one free and one data dof
        fesp.setebc(0, np.nonzero(np.abs(y) < 1e-12)[0] + 1, 2, 0.0)
        fesp.setebc(0, np.nonzero(np.abs(x) < 1e-12)[0] + 1, 1, 0.0)
    eh.numberdofs(fesp)
    dn = fesp.dofnums.copy()
    n = fesp.ndofs()
    if numbering == "perm":
        p = np.random.default_rng(seed).permutation(n) + 1
        dn = p[dn - 1]
    return mesh, dn, n


@pytest.mark.parametrize("numbering", ["plain", "split", "perm"])
@pytest.mark.parametrize("layout", [L.CSR, L.CSC])
def test_device_symbolic_elasticity(monkeypatch, numbering, layout):
    mesh, dn, n = _elast_mesh(23, numbering)
    D = W.plane_strain_D(1.0, 0.3)
    ref = O.assemble(O.make_problem(O.ELAST, O.T6, 3, mesh.conn, mesh.xy, dn, D=D), n, csc=(layout == L.CSC))
    for dev in ("1", "0"):
        monkeypatch.setenv("ELFEL_HIP_DEVICE_SYMBOLIC", dev)
        ctx = L.Context(0)
        ctx.set_mesh(0, mesh.conn, mesh.xy)
        ctx.set_dofs(0, dn, n)
        for prec in (L.FP64, L.MIXED):
            ctx.assemble(L.ELAST, L.T6, 3, precision=prec, method=L.REUSE, layout=layout, D=D)
            assert ctx.assembly_info()[:2] == (L.REUSE, prec)
            ptr, ind, val = ctx.copy_sparse(base=0)[:3]
            assert np.array_equal(ptr, ref[0]) and np.array_equal(ind, ref[1]), (dev, prec)
            if prec == L.FP64:
                assert np.array_equal(val, ref[2]), (dev, numbering)
            else:
                assert frob_rel(val, ref[2]) <= MIXED_TOL
        ctx.close()


@pytest.mark.parametrize("form", [L.STOKES_GEN, L.STOKES_VECLAP])
@pytest.mark.parametrize("perm", [False, True])
def test_device_symbolic_stokes(form, perm):
    cfg = W.stokes_t6t3(14)
    u, p, n = cfg["ufesp"], cfg["pfesp"], cfg["ndofs"]
    du, dp = u.dofnums.copy(), p.dofnums.copy()
    if perm:
        q = np.random.default_rng(5).permutation(n) + 1
        du, dp = q[du - 1], q[dp - 1]
    okind = O.STOKES_GEN if form == L.STOKES_GEN else O.STOKES_VECLAP
    ref = O.assemble(O.make_problem(okind, O.T6, 3, u.mesh.conn, u.mesh.xy, du, conn_p=p.mesh.conn, dofnums_p=dp,
                                    D=cfg["D"], mu=1.0), n, csc=True)
    ctx = L.Context(0)
    ctx.set_mesh(0, u.mesh.conn, u.mesh.xy)
    ctx.set_dofs(0, du, n)
    ctx.set_mesh(1, p.mesh.conn, p.mesh.xy)
    ctx.set_dofs(1, dp, n)
    for prec in (L.FP64, L.MIXED):
        ctx.assemble(form, L.T6, 3, precision=prec, method=L.REUSE, layout=L.CSC, D=cfg["D"], mu=1.0)
        assert ctx.assembly_info()[:2] == (L.REUSE, prec)
        ptr, ind, val = ctx.copy_sparse(base=0)[:3]
        assert np.array_equal(ptr, ref[0]) and np.array_equal(ind, ref[1])
        if prec == L.FP64:
            assert np.array_equal(val, ref[2])
        else:
            assert frob_rel(val, ref[2]) <= MIXED_TOL
    ctx.close()


def test_device_symbolic_shared_dof_falls_back():
    """Two nodes with one dof (a periodic tie): the runs of the cached sort
    order are summed in FP64, with the oracle's pattern and values."""
    mesh, dn, n = _elast_mesh(6, "plain")
    dn = dn.copy()
    dn[5] = dn[4]  # node 6 takes node 5's dofs
    used = np.unique(dn)
    remap = np.zeros(used.max() + 1, np.int64)
    remap[used] = np.arange(1, used.size + 1)
    dn = remap[dn]
    n = used.size
    D = W.plane_strain_D(1.0, 0.3)
    ref = O.assemble(O.make_problem(O.ELAST, O.T6, 3, mesh.conn, mesh.xy, dn, D=D), n, csc=True)
    ctx = L.Context(0)
    ctx.set_mesh(0, mesh.conn, mesh.xy)
    ctx.set_dofs(0, dn, n)
    ctx.assemble(L.ELAST, L.T6, 3, precision=L.MIXED, method=L.REUSE, layout=L.CSC, D=D)
    m, p, note = ctx.assembly_info()
    assert (m, p) == (L.REUSE, L.FP64) and "shared" in note, note
    ptr, ind, val = ctx.copy_sparse(base=0)[:3]
    assert np.array_equal(ptr, ref[0]) and np.array_equal(ind, ref[1]) and np.array_equal(val, ref[2])
    ctx.close()


def test_device_symbolic_first_call_time():
    """C2 (elasticity T6 400x400): the first mixed REUSE call (symbolic phase +
    numeric kernel) against the FP64 plain method; printed, and bounded by
    the plain method's time plus 20 ms (allocation and launch jitter)."""
    cfg = W.elast_t6(400, frac=0.15)
    f, n, D = cfg["fesp"], cfg["ndofs"], cfg["D"]
    ctx = L.Context(0)
    ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
    ctx.set_dofs(0, f.dofnums, n)
    ctx.assemble(L.ELAST, L.T6, 3, precision=L.FP64, method=L.SORT, layout=L.CSC, D=D)  # warm-up
    ctx.synchronize()
    import time
    t0 = time.perf_counter()
    ctx.assemble(L.ELAST, L.T6, 3, precision=L.FP64, method=L.SORT, layout=L.CSC, D=D)
    ctx.synchronize()
    t_plain = (time.perf_counter() - t0) * 1e3
    t0 = time.perf_counter()
    ctx.assemble(L.ELAST, L.T6, 3, precision=L.MIXED, method=L.REUSE, layout=L.CSC, D=D)
    ctx.synchronize()
    t_first = (time.perf_counter() - t0) * 1e3
    assert ctx.assembly_info()[:2] == (L.REUSE, L.MIXED)
    print(f"C2: FP64 plain {t_plain:.2f} ms, first mixed REUSE call {t_first:.2f} ms")
    assert t_first <= t_plain + 20.0
    ctx.close()


@pytest.mark.parametrize("N,shuffle", [(60, False), (513, False), (513, True)])
def test_device_symbolic_heat_q4(monkeypatch, N, shuffle):
    """Row-owner heat tiles built on the device (band order and runs from
    2^18 rows up) and by the host builder: FP64 REUSE equals the oracle bit
    for bit (matrix and load vector) either way, mixed within 1e-6."""
    cfg = W.heat_q4(N, sigma=1.0)
    f, n = cfg["fesp"], cfg["ndofs"]
    dn = f.dofnums.copy()
    if shuffle:
        dn = (np.random.default_rng(11).permutation(n) + 1)[dn - 1]
    p = O.make_problem(O.HEAT, O.Q4, 2, f.mesh.conn, f.mesh.xy, dn, kappa_e=cfg["kappa_e"], Q_e=cfg["Q_e"])
    ref = O.assemble(p, n, csc=False)
    for dev in ("1", "0"):
        monkeypatch.setenv("ELFEL_HIP_DEVICE_SYMBOLIC", dev)
        ctx = L.Context(0)
        ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
        ctx.set_dofs(0, dn, n)
        ctx.set_coefficients(f.mesh.nelements, cfg["kappa_e"], cfg["Q_e"], None)
        for prec in (L.FP64, L.MIXED):
            ctx.assemble(L.HEAT, L.Q4, 2, precision=prec, method=L.REUSE, layout=L.CSR)
            assert ctx.assembly_info()[:2] == (L.REUSE, prec)
            ptr, ind, val = ctx.copy_sparse(base=0)[:3]
            F = ctx.copy_vector(n)
            assert np.array_equal(ptr, ref[0]) and np.array_equal(ind, ref[1]), (dev, prec)
            if prec == L.FP64:
                assert np.array_equal(val, ref[2]) and np.array_equal(F, ref[3]), (dev, N, shuffle)
            else:
                assert frob_rel(val, ref[2]) <= MIXED_TOL and frob_rel(F, ref[3]) <= MIXED_TOL
        ctx.close()


def test_device_symbolic_heat_t3_fan_uses_task_tiles():
    """A T3 node in 12 elements: more than a row record holds, so the host
    builder makes task tiles; the result still equals the oracle."""
    K = 12
    xy = [(0.0, 0.0)] + [(np.cos(2 * np.pi * k / K), np.sin(2 * np.pi * k / K)) for k in range(K)]
    conn = np.array([(1, 2 + k, 2 + (k + 1) % K) for k in range(K)], np.int64)
    xy = np.array(xy)
    dn = np.arange(1, K + 2, dtype=np.int64).reshape(-1, 1)
    ref = O.assemble(O.make_problem(O.HEAT, O.T3, 1, conn, xy, dn, kappa=1.0, Q=1.0), K + 1, csc=False)
    ctx = L.Context(0)
    ctx.set_mesh(0, conn, xy)
    ctx.set_dofs(0, dn, K + 1)
    ctx.assemble(L.HEAT, L.T3, 1, precision=L.FP64, method=L.REUSE, layout=L.CSR, kappa=1.0, Q=1.0)
    assert ctx.assembly_info()[:2] == (L.REUSE, L.FP64)
    ptr, ind, val = ctx.copy_sparse(base=0)[:3]
    assert np.array_equal(ptr, ref[0]) and np.array_equal(ind, ref[1]) and np.array_equal(val, ref[2])
    ctx.close()
