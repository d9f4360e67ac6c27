"""Meshes generated on the device against the oracle's generators (restating
MeshSteward's T3block / Q4block / T6block / T6toT3 as the reference's drivers
call them, test/test_heat.jl:29,144, test/test_stokes.jl:32,37, pinned in
test_oracle_pins.py by the reference's fixture mesh and golden values) and
against the host package's generators: connectivity equal and coordinates
equal bit for bit.  The boundary selection (connectedv(boundary(mesh)),
test/test_heat.jl:31) is checked against the host selection on block meshes,
on the reference's fixture mesh and on a jittered mesh; an end-to-end device
pipeline (mesh, boundary, numbering, FP64 assembly) reproduces the host-built
assembly bit for bit."""
import os

import numpy as np
import pytest

import oracle as O
from elfel_hip import _lib as L
from elfel_hip import meshes as M
from elfel_hip import workloads as W
from helpers import GOLDEN

pytestmark = pytest.mark.gpu

SIZES = [(1, 1), (3, 2), (2, 5), (17, 9), (40, 40)]
BOXES = [(1.0, 1.0), (2.0, 3.0), (0.3, 7.1)]


def _gen(ctx, fe, A, B, nA, nB, field=0):
    E, Nn = ctx.block_mesh(field, fe, A, B, nA, nB)
    nen = {L.T3: 3, L.Q4: 4, L.T6: 6}[fe]
    return ctx.copy_mesh(field, E, nen, Nn)


@pytest.mark.parametrize("fe,orc,host", [(L.T3, O.t3block, M.T3block), (L.Q4, O.q4block, M.Q4block),
                                         (L.T6, O.t6block, M.T6block)])
def test_block_meshes_match_generators(ctx, fe, orc, host):
    for (nA, nB) in SIZES:
        for (A, B) in BOXES:
            conn, xy = _gen(ctx, fe, A, B, nA, nB)
            oc, oxy = orc(A, B, nA, nB)
            hm = host(A, B, nA, nB)
            assert np.array_equal(conn, oc) and np.array_equal(xy, oxy), (fe, nA, nB, A, B)
            assert np.array_equal(conn, hm.conn) and np.array_equal(xy, hm.xy), (fe, nA, nB, A, B)


def test_t6_to_t3_matches_generators(ctx):
    for (nA, nB) in SIZES:
        E6, Nn6 = ctx.block_mesh(0, L.T6, 2.0, 2.0, nA, nB)
        Nn3 = ctx.t6_to_t3(0, 1)
        c3, xy3 = ctx.copy_mesh(1, E6, 3, Nn3)
        c6, xy6 = ctx.copy_mesh(0, E6, 6, Nn6)
        oc3, oxy3 = O.t6tot3(c6, xy6)
        assert np.array_equal(c3, oc3) and np.array_equal(xy3, oxy3)
        hm = M.T6toT3(M.T6block(2.0, 2.0, nA, nB))
        assert np.array_equal(c3, hm.conn) and np.array_equal(xy3, hm.xy)


@pytest.mark.parametrize("fe,host", [(L.T3, M.T3block), (L.Q4, M.Q4block), (L.T6, M.T6block)])
def test_boundary_nodes_of_blocks(ctx, fe, host):
    for (nA, nB) in SIZES:
        E, Nn = ctx.block_mesh(0, fe, 1.0, 1.0, nA, nB)
        b = ctx.boundary_nodes(0, Nn)
        assert np.array_equal(b, host(1.0, 1.0, nA, nB).boundary_nodes())
        # closed form: the nodes of the block's perimeter
        nper = 2 * (nA + nB) * (2 if fe == L.T6 else 1)
        assert b.size == nper


def test_boundary_nodes_of_fixture_and_jittered_meshes(ctx):
    m = M.load(os.path.join(GOLDEN, "qmesh.mesh"))
    ctx.set_mesh(0, m.conn, m.xy)
    assert np.array_equal(ctx.boundary_nodes(0, m.nnodes), m.boundary_nodes())
    cfg = W.elast_t6(24, frac=0.15)
    mesh = cfg["fesp"].mesh
    ctx.set_mesh(0, mesh.conn, mesh.xy)
    assert np.array_equal(ctx.boundary_nodes(0, mesh.nnodes), mesh.boundary_nodes())


def test_device_pipeline_equals_host_built_assembly():
    """heat Q4 on 64x64: block mesh, boundary, numbering on the device, then the
    FP64 assembly; the host workload builds the same problem on the host.  Own
    contexts: the per-element coefficients set here must not reach other tests."""
    N = 64
    cfg = W.heat_q4(N, sigma=1.0)
    f = cfg["fesp"]
    host = L.Context(0)
    host.set_mesh(0, f.mesh.conn, f.mesh.xy)
    host.set_dofs(0, f.dofnums, cfg["ndofs"])
    host.set_coefficients(f.mesh.nelements, cfg["kappa_e"], cfg["Q_e"], None)
    host.assemble(L.HEAT, L.Q4, 2, precision=L.FP64, method=L.REUSE, layout=L.CSR)
    ref = host.copy_sparse(base=0)
    Fref = host.copy_vector(cfg["ndofs"])
    host.close()

    ctx = L.Context(0)
    E, Nn = ctx.block_mesh(0, L.Q4, 1.0, 1.0, N, N)
    b = ctx.boundary_nodes(0, Nn)
    isd = np.zeros((Nn, 1), bool)
    isd[b - 1] = True
    dn = ctx.number_dofs([isd])[0]
    assert np.array_equal(dn, f.dofnums)
    ctx.set_dofs(0, dn, Nn)
    ctx.set_coefficients(E, cfg["kappa_e"], cfg["Q_e"], None)
    ctx.assemble(L.HEAT, L.Q4, 2, precision=L.FP64, method=L.REUSE, layout=L.CSR)
    got = ctx.copy_sparse(base=0)
    for a, r in zip(got[:3], ref[:3]):
        assert np.array_equal(a, r)
    assert np.array_equal(ctx.copy_vector(Nn), Fref)
    ctx.close()


def test_block_mesh_at_workload_sizes(ctx):
    """T6 400x400 (the elasticity workload's mesh before jitter) and Q4
    1024x1024 against the host generators."""
    for fe, host, n in ((L.T6, M.T6block, 400), (L.Q4, M.Q4block, 1024)):
        conn, xy = _gen(ctx, fe, 1.0, 1.0, n, n)
        hm = host(1.0, 1.0, n, n)
        assert np.array_equal(conn, hm.conn) and np.array_equal(xy, hm.xy)


def test_block_mesh_rejects_bad_arguments(ctx):
    with pytest.raises(L.ElfelHipError):
        ctx.block_mesh(0, L.T6, 1.0, 1.0, 0, 4)
    with pytest.raises(L.ElfelHipError):
        ctx.block_mesh(2, L.T3, 1.0, 1.0, 4, 4)


@pytest.mark.parametrize("name", ["qmesh.mesh", "mt3gen3.mesh"])
def test_load_mesh_matches_host_reader(ctx, name):
    """The reference's fixture meshes (test/qmesh.mesh, test/mt3gen3.mesh) read
    by the library equal the host reader's arrays."""
    path = os.path.join(GOLDEN, name)
    nen, E, Nn = ctx.load_mesh(0, path)
    conn, xy = ctx.copy_mesh(0, E, nen, Nn)
    m = M.load(path)
    assert np.array_equal(conn, m.conn) and np.array_equal(xy, m.xy)
    with pytest.raises(L.ElfelHipError):
        ctx.load_mesh(0, os.path.join(GOLDEN, "no-such.mesh"))
