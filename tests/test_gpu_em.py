"""Element matrices of plane elasticity on T6 (3-point rule, mixed arithmetic)
from the two device kernels of elfel_hip_element_matrices: the vector form
(one work-item per element) and the matrix-core form (v_mfma_f32_4x4x1f32,
16 elements per wave), against the oracle's FP64 element blocks (the ke of
each element in the oracle's COO stream, column-major, the order of
LocalMatrixAssembler.M) and against each other.

Tolerance: the gradients are FP32, so each block is within 2e-6 of the FP64
block in the Frobenius norm, relative to that block; the two kernels sum the
same FP32 products in different orders and agree to 1e-6.
"""
import numpy as np
import pytest

import oracle as O
from elfel_hip import _lib as L
from elfel_hip import workloads as W
from helpers import oracle_problem

pytestmark = pytest.mark.gpu

TOL_ORACLE = 2e-6
TOL_PAIR = 1e-6


def _blockwise(a, ref):
    num = np.linalg.norm(a.astype(np.float64) - ref, axis=1)
    return float((num / np.linalg.norm(ref, axis=1)).max())


@pytest.mark.parametrize("N,E_cut", [(60, None), (61, 2 * 61 * 61 - 5)])
def test_element_matrices_vs_oracle(N, E_cut):
    """Jittered T6 mesh; the second case cuts the mesh to an element count
    that is not a multiple of 16 or 64 (a partial last wave)."""
    cfg = W.elast_t6(N, frac=0.15)
    f, D = cfg["fesp"], cfg["D"]
    conn = f.mesh.conn if E_cut is None else f.mesh.conn[:E_cut]
    E = conn.shape[0]
    I, J, V, _, _ = O.coo(O.make_problem(O.ELAST, O.T6, 3, conn, f.mesh.xy, f.dofnums, D=D))
    ref = V.reshape(E, 144)
    ctx = L.Context(0)
    ctx.set_mesh(0, conn, f.mesh.xy)
    kv, tv = ctx.element_matrices(E, D, L.EM_VALU)
    km, tm = ctx.element_matrices(E, D, L.EM_MFMA)
    ctx.close()
    ev, em, ep = _blockwise(kv, ref), _blockwise(km, ref), _blockwise(km, kv.astype(np.float64))
    print(f"E={E}: VALU vs oracle {ev:.2e}, MFMA vs oracle {em:.2e}, MFMA vs VALU {ep:.2e}; "
          f"{tv * 1e3:.1f} / {tm * 1e3:.1f} us")
    assert ev <= TOL_ORACLE and em <= TOL_ORACLE and ep <= TOL_PAIR
    # symmetry of every block (B^T D B with symmetric D)
    for k in (kv, km):
        kk = k.reshape(E, 12, 12)
        assert np.abs(kk - kk.transpose(0, 2, 1)).max() <= 1e-6 * np.abs(kk).max()


def test_element_matrices_refusals():
    cfg = W.heat_t3(8)
    f = cfg["fesp"]
    ctx = L.Context(0)
    ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
    with pytest.raises(L.ElfelHipError) as e:
        ctx.element_matrices(f.mesh.nelements, np.eye(3), L.EM_MFMA)
    assert e.value.code == L.ESTATE
    ctx.close()
    cfg = W.elast_t6(4, frac=0.0)
    f = cfg["fesp"]
    ctx = L.Context(0)
    ctx.set_mesh(0, f.mesh.conn, f.mesh.xy)
    with pytest.raises(L.ElfelHipError) as e:
        ctx.element_matrices(f.mesh.nelements + 1, cfg["D"], L.EM_MFMA)
    assert e.value.code == L.EINVAL
    with pytest.raises(L.ElfelHipError) as e:
        ctx.element_matrices(f.mesh.nelements, cfg["D"], 7)
    assert e.value.code == L.EINVAL
    ctx.close()
