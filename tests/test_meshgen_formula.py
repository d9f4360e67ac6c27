"""CPU check of the closed form the device T6 block generator numbers its
midside nodes by (elfel.jl_amd/csrc/meshgen.hip): cell (i, j) of an nA x nB
block (i outer, j inner) creates, in this order, its bottom edge if j = 0,
its right edge, its diagonal, its top edge and its left edge if i = 0, so
3 c + i + [j > 0] + (i = 0 ? j : nB) nodes precede its own, c = i nB + j.
The formula, restated here, must reproduce the oracle's first-appearance
numbering (orc_t6block, a restatement of MeshSteward's T6block) on every
block shape, including the thin ones where a plausible off-by-one in the
i = 0 / j = 0 terms would show."""
import numpy as np
import pytest

import oracle as O


def _created_before(i, j, nB):
    return 3 * (i * nB + j) + (i + (1 if j > 0 else 0)) + (j if i == 0 else nB)


def _t6_by_formula(nA, nB):
    W = nA + 1
    Nv = W * (nB + 1)
    conn = np.zeros((2 * nA * nB, 6), np.int64)
    for i in range(nA):
        for j in range(nB):
            c = i * nB + j
            idx = Nv + _created_before(i, j, nB)
            if j == 0:
                hb = idx
                idx += 1
            else:
                hb = Nv + _created_before(i, j - 1, nB) + (1 if j - 1 == 0 else 0) + 2
            vr, dg, ht = idx, idx + 1, idx + 2
            idx += 3
            vl = idx if i == 0 else Nv + _created_before(i - 1, j, nB) + (1 if j == 0 else 0)
            f = j * W + i
            v0, v1, v2, v3 = f, f + 1, f + W + 1, f + W
            conn[2 * c] = [v0, v1, v2, hb, vr, dg]
            conn[2 * c + 1] = [v0, v2, v3, dg, ht, vl]
    return conn + 1


@pytest.mark.parametrize("nA,nB", [(1, 1), (1, 5), (5, 1), (2, 3), (3, 2), (7, 4), (12, 13)])
def test_t6_closed_form_matches_first_appearance(nA, nB):
    oc, _ = O.t6block(1.0, 1.0, nA, nB)
    assert np.array_equal(_t6_by_formula(nA, nB), oc)


def test_t6_closed_form_counts_every_edge_once():
    for nA, nB in [(3, 4), (9, 2)]:
        c = _t6_by_formula(nA, nB)
        Nv = (nA + 1) * (nB + 1)
        mids = np.unique(c[:, 3:])
        nedges = nA * (nB + 1) + (nA + 1) * nB + nA * nB
        assert mids.size == nedges and mids.min() == Nv + 1 and mids.max() == Nv + nedges
