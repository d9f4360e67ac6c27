"""Shared helpers of the test suite: build the same problem for the device and
the oracle from one set of host arrays, and compare sparse results."""
import json
import os

import numpy as np

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")


def pins():
    with open(os.path.join(GOLDEN, "reference_pins.json")) as fh:
        return json.load(fh)


def oracle_problem(O, kind, fesp, qsettings, pfesp=None, **kw):
    fe = fesp.fe
    rule = qsettings.get("npts", 1) if fe.refshape == "triangle" else qsettings.get("order", 1)
    code = {"T3": O.T3, "T6": O.T6, "Q4": O.Q4}[fe.name]
    extra = {}
    if pfesp is not None:
        extra = dict(conn_p=pfesp.mesh.conn, dofnums_p=pfesp.dofnums)
    return O.make_problem(kind, code, rule, fesp.mesh.conn, fesp.mesh.xy, fesp.dofnums, **extra, **kw)


def compare_sparse(dev, ref, rtol=1e-12, exact_pattern=True):
    """dev, ref: (ptr, ind, val) 0-based.  Returns the max relative value error
    (max |dv - rv| / max |rv|) after checking the pattern."""
    dptr, dind, dval = dev[:3]
    rptr, rind, rval = ref[:3]
    assert dptr.shape == rptr.shape, "major dimension differs"
    assert np.array_equal(dptr, rptr), "pointer array differs"
    assert np.array_equal(dind, rind), "index array differs"
    scale = max(np.abs(rval).max(), 1e-300) if rval.size else 1.0
    err = float(np.abs(dval - rval).max() / scale) if rval.size else 0.0
    assert err <= rtol, f"relative value error {err:.3e} > {rtol:.1e}"
    return err


def frob_rel(dval, rval):
    return float(np.linalg.norm(dval - rval) / max(np.linalg.norm(rval), 1e-300))
