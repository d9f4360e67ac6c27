"""Device-integrated weak forms: the element loop moved onto the GPU.

Each function replaces one of the reference's user `integrate!` closures
(the hot loop) together with its LocalMatrixAssembler / SysmatAssemblerSparse
calls, and leaves the global matrix on the device for finish(am):

* integrate_heat        test/test_heat.jl:36-71, examples/heat/poisson/q4.jl:31-54
* integrate_elasticity  test/test_elasticity.jl:149-181
* integrate_stokes      examples/stokes/colliding_flow/ht_p2_p1_gen.jl:46-90 (form="gen"),
                        test/test_stokes.jl:629-668 (form="veclap")
* integrate_stokes_reddy  three scalar spaces ux, uy, p: test/test_stokes.jl:373-431
                        (T6/T3), examples/stokes/colliding_flow/p1b_p1.jl:50-108
                        (FEH1_T3_BUBBLE/T3), q1_q0.jl:52-110 (Q4/FEL2_Q4)
* integrate_mass        examples/klein_gordon/klein_gordon.jl:34-51 (M), :94-112 (m^2 M)

integrate_heat also takes FEH1_T4 (examples/heat/poisson/t4.jl:33-55) and
FEH1_L2 spaces (klein_gordon.jl:64-81 is the heat form with kappa = 1, Q = 0).
The element families, the Reddy form and the mass form run on the plain FP64
method.

precision: "fp64" (the plain method: every operation as the reference does it)
or "mixed" (FP32 element arithmetic, FP64 sums into the matrix).
method: "sort" (COO -> stable radix sort -> segmented reduce, every call) or
"reuse" (pattern built on the first call, then one fused kernel per call).
"""
import numpy as np

from . import _lib
from .refshapes import rule_id

_PREC = {"fp64": _lib.FP64, "mixed": _lib.MIXED}
_METH = {"sort": _lib.SORT, "reuse": _lib.REUSE}
_LAYOUT = {"csc": _lib.CSC, "csr": _lib.CSR}


def _bind(ctx, field, fesp):
    """Uploads mesh and dof map of a space unless the context already holds them."""
    key = (id(fesp.mesh), fesp.mesh.conn.ctypes.data, fesp.mesh.xy.ctypes.data, id(fesp), fesp.dofnums.ctypes.data)
    bound = getattr(ctx, "_bound", {})
    if bound.get(field) == key and not getattr(fesp, "_dirty", False):
        return
    ctx.set_mesh(field, fesp.mesh.conn, fesp.mesh.xy)
    ctx.set_dofs(field, fesp.dofnums, getattr(ctx, "_system_ndofs", fesp.ndofs()))
    bound[field] = key
    ctx._bound = bound


def _run(am, spaces, kind, fe, rule, precision, method, layout, coef, **kw):
    ctx = am.context()
    n = am.nrow
    if n <= 0 or am.ncol != n:
        raise ValueError("start(am, ndofs, ndofs) before integrating")
    ctx._system_ndofs = n
    for f, sp in enumerate(spaces):
        _bind(ctx, f, sp)
    E = spaces[0].mesh.nelements
    ctx.set_coefficients(E, *coef)
    nnz = ctx.assemble(kind, fe, rule, precision=_PREC[precision], method=_METH[method], layout=_LAYOUT[layout],
                       **kw)
    am._device_result = True
    return nnz


def integrate_heat(am, av, fesp, qsettings, kappa=1.0, Q=0.0, kappa_e=None, Q_e=None, precision="fp64",
                   method="sort", layout="csc"):
    fe = fesp.fe
    nnz = _run(am, [fesp], _lib.HEAT, fe.code, rule_id(fe.refshape, qsettings), precision, method, layout,
               (kappa_e, Q_e, None), kappa=float(kappa), Q=float(Q))
    if av is not None:
        av._device_src = (am.context(), am.nrow)
    return nnz


def integrate_elasticity(am, fesp, qsettings, D, scale_e=None, precision="fp64", method="sort", layout="csc"):
    fe = fesp.fe
    return _run(am, [fesp], _lib.ELAST, fe.code, rule_id(fe.refshape, qsettings), precision, method, layout,
                (None, None, scale_e), D=np.asarray(D, np.float64))


def integrate_stokes(am, ufesp, pfesp, qsettings, form="gen", D=None, mu=1.0, precision="fp64", method="sort",
                     layout="csc"):
    kind = _lib.STOKES_GEN if form == "gen" else _lib.STOKES_VECLAP
    if D is None:
        D = np.diag([2 * mu, 2 * mu, mu])
    return _run(am, [ufesp, pfesp], kind, ufesp.fe.code, rule_id("triangle", qsettings), precision, method, layout,
                (None, None, None), D=np.asarray(D, np.float64), mu=float(mu))


def integrate_mass(am, fesp, qsettings, coef=1.0, layout="csc"):
    """me[i, j] += (N_i N_j) * (coef JxW) over every element."""
    fe = fesp.fe
    return _run(am, [fesp], _lib.MASS, fe.code, rule_id(fe.refshape, qsettings), "fp64", "sort", layout,
                (None, None, None), kappa=float(coef))


def integrate_stokes_reddy(am, uxfesp, uyfesp, pfesp, qsettings, mu=1.0, layout="csc"):
    """The Reddy-form Stokes blocks kxx, kxy, kxy', kyy, kxp, kxp', kyp, kyp'
    of the spaces ux, uy (one mesh, one element) and p."""
    if uxfesp.mesh is not uyfesp.mesh or type(uxfesp.fe) is not type(uyfesp.fe):
        raise ValueError("ux and uy must be spaces of one element type on one mesh")
    ctx = am.context()
    n = am.nrow
    if n <= 0 or am.ncol != n:
        raise ValueError("start(am, ndofs, ndofs) before integrating")
    mesh = uxfesp.mesh
    ctx.set_mesh(0, mesh.conn, mesh.xy)
    ctx.set_dofs(0, np.stack([uxfesp.dofnums[:, 0], uyfesp.dofnums[:, 0]], 1), n)
    if uxfesp.cell_dofnums.shape[0]:
        ctx.set_element_dofs(0, np.stack([uxfesp.cell_dofnums[:, 0], uyfesp.cell_dofnums[:, 0]], 1), n)
    if pfesp.dofnums.shape[0]:
        ctx.set_mesh(1, pfesp.mesh.conn, pfesp.mesh.xy)
        ctx.set_dofs(1, pfesp.dofnums, n)
    else:
        ctx.set_element_dofs(1, pfesp.cell_dofnums, n)
    ctx._bound = {}  # fields 0 and 1 now hold these spaces
    ctx.set_coefficients(mesh.nelements, None, None, None)
    nnz = ctx.assemble(_lib.STOKES_REDDY, uxfesp.fe.code, rule_id(uxfesp.fe.refshape, qsettings),
                       layout=_LAYOUT[layout], mu=float(mu), fe_p=pfesp.fe.code)
    am._device_result = True
    return nnz


def plane_strain_D(E, nu):
    """Plane-strain material matrix E/((1+nu)(1-2nu)) [[1-nu, nu, 0], [nu, 1-nu, 0], [0, 0, (1-2nu)/2]]."""
    f = E / ((1 + nu) * (1 - 2 * nu))
    return f * np.array([[1 - nu, nu, 0], [nu, 1 - nu, 0], [0, 0, (1 - 2 * nu) / 2]])


def plane_stress_D(E, nu):
    """Plane-stress D of test/test_elasticity.jl:25-27: E/(1-nu^2) [[1, nu, 0], [nu, 1, 0], [0, 0, (1-nu)/2]]."""
    return E / (1 - nu ** 2) * np.array([[1, nu, 0], [nu, 1, 0], [0, 0, (1 - nu) / 2]])
