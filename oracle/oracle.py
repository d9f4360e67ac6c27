"""ctypes front end of the CPU oracle (liborc.so).

TEST INFRASTRUCTURE ONLY: used by tests/, __graft_entry__.smoke() and the
cpu_baseline leg of bench.py as the checker and the CPU baseline.  The product
path (the HIP library and the elfel_hip package) never imports this module.

Every function mirrors a routine of elfel_oracle.c, which cites the Elfel.jl
source lines it restates.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIBPATH = os.path.join(_HERE, "liborc.so")

T3, T6, Q4, T3B, Q4L2, T4, L2 = 0, 1, 2, 3, 4, 5, 6
HEAT, ELAST, STOKES_GEN, STOKES_VECLAP, STOKES_REDDY, MASS = 0, 1, 2, 3, 4, 5

_i64p = ctypes.POINTER(ctypes.c_int64)
_f64p = ctypes.POINTER(ctypes.c_double)


class Problem(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int), ("fe", ctypes.c_int), ("rule", ctypes.c_int),
        ("E", ctypes.c_int64), ("nen", ctypes.c_int),
        ("conn", _i64p), ("xy", _f64p), ("Nn", ctypes.c_int64),
        ("ncomp", ctypes.c_int), ("dofnums", _i64p),
        ("conn_p", _i64p), ("dofnums_p", _i64p), ("Np", ctypes.c_int64),
        ("kappa", ctypes.c_double), ("Q", ctypes.c_double),
        ("kappa_e", _f64p), ("Q_e", _f64p),
        ("D", ctypes.c_double * 9), ("scale_e", _f64p), ("mu", ctypes.c_double),
    ]


def _load():
    if not os.path.exists(_LIBPATH):
        raise RuntimeError("oracle library missing: run `make -C oracle` (or __graft_entry__.build())")
    lib = ctypes.CDLL(_LIBPATH)
    lib.orc_coo_len.restype = ctypes.c_int64
    lib.orc_sparse.restype = ctypes.c_int64
    lib.orc_t3block.restype = ctypes.c_int64
    lib.orc_q4block.restype = ctypes.c_int64
    lib.orc_t6block.restype = ctypes.c_int64
    lib.orc_t6tot3.restype = ctypes.c_int64
    lib.orc_stokes_reddy_coo_len.restype = ctypes.c_int64
    for f in ("orc_fam_scalar_coo_len", "orc_fam_reddy_coo_len", "orc_l2block", "orc_t4block"):
        getattr(lib, f).restype = ctypes.c_int64
    for f in ("orc_colliding_p_error", "orc_colliding_v_error_scalar", "orc_colliding_v_error_vector"):
        getattr(lib, f).restype = ctypes.c_double
    return lib


lib = _load()


def _p(a, t):
    return a.ctypes.data_as(t) if a is not None else ctypes.cast(None, t)


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


# ---- quadrature and basis -------------------------------------------------
def gauss1(order):
    pc, w = np.zeros(8), np.zeros(8)
    n = lib.orc_gauss1(order, _p(pc, _f64p), _p(w, _f64p))
    return pc[:n].copy(), w[:n].copy()


def quad_triangle(npts):
    pc, w = np.zeros(32), np.zeros(16)
    n = lib.orc_quad_triangle(npts, _p(pc, _f64p), _p(w, _f64p))
    if n < 0:
        raise ValueError(f"Unknown number of integration points {npts}")
    return pc[: 2 * n].reshape(n, 2).copy(), w[:n].copy()


def quad_square(order):
    pc, w = np.zeros(64), np.zeros(32)
    n = lib.orc_quad_square(order, _p(pc, _f64p), _p(w, _f64p))
    return pc[: 2 * n].reshape(n, 2).copy(), w[:n].copy()


def nen(fe):
    return lib.orc_nen(fe)


def bfun(fe, pc):
    N = np.zeros(8)
    lib.orc_bfun(fe, _p(_f64(pc), _f64p), _p(N, _f64p))
    return N[: nen(fe)].copy()


def bfungradpar(fe, pc):
    g = np.zeros(16)
    lib.orc_bfungradpar(fe, _p(_f64(pc), _f64p), _p(g, _f64p))
    return g[: 2 * nen(fe)].reshape(-1, 2).copy()


# ---- meshes ---------------------------------------------------------------
def t3block(A, B, nA, nB):
    E, Nn = 2 * nA * nB, (nA + 1) * (nB + 1)
    conn, xy = np.zeros((E, 3), np.int64), np.zeros((Nn, 2))
    lib.orc_t3block(ctypes.c_double(A), ctypes.c_double(B), nA, nB, _p(conn, _i64p), _p(xy, _f64p))
    return conn, xy


def q4block(A, B, nA, nB):
    E, Nn = nA * nB, (nA + 1) * (nB + 1)
    conn, xy = np.zeros((E, 4), np.int64), np.zeros((Nn, 2))
    lib.orc_q4block(ctypes.c_double(A), ctypes.c_double(B), nA, nB, _p(conn, _i64p), _p(xy, _f64p))
    return conn, xy


def t6block(A, B, nA, nB):
    E = 2 * nA * nB
    Nn = (nA + 1) * (nB + 1) + nA * (nB + 1) + (nA + 1) * nB + nA * nB
    conn, xy = np.zeros((E, 6), np.int64), np.zeros((Nn, 2))
    n = lib.orc_t6block(ctypes.c_double(A), ctypes.c_double(B), nA, nB, _p(conn, _i64p), _p(xy, _f64p))
    assert n == Nn
    return conn, xy


def t6tot3(conn6, xy6):
    conn6, xy6 = _i64(conn6), _f64(xy6)
    E = conn6.shape[0]
    conn3, xy3 = np.zeros((E, 3), np.int64), np.zeros_like(xy6)
    n = lib.orc_t6tot3(E, _p(conn6, _i64p), xy6.shape[0], _p(xy6, _f64p), _p(conn3, _i64p), _p(xy3, _f64p))
    return conn3, xy3[:n].copy()


# ---- numbering ------------------------------------------------------------
def numberdofs(isdatum_list):
    """isdatum_list: list of (nterms, ncomp) bool arrays, one per space."""
    nterms = np.array([a.shape[0] for a in isdatum_list], np.int64)
    ncomp = np.array([a.shape[1] for a in isdatum_list], np.int32)
    flat = np.ascontiguousarray(np.concatenate([np.asarray(a, np.uint8).ravel() for a in isdatum_list]))
    out = np.zeros(flat.size, np.int64)
    lib.orc_numberdofs(len(isdatum_list), _p(nterms, _i64p), ncomp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                       flat.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)), _p(out, _i64p))
    res, o = [], 0
    for a in isdatum_list:
        res.append(out[o:o + a.size].reshape(a.shape).copy())
        o += a.size
    return res


# ---- assembly -------------------------------------------------------------
def make_problem(kind, fe, rule, conn, xy, dofnums, conn_p=None, dofnums_p=None,
                 kappa=1.0, Q=0.0, kappa_e=None, Q_e=None, D=None, scale_e=None, mu=1.0):
    keep = {}
    keep["conn"], keep["xy"], keep["dofnums"] = _i64(conn), _f64(xy), _i64(dofnums)
    p = Problem()
    p.kind, p.fe, p.rule = kind, fe, rule
    p.E, p.nen = keep["conn"].shape[0], keep["conn"].shape[1]
    p.conn, p.xy, p.Nn = _p(keep["conn"], _i64p), _p(keep["xy"], _f64p), keep["xy"].shape[0]
    p.ncomp = keep["dofnums"].shape[1] if keep["dofnums"].ndim == 2 else 1
    p.dofnums = _p(keep["dofnums"], _i64p)
    if conn_p is not None:
        keep["conn_p"], keep["dofnums_p"] = _i64(conn_p), _i64(dofnums_p)
        p.conn_p, p.dofnums_p, p.Np = _p(keep["conn_p"], _i64p), _p(keep["dofnums_p"], _i64p), keep["dofnums_p"].size
    p.kappa, p.Q, p.mu = kappa, Q, mu
    for name in ("kappa_e", "Q_e", "scale_e"):
        v = {"kappa_e": kappa_e, "Q_e": Q_e, "scale_e": scale_e}[name]
        if v is not None:
            keep[name] = _f64(v)
            setattr(p, name, _p(keep[name], _f64p))
    if D is not None:
        Dc = np.asarray(D, np.float64).reshape(3, 3).ravel(order="F")
        for i in range(9):
            p.D[i] = Dc[i]
    p._keep = keep
    return p


def coo(p, nthreads=1):
    nv = ctypes.c_int64(0)
    n = lib.orc_coo_len(ctypes.byref(p), ctypes.byref(nv))
    I, J, V = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n)
    FI, FV = np.zeros(nv.value, np.int64), np.zeros(nv.value)
    rc = lib.orc_coo_parallel(ctypes.byref(p), _p(I, _i64p), _p(J, _i64p), _p(V, _f64p),
                              _p(FI, _i64p), _p(FV, _f64p), nthreads)
    if rc != 0:
        raise ValueError(f"orc_coo failed ({rc})")
    return I, J, V, FI, FV


def stokes_reddy_coo(conn6, xy6, conn3, xy3, dn_ux, dn_uy, dn_p, mu):
    conn6, xy6, conn3, xy3 = _i64(conn6), _f64(xy6), _i64(conn3), _f64(xy3)
    dn_ux, dn_uy, dn_p = _i64(dn_ux), _i64(dn_uy), _i64(dn_p)
    E = conn6.shape[0]
    n = lib.orc_stokes_reddy_coo_len(E)
    I, J, V = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n)
    lib.orc_stokes_reddy_coo(E, _p(conn6, _i64p), _p(xy6, _f64p), _p(conn3, _i64p), _p(xy3, _f64p),
                             _p(dn_ux, _i64p), _p(dn_uy, _i64p), _p(dn_p, _i64p), ctypes.c_double(mu),
                             _p(I, _i64p), _p(J, _i64p), _p(V, _f64p))
    return I, J, V


# ---- element families (elfel_oracle_fam.c) --------------------------------
def fam_nbf(fe):
    return lib.orc_fam_nbf(fe)


def fam_mdim(fe):
    return lib.orc_fam_mdim(fe)


def fam_quad(fe, rule):
    pc, w = np.zeros(48), np.zeros(16)
    n = lib.orc_fam_quad(fe, rule, _p(pc, _f64p), _p(w, _f64p))
    if n < 0:
        raise ValueError(f"no rule {rule} for element {fe}")
    return pc[: 3 * n].reshape(n, 3)[:, : fam_mdim(fe)].copy(), w[:n].copy()


def fam_bfun(fe, pc):
    pcx = np.zeros(3)
    pcx[: len(pc)] = pc
    N = np.zeros(8)
    lib.orc_fam_bfun(fe, _p(pcx, _f64p), _p(N, _f64p))
    return N[: fam_nbf(fe)].copy()


def fam_bfungradpar(fe, pc):
    pcx = np.zeros(3)
    pcx[: len(pc)] = pc
    g = np.zeros(24)
    lib.orc_fam_bfungradpar(fe, _p(pcx, _f64p), _p(g, _f64p))
    m = fam_mdim(fe)
    return g[: m * fam_nbf(fe)].reshape(-1, m).copy()


def fam_scalar_coo(kind, fe, rule, conn, x, dofnums, kappa=1.0, Q=0.0):
    """Heat (grad-grad + source) or mass on a nodal family; x is (Nn, mdim)."""
    conn, x, dofnums = _i64(conn), _f64(x), _i64(dofnums).ravel()
    E = conn.shape[0]
    nv = ctypes.c_int64(0)
    n = lib.orc_fam_scalar_coo_len(fe, E, ctypes.byref(nv))
    I, J, V = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n)
    FI, FV = np.zeros(nv.value, np.int64), np.zeros(nv.value)
    rc = lib.orc_fam_scalar_coo(kind, fe, rule, E, _p(conn, _i64p), _p(x, _f64p), _p(dofnums, _i64p),
                                ctypes.c_double(kappa), ctypes.c_double(Q), _p(I, _i64p), _p(J, _i64p),
                                _p(V, _f64p), _p(FI, _i64p), _p(FV, _f64p))
    if rc != 0:
        raise ValueError(f"orc_fam_scalar_coo failed ({rc})")
    if kind != HEAT:
        FI, FV = FI[:0], FV[:0]
    return I, J, V, FI, FV


def fam_reddy_coo(vfe, pfe, rule, conn_v, xy_v, dn_ux, dn_uy, mu=1.0, de_ux=None, de_uy=None,
                  conn_p=None, xy_p=None, dn_p=None, de_p=None):
    """Reddy-form Stokes over the spaces ux, uy (node dofs dn_*, cell dofs de_*) and p."""
    k = {"conn_v": _i64(conn_v), "xy_v": _f64(xy_v), "dn_ux": _i64(dn_ux), "dn_uy": _i64(dn_uy)}
    for name, v in (("de_ux", de_ux), ("de_uy", de_uy), ("dn_p", dn_p), ("de_p", de_p)):
        k[name] = None if v is None else _i64(v)
    k["conn_p"] = None if conn_p is None else _i64(conn_p)
    k["xy_p"] = None if xy_p is None else _f64(xy_p)
    E = k["conn_v"].shape[0]
    n = lib.orc_fam_reddy_coo_len(vfe, pfe, E)
    I, J, V = np.zeros(n, np.int64), np.zeros(n, np.int64), np.zeros(n)
    rc = lib.orc_fam_reddy_coo(vfe, pfe, rule, E, _p(k["conn_v"], _i64p), _p(k["xy_v"], _f64p),
                               _p(k["conn_p"], _i64p), _p(k["xy_p"], _f64p), _p(k["dn_ux"], _i64p),
                               _p(k["dn_uy"], _i64p), _p(k["de_ux"], _i64p), _p(k["de_uy"], _i64p),
                               _p(k["dn_p"], _i64p), _p(k["de_p"], _i64p), ctypes.c_double(mu),
                               _p(I, _i64p), _p(J, _i64p), _p(V, _f64p))
    if rc != 0:
        raise ValueError(f"orc_fam_reddy_coo failed ({rc})")
    return I, J, V


def l2block(L, n):
    conn, x = np.zeros((n, 2), np.int64), np.zeros((n + 1, 1))
    lib.orc_l2block(ctypes.c_double(L), n, _p(conn, _i64p), _p(x, _f64p))
    return conn, x


def t4block(A, B, C, nA, nB, nC):
    E, Nn = 6 * nA * nB * nC, (nA + 1) * (nB + 1) * (nC + 1)
    conn, xyz = np.zeros((E, 4), np.int64), np.zeros((Nn, 3))
    lib.orc_t4block(ctypes.c_double(A), ctypes.c_double(B), ctypes.c_double(C), nA, nB, nC,
                    _p(conn, _i64p), _p(xyz, _f64p))
    return conn, xyz


def sparse(I, J, V, nrow, ncol, csc=True):
    """Returns (ptr, ind, val), 0-based, sorted minor indices (SparseArrays.sparse)."""
    I, J, V = _i64(I), _i64(J), _f64(V)
    n = I.size
    nmaj = ncol if csc else nrow
    ptr = np.zeros(nmaj + 1, np.int64)
    ind, val = np.zeros(max(n, 1), np.int64), np.zeros(max(n, 1))
    nnz = lib.orc_sparse(n, _p(I, _i64p), _p(J, _i64p), _p(V, _f64p), nrow, ncol, 1 if csc else 0,
                         _p(ptr, _i64p), _p(ind, _i64p), _p(val, _f64p))
    if nnz < 0:
        raise ValueError("row or column index out of range")
    return ptr, ind[:nnz].copy(), val[:nnz].copy()


def sparse_scipy(I, J, V, nrow, ncol, csc=True):
    import scipy.sparse as sp
    ptr, ind, val = sparse(I, J, V, nrow, ncol, csc)
    if csc:
        return sp.csc_matrix((val, ind, ptr), shape=(nrow, ncol))
    return sp.csr_matrix((val, ind, ptr), shape=(nrow, ncol))


def vec_assemble(FI, FV, ndofs):
    FI, FV = _i64(FI), _f64(FV)
    out = np.zeros(ndofs)
    rc = lib.orc_vec_assemble(FI.size, _p(FI, _i64p), _p(FV, _f64p), ndofs, _p(out, _f64p))
    if rc != 0:
        raise IndexError("dof number out of range")
    return out


def assemble(p, ndofs, csc=False):
    """Full oracle assembly: COO in stream order, then sparse(); returns (ptr, ind, val, F)."""
    I, J, V, FI, FV = coo(p)
    ptr, ind, val = sparse(I, J, V, ndofs, ndofs, csc)
    F = vec_assemble(FI, FV, ndofs) if FI.size else np.zeros(ndofs)
    return ptr, ind, val, F


# ---- post-processing and solve ---------------------------------------------
def colliding_p_error(conn3, xy3, dn_p, U):
    conn3, xy3, dn_p, U = _i64(conn3), _f64(xy3), _i64(dn_p), _f64(U)
    return lib.orc_colliding_p_error(conn3.shape[0], _p(conn3, _i64p), _p(xy3, _f64p), _p(dn_p, _i64p), _p(U, _f64p))


def colliding_v_error_scalar(conn6, xy6, dn_ux, dn_uy, U):
    conn6, xy6, dn_ux, dn_uy, U = _i64(conn6), _f64(xy6), _i64(dn_ux), _i64(dn_uy), _f64(U)
    return lib.orc_colliding_v_error_scalar(conn6.shape[0], _p(conn6, _i64p), _p(xy6, _f64p),
                                            _p(dn_ux, _i64p), _p(dn_uy, _i64p), _p(U, _f64p))


def colliding_v_error_vector(conn6, xy6, dn_u, U):
    conn6, xy6, dn_u, U = _i64(conn6), _f64(xy6), _i64(dn_u), _f64(U)
    return lib.orc_colliding_v_error_vector(conn6.shape[0], _p(conn6, _i64p), _p(xy6, _f64p),
                                            _p(dn_u, _i64p), _p(U, _f64p))


def pcg(indptr, indices, vals, b, nfree, tol=1e-10, maxit=100000, x0=None):
    ip, ix, v = _i64(indptr), _i64(indices), _f64(vals)
    b = _f64(b[:nfree])
    x = np.zeros(nfree) if x0 is None else _f64(x0).copy()
    rr = ctypes.c_double(0)
    it = lib.orc_pcg(nfree, _p(ip, _i64p), _p(ix, _i64p), _p(v, _f64p), _p(b, _f64p), _p(x, _f64p),
                     ctypes.c_double(tol), maxit, ctypes.byref(rr))
    return x, it, rr.value


def time_assembly(p, reps=1, nthreads=1):
    te, ts, nnz = ctypes.c_double(0), ctypes.c_double(0), ctypes.c_int64(0)
    rc = lib.orc_time_assembly(ctypes.byref(p), reps, nthreads, ctypes.byref(te), ctypes.byref(ts), ctypes.byref(nnz))
    if rc != 0:
        raise RuntimeError(f"orc_time_assembly failed ({rc})")
    return te.value, ts.value, nnz.value
