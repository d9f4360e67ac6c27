"""ctypes binding of libelfel_hip.so (the C ABI declared in include/elfel_hip.h).

There is no CPU fallback: if the library is missing or no HIP device is present
the constructors raise, so nothing silently runs elsewhere.
"""
import ctypes
import os

import numpy as np

_PKG = os.path.dirname(os.path.abspath(__file__))
LIBPATH = os.environ.get("ELFEL_HIP_LIB", os.path.join(os.path.dirname(_PKG), "libelfel_hip.so"))

# status codes
OK, EINVAL, ERANGE, ENOMEM, EHIP, ESTATE, EGEOM, ENOTSUP = 0, -1, -2, -3, -4, -5, -6, -7
CHECK_GEOM = 1  # elfel_hip_problem.flags
# enums
HEAT, ELAST, STOKES_GEN, STOKES_VECLAP, STOKES_REDDY, MASS = 0, 1, 2, 3, 4, 5
T3, T6, Q4, T3B, Q4L2, T4, L2 = 0, 1, 2, 3, 4, 5, 6
FP64, MIXED = 0, 1
SORT, REUSE = 0, 1
CSR, CSC = 0, 1
PH_GATHER, PH_ELEMENT, PH_SORT, PH_REDUCE, PH_NUMERIC, PH_SYMBOLIC, PH_TOTAL, PH_SOLVE = range(8)
NPHASES = 8
PHASE_NAMES = ["gather", "element", "sort", "reduce", "numeric", "symbolic", "total", "solve"]

EXPORTED = [
    "elfel_hip_init", "elfel_hip_free", "elfel_hip_last_error", "elfel_hip_device_name", "elfel_hip_copy_bandwidth",
    "elfel_hip_sparse", "elfel_hip_copy_sparse", "elfel_hip_result_info", "elfel_hip_vector",
    "elfel_hip_set_mesh", "elfel_hip_set_dofs", "elfel_hip_set_coefficients", "elfel_hip_assemble",
    "elfel_hip_copy_vector", "elfel_hip_synchronize", "elfel_hip_geometry_stats", "elfel_hip_pcg",
    "elfel_hip_timing", "elfel_hip_timing_read",
    "elfel_hip_batch_begin", "elfel_hip_batch_add", "elfel_hip_batch_build", "elfel_hip_batch_assemble",
    "elfel_hip_batch_info", "elfel_hip_batch_copy", "elfel_hip_batch_stats",
    "elfel_hip_number_dofs", "elfel_hip_gather_sysvec", "elfel_hip_scatter_sysvec",
    "elfel_hip_colliding_errors",
    "elfel_hip_block_mesh", "elfel_hip_t6_to_t3", "elfel_hip_boundary_nodes", "elfel_hip_copy_mesh",
    "elfel_hip_load_mesh", "elfel_hip_set_mesh_sdim", "elfel_hip_set_element_dofs",
    "elfel_hip_assembly_info", "elfel_hip_sparse_noalloc", "elfel_hip_element_matrices",
]
EM_VALU, EM_MFMA = 0, 1
NOALLOC_MAX_COO = 1 << 20  # ELFEL_NOALLOC_MAX_COO


class ElfelHipError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"[{code}] {msg}")
        self.code = code


class ElfelRangeError(ElfelHipError, IndexError):
    """Index out of range: Elfel raises BoundsError / ArgumentError here."""


class Problem(ctypes.Structure):
    _fields_ = [
        ("kind", ctypes.c_int32), ("fe", ctypes.c_int32), ("rule", ctypes.c_int32),
        ("precision", ctypes.c_int32), ("method", ctypes.c_int32), ("layout", ctypes.c_int32),
        ("kappa", ctypes.c_double), ("Q", ctypes.c_double), ("mu", ctypes.c_double),
        ("D", ctypes.c_double * 9), ("fe_p", ctypes.c_int32), ("flags", ctypes.c_int32),
    ]


_lib = None


def load():
    """Loads libelfel_hip.so once; raises if it is not built."""
    global _lib
    if _lib is not None:
        return _lib
    if not os.path.exists(LIBPATH):
        raise ElfelHipError(ESTATE, f"{LIBPATH} is not built; run __graft_entry__.build() or make -C elfel.jl_amd/csrc")
    lib = ctypes.CDLL(LIBPATH)
    P = ctypes.c_void_p
    i64p = ctypes.POINTER(ctypes.c_int64)
    f64p = ctypes.POINTER(ctypes.c_double)
    lib.elfel_hip_init.argtypes = [ctypes.c_int, ctypes.POINTER(P)]
    lib.elfel_hip_free.argtypes = [P]
    lib.elfel_hip_last_error.argtypes = [P]
    lib.elfel_hip_last_error.restype = ctypes.c_char_p
    lib.elfel_hip_device_name.argtypes = [P, ctypes.c_char_p, ctypes.c_int, ctypes.POINTER(ctypes.c_int)]
    lib.elfel_hip_copy_bandwidth.argtypes = [P, ctypes.c_int64, ctypes.c_int, f64p]
    lib.elfel_hip_sparse.argtypes = [P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, i64p, i64p, f64p,
                                     ctypes.c_int, i64p]
    lib.elfel_hip_copy_sparse.argtypes = [P, ctypes.c_int, i64p, i64p, f64p]
    lib.elfel_hip_sparse_noalloc.argtypes = [P, ctypes.c_int64, ctypes.c_int64, ctypes.c_int64, P, P, P, ctypes.c_int,
                                             P, P, P, i64p]
    lib.elfel_hip_result_info.argtypes = [P, i64p, i64p, i64p, ctypes.POINTER(ctypes.c_int)]
    lib.elfel_hip_vector.argtypes = [P, ctypes.c_int64, ctypes.c_int64, i64p, f64p, f64p]
    lib.elfel_hip_set_mesh.argtypes = [P, ctypes.c_int, ctypes.c_int32, ctypes.c_int64, i64p, ctypes.c_int64, f64p]
    lib.elfel_hip_set_dofs.argtypes = [P, ctypes.c_int, ctypes.c_int32, ctypes.c_int64, i64p, ctypes.c_int64]
    lib.elfel_hip_set_mesh_sdim.argtypes = [P, ctypes.c_int, ctypes.c_int32, ctypes.c_int64, i64p, ctypes.c_int64,
                                            ctypes.c_int32, f64p]
    lib.elfel_hip_set_element_dofs.argtypes = [P, ctypes.c_int, ctypes.c_int32, ctypes.c_int64, i64p,
                                               ctypes.c_int64]
    lib.elfel_hip_set_coefficients.argtypes = [P, ctypes.c_int64, f64p, f64p, f64p]
    lib.elfel_hip_assemble.argtypes = [P, ctypes.POINTER(Problem), i64p]
    lib.elfel_hip_copy_vector.argtypes = [P, f64p, ctypes.c_int64]
    lib.elfel_hip_element_matrices.argtypes = [P, ctypes.POINTER(Problem), ctypes.c_int, ctypes.POINTER(ctypes.c_float),
                                               ctypes.c_int64, f64p]
    lib.elfel_hip_assembly_info.argtypes = [P, ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int),
                                            ctypes.c_char_p, ctypes.c_int]
    lib.elfel_hip_synchronize.argtypes = [P]
    lib.elfel_hip_geometry_stats.argtypes = [P, f64p, i64p]
    lib.elfel_hip_number_dofs.argtypes = [P, ctypes.c_int32, i64p, ctypes.POINTER(ctypes.c_int32),
                                          ctypes.POINTER(ctypes.c_uint8), i64p]
    lib.elfel_hip_gather_sysvec.argtypes = [P, ctypes.c_int, ctypes.c_int64, f64p, ctypes.c_int64, f64p]
    lib.elfel_hip_scatter_sysvec.argtypes = [P, ctypes.c_int, ctypes.c_int64, f64p, ctypes.c_int64, f64p]
    lib.elfel_hip_colliding_errors.argtypes = [P, ctypes.c_int64, f64p, f64p, f64p]
    lib.elfel_hip_block_mesh.argtypes = [P, ctypes.c_int, ctypes.c_int, ctypes.c_double, ctypes.c_double,
                                         ctypes.c_int64, ctypes.c_int64, i64p, i64p]
    lib.elfel_hip_t6_to_t3.argtypes = [P, ctypes.c_int, ctypes.c_int, i64p]
    lib.elfel_hip_boundary_nodes.argtypes = [P, ctypes.c_int, ctypes.c_int64, ctypes.POINTER(ctypes.c_uint8), i64p]
    lib.elfel_hip_copy_mesh.argtypes = [P, ctypes.c_int, ctypes.c_int64, i64p, ctypes.c_int64, f64p]
    lib.elfel_hip_load_mesh.argtypes = [P, ctypes.c_int, ctypes.c_char_p, ctypes.POINTER(ctypes.c_int32), i64p, i64p]
    lib.elfel_hip_pcg.argtypes = [P, ctypes.c_int64, f64p, f64p, ctypes.c_int, ctypes.c_double, ctypes.c_int,
                                  ctypes.POINTER(ctypes.c_int), ctypes.POINTER(ctypes.c_int), f64p]
    lib.elfel_hip_batch_begin.argtypes = [P]
    lib.elfel_hip_batch_add.argtypes = [P, ctypes.POINTER(Problem), ctypes.c_int32, ctypes.c_int64, i64p,
                                        ctypes.c_int64, f64p, ctypes.c_int32, i64p, ctypes.c_int64, f64p]
    lib.elfel_hip_batch_build.argtypes = [P, ctypes.c_int]
    lib.elfel_hip_batch_assemble.argtypes = [P, ctypes.c_int]
    ip = ctypes.POINTER(ctypes.c_int)
    lib.elfel_hip_batch_info.argtypes = [P, ctypes.c_int32, i64p, i64p, i64p, ip, ip]
    lib.elfel_hip_batch_copy.argtypes = [P, ctypes.c_int32, ctypes.c_int, i64p, i64p, f64p, f64p]
    lib.elfel_hip_batch_stats.argtypes = [P, i64p, i64p, i64p, i64p, f64p]
    lib.elfel_hip_timing.argtypes = [P, ctypes.c_int]
    lib.elfel_hip_timing_read.argtypes = [P, f64p, i64p]
    for name in EXPORTED:
        if name != "elfel_hip_last_error":
            getattr(lib, name).restype = ctypes.c_int
    _lib = lib
    return lib


def _i64(a):
    return np.ascontiguousarray(a, dtype=np.int64)


def _f64(a):
    return np.ascontiguousarray(a, dtype=np.float64)


def _ptr(a, t):
    if a is None:
        return ctypes.cast(None, ctypes.POINTER(t))
    return a.ctypes.data_as(ctypes.POINTER(t))


class Context:
    """One device context = one HIP stream on one GPU (elfel_hip_init)."""

    def __init__(self, device=0):
        self.lib = load()
        h = ctypes.c_void_p()
        rc = self.lib.elfel_hip_init(device, ctypes.byref(h))
        if rc != OK:
            raise ElfelHipError(rc, f"elfel_hip_init({device}) failed: no usable HIP device")
        self.h = h

    def close(self):
        if getattr(self, "h", None):
            self.lib.elfel_hip_free(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != OK:
            msg = self.lib.elfel_hip_last_error(self.h).decode(errors="replace")
            cls = ElfelRangeError if rc == ERANGE else ElfelHipError
            raise cls(rc, f"{what}: {msg}")

    def device_name(self):
        buf = ctypes.create_string_buffer(256)
        cus = ctypes.c_int(0)
        self._chk(self.lib.elfel_hip_device_name(self.h, buf, 256, ctypes.byref(cus)), "device_name")
        return buf.value.decode(), cus.value

    def copy_bandwidth(self, nbytes=1 << 30, reps=10):
        """GB/s (read + write) of a plain device copy of nbytes."""
        g = ctypes.c_double()
        self._chk(self.lib.elfel_hip_copy_bandwidth(self.h, nbytes, reps, ctypes.byref(g)), "copy_bandwidth")
        return g.value

    # ---- generic drop-in --------------------------------------------------
    def sparse(self, I, J, V, nrow, ncol, layout=CSC):
        I, J, V = _i64(I), _i64(J), _f64(V)
        if not (I.size == J.size == V.size):
            raise ValueError("row, col and val must have the same length")
        nnz = ctypes.c_int64(0)
        self._chk(self.lib.elfel_hip_sparse(self.h, nrow, ncol, I.size, _ptr(I, ctypes.c_int64),
                                            _ptr(J, ctypes.c_int64), _ptr(V, ctypes.c_double), layout,
                                            ctypes.byref(nnz)), "sparse")
        return nnz.value

    def result_info(self):
        nr, nc, nnz, lay = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int()
        self._chk(self.lib.elfel_hip_result_info(self.h, ctypes.byref(nr), ctypes.byref(nc), ctypes.byref(nnz),
                                                 ctypes.byref(lay)), "result_info")
        return nr.value, nc.value, nnz.value, lay.value

    def copy_sparse(self, base=0):
        nr, nc, nnz, lay = self.result_info()
        nmaj = nc if lay == CSC else nr
        ptr = np.empty(nmaj + 1, np.int64)
        ind = np.empty(nnz, np.int64)
        val = np.empty(nnz, np.float64)
        self._chk(self.lib.elfel_hip_copy_sparse(self.h, base, _ptr(ptr, ctypes.c_int64), _ptr(ind, ctypes.c_int64),
                                                 _ptr(val, ctypes.c_double)), "copy_sparse")
        return ptr, ind, val, (nr, nc), lay

    def sparse_noalloc(self, nrow, ncol, n, row, col, val, layout, ptr, ind, out):
        """elfel_hip_sparse_noalloc on device arrays given by address (row, col:
        1-based int64; val float64; ptr, ind: int32; out: float64); returns nnz."""
        nnz = ctypes.c_int64(0)
        self._chk(self.lib.elfel_hip_sparse_noalloc(self.h, nrow, ncol, n, row, col, val, layout, ptr, ind, out,
                                                    ctypes.byref(nnz)), "sparse_noalloc")
        return nnz.value

    def vector(self, idx, val, ndofs):
        idx, val = _i64(idx), _f64(val)
        out = np.empty(ndofs, np.float64)
        self._chk(self.lib.elfel_hip_vector(self.h, ndofs, idx.size, _ptr(idx, ctypes.c_int64),
                                            _ptr(val, ctypes.c_double), _ptr(out, ctypes.c_double)), "vector")
        return out

    # ---- device-resident problem ----------------------------------------------
    def set_mesh(self, field, conn, xy):
        """conn: E x nen (1-based); xy: Nn x sdim coordinates (sdim 1, 2 or 3)."""
        conn, xy = _i64(conn), _f64(xy)
        if xy.ndim == 1:
            xy = xy.reshape(-1, 1)
        E, nen = conn.shape
        Nn, sdim = xy.shape
        if sdim == 2:
            rc = self.lib.elfel_hip_set_mesh(self.h, field, nen, E, _ptr(conn, ctypes.c_int64), Nn,
                                             _ptr(xy, ctypes.c_double))
        else:
            rc = self.lib.elfel_hip_set_mesh_sdim(self.h, field, nen, E, _ptr(conn, ctypes.c_int64), Nn, sdim,
                                                  _ptr(xy, ctypes.c_double))
        self._chk(rc, "set_mesh")
        self._set_sdim(field, sdim)

    def _set_sdim(self, field, sdim):
        self._sdim = getattr(self, "_sdim", {})
        self._sdim[field] = sdim

    def set_element_dofs(self, field, dofnums, ndofs):
        """Cell dofs (bubble, FEL2_Q4): E x ncomp 1-based numbers."""
        d = _i64(dofnums)
        if d.ndim == 1:
            d = d.reshape(-1, 1)
        self._chk(self.lib.elfel_hip_set_element_dofs(self.h, field, d.shape[1], d.shape[0],
                                                      _ptr(d, ctypes.c_int64), ndofs), "set_element_dofs")

    def set_dofs(self, field, dofnums, ndofs):
        d = _i64(dofnums)
        if d.ndim == 1:
            d = d.reshape(-1, 1)
        self._chk(self.lib.elfel_hip_set_dofs(self.h, field, d.shape[1], d.shape[0], _ptr(d, ctypes.c_int64), ndofs),
                  "set_dofs")

    def set_coefficients(self, E, kappa_e=None, Q_e=None, scale_e=None):
        ks = [None if a is None else _f64(a) for a in (kappa_e, Q_e, scale_e)]
        for a in ks:
            if a is not None and a.size != E:
                raise ValueError("per-element coefficient array has the wrong length")
        self._keep_coef = ks
        self._chk(self.lib.elfel_hip_set_coefficients(self.h, E, *[_ptr(a, ctypes.c_double) for a in ks]),
                  "set_coefficients")

    def assemble(self, kind, fe, rule, precision=FP64, method=SORT, layout=CSR, kappa=1.0, Q=0.0, mu=1.0,
                 D=None, wait=True, fe_p=T3, flags=0):
        pb = Problem()
        pb.kind, pb.fe, pb.rule, pb.fe_p, pb.flags = kind, fe, rule, fe_p, flags
        pb.precision, pb.method, pb.layout = precision, method, layout
        pb.kappa, pb.Q, pb.mu = kappa, Q, mu
        if D is not None:
            Dc = np.asarray(D, np.float64).reshape(3, 3).ravel(order="F")
            for i in range(9):
                pb.D[i] = Dc[i]
        self._pb = pb
        nnz = ctypes.c_int64(0)
        self._chk(self.lib.elfel_hip_assemble(self.h, ctypes.byref(pb), ctypes.byref(nnz) if wait else None),
                  "assemble")
        return nnz.value if wait else None

    def element_matrices(self, E, D, method=EM_VALU, copy=True):
        """ke of the E elements of field 0 (T6 plane elasticity, 3-point rule,
        mixed): (E, 144) float32, column-major per element, and the kernel ms
        (copy False: the kernel alone, returns (None, ms))."""
        pb = Problem()
        pb.kind, pb.fe, pb.rule, pb.precision, pb.method, pb.layout = ELAST, T6, 3, MIXED, REUSE, CSC
        Dc = np.asarray(D, np.float64).reshape(3, 3).ravel(order="F")
        for i in range(9):
            pb.D[i] = Dc[i]
        ke = np.empty((E, 144), np.float32) if copy else None
        ms = ctypes.c_double()
        kp = ke.ctypes.data_as(ctypes.POINTER(ctypes.c_float)) if copy else None
        self._chk(self.lib.elfel_hip_element_matrices(self.h, ctypes.byref(pb), method, kp, ke.size if copy else 0,
                                                      ctypes.byref(ms)), "element_matrices")
        return ke, ms.value

    def assembly_info(self):
        """(method, precision, note) of what the last assemble ran: on meshes
        its tiles do not hold, the REUSE method sums the runs of the plain
        method's stream in its cached sort order, in FP64 (REUSE, FP64; the
        note says why)."""
        m, p = ctypes.c_int(), ctypes.c_int()
        buf = ctypes.create_string_buffer(256)
        self._chk(self.lib.elfel_hip_assembly_info(self.h, ctypes.byref(m), ctypes.byref(p), buf, 256),
                  "assembly_info")
        return m.value, p.value, buf.value.decode()

    def copy_vector(self, n):
        F = np.empty(n, np.float64)
        self._chk(self.lib.elfel_hip_copy_vector(self.h, _ptr(F, ctypes.c_double), n), "copy_vector")
        return F

    def synchronize(self):
        self._chk(self.lib.elfel_hip_synchronize(self.h), "synchronize")

    def geometry_stats(self):
        md, nf = ctypes.c_double(), ctypes.c_int64()
        self._chk(self.lib.elfel_hip_geometry_stats(self.h, ctypes.byref(md), ctypes.byref(nf)), "geometry_stats")
        return md.value, nf.value

    # ---- dof numbering and system vectors -------------------------------------
    def number_dofs(self, isdatum_list):
        """numberdofs! on the device.  isdatum_list: one (nterms, ncomp) bool
        array per space; returns the 1-based dof numbers in the same shapes."""
        nterms = np.array([a.shape[0] for a in isdatum_list], np.int64)
        ncomp = np.array([a.shape[1] for a in isdatum_list], np.int32)
        flat = np.ascontiguousarray(np.concatenate([np.asarray(a, np.uint8).ravel() for a in isdatum_list]))
        out = np.zeros(flat.size, np.int64)
        self._chk(self.lib.elfel_hip_number_dofs(self.h, len(isdatum_list), _ptr(nterms, ctypes.c_int64),
                                                 ncomp.ctypes.data_as(ctypes.POINTER(ctypes.c_int32)),
                                                 flat.ctypes.data_as(ctypes.POINTER(ctypes.c_uint8)),
                                                 _ptr(out, ctypes.c_int64)), "number_dofs")
        res, o = [], 0
        for a in isdatum_list:
            res.append(out[o:o + a.size].reshape(a.shape).copy())
            o += a.size
        return res

    def gather_sysvec(self, field, dofvals, v):
        """v[dofnums - 1] = dofvals for the field bound by set_dofs (in place)."""
        d = _f64(dofvals).ravel()
        assert v.dtype == np.float64 and v.flags.c_contiguous
        self._chk(self.lib.elfel_hip_gather_sysvec(self.h, field, d.size, _ptr(d, ctypes.c_double), v.size,
                                                   _ptr(v, ctypes.c_double)), "gather_sysvec")
        return v

    def scatter_sysvec(self, field, v, dofvals):
        """dofvals = v[dofnums - 1] where the dof lies within v (in place)."""
        vv = _f64(v)
        assert dofvals.dtype == np.float64 and dofvals.flags.c_contiguous
        self._chk(self.lib.elfel_hip_scatter_sysvec(self.h, field, vv.size, _ptr(vv, ctypes.c_double), dofvals.size,
                                                    _ptr(dofvals, ctypes.c_double)), "scatter_sysvec")
        return dofvals

    def block_mesh(self, field, fe, A, B, nA, nB):
        """T3block / Q4block / T6block generated on the device into `field`;
        returns (E, Nn)."""
        E, Nn = ctypes.c_int64(), ctypes.c_int64()
        self._chk(self.lib.elfel_hip_block_mesh(self.h, field, fe, A, B, nA, nB, ctypes.byref(E), ctypes.byref(Nn)),
                  "block_mesh")
        self._set_sdim(field, 2)
        return E.value, Nn.value

    def t6_to_t3(self, src, dst):
        """T6toT3 of field src's mesh into field dst; returns the node count."""
        n = ctypes.c_int64()
        self._chk(self.lib.elfel_hip_t6_to_t3(self.h, src, dst, ctypes.byref(n)), "t6_to_t3")
        self._set_sdim(dst, 2)
        return n.value

    def boundary_nodes(self, field, Nn):
        """1-based numbers of the boundary nodes of `field`'s mesh (connectedv(boundary))."""
        flags = np.zeros(Nn, np.uint8)
        cnt = ctypes.c_int64()
        self._chk(self.lib.elfel_hip_boundary_nodes(self.h, field, Nn, _ptr(flags, ctypes.c_uint8),
                                                    ctypes.byref(cnt)), "boundary_nodes")
        b = np.nonzero(flags)[0] + 1
        assert b.size == cnt.value
        return b

    def load_mesh(self, field, path):
        """Reads a .mesh manifest into `field`; returns (nen, E, Nn)."""
        nen, E, Nn = ctypes.c_int32(), ctypes.c_int64(), ctypes.c_int64()
        self._chk(self.lib.elfel_hip_load_mesh(self.h, field, os.fsencode(path), ctypes.byref(nen), ctypes.byref(E),
                                               ctypes.byref(Nn)), "load_mesh")
        self._set_sdim(field, 2)
        return nen.value, E.value, Nn.value

    def copy_mesh(self, field, E, nen, Nn, sdim=None):
        """(conn E x nen 1-based int64, xy Nn x sdim) of `field`'s mesh."""
        conn = np.zeros((E, nen), np.int64)
        if sdim is None:
            sdim = getattr(self, "_sdim", {}).get(field, 2)
        xy = np.zeros((Nn, sdim))
        self._chk(self.lib.elfel_hip_copy_mesh(self.h, field, E, _ptr(conn, ctypes.c_int64), Nn,
                                               _ptr(xy, ctypes.c_double)), "copy_mesh")
        return conn, xy

    def colliding_errors(self, U):
        """(pressure, velocity) L2 errors of the system vector U against the
        colliding-flow exact fields; field 0 vector T6, field 1 scalar T3."""
        U = _f64(U)
        ep, ev = ctypes.c_double(), ctypes.c_double()
        self._chk(self.lib.elfel_hip_colliding_errors(self.h, U.size, _ptr(U, ctypes.c_double), ctypes.byref(ep),
                                                      ctypes.byref(ev)), "colliding_errors")
        return ep.value, ev.value

    def pcg(self, nfree, F, x, precision=FP64, tol=1e-10, maxit=100000):
        F = _f64(F)
        x = _f64(x).copy()
        it, outer, rr = ctypes.c_int(), ctypes.c_int(), ctypes.c_double()
        self._chk(self.lib.elfel_hip_pcg(self.h, nfree, _ptr(F, ctypes.c_double), _ptr(x, ctypes.c_double), precision,
                                         tol, maxit, ctypes.byref(it), ctypes.byref(outer), ctypes.byref(rr)), "pcg")
        return x, it.value, outer.value, rr.value

    # ---- batched problems ----------------------------------------------------
    def batch_begin(self):
        self._chk(self.lib.elfel_hip_batch_begin(self.h), "batch_begin")

    def batch_add(self, kind, fe, rule, conn, xy, dofnums, ndofs, coef_e=None, kappa=1.0, Q=0.0, D=None,
                  layout=CSC, precision=MIXED):
        pb = Problem()
        pb.kind, pb.fe, pb.rule, pb.precision, pb.method, pb.layout = kind, fe, rule, precision, REUSE, layout
        pb.kappa, pb.Q, pb.mu = kappa, Q, 1.0
        if D is not None:
            Dc = np.asarray(D, np.float64).reshape(3, 3).ravel(order="F")
            for i in range(9):
                pb.D[i] = Dc[i]
        conn, xy, d = _i64(conn), _f64(xy), _i64(dofnums)
        if d.ndim == 1:
            d = d.reshape(-1, 1)
        ce = None if coef_e is None else _f64(coef_e)
        self._chk(self.lib.elfel_hip_batch_add(self.h, ctypes.byref(pb), conn.shape[1], conn.shape[0],
                                               _ptr(conn, ctypes.c_int64), xy.shape[0], _ptr(xy, ctypes.c_double),
                                               d.shape[1], _ptr(d, ctypes.c_int64), ndofs,
                                               _ptr(ce, ctypes.c_double)), "batch_add")

    def batch_build(self, autotune=False):
        self._chk(self.lib.elfel_hip_batch_build(self.h, 1 if autotune else 0), "batch_build")

    def batch_assemble(self, wait=True):
        self._chk(self.lib.elfel_hip_batch_assemble(self.h, 1 if wait else 0), "batch_assemble")

    def batch_info(self, k):
        nd, nnz, ne = ctypes.c_int64(), ctypes.c_int64(), ctypes.c_int64()
        bl, gr = ctypes.c_int(), ctypes.c_int()
        self._chk(self.lib.elfel_hip_batch_info(self.h, k, ctypes.byref(nd), ctypes.byref(nnz), ctypes.byref(ne),
                                                ctypes.byref(bl), ctypes.byref(gr)), "batch_info")
        return dict(ndofs=nd.value, nnz=nnz.value, nelem=ne.value, block=bl.value, grid=gr.value)

    def batch_copy(self, k, base=0):
        info = self.batch_info(k)
        n, nnz = info["ndofs"], info["nnz"]
        ptr, ind, val, F = np.empty(n + 1, np.int64), np.empty(nnz, np.int64), np.empty(nnz), np.zeros(n)
        self._chk(self.lib.elfel_hip_batch_copy(self.h, k, base, _ptr(ptr, ctypes.c_int64), _ptr(ind, ctypes.c_int64),
                                                _ptr(val, ctypes.c_double), _ptr(F, ctypes.c_double)), "batch_copy")
        return ptr, ind, val, F

    def batch_stats(self):
        a = [ctypes.c_int64() for _ in range(4)]
        t = ctypes.c_double()
        self._chk(self.lib.elfel_hip_batch_stats(self.h, *[ctypes.byref(x) for x in a], ctypes.byref(t)),
                  "batch_stats")
        return dict(nprob=a[0].value, nelem=a[1].value, nnz=a[2].value, ntiles=a[3].value, tuned_ms=t.value)

    def timing(self, enable=True):
        self._chk(self.lib.elfel_hip_timing(self.h, 1 if enable else 0), "timing")

    def timing_region(self):
        """One hipEvent pair around everything submitted until timing_read()
        (whose "total" entry then holds the elapsed ms and the assembly count)."""
        self._chk(self.lib.elfel_hip_timing(self.h, 2), "timing")

    def timing_read(self):
        ms = np.zeros(NPHASES)
        n = np.zeros(NPHASES, np.int64)
        self._chk(self.lib.elfel_hip_timing_read(self.h, _ptr(ms, ctypes.c_double), _ptr(n, ctypes.c_int64)),
                  "timing_read")
        return {PHASE_NAMES[i]: (float(ms[i]), int(n[i])) for i in range(NPHASES)}
