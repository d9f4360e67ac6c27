// common.h -- shared host/device declarations of libelfel_hip.so (gfx950).
//
// The library integrates finite elements and assembles the global sparse
// matrix on one MI355X.  Layout in HBM (all 0-based int32 indices):
//   conn   [E][nen]        element-major node numbers (the Julia layout)
//   xy     [Nn] double2    node coordinates
//   dof    [Nn][ncomp]     dof numbers (FEField.dofnums, 0-based)
//   geom   [E][nen][2]     per-element packed coordinates written by the
//                          gather kernel (FP64 absolute, or FP32 relative to
//                          the element's first node for the mixed kernels)
//   CSR/CSC rowptr [nmaj+1], colind [nnz], values [nnz] double.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <exception>
#include <functional>
#include <new>
#include <string>
#include <thread>
#include <vector>

#include "../../include/elfel_hip.h"
#include "qtable.h"

// Quadrature / basis tables (qtable.h) reach the FP64 kernels by value: every
// index into the table is wave-uniform, so the loads are scalar (SGPR) loads
// from the kernel argument segment.
static_assert(QT_T3 == ELFEL_T3 && QT_T6 == ELFEL_T6 && QT_Q4 == ELFEL_Q4, "element codes");

// Padded storage of the large element- and node-indexed arrays that the
// assembly kernels gather from: after every 4096 entries the array skips a
// gap of a few entries, so entry i sits at padded(i, gap).  On a structured
// mesh whose rows hold a power of two of elements (4096^2 quads) the element
// rows of an array are otherwise an exact power of two of bytes apart, and the
// rows that concurrent tiles read map to the same L2 sets: measured 23 against
// 73 % L2 hits and 1.17 against 0.90 ms per heat assembly on 4096^2 quads.
// Each array kind has its own gap (in entries), so that the rows of different
// arrays do not collide with one another either.
#define PAD_LOG2 12
enum { COEF_GAP = 16, CONN_GAP = 24, XY_GAP = 40 };
__host__ __device__ __forceinline__ int64_t padded(int64_t i, int gap) { return i + (i >> PAD_LOG2) * gap; }
__host__ __device__ __forceinline__ int32_t padded32(int32_t i, int gap) { return i + (i >> PAD_LOG2) * gap; }
// entries to allocate for n padded entries
__host__ __device__ __forceinline__ int64_t padded_len(int64_t n, int gap) { return padded(n, gap) + 1; }

// Coefficients of the weak form (uniform values + optional per-element arrays,
// stored padded with gap cgap: 0 for plain arrays).
struct Coef {
    double kappa, Q, mu;
    double D[9];  // column-major
    const double *kappa_e, *Q_e, *scale_e;
    int cgap;
};

// ---- device buffers ---------------------------------------------------------
struct Buf {
    void *p = nullptr;     // start of the usable bytes (base + colour offset)
    void *base = nullptr;  // what hipMalloc returned
    size_t bytes = 0;
};

struct Field {
    int nen = 0, ncomp = 0;
    int sdim = 2;                // space dimension of the coordinates (xy holds sdim doubles per node)
    int64_t E = 0, Nn = 0, ndofs = 0;
    int64_t dof_Nn = 0;          // node count of the dof map (set_dofs)
    bool has_mesh = false, has_dofs = false;
    Buf conn, xy, dof;
    // cell dofs [E][edof_ncomp] (bubble, FEL2_Q4), 0-based
    int edof_ncomp = 0;
    int64_t edof_E = 0, edof_ndofs = 0;
    bool has_edofs = false;
    uint64_t edof_version = 0;
    Buf edof;
    std::vector<int32_t> h_conn, h_dof;  // host copies (0-based) for the symbolic phase
    uint64_t mesh_version = 0, dof_version = 0;
};

// A new mesh of a field: a dof map of another node count, or cell dofs of
// another element count, no longer describe it and must be set again (the
// kernels would index them out of range).
inline void field_mesh_changed(Field &f, int64_t E, int64_t Nn)
{
    if (f.has_dofs && f.dof_Nn != Nn) f.has_dofs = false;
    if (f.has_edofs && f.edof_E != E) f.has_edofs = false;
}

// Pattern cached for the REUSE method (built once per mesh + dof map).
struct Pattern {
    bool valid = false;
    int kind = -1, fe = -1, layout = -1;
    uint64_t key_mesh0 = 0, key_dof0 = 0, key_mesh1 = 0, key_dof1 = 0;
    int64_t nrow = 0, ncol = 0, nnz = 0;
    int64_t ntiles = 0, ntasks = 0, ngroups = 0;
    int ncolmax = 0;             // bytes of pos per task
    int maxrank = 0;
    int lds_vals = 0;            // largest tile value count
    int block = 0;               // task slots per tile (threads per workgroup x tasks per thread)
    double symbolic_ms = 0;      // host wall time of the last symbolic build
    Buf ptr, ind;                // CSR pattern (int32)
    Buf tile;                    // int4 per tile: task0, ntasks, v0, nvals | (row0 in .w)
    Buf tile_rows;               // int2 per tile: row0, nrows
    Buf task_elem;               // int32 per task (uint8 for element-staged tiles, elem8)
    bool elem8 = false;
    Buf task_info;               // uint32 per task
    Buf task_pos;                // uint8 [ntasks][ncolmax]
    Buf tile_elem;               // int32 [ntiles][slots]: distinct elements of a tile (heat)
    Buf conn_pad, xy_pad;        // field 0's connectivity and coordinates, padded (CONN_GAP, XY_GAP)
    Buf truns, runs;             // patch tiles: per-tile run ranges and the runs (empty: contiguous tiles)
    bool has_runs = false;
    int maxu = 0;                // most distinct elements in one tile
    // the mesh does not fit the tiles (a node in more than 16 elements, a row
    // of more than 128 entries, a dof shared by two nodes, an element that
    // repeats a dof): assemblies of this pattern sum the runs of the plain
    // method's stream in its cached sort order, in FP64 (stream_reuse)
    bool fallback = false;
    std::string fallback_note;
    // row-owner heat tiles: one record of rw words per row (rw 0: task tables),
    // tile_elem with tstride entries per tile
    Buf rrec;
    Buf tconn, txy;              // row-owner heat tiles: tile-local nodes (ushort4 per slot) and coordinates
    Buf sym, sym2;               // scratch of the device symbolic phase (symbolic.hip)
    bool tgeom = false;          // tconn / txy hold the tile-local geometry
    int rw = 0, tstride = 0;
};

// Pattern cached for the element families on the REUSE method (families.hip):
// the stable sort of the COO stream done once, kept as a permutation -- for
// each sorted position, the stream position of its entry -- and the start of
// each matrix entry's run in sorted order (the load vector likewise, one run
// per dof).  Built once per mesh + dof map + weak form.
struct FamPattern {
    bool valid = false;
    int kind = -1, fe = -1, fe_p = -1, layout = -1;
    uint64_t key[6] = {0, 0, 0, 0, 0, 0};  // mesh, dof and cell-dof versions of fields 0 and 1
    int64_t n = 0, nv = 0, nnz = 0, ndofs = 0;
    double symbolic_ms = 0;
    Buf ptr, ind;     // CSR / CSC pattern (int32)
    Buf perm, seg;    // int32 [n] stream position per sorted position; int32 [nnz + 1] run starts
    Buf fperm, fbeg;  // load vector: int32 [nv]; int32 [ndofs + 1]
    Buf coo, fcoo;    // the element values of one assembly, in stream order (FP64, or FP32 for MIXED)
};

struct Timing {
    bool on = false;
    // region mode: one event pair around everything submitted between
    // elfel_hip_timing(ctx, 2) and elfel_hip_timing_read; region_n counts the
    // assemblies (single or batched) in between
    bool region = false;
    hipEvent_t r0 = nullptr, r1 = nullptr;
    int64_t region_n = 0;
    double ms[ELFEL_NPHASES] = {0};
    int64_t n[ELFEL_NPHASES] = {0};
    std::vector<hipEvent_t> pool;
    struct Pending { int phase; hipEvent_t a, b; };
    std::vector<Pending> pending;
};

struct elfel_hip_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    std::string err;
    Field f[2];
    Buf kappa_e, Q_e, scale_e;
    int64_t coef_E = 0;
    bool has_kappa_e = false, has_Q_e = false, has_scale_e = false;
    // scratch
    Buf geom;                     // gathered element geometry
    Buf keys[2], vals[2];         // COO keys/values (double-buffered for the sort)
    Buf hist, scan_tmp, flags, scal;
    Buf fkeys[2], fvals[2];
    Buf stage64;                  // int64 staging for copy-out
    Buf segpos;                   // segment head positions of the sorted COO stream (+ the count)
    // result
    Buf ptr, ind, val, F;
    int64_t nmaj = 0, nmin = 0, nnz = -1, nrow = 0, ncol = 0, nF = 0;
    int layout = 0;
    bool has_result = false, has_F = false;
    int result_src = 0;   // where ptr/ind of the result live: RES_PLAIN, RES_PATTERN, RES_FAMILY
    int used_method = -1, used_precision = -1;  // what the last assembly ran (elfel_hip_assembly_info)
    std::string used_note;
    Pattern pat;
    FamPattern fpat;
    // device scalars: [0] nnz (int32), [1] error flags; int64 counts of badly shaped elements (byte 24) and
    // of elements with det J <= 0 (byte 32)
    Buf dscal;
    Timing tm;
    // solver scratch: vectors, then the SELL-64 copy of the free block
    Buf pcg, pcg_sell;
    // batched problems (batch.hip)
    void *batch = nullptr;
};

enum { RES_PLAIN = 0, RES_PATTERN = 1, RES_FAMILY = 2 };
// The device ptr / ind of the last assembly's matrix.
inline const int32_t *result_ptr(const elfel_hip_ctx *c)
{
    return (const int32_t *)(c->result_src == RES_PATTERN ? c->pat.ptr.p
                             : c->result_src == RES_FAMILY ? c->fpat.ptr.p : c->ptr.p);
}
inline const int32_t *result_ind(const elfel_hip_ctx *c)
{
    return (const int32_t *)(c->result_src == RES_PATTERN ? c->pat.ind.p
                             : c->result_src == RES_FAMILY ? c->fpat.ind.p : c->ind.p);
}

// ---- error handling -------------------------------------------------------
int set_err(elfel_hip_ctx *c, int code, const std::string &msg);
#define HIPCHK(c, call)                                                        \
    do {                                                                       \
        hipError_t e_ = (call);                                                \
        if (e_ != hipSuccess)                                                  \
            return set_err((c), ELFEL_EHIP, std::string(#call) + ": " + hipGetErrorString(e_)); \
    } while (0)

// No C++ exception crosses the C ABI or ends a host worker thread: the host
// tables are std containers, and a failed host allocation (std::bad_alloc) or
// any other exception becomes an error code and message.
inline int exception_code(std::string *msg) noexcept
{
    try {
        throw;
    } catch (const std::bad_alloc &) {
        if (msg) *msg = "host allocation failed";
        return ELFEL_ENOMEM;
    } catch (const std::exception &e) {
        if (msg) *msg = std::string("host exception: ") + e.what();
        return ELFEL_EINVAL;
    } catch (...) {
        if (msg) *msg = "host exception";
        return ELFEL_EINVAL;
    }
}
template <typename F>
int host_guard(elfel_hip_ctx *c, F &&f) noexcept
{
    try {
        return f();
    } catch (...) {
        std::string m;
        const int code = exception_code(&m);
        return set_err(c, code, m);
    }
}
// Runs `before` (wakes waiting threads) and joins a set of host threads on
// every way out of a scope, an exception included: no joinable std::thread is
// ever destroyed.
template <typename F>
struct ThreadJoiner {
    std::vector<std::thread> &ts;
    F before;
    ~ThreadJoiner()
    {
        before();
        for (auto &t : ts)
            if (t.joinable()) t.join();
    }
};
template <typename F>
ThreadJoiner<F> join_on_exit(std::vector<std::thread> &ts, F before) { return ThreadJoiner<F>{ts, before}; }

// Grows a buffer to at least `bytes` (contents are not preserved).
int ensure(elfel_hip_ctx *c, Buf &b, size_t bytes);
int copy_padded(elfel_hip_ctx *c, void *dst, const void *src, int64_t n, size_t esize, int gap);
void release(Buf &b);

// ---- timing -----------------------------------------------------------------
void tm_begin(elfel_hip_ctx *c, int phase, hipEvent_t *a);
void tm_end(elfel_hip_ctx *c, int phase, hipEvent_t a);
void tm_collect(elfel_hip_ctx *c);

// ---- device error flags (dscal[1]) ----------------------------------------
enum { DERR_RANGE = 1, DERR_GEOM = 2 };

// ---- quadrature tables (tables.cpp) ---------------------------------------
int make_qtable(int fe, int rule, QTable *t);

// ---- sort.hip ---------------------------------------------------------------
// Scratch of the sort and the segmented reduce handed in by the caller (the
// allocation-free entry point); null: the context's buffers, grown on demand.
struct SortScratch {
    uint32_t *hist = nullptr;    // radix_hist_entries(n)
    int64_t hist_cap = 0;
    uint32_t *scan = nullptr;    // scan_scratch_entries of the longest scan
    int64_t scan_cap = 0;
    uint32_t *flags = nullptr;   // n entries
    uint32_t *segpos = nullptr;  // n + 1 entries
    int64_t n_cap = 0;
};
int bits_for(int64_t n);  // bits needed to hold 0..n-1
int64_t radix_hist_entries(int64_t n);
int64_t scan_scratch_entries(int64_t n);
int radix_sort_pairs(elfel_hip_ctx *c, uint64_t *&keys, uint64_t *&vals, uint64_t *&keys_alt,
                     uint64_t *&vals_alt, int64_t n, int nbits, const SortScratch *s = nullptr);
int exclusive_scan_u32(elfel_hip_ctx *c, uint32_t *data, int64_t n, uint32_t *total_dev,
                       const SortScratch *s = nullptr);
// From sorted (key, value) pairs build CSR: key = (major << minor_bits) | minor.
int build_csr_from_sorted(elfel_hip_ctx *c, const uint64_t *keys, const double *vals, int64_t n,
                          int minor_bits, int64_t nmaj, int32_t *ptr, int32_t *ind, double *val,
                          int32_t *nnz_dev, bool values, const SortScratch *s = nullptr);
// Sorted (row, value) pairs -> dense vector (rows without entries = 0).
int build_dense_from_sorted(elfel_hip_ctx *c, const uint64_t *keys, const double *vals, int64_t n,
                            int64_t ndofs, double *out);

// ---- plain.hip (FP64, unfused arithmetic) -------------------------------------
int plain_assemble(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf);
// COO buffers for n matrix entries and nv vector entries (plain method), and
// the sort + reduce that turns the keys written into keys[0]/vals[0] (and
// fkeys[0]/fvals[0]) into the CSR/CSC result and the load vector.
// vec: the problem has a load vector (nv pairs, possibly none).
int plain_buffers(elfel_hip_ctx *c, int64_t n, int64_t nv, int64_t ndofs, bool vec);
// ---- families.hip: the REUSE method over a cached sort of the COO stream -----
// What an element kernel writes: keys of the matrix and load-vector streams
// (null: values only) and their values (FP64, or FP32 when asked for f32).
struct StreamOut {
    uint64_t *keys, *fkeys;
    void *vals, *fvals;
};
typedef std::function<int(const StreamOut &, bool f32)> StreamElements;
// Assembles pb's problem by summing the runs of its stream through the cached
// sort permutation (c->fpat; built on the first call for this mesh, dof map
// and weak form): n matrix and nv load-vector stream entries (nv 0: no load
// vector), ndofs rows, keys with minor_bits of minor index.  The result: the
// sparse matrix (and load vector) in the context, FP64 values equal to the
// plain method's bit for bit.
int stream_reuse(elfel_hip_ctx *c, const elfel_hip_problem *pb, int64_t n, int64_t nv, int64_t ndofs, int minor_bits,
                 bool f32, const StreamElements &elements);
// The 2D weak forms (plain.hip) on stream_reuse, FP64: the REUSE method's
// path for meshes its tiles do not hold.
int plain_stream_reuse(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf);
int plain_finish(elfel_hip_ctx *c, int layout, int64_t n, int64_t nv, int64_t ndofs, int minor_bits, bool vec);

// ---- families.hip (element families, plain FP64 method) -------------------------
bool family_problem(const elfel_hip_problem *pb);
int family_validate(elfel_hip_ctx *c, const elfel_hip_problem *pb);  // sets c->nrow, c->ncol
int family_assemble(elfel_hip_ctx *c, const elfel_hip_problem *pb);
int convert_indices(elfel_hip_ctx *c, const int64_t *src_dev, int32_t *dst, int64_t n, int64_t lo,
                    int64_t hi);

// ---- reuse.hip ----------------------------------------------------------------
int reuse_assemble(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf);

// ---- batch.hip ----
void batch_free(elfel_hip_ctx *c);

// ---- solve.hip ----------------------------------------------------------------
int pcg_solve(elfel_hip_ctx *c, int64_t nfree, const double *F, double *x, int precision, double tol,
              int maxit, int *iters, int *outer, double *relres);

static inline int ceil_div(int64_t a, int64_t b) { return (int)((a + b - 1) / b); }
