// api.hip -- the extern "C" entry points of libelfel_hip.so (include/elfel_hip.h).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <new>

#include "common.h"

int plain_coo_keys(elfel_hip_ctx *c, const int64_t *I, const int64_t *J, int64_t n, int64_t nrow, int64_t ncol,
                   int minor_bits, int csc, uint64_t *keys);
int plain_vec_keys(elfel_hip_ctx *c, const int64_t *I, int64_t n, int64_t ndofs, uint64_t *keys);
int geometry_stats_run(elfel_hip_ctx *c, double *min_det, int64_t *nbad);
int geometry_check(elfel_hip_ctx *c, int rule, double *min_det, int64_t *nbad, int64_t *nneg);
int noalloc_init_device(int device);

// ---------------------------------------------------------------------------
// helpers
// ---------------------------------------------------------------------------
int set_err(elfel_hip_ctx *c, int code, const std::string &msg)
{
    if (c) {
        try {
            c->err = msg;
        } catch (...) {
            c->err.clear();
        }
    }
    return code;
}

// Every buffer starts at its own "colour": an offset of a few KiB past the
// allocation's (2 MiB-aligned) base, different for successive allocations.
// Arrays indexed by the same element or node number (coefficients, node
// numbers, coordinates) are read together; with power-of-two sizes (4096^2
// elements) their bases would otherwise agree in every low address bit and
// the simultaneous reads would fall on the same HBM channels (the heat
// assembly on 4096^2 quads: 1.24 against 1.19 ms per launch with colours).
static size_t next_colour()
{
    static const bool on = [] {
        const char *v = getenv("ELFEL_HIP_BUF_COLOUR");
        return !(v && *v && atoi(v) == 0);
    }();
    static std::atomic<unsigned> k{0};
    if (!on) return 0;
    return (size_t)((((k++) * 37u) % 61u) + 1u) * 768u;  // multiples of 256 B, up to ~46 KiB
}

// dst[padded(i, gap)] = src[i] for n entries of esize bytes (device to device,
// on the context's stream): full blocks of 4096 entries by one strided copy,
// the last partial block by a plain one.
int copy_padded(elfel_hip_ctx *c, void *dst, const void *src, int64_t n, size_t esize, int gap)
{
    const int64_t blk = (int64_t)1 << PAD_LOG2, rows = n >> PAD_LOG2, rem = n & (blk - 1);
    if (rows)
        HIPCHK(c, hipMemcpy2DAsync(dst, esize * (size_t)(blk + gap), src, esize * (size_t)blk, esize * (size_t)blk,
                                   (size_t)rows, hipMemcpyDeviceToDevice, c->stream));
    if (rem)
        HIPCHK(c, hipMemcpyAsync((char *)dst + esize * (size_t)(rows * (blk + gap)), (const char *)src + esize * (size_t)(rows * blk),
                                 esize * (size_t)rem, hipMemcpyDeviceToDevice, c->stream));
    return 0;
}

int ensure(elfel_hip_ctx *c, Buf &b, size_t bytes)
{
    if (bytes == 0) bytes = 16;
    if (b.bytes >= bytes) return 0;
    if (b.base) {
        // The stream may still use the old buffer.
        hipError_t e = hipStreamSynchronize(c->stream);
        if (e != hipSuccess) return set_err(c, ELFEL_EHIP, std::string("hipStreamSynchronize: ") + hipGetErrorString(e));
        hipFree(b.base);
        b.p = b.base = nullptr;
        b.bytes = 0;
    }
    const size_t colour = next_colour();
    size_t want = bytes + bytes / 8;
    hipError_t e = hipMalloc(&b.base, want + colour);
    if (e != hipSuccess) {
        (void)hipGetLastError();
        b.base = nullptr;
        e = hipMalloc(&b.base, bytes + colour);
        if (e != hipSuccess) {
            (void)hipGetLastError();
            b.base = nullptr;
            char msg[128];
            snprintf(msg, sizeof msg, "hipMalloc of %zu bytes failed", bytes);
            return set_err(c, ELFEL_ENOMEM, msg);
        }
        want = bytes;
    }
    b.p = (char *)b.base + colour;
    b.bytes = want;
    return 0;
}

void release(Buf &b)
{
    if (b.base) hipFree(b.base);
    b.p = b.base = nullptr;
    b.bytes = 0;
}

// ---- timing ----------------------------------------------------------------
static hipEvent_t ev_get(elfel_hip_ctx *c)
{
    if (!c->tm.pool.empty()) {
        hipEvent_t e = c->tm.pool.back();
        c->tm.pool.pop_back();
        return e;
    }
    hipEvent_t e = nullptr;
    if (hipEventCreate(&e) != hipSuccess) return nullptr;
    return e;
}

void tm_begin(elfel_hip_ctx *c, int phase, hipEvent_t *a)
{
    (void)phase;
    *a = nullptr;
    if (!c->tm.on) return;
    *a = ev_get(c);
    if (*a) hipEventRecord(*a, c->stream);
}

void tm_end(elfel_hip_ctx *c, int phase, hipEvent_t a)
{
    if (!c->tm.on || !a) return;
    hipEvent_t b = ev_get(c);
    if (!b) return;
    hipEventRecord(b, c->stream);
    c->tm.pending.push_back({phase, a, b});
    if (c->tm.pending.size() >= 240) tm_collect(c);
}

void tm_collect(elfel_hip_ctx *c)
{
    if (c->tm.pending.empty()) return;
    hipStreamSynchronize(c->stream);
    for (auto &p : c->tm.pending) {
        float ms = 0.f;
        if (hipEventElapsedTime(&ms, p.a, p.b) == hipSuccess) {
            c->tm.ms[p.phase] += ms;
            c->tm.n[p.phase] += 1;
        }
        c->tm.pool.push_back(p.a);
        c->tm.pool.push_back(p.b);
    }
    c->tm.pending.clear();
}

// Reads and clears the device error flags.
static int take_flags(elfel_hip_ctx *c, int *flags)
{
    int h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, (int *)c->dscal.p + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (h) HIPCHK(c, hipMemsetAsync((int *)c->dscal.p + 1, 0, sizeof(int), c->stream));
    *flags = h;
    return 0;
}

#define CHECK_CTX(c)                                  \
    do {                                              \
        if (!(c)) return ELFEL_EINVAL;                \
        c->err.clear();                               \
        (void)hipSetDevice(c->device);                \
    } while (0)

// ---------------------------------------------------------------------------
// context
// ---------------------------------------------------------------------------
extern "C" int elfel_hip_init(int device, elfel_hip_ctx **out)
{
    return host_guard(nullptr, [&]() -> int {
        if (!out) return ELFEL_EINVAL;
        *out = nullptr;
        int ndev = 0;
        if (hipGetDeviceCount(&ndev) != hipSuccess || ndev <= 0) return ELFEL_EHIP;
        if (device < 0 || device >= ndev) return ELFEL_EINVAL;
        elfel_hip_ctx *c = new (std::nothrow) elfel_hip_ctx();
        if (!c) return ELFEL_ENOMEM;
        c->device = device;
        if (hipSetDevice(device) != hipSuccess || hipStreamCreate(&c->stream) != hipSuccess) {
            delete c;
            return ELFEL_EHIP;
        }
        if (ensure(c, c->dscal, 64) || hipMemset(c->dscal.p, 0, 64) != hipSuccess) {
            hipStreamDestroy(c->stream);
            delete c;
            return ELFEL_ENOMEM;
        }
        // this device's address of the allocation-free entry point's scratch
        if (int rc = noalloc_init_device(device)) {
            release(c->dscal);
            hipStreamDestroy(c->stream);
            delete c;
            return rc;
        }
        *out = c;
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_free(elfel_hip_ctx *c)
{
    if (!c) return ELFEL_EINVAL;
    hipSetDevice(c->device);
    hipStreamSynchronize(c->stream);
    for (auto &f : c->f) { release(f.conn); release(f.xy); release(f.dof); release(f.edof); }
    Buf *bufs[] = {&c->kappa_e, &c->Q_e, &c->scale_e, &c->geom, &c->keys[0], &c->keys[1], &c->vals[0], &c->vals[1],
                   &c->hist, &c->scan_tmp, &c->flags, &c->scal, &c->fkeys[0], &c->fkeys[1], &c->fvals[0],
                   &c->fvals[1], &c->stage64, &c->segpos, &c->ptr, &c->ind, &c->val, &c->F, &c->dscal, &c->pcg, &c->pcg_sell,
                   &c->pat.ptr, &c->pat.ind, &c->pat.tile, &c->pat.tile_rows, &c->pat.task_elem,
                   &c->pat.task_info, &c->pat.task_pos, &c->pat.tile_elem, &c->pat.conn_pad, &c->pat.xy_pad,
                   &c->pat.truns, &c->pat.runs, &c->pat.rrec, &c->pat.tconn, &c->pat.txy, &c->pat.sym, &c->pat.sym2,
                   &c->fpat.ptr, &c->fpat.ind, &c->fpat.perm, &c->fpat.seg, &c->fpat.fperm, &c->fpat.fbeg,
                   &c->fpat.coo, &c->fpat.fcoo};
    for (Buf *b : bufs) release(*b);
    batch_free(c);
    for (auto &p : c->tm.pending) { hipEventDestroy(p.a); hipEventDestroy(p.b); }
    if (c->tm.r0) hipEventDestroy(c->tm.r0);
    if (c->tm.r1) hipEventDestroy(c->tm.r1);
    for (auto e : c->tm.pool) hipEventDestroy(e);
    hipStreamDestroy(c->stream);
    delete c;
    return ELFEL_OK;
}

extern "C" const char *elfel_hip_last_error(const elfel_hip_ctx *c)
{
    if (!c) return "null context";
    return c->err.c_str();
}

extern "C" int elfel_hip_device_name(elfel_hip_ctx *c, char *buf, int len, int *num_cus)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        hipDeviceProp_t p;
        HIPCHK(c, hipGetDeviceProperties(&p, c->device));
        if (buf && len > 0) {
            snprintf(buf, (size_t)len, "%s (%s)", p.name, p.gcnArchName);
        }
        if (num_cus) *num_cus = p.multiProcessorCount;
        return ELFEL_OK;
    });
}

// A plain device copy, 16 bytes per work-item: the measured bandwidth that the
// roofline fractions are read beside (the published 8 TB/s is the peak).
__global__ __launch_bounds__(256) void k_copy_u4(const uint4 *__restrict__ a, uint4 *__restrict__ b, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = a[i];
}

extern "C" int elfel_hip_copy_bandwidth(elfel_hip_ctx *c, int64_t bytes, int reps, double *gbs)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (bytes < 16 || (bytes & 15) || reps < 1 || !gbs) return set_err(c, ELFEL_EINVAL, "elfel_hip_copy_bandwidth: bad arguments");
        void *a = nullptr, *b = nullptr;
        hipEvent_t e0 = nullptr, e1 = nullptr;
        const int64_t n = bytes / 16;
        int rc = ELFEL_OK;
        float ms = 0;
        if (hipMalloc(&a, bytes) != hipSuccess || hipMalloc(&b, bytes) != hipSuccess) {
            rc = set_err(c, ELFEL_ENOMEM, "elfel_hip_copy_bandwidth: hipMalloc failed");
        } else if (hipMemsetAsync(a, 1, bytes, c->stream) != hipSuccess || hipEventCreate(&e0) != hipSuccess ||
                   hipEventCreate(&e1) != hipSuccess) {
            rc = set_err(c, ELFEL_EHIP, "elfel_hip_copy_bandwidth: set-up failed");
        } else {
            const dim3 grid((unsigned)ceil_div(n, 256));
            hipLaunchKernelGGL(k_copy_u4, grid, dim3(256), 0, c->stream, (const uint4 *)a, (uint4 *)b, n);  // warm-up
            hipEventRecord(e0, c->stream);
            for (int r = 0; r < reps; r++)
                hipLaunchKernelGGL(k_copy_u4, grid, dim3(256), 0, c->stream, (const uint4 *)a, (uint4 *)b, n);
            hipEventRecord(e1, c->stream);
            if (hipEventSynchronize(e1) != hipSuccess || hipEventElapsedTime(&ms, e0, e1) != hipSuccess || ms <= 0)
                rc = set_err(c, ELFEL_EHIP, "elfel_hip_copy_bandwidth: timing failed");
            else
                *gbs = 2.0 * (double)bytes * reps / (ms * 1e-3) / 1e9;
        }
        if (e0) hipEventDestroy(e0);
        if (e1) hipEventDestroy(e1);
        if (a) hipFree(a);
        if (b) hipFree(b);
        return rc;
    });
}

extern "C" int elfel_hip_synchronize(elfel_hip_ctx *c)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return ELFEL_OK;
    });
}

// ---------------------------------------------------------------------------
// generic drop-in: host COO -> sparse matrix
// ---------------------------------------------------------------------------
extern "C" int elfel_hip_sparse(elfel_hip_ctx *c, int64_t nrow, int64_t ncol, int64_t n, const int64_t *row,
                                const int64_t *col, const double *val, int layout, int64_t *nnz)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (nrow < 0 || ncol < 0 || n < 0 || (n > 0 && (!row || !col || !val)) ||
            (layout != ELFEL_CSR && layout != ELFEL_CSC))
            return set_err(c, ELFEL_EINVAL, "elfel_hip_sparse: bad arguments");
        if (nrow >= INT32_MAX || ncol >= INT32_MAX || n >= INT32_MAX)
            return set_err(c, ELFEL_ENOTSUP, "elfel_hip_sparse: dimensions beyond int32");
        const int csc = layout == ELFEL_CSC;
        const int64_t nmaj = csc ? ncol : nrow, nmin = csc ? nrow : ncol;
        const int minor_bits = bits_for(nmin), major_bits = bits_for(nmaj);
        int rc;
        if ((rc = ensure(c, c->keys[0], 8 * n)) || (rc = ensure(c, c->keys[1], 8 * n)) || (rc = ensure(c, c->vals[0], 8 * n)) ||
            (rc = ensure(c, c->vals[1], 8 * n)) || (rc = ensure(c, c->stage64, 16 * n)) ||
            (rc = ensure(c, c->ptr, 4 * (nmaj + 1))) || (rc = ensure(c, c->ind, 4 * n)) || (rc = ensure(c, c->val, 8 * n)))
            return rc;
        c->has_result = false;
        int64_t *dI = (int64_t *)c->stage64.p, *dJ = dI + n;
        if (n > 0) {
            HIPCHK(c, hipMemcpyAsync(dI, row, 8 * n, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(dJ, col, 8 * n, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(c->vals[0].p, val, 8 * n, hipMemcpyHostToDevice, c->stream));
        }
        uint64_t *k0 = (uint64_t *)c->keys[0].p, *k1 = (uint64_t *)c->keys[1].p;
        uint64_t *v0 = (uint64_t *)c->vals[0].p, *v1 = (uint64_t *)c->vals[1].p;
        if ((rc = plain_coo_keys(c, dI, dJ, n, nrow, ncol, minor_bits, csc, k0))) return rc;
        int flags = 0;
        if ((rc = take_flags(c, &flags))) return rc;
        if (flags & DERR_RANGE) return set_err(c, ELFEL_ERANGE, "row or column index out of range");
        if ((rc = radix_sort_pairs(c, k0, v0, k1, v1, n, major_bits + minor_bits))) return rc;
        if ((rc = build_csr_from_sorted(c, k0, (const double *)v0, n, minor_bits, nmaj, (int32_t *)c->ptr.p,
                                        (int32_t *)c->ind.p, (double *)c->val.p, (int32_t *)c->dscal.p, true)))
            return rc;
        int32_t h = 0;
        HIPCHK(c, hipMemcpyAsync(&h, c->dscal.p, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        c->nnz = h;
        c->nrow = nrow; c->ncol = ncol; c->nmaj = nmaj; c->nmin = nmin; c->layout = layout;
        c->has_result = true;
        c->result_src = RES_PLAIN;
        c->used_method = ELFEL_SORT;
        c->used_precision = ELFEL_FP64;
        c->used_note.clear();
        if (nnz) *nnz = h;
        return ELFEL_OK;
    });
}

__global__ void k_i32_to_i64(const int32_t *__restrict__ s, int64_t *__restrict__ d, int64_t n, int64_t base)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (int64_t)s[i] + base;
}

static int fetch_nnz(elfel_hip_ctx *c)
{
    if (c->nnz >= 0) return 0;
    int32_t h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, c->dscal.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    c->nnz = h;
    return 0;
}

extern "C" int elfel_hip_result_info(elfel_hip_ctx *c, int64_t *nrow, int64_t *ncol, int64_t *nnz, int *layout)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (!c->has_result) return set_err(c, ELFEL_ESTATE, "no matrix has been built");
        int rc = fetch_nnz(c);
        if (rc) return rc;
        if (nrow) *nrow = c->nrow;
        if (ncol) *ncol = c->ncol;
        if (nnz) *nnz = c->nnz;
        if (layout) *layout = c->layout;
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_copy_sparse(elfel_hip_ctx *c, int base, int64_t *ptr, int64_t *ind, double *val)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (!c->has_result) return set_err(c, ELFEL_ESTATE, "no matrix has been built");
        if (base != 0 && base != 1) return set_err(c, ELFEL_EINVAL, "base must be 0 or 1");
        int rc = fetch_nnz(c);
        if (rc) return rc;
        const int64_t nnz = c->nnz, np = c->nmaj + 1;
        const int32_t *dptr = result_ptr(c);
        const int32_t *dind = result_ind(c);
        if ((rc = ensure(c, c->stage64, 8 * std::max(np, nnz)))) return rc;
        int64_t *st = (int64_t *)c->stage64.p;
        if (ptr) {
            hipLaunchKernelGGL(k_i32_to_i64, dim3(ceil_div(np, 256)), dim3(256), 0, c->stream, dptr, st, np, (int64_t)base);
            HIPCHK(c, hipMemcpyAsync(ptr, st, 8 * np, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        if (ind && nnz > 0) {
            hipLaunchKernelGGL(k_i32_to_i64, dim3(ceil_div(nnz, 256)), dim3(256), 0, c->stream, dind, st, nnz, (int64_t)base);
            HIPCHK(c, hipMemcpyAsync(ind, st, 8 * nnz, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        if (val && nnz > 0) HIPCHK(c, hipMemcpyAsync(val, c->val.p, 8 * nnz, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return ELFEL_OK;
    });
}

// ---------------------------------------------------------------------------
// generic drop-in: SysvecAssembler
// ---------------------------------------------------------------------------
extern "C" int elfel_hip_vector(elfel_hip_ctx *c, int64_t ndofs, int64_t n, const int64_t *idx, const double *v,
                                double *out)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (ndofs < 0 || n < 0 || (n > 0 && (!idx || !v)) || (ndofs > 0 && !out))
            return set_err(c, ELFEL_EINVAL, "elfel_hip_vector: bad arguments");
        if (n >= INT32_MAX) return set_err(c, ELFEL_ENOTSUP, "elfel_hip_vector: too many entries");
        int rc;
        if ((rc = ensure(c, c->fkeys[0], 8 * n)) || (rc = ensure(c, c->fkeys[1], 8 * n)) ||
            (rc = ensure(c, c->fvals[0], 8 * n)) || (rc = ensure(c, c->fvals[1], 8 * n)) ||
            (rc = ensure(c, c->stage64, 8 * n)) || (rc = ensure(c, c->scal, 8 * ndofs)))
            return rc;
        uint64_t *k0 = (uint64_t *)c->fkeys[0].p, *k1 = (uint64_t *)c->fkeys[1].p;
        uint64_t *v0 = (uint64_t *)c->fvals[0].p, *v1 = (uint64_t *)c->fvals[1].p;
        if (n > 0) {
            HIPCHK(c, hipMemcpyAsync(c->stage64.p, idx, 8 * n, hipMemcpyHostToDevice, c->stream));
            HIPCHK(c, hipMemcpyAsync(v0, v, 8 * n, hipMemcpyHostToDevice, c->stream));
        }
        if ((rc = plain_vec_keys(c, (const int64_t *)c->stage64.p, n, ndofs, k0))) return rc;
        int flags = 0;
        if ((rc = take_flags(c, &flags))) return rc;
        if (flags & DERR_RANGE) return set_err(c, ELFEL_ERANGE, "dof number out of range (BoundsError)");
        if ((rc = radix_sort_pairs(c, k0, v0, k1, v1, n, bits_for(ndofs)))) return rc;
        if ((rc = build_dense_from_sorted(c, k0, (const double *)v0, n, ndofs, (double *)c->scal.p))) return rc;
        if (ndofs > 0) HIPCHK(c, hipMemcpyAsync(out, c->scal.p, 8 * ndofs, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return ELFEL_OK;
    });
}

// ---------------------------------------------------------------------------
// device-resident problem
// ---------------------------------------------------------------------------
extern "C" int elfel_hip_set_mesh_sdim(elfel_hip_ctx *c, int field, int32_t nen, int64_t E, const int64_t *conn,
                                       int64_t Nn, int32_t sdim, const double *x)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (field < 0 || field > 1 || nen < 1 || nen > ELFEL_MAXBF || E < 0 || Nn < 0 || sdim < 1 || sdim > 3 ||
            (E > 0 && !conn) || (Nn > 0 && !x))
            return set_err(c, ELFEL_EINVAL, "elfel_hip_set_mesh: bad arguments");
        if (E * nen >= INT32_MAX || Nn >= INT32_MAX) return set_err(c, ELFEL_ENOTSUP, "mesh beyond int32 indexing");
        Field &f = c->f[field];
        int rc;
        if ((rc = ensure(c, f.conn, 4 * E * nen)) || (rc = ensure(c, f.xy, 8 * sdim * Nn)) ||
            (rc = ensure(c, c->stage64, 8 * E * nen)))
            return rc;
        f.has_mesh = false;
        if (E > 0) HIPCHK(c, hipMemcpyAsync(c->stage64.p, conn, 8 * E * nen, hipMemcpyHostToDevice, c->stream));
        if (Nn > 0) HIPCHK(c, hipMemcpyAsync(f.xy.p, x, 8 * sdim * Nn, hipMemcpyHostToDevice, c->stream));
        if ((rc = convert_indices(c, (const int64_t *)c->stage64.p, (int32_t *)f.conn.p, E * nen, 1, Nn))) return rc;
        int flags = 0;
        if ((rc = take_flags(c, &flags))) return rc;
        if (flags & DERR_RANGE) return set_err(c, ELFEL_ERANGE, "connectivity refers to a node outside 1..Nn");
        field_mesh_changed(f, E, Nn);
        f.nen = nen; f.E = E; f.Nn = Nn; f.sdim = sdim;
        f.h_conn.resize((size_t)(E * nen));
        for (int64_t i = 0; i < E * nen; i++) f.h_conn[i] = (int32_t)(conn[i] - 1);
        f.has_mesh = true;
        f.mesh_version++;
        c->pat.valid = false;
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_set_mesh(elfel_hip_ctx *c, int field, int32_t nen, int64_t E, const int64_t *conn, int64_t Nn,
                                  const double *xy)
{
    return host_guard(c, [&]() -> int {
        return elfel_hip_set_mesh_sdim(c, field, nen, E, conn, Nn, 2, xy);
    });
}

extern "C" int elfel_hip_set_element_dofs(elfel_hip_ctx *c, int field, int32_t ncomp, int64_t E, const int64_t *dofnums,
                                          int64_t ndofs)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (field < 0 || field > 1 || ncomp < 1 || ncomp > 2 || E < 0 || ndofs < 0 || (E > 0 && !dofnums))
            return set_err(c, ELFEL_EINVAL, "elfel_hip_set_element_dofs: bad arguments");
        if (ndofs >= INT32_MAX / 2 || E * ncomp >= INT32_MAX) return set_err(c, ELFEL_ENOTSUP, "beyond int32 indexing");
        Field &f = c->f[field];
        const int64_t n = E * ncomp;
        int rc;
        if ((rc = ensure(c, f.edof, 4 * n)) || (rc = ensure(c, c->stage64, 8 * n))) return rc;
        f.has_edofs = false;
        if (n > 0) HIPCHK(c, hipMemcpyAsync(c->stage64.p, dofnums, 8 * n, hipMemcpyHostToDevice, c->stream));
        if ((rc = convert_indices(c, (const int64_t *)c->stage64.p, (int32_t *)f.edof.p, n, 1, ndofs))) return rc;
        int flags = 0;
        if ((rc = take_flags(c, &flags))) return rc;
        if (flags & DERR_RANGE) return set_err(c, ELFEL_ERANGE, "cell dof number outside 1..ndofs (unnumbered dof 0?)");
        f.edof_ncomp = ncomp; f.edof_E = E; f.edof_ndofs = ndofs;
        f.has_edofs = true;
        f.edof_version++;
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_set_dofs(elfel_hip_ctx *c, int field, int32_t ncomp, int64_t Nn, const int64_t *dofnums,
                                  int64_t ndofs)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (field < 0 || field > 1 || ncomp < 1 || ncomp > 2 || Nn < 0 || ndofs < 0 || (Nn > 0 && !dofnums))
            return set_err(c, ELFEL_EINVAL, "elfel_hip_set_dofs: bad arguments");
        if (ndofs >= INT32_MAX / 2) return set_err(c, ELFEL_ENOTSUP, "dof count beyond int32 indexing");
        Field &f = c->f[field];
        const int64_t n = Nn * ncomp;
        int rc;
        if ((rc = ensure(c, f.dof, 4 * n)) || (rc = ensure(c, c->stage64, 8 * n))) return rc;
        f.has_dofs = false;
        if (n > 0) HIPCHK(c, hipMemcpyAsync(c->stage64.p, dofnums, 8 * n, hipMemcpyHostToDevice, c->stream));
        if ((rc = convert_indices(c, (const int64_t *)c->stage64.p, (int32_t *)f.dof.p, n, 1, ndofs))) return rc;
        int flags = 0;
        if ((rc = take_flags(c, &flags))) return rc;
        if (flags & DERR_RANGE)
            return set_err(c, ELFEL_ERANGE, "dof number outside 1..ndofs (unnumbered dof 0?)");
        if (f.Nn != Nn && f.has_mesh) return set_err(c, ELFEL_EINVAL, "dof map size does not match the mesh");
        f.ncomp = ncomp; f.ndofs = ndofs; f.dof_Nn = Nn;
        f.h_dof.resize((size_t)n);
        for (int64_t i = 0; i < n; i++) f.h_dof[i] = (int32_t)(dofnums[i] - 1);
        f.has_dofs = true;
        f.dof_version++;
        c->pat.valid = false;
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_set_coefficients(elfel_hip_ctx *c, int64_t E, const double *kappa_e, const double *Q_e,
                                          const double *scale_e)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (E < 0) return set_err(c, ELFEL_EINVAL, "elfel_hip_set_coefficients: bad E");
        struct { const double *src; Buf *dst; bool *flag; } items[3] = {
            {kappa_e, &c->kappa_e, &c->has_kappa_e}, {Q_e, &c->Q_e, &c->has_Q_e}, {scale_e, &c->scale_e, &c->has_scale_e}};
        for (auto &it : items) {
            *it.flag = false;
            if (!it.src) continue;
            // padded storage: one host-to-device copy into the staging buffer,
            // then a strided device copy into blocks of 4096 entries + COEF_GAP
            int rc = ensure(c, *it.dst, 8 * padded_len(E, COEF_GAP));
            if (!rc) rc = ensure(c, c->stage64, 8 * E);
            if (rc) return rc;
            if (E > 0) {
                HIPCHK(c, hipMemcpyAsync(c->stage64.p, it.src, 8 * E, hipMemcpyHostToDevice, c->stream));
                if ((rc = copy_padded(c, it.dst->p, c->stage64.p, E, 8, COEF_GAP))) return rc;
            }
            *it.flag = true;
        }
        c->coef_E = E;
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return ELFEL_OK;
    });
}

static int validate_problem(elfel_hip_ctx *c, const elfel_hip_problem *pb, QTable *tab, Coef *cf)
{
    if (!pb) return set_err(c, ELFEL_EINVAL, "null problem");
    const Field &F = c->f[0];
    if (!F.has_mesh || !F.has_dofs) return set_err(c, ELFEL_ESTATE, "set_mesh/set_dofs of field 0 first");
    if (F.sdim != 2) return set_err(c, ELFEL_EINVAL, "this weak form needs a mesh in two space dimensions");
    if (pb->kind < ELFEL_HEAT || pb->kind > ELFEL_STOKES_VECLAP) return set_err(c, ELFEL_EINVAL, "bad kind");
    const int want_nen = pb->fe == ELFEL_T3 ? 3 : pb->fe == ELFEL_T6 ? 6 : pb->fe == ELFEL_Q4 ? 4 : -1;
    if (want_nen < 0 || want_nen != F.nen) return set_err(c, ELFEL_EINVAL, "element type does not match the mesh");
    const bool stokes = pb->kind == ELFEL_STOKES_GEN || pb->kind == ELFEL_STOKES_VECLAP;
    const int want_nc = pb->kind == ELFEL_HEAT ? 1 : 2;
    if (F.ncomp != want_nc) return set_err(c, ELFEL_EINVAL, "number of field components does not match the weak form");
    if (stokes) {
        const Field &P = c->f[1];
        if (pb->fe != ELFEL_T6) return set_err(c, ELFEL_ENOTSUP, "Taylor-Hood Stokes needs T6 velocity");
        if (!P.has_mesh || !P.has_dofs || P.nen != 3 || P.E != F.E || P.ncomp != 1 || P.sdim != 2)
            return set_err(c, ELFEL_ESTATE, "Stokes needs the T3 pressure field (field 1) with the same elements");
    }
    if (pb->precision != ELFEL_FP64 && pb->precision != ELFEL_MIXED) return set_err(c, ELFEL_EINVAL, "bad precision");
    if (pb->method != ELFEL_SORT && pb->method != ELFEL_REUSE) return set_err(c, ELFEL_EINVAL, "bad method");
    if (pb->method == ELFEL_SORT && pb->precision != ELFEL_FP64)
        return set_err(c, ELFEL_ENOTSUP, "the sort method is the plain FP64 method");
    if (pb->layout != ELFEL_CSR && pb->layout != ELFEL_CSC) return set_err(c, ELFEL_EINVAL, "bad layout");
    if (make_qtable(pb->fe, pb->rule, tab)) return set_err(c, ELFEL_ENOTSUP, "unsupported quadrature rule");
    if ((c->has_kappa_e || c->has_Q_e || c->has_scale_e) && c->coef_E != F.E)
        return set_err(c, ELFEL_EINVAL, "per-element coefficients do not match the element count");
    cf->kappa = pb->kappa; cf->Q = pb->Q; cf->mu = pb->mu;
    for (int i = 0; i < 9; i++) cf->D[i] = pb->D[i];
    cf->kappa_e = c->has_kappa_e ? (const double *)c->kappa_e.p : nullptr;
    cf->Q_e = c->has_Q_e ? (const double *)c->Q_e.p : nullptr;
    cf->scale_e = c->has_scale_e ? (const double *)c->scale_e.p : nullptr;
    cf->cgap = COEF_GAP;
    int64_t ndofs = F.ndofs;
    if (stokes) ndofs = std::max(ndofs, c->f[1].ndofs);
    c->nrow = c->ncol = ndofs;
    return 0;
}

extern "C" int elfel_hip_assemble(elfel_hip_ctx *c, const elfel_hip_problem *pb, int64_t *nnz)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        QTable tab;
        Coef cf;
        const bool fam = pb && family_problem(pb);
        int rc = fam ? family_validate(c, pb) : validate_problem(c, pb, &tab, &cf);
        if (rc) return rc;
        if (pb->flags & ~ELFEL_CHECK_GEOM) return set_err(c, ELFEL_EINVAL, "unknown problem flags");
        if (pb->flags & ELFEL_CHECK_GEOM) {
            double mn = 0;
            int64_t nbad = 0, nneg = 0;
            if ((rc = geometry_check(c, pb->rule, &mn, &nbad, &nneg))) return rc;
            if (nneg > 0) {
                char msg[160];
                snprintf(msg, sizeof msg, "%lld element(s) with det J <= 0 at a quadrature point (smallest %.3e)",
                         (long long)nneg, mn);
                return set_err(c, ELFEL_EGEOM, msg);
            }
        }
        c->has_result = false;
        c->used_method = pb->method;
        c->used_precision = pb->precision;
        c->used_note.clear();
        hipEvent_t ev = nullptr;
        tm_begin(c, ELFEL_PH_TOTAL, &ev);
        if (c->tm.region) c->tm.region_n++;
        if (fam) {
            c->nnz = -1;
            rc = family_assemble(c, pb);
        } else if (pb->method == ELFEL_SORT) {
            c->nnz = -1;
            rc = plain_assemble(c, pb, tab, cf);
        } else {
            rc = reuse_assemble(c, pb, tab, cf);
        }
        if (rc) return rc;
        tm_end(c, ELFEL_PH_TOTAL, ev);
        c->has_result = true;
        if (nnz) {
            if ((rc = fetch_nnz(c))) return rc;
            *nnz = c->nnz;
        }
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_assembly_info(elfel_hip_ctx *c, int *method, int *precision, char *note, int len)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (!c->has_result || c->used_method < 0) return set_err(c, ELFEL_ESTATE, "no assembly has run");
        if (method) *method = c->used_method;
        if (precision) *precision = c->used_precision;
        if (note && len > 0) snprintf(note, (size_t)len, "%s", c->used_note.c_str());
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_copy_vector(elfel_hip_ctx *c, double *F, int64_t n)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (!c->has_F) return set_err(c, ELFEL_ESTATE, "the last assembly produced no load vector");
        if (!F || n != c->nF) return set_err(c, ELFEL_EINVAL, "load vector size mismatch");
        HIPCHK(c, hipMemcpyAsync(F, c->F.p, 8 * n, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_geometry_stats(elfel_hip_ctx *c, double *min_det, int64_t *nbad)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        return geometry_stats_run(c, min_det, nbad);
    });
}

extern "C" int elfel_hip_pcg(elfel_hip_ctx *c, int64_t nfree, const double *F, double *x, int precision, double tol,
                             int maxit, int *iters, int *outer, double *relres)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (!c->has_result || c->layout != ELFEL_CSR) return set_err(c, ELFEL_ESTATE, "assemble a CSR matrix first");
        if (nfree < 0 || nfree > c->nrow || !F || !x || tol <= 0 || maxit < 1)
            return set_err(c, ELFEL_EINVAL, "elfel_hip_pcg: bad arguments");
        hipEvent_t ev = nullptr;
        tm_begin(c, ELFEL_PH_SOLVE, &ev);
        int rc = pcg_solve(c, nfree, F, x, precision, tol, maxit, iters, outer, relres);
        tm_end(c, ELFEL_PH_SOLVE, ev);
        return rc;
    });
}

extern "C" int elfel_hip_timing(elfel_hip_ctx *c, int enable)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        tm_collect(c);
        c->tm.region = false;
        if (enable == 2) {
            // region mode: no per-phase events, one pair around the whole region
            c->tm.on = false;
            for (int i = 0; i < ELFEL_NPHASES; i++) { c->tm.ms[i] = 0; c->tm.n[i] = 0; }
            if (!c->tm.r0) HIPCHK(c, hipEventCreate(&c->tm.r0));
            if (!c->tm.r1) HIPCHK(c, hipEventCreate(&c->tm.r1));
            HIPCHK(c, hipEventRecord(c->tm.r0, c->stream));
            c->tm.region = true;
            c->tm.region_n = 0;
            return ELFEL_OK;
        }
        c->tm.on = enable != 0;
        for (int i = 0; i < ELFEL_NPHASES; i++) { c->tm.ms[i] = 0; c->tm.n[i] = 0; }
        // events are created here, not inside timed launches
        while (c->tm.on && c->tm.pool.size() < 512) {
            hipEvent_t e = nullptr;
            if (hipEventCreate(&e) != hipSuccess) break;
            c->tm.pool.push_back(e);
        }
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_timing_read(elfel_hip_ctx *c, double *ms, int64_t *counts)
{
    return host_guard(c, [&]() -> int {
        CHECK_CTX(c);
        if (c->tm.region) {
            HIPCHK(c, hipEventRecord(c->tm.r1, c->stream));
            HIPCHK(c, hipEventSynchronize(c->tm.r1));
            float t = 0.f;
            HIPCHK(c, hipEventElapsedTime(&t, c->tm.r0, c->tm.r1));
            c->tm.region = false;
            for (int i = 0; i < ELFEL_NPHASES; i++) {
                if (ms) ms[i] = i == ELFEL_PH_TOTAL ? (double)t : 0.0;
                if (counts) counts[i] = i == ELFEL_PH_TOTAL ? c->tm.region_n : 0;
            }
            return ELFEL_OK;
        }
        tm_collect(c);
        for (int i = 0; i < ELFEL_NPHASES; i++) {
            if (ms) ms[i] = c->tm.ms[i];
            if (counts) counts[i] = c->tm.n[i];
        }
        return ELFEL_OK;
    });
}
