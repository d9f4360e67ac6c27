// plain.hip -- the plain FP64 method: gather, per-element kernels that write
// the COO stream in Elfel's order, stable radix sort, in-order segmented
// reduce.  Compiled with -ffp-contract=off so every product and sum rounds
// exactly like the reference's unfused Julia arithmetic.
//
// COO stream order (what SysmatAssemblerSparse's append! produces):
//   per element, in element order (for el in elit, src/FEIterators.jl:92-98),
//   the element block column-major (LocalMatrixAssembler init!,
//   src/LocalAssemblers.jl:68-83; assemble!, src/Assemblers.jl:97-102);
//   Stokes adds kup column-major and transpose(kup) row-major of kup
//   (src/Assemblers.jl:109-114), as test/test_stokes.jl:654-656 does.
#include "elem.h"

// ---- index conversion: 1-based int64 -> 0-based int32 with range check ----
__global__ void k_convert(const int64_t *__restrict__ src, int32_t *__restrict__ dst, int64_t n, int64_t lo,
                          int64_t hi, int *__restrict__ err)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int64_t v = src[i];
    if (v < lo || v > hi) {
        atomicOr(err, DERR_RANGE);
        v = lo;
    }
    dst[i] = (int32_t)(v - 1);
}

int convert_indices(elfel_hip_ctx *c, const int64_t *src_dev, int32_t *dst, int64_t n, int64_t lo, int64_t hi)
{
    if (n == 0) return 0;
    int *err = (int *)c->dscal.p + 1;
    hipLaunchKernelGGL(k_convert, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, src_dev, dst, n, lo, hi, err);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// ---- gather: element-major SoA copies of coordinates and dofs -------------
// gxy[k*E + e] = xy[conn[e][k]], gd[j*E + e] = dof of element dof j, so the
// element kernels read every operand with unit stride across a wavefront.
template <int NEN, int NC>
__global__ __launch_bounds__(256) void k_gather(int64_t E, const int32_t *__restrict__ conn,
                                                const double2 *__restrict__ xy, const int32_t *__restrict__ dof,
                                                const int32_t *__restrict__ connp, const int32_t *__restrict__ dofp,
                                                double2 *__restrict__ gxy, int32_t *__restrict__ gd)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
#pragma unroll
    for (int k = 0; k < NEN; k++) {
        const int32_t n = conn[e * NEN + k];
        gxy[k * E + e] = xy[n];
#pragma unroll
        for (int c = 0; c < NC; c++) gd[(k * NC + c) * E + e] = dof[(int64_t)n * NC + c];
    }
    if (connp) {
#pragma unroll
        for (int k = 0; k < 3; k++) gd[(NEN * NC + k) * E + e] = dofp[connp[e * 3 + k]];
    }
}

// ---- element kernel: one work-item per element, COO in stream order -------
template <int KIND, int NEN, int NC, int NQ>
__global__ __launch_bounds__(256) void k_elem_coo(int64_t E, const double2 *__restrict__ gxy,
                                                  const int32_t *__restrict__ gd, Coef cf, QTable tab,
                                                  int minor_bits, int csc, uint64_t *__restrict__ keys,
                                                  double *__restrict__ vals, uint64_t *__restrict__ fkeys,
                                                  double *__restrict__ fvals)
{
    constexpr int ND = NEN * NC;
    constexpr bool STOKES = (KIND == ELFEL_STOKES_GEN || KIND == ELFEL_STOKES_VECLAP);
    constexpr int NBLK = ND * ND + (STOKES ? 2 * ND * 3 : 0);
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double X[NEN], Y[NEN];
#pragma unroll
    for (int k = 0; k < NEN; k++) {
        const double2 p = gxy[k * E + e];
        X[k] = p.x; Y[k] = p.y;
    }
    int32_t d[ND + (STOKES ? 3 : 0)];
#pragma unroll
    for (int j = 0; j < ND + (STOKES ? 3 : 0); j++) d[j] = gd[j * E + e];
    Geo<double, NEN, NQ> g;
    geometry<double, NEN, NQ>(tab, X, Y, g);
    double D[9];
    const double sc = cf.scale_e ? cf.scale_e[padded(e, cf.cgap)] : 1.0;
#pragma unroll
    for (int i = 0; i < 9; i++) D[i] = cf.scale_e ? sc * cf.D[i] : cf.D[i];
    const double kap = cf.kappa_e ? cf.kappa_e[padded(e, cf.cgap)] : cf.kappa;
    auto key = [&](int32_t r, int32_t col) -> uint64_t {
        const uint64_t mj = (uint64_t)(csc ? col : r), mn = (uint64_t)(csc ? r : col);
        return (mj << minor_bits) | mn;
    };
    int64_t o = e * NBLK;
    Entry<KIND, double, NEN, NC, NQ> ent;
#pragma unroll
    for (int j = 0; j < ND; j++) {
        ent.column(g, D, j);
#pragma unroll
        for (int i = 0; i < ND; i++) {
            if (keys) keys[o] = key(d[i], d[j]);
            vals[o] = ent.value(g, kap, cf.mu, i, j);
            o++;
        }
    }
    if constexpr (STOKES) {
        double kup[ND][3];
#pragma unroll
        for (int jp = 0; jp < 3; jp++)
#pragma unroll
            for (int i = 0; i < ND; i++) kup[i][jp] = kup_value<double, NEN, NC, NQ>(tab, g, i, jp);
#pragma unroll
        for (int jp = 0; jp < 3; jp++)
#pragma unroll
            for (int i = 0; i < ND; i++) {
                if (keys) keys[o] = key(d[i], d[ND + jp]);
                vals[o] = kup[i][jp];
                o++;
            }
#pragma unroll
        for (int b = 0; b < ND; b++)
#pragma unroll
            for (int a = 0; a < 3; a++) {
                if (keys) keys[o] = key(d[ND + a], d[b]);
                vals[o] = kup[b][a];
                o++;
            }
    }
    if constexpr (KIND == ELFEL_HEAT) {
        if (fvals) {
            const double Qe = cf.Q_e ? cf.Q_e[padded(e, cf.cgap)] : cf.Q;
#pragma unroll
            for (int j = 0; j < ND; j++) {
                if (fkeys) fkeys[e * ND + j] = (uint64_t)d[j];
                fvals[e * ND + j] = fe_value<double, NEN, NQ>(tab, g, Qe, j);
            }
        }
    }
}

// ---- generic COO (host stream) -> keys with range check --------------------
__global__ void k_coo_keys(const int64_t *__restrict__ I, const int64_t *__restrict__ J, int64_t n, int64_t nrow,
                           int64_t ncol, int minor_bits, int csc, uint64_t *__restrict__ keys, int *__restrict__ err)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int64_t r = I[k], cc = J[k];
    if (r < 1 || r > nrow || cc < 1 || cc > ncol) {
        atomicOr(err, DERR_RANGE);
        r = 1; cc = 1;
    }
    const uint64_t mj = (uint64_t)(csc ? cc - 1 : r - 1), mn = (uint64_t)(csc ? r - 1 : cc - 1);
    keys[k] = (mj << minor_bits) | mn;
}

__global__ void k_vec_keys(const int64_t *__restrict__ I, int64_t n, int64_t ndofs, uint64_t *__restrict__ keys,
                           int *__restrict__ err)
{
    int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= n) return;
    int64_t r = I[k];
    if (r < 1 || r > ndofs) {
        atomicOr(err, DERR_RANGE);
        r = 1;
    }
    keys[k] = (uint64_t)(r - 1);
}

// Host wrappers used by api.hip for the generic drop-in.
int plain_coo_keys(elfel_hip_ctx *c, const int64_t *I, const int64_t *J, int64_t n, int64_t nrow, int64_t ncol,
                   int minor_bits, int csc, uint64_t *keys)
{
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_coo_keys, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, I, J, n, nrow, ncol, minor_bits,
                       csc, keys, (int *)c->dscal.p + 1);
    HIPCHK(c, hipGetLastError());
    return 0;
}

int plain_vec_keys(elfel_hip_ctx *c, const int64_t *I, int64_t n, int64_t ndofs, uint64_t *keys)
{
    if (n == 0) return 0;
    hipLaunchKernelGGL(k_vec_keys, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, I, n, ndofs, keys,
                       (int *)c->dscal.p + 1);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// ---- orchestration --------------------------------------------------------
template <int KIND, int NEN, int NC, int NQ>
static int launch_plain(elfel_hip_ctx *c, const QTable &tab, const Coef &cf, int minor_bits, int csc,
                        uint64_t *keys, double *vals, uint64_t *fkeys, double *fvals)
{
    const Field &F = c->f[0];
    const int64_t E = F.E;
    constexpr bool STOKES = (KIND == ELFEL_STOKES_GEN || KIND == ELFEL_STOKES_VECLAP);
    const int nd = NEN * NC + (STOKES ? 3 : 0);
    int rc = ensure(c, c->geom, (size_t)E * (NEN * sizeof(double2) + nd * sizeof(int32_t)));
    if (rc) return rc;
    double2 *gxy = (double2 *)c->geom.p;
    int32_t *gd = (int32_t *)(gxy + (size_t)NEN * E);
    if (E == 0) return 0;  // no elements: an empty stream
    hipEvent_t ev = nullptr;
    tm_begin(c, ELFEL_PH_GATHER, &ev);
    hipLaunchKernelGGL((k_gather<NEN, NC>), dim3(ceil_div(E, 256)), dim3(256), 0, c->stream, E,
                       (const int32_t *)F.conn.p, (const double2 *)F.xy.p, (const int32_t *)F.dof.p,
                       STOKES ? (const int32_t *)c->f[1].conn.p : nullptr,
                       STOKES ? (const int32_t *)c->f[1].dof.p : nullptr, gxy, gd);
    HIPCHK(c, hipGetLastError());
    tm_end(c, ELFEL_PH_GATHER, ev);
    tm_begin(c, ELFEL_PH_ELEMENT, &ev);
    hipLaunchKernelGGL((k_elem_coo<KIND, NEN, NC, NQ>), dim3(ceil_div(E, 256)), dim3(256), 0, c->stream, E, gxy, gd,
                       cf, tab, minor_bits, csc, keys, vals, fkeys, fvals);
    HIPCHK(c, hipGetLastError());
    tm_end(c, ELFEL_PH_ELEMENT, ev);
    return 0;
}

typedef int (*plain_fn)(elfel_hip_ctx *, const QTable &, const Coef &, int, int, uint64_t *, double *, uint64_t *,
                        double *);

static plain_fn pick_plain(int kind, int fe, int nq)
{
#define P(K, N, C, Q) return launch_plain<K, N, C, Q>
    if (kind == ELFEL_HEAT) {
        if (fe == ELFEL_T3 && nq == 1) P(ELFEL_HEAT, 3, 1, 1);
        if (fe == ELFEL_T3 && nq == 3) P(ELFEL_HEAT, 3, 1, 3);
        if (fe == ELFEL_T6 && nq == 3) P(ELFEL_HEAT, 6, 1, 3);
        if (fe == ELFEL_Q4 && nq == 1) P(ELFEL_HEAT, 4, 1, 1);
        if (fe == ELFEL_Q4 && nq == 4) P(ELFEL_HEAT, 4, 1, 4);
    } else if (kind == ELFEL_ELAST) {
        if (fe == ELFEL_T3 && nq == 1) P(ELFEL_ELAST, 3, 2, 1);
        if (fe == ELFEL_T3 && nq == 3) P(ELFEL_ELAST, 3, 2, 3);
        if (fe == ELFEL_T6 && nq == 3) P(ELFEL_ELAST, 6, 2, 3);
        if (fe == ELFEL_Q4 && nq == 4) P(ELFEL_ELAST, 4, 2, 4);
    } else if (kind == ELFEL_STOKES_GEN) {
        if (fe == ELFEL_T6 && nq == 3) P(ELFEL_STOKES_GEN, 6, 2, 3);
    } else if (kind == ELFEL_STOKES_VECLAP) {
        if (fe == ELFEL_T6 && nq == 3) P(ELFEL_STOKES_VECLAP, 6, 2, 3);
    }
#undef P
    return nullptr;
}

int plain_buffers(elfel_hip_ctx *c, int64_t n, int64_t nv, int64_t ndofs, bool vec)
{
    int rc;
    if ((rc = ensure(c, c->keys[0], sizeof(uint64_t) * n)) || (rc = ensure(c, c->keys[1], sizeof(uint64_t) * n)) ||
        (rc = ensure(c, c->vals[0], sizeof(double) * n)) || (rc = ensure(c, c->vals[1], sizeof(double) * n)))
        return rc;
    if (vec) {
        if ((rc = ensure(c, c->fkeys[0], sizeof(uint64_t) * nv)) || (rc = ensure(c, c->fkeys[1], sizeof(uint64_t) * nv)) ||
            (rc = ensure(c, c->fvals[0], sizeof(double) * nv)) || (rc = ensure(c, c->fvals[1], sizeof(double) * nv)) ||
            (rc = ensure(c, c->F, sizeof(double) * ndofs)))
            return rc;
    }
    if ((rc = ensure(c, c->ptr, sizeof(int32_t) * (ndofs + 1))) || (rc = ensure(c, c->ind, sizeof(int32_t) * n)) ||
        (rc = ensure(c, c->val, sizeof(double) * n)))
        return rc;
    return 0;
}

int plain_finish(elfel_hip_ctx *c, int layout, int64_t n, int64_t nv, int64_t ndofs, int minor_bits, bool vec)
{
    uint64_t *k0 = (uint64_t *)c->keys[0].p, *k1 = (uint64_t *)c->keys[1].p;
    uint64_t *v0 = (uint64_t *)c->vals[0].p, *v1 = (uint64_t *)c->vals[1].p;
    uint64_t *fk0 = vec ? (uint64_t *)c->fkeys[0].p : nullptr, *fk1 = vec ? (uint64_t *)c->fkeys[1].p : nullptr;
    uint64_t *fv0 = vec ? (uint64_t *)c->fvals[0].p : nullptr, *fv1 = vec ? (uint64_t *)c->fvals[1].p : nullptr;
    hipEvent_t ev = nullptr;
    tm_begin(c, ELFEL_PH_SORT, &ev);
    int rc = radix_sort_pairs(c, k0, v0, k1, v1, n, bits_for(ndofs) + minor_bits);
    if (rc) return rc;
    if (vec) {
        rc = radix_sort_pairs(c, fk0, fv0, fk1, fv1, nv, bits_for(ndofs));
        if (rc) return rc;
    }
    tm_end(c, ELFEL_PH_SORT, ev);
    tm_begin(c, ELFEL_PH_REDUCE, &ev);
    rc = build_csr_from_sorted(c, k0, (const double *)v0, n, minor_bits, ndofs, (int32_t *)c->ptr.p,
                               (int32_t *)c->ind.p, (double *)c->val.p, (int32_t *)c->dscal.p, true);
    if (rc) return rc;
    if (vec) {
        rc = build_dense_from_sorted(c, fk0, (const double *)fv0, nv, ndofs, (double *)c->F.p);
        if (rc) return rc;
    }
    tm_end(c, ELFEL_PH_REDUCE, ev);
    c->has_F = vec;
    c->nF = vec ? ndofs : 0;
    c->nmaj = ndofs;
    c->nmin = ndofs;
    c->layout = layout;
    c->result_src = RES_PLAIN;
    return 0;
}

// The 2D weak forms on the stream REUSE method (FP64): the element kernel of
// the plain method writes keys on the first call, values only afterwards, and
// every entry is the sum of its run in the cached sort order.
int plain_stream_reuse(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf)
{
    const Field &F = c->f[0];
    const bool stokes = pb->kind == ELFEL_STOKES_GEN || pb->kind == ELFEL_STOKES_VECLAP;
    plain_fn fn = pick_plain(pb->kind, pb->fe, tab.nq);
    if (!fn) return set_err(c, ELFEL_ENOTSUP, "unsupported (kind, element, rule) for the plain method");
    const int64_t nd = (int64_t)F.nen * F.ncomp;
    const int64_t n = F.E * (nd * nd + (stokes ? 2 * nd * 3 : 0));
    const bool heat = pb->kind == ELFEL_HEAT;
    const int64_t nv = heat ? F.E * nd : 0;
    const int64_t ndofs = c->nrow;
    const int minor_bits = bits_for(ndofs);
    const int csc = pb->layout == ELFEL_CSC ? 1 : 0;
    elfel_hip_problem p = *pb;
    p.precision = ELFEL_FP64;
    if (n >= INT32_MAX || nv >= INT32_MAX) {  // the cached permutation is int32
        p.method = ELFEL_SORT;
        c->used_method = ELFEL_SORT;
        c->used_note += " (a COO stream of 2^31 entries or more: the plain FP64 method)";
        c->nnz = -1;
        return plain_assemble(c, &p, tab, cf);
    }
    return stream_reuse(c, &p, n, nv, ndofs, minor_bits, false, [&](const StreamOut &o, bool) {
        return fn(c, tab, cf, minor_bits, csc, o.keys, (double *)o.vals, o.fkeys, (double *)o.fvals);
    });
}

int plain_assemble(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf)
{
    const Field &F = c->f[0];
    const bool stokes = pb->kind == ELFEL_STOKES_GEN || pb->kind == ELFEL_STOKES_VECLAP;
    plain_fn fn = pick_plain(pb->kind, pb->fe, tab.nq);
    if (!fn) return set_err(c, ELFEL_ENOTSUP, "unsupported (kind, element, rule) for the plain method");
    const int nc = F.ncomp, nen = F.nen;
    const int64_t nd = (int64_t)nen * nc;
    const int64_t nblk = nd * nd + (stokes ? 2 * nd * 3 : 0);
    const int64_t n = F.E * nblk;
    const int64_t ndofs = c->nrow;
    const int csc = pb->layout == ELFEL_CSC ? 1 : 0;
    const int minor_bits = bits_for(ndofs);
    const bool heat = pb->kind == ELFEL_HEAT;
    const int64_t nv = heat ? F.E * nd : 0;
    int rc = plain_buffers(c, n, nv, ndofs, heat);
    if (rc) return rc;
    rc = fn(c, tab, cf, minor_bits, csc, (uint64_t *)c->keys[0].p, (double *)c->vals[0].p,
            heat ? (uint64_t *)c->fkeys[0].p : nullptr, heat ? (double *)c->fvals[0].p : nullptr);
    if (rc) return rc;
    return plain_finish(c, pb->layout, n, nv, ndofs, minor_bits, heat);
}
