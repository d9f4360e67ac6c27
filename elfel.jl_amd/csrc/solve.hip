// solve.hip -- Jacobi-preconditioned conjugate gradients on the free block of
// the assembled matrix, replacing the driver's `K[1:nu,1:nu] \ (F - K*T)[1:nu]`
// (test/test_heat.jl:73-76) for SPD problems.
//
// FP64 PCG: textbook PCG (Saad, Iterative Methods for Sparse Linear Systems,
// Alg. 9.1) with every vector and the matrix in FP64; when the recurrence
// residual reaches the tolerance, the true residual b - A x is formed and
// replaces it (the iteration goes on if the true one is not yet small enough).
// Mixed: the same PCG with an FP32 matrix copy, FP32 r, z, p and q, the FP64
// iterate x (x += alpha p in FP64), FP64 dot products, and reliable updates
// (Sleijpen & van der Vorst, Computing 56, 1996): whenever the recurrence
// residual has fallen by RU_DROP since the last update, the FP64 residual
// b - A x is formed with the FP64 matrix and replaces the FP32 one, and the
// search direction is kept unless the recurrence has drifted below half the
// true residual (then p restarts from z).  An iteration of the mixed solve
// moves about 0.55 of the bytes of an FP64 one.  With an FP32 iterate (added
// into the FP64 solution at each update) the iteration count rose with the
// mesh: on heat Q4 with lognormal(0, 1) conductivity, +9 % over FP64 PCG at
// 256x256 and +69 % at 1024x1024 (numpy emulation of this code), +45 % at
// 4096x4096 (measured: 19224 against 13280); the FP64 iterate keeps the late,
// small corrections that an FP32 one rounds away (+9 % at 1024x1024).
// Updates at every 10x drop, or never restarting, failed to converge at
// 4096x4096 (FP32 iterate).
// Convergence is declared on the FP64 residual: ||b - A x|| / ||b|| <= tol.
//
// Storage: the free-free block in SELL-64 (one slice per 64 consecutive rows,
// each slice as wide as its longest row, entries column-major inside the
// slice), built once per solve from the CSR result: one thread per row, every
// load coalesced, data columns left out (their part of b is formed once with
// the CSR matrix).  When every free column lies within a 16-bit offset of its
// row (banded numberings: structured grids, RCM-ordered meshes) the columns
// are stored as int16 offsets, 2 bytes per entry instead of 4.
// Per iteration three kernels on a fixed grid of 4 workgroups per CU
// (grid-stride loops): q = A p with one partial of p.q per workgroup; the
// x / r / z update (every workgroup first sums the p.q and previous r.z
// partials itself, in the same order) with partials of r.z and r.r; and
// p = z + beta p.  No atomics and no extra reduction launches; every run gives
// the same bits.  The CG scalars stay on the device and the host reads the
// residual every CHECK iterations.
#include <math.h>

#include <algorithm>
#include <type_traits>

#include "common.h"

int env_int(const char *name, int dflt);  // reuse.hip

#define SV_BLOCK 256
#define CHECK 8
#define SELL_UNROLL 12  // slices at most this wide are read with every load in flight at once
#ifndef RU_DROP
#define RU_DROP 1e-2   // reliable update when the recurrence residual has dropped by this factor
#endif
#ifndef RU_KEEP
#define RU_KEEP 0.5    // keep p at an update while ||r recurrence|| >= RU_KEEP ||r true||
#endif

namespace {

// Scalars (double) on the device, for the host's convergence checks.
enum { S_RZ = 0, S_RR = 1, S_BB = 2, S_RR64 = 3, S_NS = 8 };

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();  // sh may still be read by a previous call
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sh[k];
    return s;  // the same value in every thread
}

// Fixed-order sum of the g partials of the previous kernel (every block of
// the grid computes it, in the same order, so all get the same bits).
__device__ __forceinline__ double sum_partials(const double *p, int g, double *sh)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < g; i += blockDim.x) v += p[i];
    return block_sum(v, sh);
}

// ---- SELL-64 construction ----------------------------------------------------
// Free entries per row, and whether a free column lies farther than a 16-bit
// offset from the row (then the column indices stay 32-bit).
__global__ void k_sell_rowlen(int64_t nfree, const int32_t *__restrict__ ptr, const int32_t *__restrict__ ind,
                              uint32_t *__restrict__ len, uint32_t *__restrict__ far)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nfree) return;
    uint32_t k = 0, f = 0;
    for (int32_t j = ptr[i]; j < ptr[i + 1]; j++) {
        if (ind[j] >= nfree) continue;
        k++;
        const int64_t d = (int64_t)ind[j] - i;
        f |= (d < -32768 || d > 32767) ? 1u : 0u;
    }
    len[i] = k;
    far[i] = f;
}

// Column of entry k of a SELL row: 32-bit index, or 16-bit offset from the row.
template <typename C>
__device__ __forceinline__ int32_t sell_col(C c, int64_t row);
template <>
__device__ __forceinline__ int32_t sell_col<int32_t>(int32_t c, int64_t row) { return c; }
template <>
__device__ __forceinline__ int32_t sell_col<int16_t>(int16_t c, int64_t row) { return (int32_t)row + (int32_t)c; }

// One wave per slice: width = longest row; writes 64 * width (the slice size).
__global__ void k_sell_width(int64_t nfree, int64_t nslices, const uint32_t *__restrict__ len,
                             uint32_t *__restrict__ size)
{
    const int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (s >= nslices) return;
    const int64_t row = s * 64 + lane;
    uint32_t w = row < nfree ? len[row] : 0u;
    for (int d = 32; d > 0; d >>= 1) w = max(w, (uint32_t)__shfl_down((int)w, d, 64));
    if (lane == 0) size[s] = 64u * w;  // lane 0 holds the maximum
}

template <typename C>
__global__ void k_sell_fill(int64_t nfree, const int32_t *__restrict__ ptr, const int32_t *__restrict__ ind,
                            const double *__restrict__ val, const uint32_t *__restrict__ off,
                            C *__restrict__ col, double *__restrict__ v64, float *__restrict__ v32)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t s = i >> 6;
    const int lane = (int)(i & 63);
    if (s * 64 >= nfree) return;
    const uint32_t o = off[s], w = (off[s + 1] - o) >> 6;
    uint32_t k = 0;
    if (i < nfree) {
        for (int32_t j = ptr[i]; j < ptr[i + 1]; j++) {
            const int32_t cj = ind[j];
            if (cj >= nfree) continue;
            const size_t at = (size_t)o + (size_t)k * 64 + lane;
            col[at] = sizeof(C) == 4 ? (C)cj : (C)(cj - i);
            v64[at] = val[j];
            if (v32) v32[at] = (float)val[j];
            k++;
        }
    }
    for (; k < w; k++) {  // padding: a zero on the row's own column (or row 0 beyond nfree)
        const size_t at = (size_t)o + (size_t)k * 64 + lane;
        col[at] = sizeof(C) == 4 ? (C)(i < nfree ? i : 0) : (C)0;  // rows past nfree are never read
        v64[at] = 0.0;
        if (v32) v32[at] = 0.f;
    }
}

// Slices whose every slot holds one column offset for all its rows (each row
// of the slice has its k-th entry at row + d_k: the stencil of a structured
// grid away from its boundary) are flagged, and their offsets d_k are kept
// once per slot; such a slice reads no per-entry column at all.  One wave per
// slice.
template <typename C>
__global__ void k_sell_uniform(int64_t nfree, int64_t nslices, const uint32_t *__restrict__ off,
                               const C *__restrict__ col, uint32_t *__restrict__ uflag, int32_t *__restrict__ uoff)
{
    const int64_t s = ((int64_t)blockIdx.x * blockDim.x + threadIdx.x) >> 6;
    const int lane = threadIdx.x & 63;
    if (s >= nslices) return;
    const int64_t i = s * 64 + lane;
    const uint32_t o = off[s], w = (off[s + 1] - o) >> 6, sb = o >> 6;
    bool uniform = true;
    for (uint32_t k = 0; k < w; k++) {
        const int32_t d = sell_col<C>(col[(size_t)o + (size_t)k * 64 + lane], i) - (int32_t)i;
        const int32_t d0 = __shfl(d, 0, 64);  // row s*64 < nfree: lane 0 is a real row
        if (__ballot((i < nfree) && (d != d0)) != 0ull) uniform = false;
        if (lane == 0) uoff[sb + k] = d0;
    }
    if (lane == 0) uflag[s] = uniform ? 1u : 0u;
}

struct SellU {
    const uint32_t *uflag;  // per slice: 1 when every slot has one column offset for all rows
    const int32_t *uoff;    // per slot: that offset
};

// A slice's start, width and uniform flag (the first loads of its row).
struct SliceMeta {
    uint32_t o, w;
    bool uni;
};
__device__ __forceinline__ SliceMeta slice_meta(int64_t i, const uint32_t *__restrict__ off, SellU U)
{
    // one wave holds one slice: s is the same on every lane
    const int64_t s = __builtin_amdgcn_readfirstlane((int)(i >> 6));
    const uint32_t o = off[s];
    return SliceMeta{o, (off[s + 1] - o) >> 6, U.uflag[s] != 0u};
}

template <typename T, typename C>
__device__ __forceinline__ T sell_row(int64_t i, const uint32_t *__restrict__ off, const C *__restrict__ col,
                                      const T *__restrict__ val, const T *__restrict__ x, SellU U,
                                      SliceMeta m)
{
    const int lane = (int)(i & 63);
    const uint32_t o = m.o, w = m.w;
    T acc = T(0);
    const T *v = val + o + lane;
    const C *c = col + o + lane;
    const bool uni = m.uni;
    const int32_t *d = U.uoff + (o >> 6);
    const int32_t ri = (int32_t)i;
    if (w <= SELL_UNROLL) {
        // every load of the slice is issued before the first product: the
        // values and (uniform slices) the gathers at once, else the columns,
        // then the gathers.  Slots past w read slot 0 again (in bounds) and
        // add nothing; in uniform slices they are skipped (w is the same on
        // every lane), and the values, read once per product, bypass the
        // caches that hold the vector
        T vv[SELL_UNROLL], xx[SELL_UNROLL];
        if (uni) {
#pragma unroll
            for (int k = 0; k < SELL_UNROLL; k++) {
                vv[k] = T(0);
                xx[k] = T(0);
                if (k < (int)w) {
                    vv[k] = __builtin_nontemporal_load(v + (size_t)k * 64);
                    xx[k] = x[ri + d[k]];
                }
            }
        } else {
            int32_t cj[SELL_UNROLL];
#pragma unroll
            for (int k = 0; k < SELL_UNROLL; k++) {
                const int kk = k < (int)w ? k : 0;
                vv[k] = v[(size_t)kk * 64];
                cj[k] = sell_col<C>(c[(size_t)kk * 64], i);
            }
#pragma unroll
            for (int k = 0; k < SELL_UNROLL; k++) xx[k] = x[cj[k]];
        }
#pragma unroll
        for (int k = 0; k < SELL_UNROLL; k++)
            if (k < (int)w) acc += vv[k] * xx[k];
        return acc;
    }
    if (uni) {
#pragma unroll 4
        for (uint32_t k = 0; k < w; k++) acc += v[(size_t)k * 64] * x[ri + d[k]];
    } else {
#pragma unroll 4
        for (uint32_t k = 0; k < w; k++) acc += v[(size_t)k * 64] * x[sell_col<C>(c[(size_t)k * 64], i)];
    }
    return acc;
}

// The CG kernels run on a fixed grid of G workgroups with grid-stride loops
// over the rows; each writes one partial per workgroup per dot product.
//   parts layout: pq[G], rz[2][G] (by iteration parity), rr[G], aux[G]

// q = A p; partial p.q
template <typename T, typename C>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_spmv(int64_t n, const uint32_t *__restrict__ off,
                                                      const C *__restrict__ col, const T *__restrict__ val,
                                                      SellU U, const T *__restrict__ p, T *__restrict__ q,
                                                      double *__restrict__ part_pq)
{
    __shared__ double sh[SV_BLOCK / 64];
    double d = 0.0;
    // (Loading the next slice's start, width and flag while this slice's row
    // is summed measured slower: 544 against 373 us per mixed iteration on
    // heat Q4 4096^2.)
    for (int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * SV_BLOCK) {
        const T y = sell_row<T, C>(i, off, col, val, p, U, slice_meta(i, off, U));
        q[i] = y;
        d += (double)p[i] * (double)y;
    }
    const double s = block_sum(d, sh);
    if (threadIdx.x == 0) part_pq[blockIdx.x] = s;
}

// alpha = rz_old / pq; x += alpha p; r -= alpha q; z = M r; partials r.z (into
// rz_new) and r.r
// The FP64 solve updates its FP64 iterate here (TX = double).  The mixed
// solve adds alpha p to an FP32 accumulator of the last few iterations' steps
// (TX = float), which k_cg_flush folds into the FP64 iterate every CHECK
// iterations: an FP32 iterate itself absorbs the late, small corrections
// (x + alpha p rounds back to x; the mixed solve then needed 69 % more
// iterations than FP64 PCG on heat Q4 1024^2 against 9 % with an FP64
// iterate, numpy emulation of this code), while the accumulator holds only
// corrections, so each keeps its FP32 relative precision.
template <typename T, typename TX>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_update(int64_t n, const T *__restrict__ dinv, TX *__restrict__ x,
                                                        T *__restrict__ r, T *__restrict__ z,
                                                        const T *__restrict__ p, const T *__restrict__ q,
                                                        const double *__restrict__ part_pq,
                                                        const double *__restrict__ rz_old,
                                                        double *__restrict__ rz_new, double *__restrict__ part_rr)
{
    __shared__ double sh[SV_BLOCK / 64];
    const int G = gridDim.x;
    const double pq = sum_partials(part_pq, G, sh);
    const double rz = sum_partials(rz_old, G, sh);
    const double alpha64 = pq != 0.0 ? rz / pq : 0.0;
    const T alpha = (T)alpha64;
    const TX alphax = (TX)alpha64;
    double a = 0.0, b = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x; i < n; i += (int64_t)G * SV_BLOCK) {
        x[i] += alphax * (TX)p[i];
        const T ri = r[i] - alpha * q[i];
        const T zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        a += (double)ri * (double)zi;
        b += (double)ri * (double)ri;
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) { rz_new[blockIdx.x] = a; part_rr[blockIdx.x] = b; }
}

// x += xlo, xlo = 0: the mixed solve's accumulated steps into its FP64 iterate
__global__ void k_cg_flush(int64_t n, double *__restrict__ x, float *__restrict__ xlo)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x[i] += (double)xlo[i];
        xlo[i] = 0.f;
    }
}

// beta = rz_new / rz_old; p = z + beta p.  Workgroup 0 also leaves rz and
// ||r||^2 for the host.
template <typename T>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_dir(int64_t n, const double *__restrict__ rz_new,
                                                     const double *__restrict__ rz_old,
                                                     const double *__restrict__ part_rr, const T *__restrict__ z,
                                                     T *__restrict__ p, double *__restrict__ sc)
{
    __shared__ double sh[SV_BLOCK / 64];
    const int G = gridDim.x;
    const double a = sum_partials(rz_new, G, sh);
    const double b = sum_partials(rz_old, G, sh);
    const T beta = (T)(b != 0.0 ? a / b : 0.0);
    for (int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x; i < n; i += (int64_t)G * SV_BLOCK)
        p[i] = z[i] + beta * p[i];
    if (blockIdx.x == 0) {
        const double rr = sum_partials(part_rr, G, sh);
        if (threadIdx.x == 0) { sc[S_RZ] = a; sc[S_RR] = rr; }
    }
}

// r = (T) r64, z = M r (p = z when `restart`); partials r.z (into rz_out) and r.r
template <typename T>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_reset(int64_t n, const double *__restrict__ r64,
                                                       const T *__restrict__ dinv, T *__restrict__ r,
                                                       T *__restrict__ z, T *__restrict__ p, int restart,
                                                       double *__restrict__ rz_out, double *__restrict__ part_rr)
{
    __shared__ double sh[SV_BLOCK / 64];
    double a = 0.0, b = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * SV_BLOCK) {
        const T ri = (T)r64[i];
        const T zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        if (restart) p[i] = zi;
        a += (double)ri * (double)zi;
        b += (double)ri * (double)ri;
    }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) { rz_out[blockIdx.x] = a; part_rr[blockIdx.x] = b; }
}

// One workgroup: sc[slot] = sum of g partials (and sc[slot2] of a second set).
__global__ void k_reduce(const double *__restrict__ p, int g, double *__restrict__ sc, int slot,
                         const double *__restrict__ p2, int slot2)
{
    __shared__ double sh[1024 / 64];
    const double a = sum_partials(p, g, sh);
    const double b = p2 ? sum_partials(p2, g, sh) : 0.0;
    if (threadIdx.x == 0) {
        sc[slot] = a;
        if (p2) sc[slot2] = b;
    }
}


// r = b - A x over the free block (SELL, FP64); partial ||r||^2
template <typename C>
__global__ __launch_bounds__(SV_BLOCK) void k_sell_resid(int64_t n, const uint32_t *__restrict__ off,
                                                         const C *__restrict__ col,
                                                         const double *__restrict__ val, SellU U,
                                                         const double *__restrict__ x, const double *__restrict__ b,
                                                         double *__restrict__ r, double *__restrict__ part)
{
    __shared__ double sh[SV_BLOCK / 64];
    double rr = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x; i < n; i += (int64_t)gridDim.x * SV_BLOCK) {
        const double v = b[i] - sell_row<double, C>(i, off, col, val, x, U, slice_meta(i, off, U));
        r[i] = v;
        rr += v * v;
    }
    const double s = block_sum(rr, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

// b = F_f - K_fd x_d with the CSR matrix (data columns only); partial ||b||^2
__global__ __launch_bounds__(SV_BLOCK) void k_rhs(int64_t nfree, const int32_t *__restrict__ ptr,
                                                  const int32_t *__restrict__ ind, const double *__restrict__ val,
                                                  const double *__restrict__ xfull, const double *F, double *b,
                                                  double *__restrict__ part)
{
    __shared__ double sh[SV_BLOCK / 64];
    double bb = 0.0;
    for (int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x; i < nfree; i += (int64_t)gridDim.x * SV_BLOCK) {
        double acc = 0.0;
        for (int32_t j = ptr[i]; j < ptr[i + 1]; j++)
            if (ind[j] >= nfree) acc += val[j] * xfull[ind[j]];
        const double v = F[i] - acc;
        b[i] = v;
        bb += v * v;
    }
    const double s = block_sum(bb, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
}

__global__ void k_diag_inv(int64_t n, const int32_t *__restrict__ ptr, const int32_t *__restrict__ ind,
                           const double *__restrict__ val, double *__restrict__ d64, float *__restrict__ d32)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    double d = 0.0;
    int32_t lo = ptr[i], hi = ptr[i + 1];
    while (lo < hi) {  // sorted indices: binary search for the diagonal
        const int32_t m = (lo + hi) >> 1;
        if (ind[m] < i) lo = m + 1;
        else hi = m;
    }
    if (lo < ptr[i + 1] && ind[lo] == i) d = val[lo];
    const double inv = d != 0.0 ? 1.0 / d : 1.0;
    d64[i] = inv;
    if (d32) d32[i] = (float)inv;
}

template <typename T>
__global__ void k_fill(int64_t n, T *__restrict__ a, T v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

template <typename T>
struct Vecs {
    double *x;  // the iterate: FP64 in both solves
    T *r, *z, *p, *q, *dinv;
    float *xlo;  // mixed solve: the steps since the last k_cg_flush (null: x updated directly)
};

struct Sell {
    const uint32_t *off;
    const void *col;  // int32 columns, or int16 offsets from the row when c16
    bool c16;
    const double *v64;
    const float *v32;
    SellU U;
};

template <typename T>
const T *sell_vals(const Sell &S);
template <>
const double *sell_vals<double>(const Sell &S) { return S.v64; }
template <>
const float *sell_vals<float>(const Sell &S) { return S.v32; }

struct Parts {
    double *pq, *rz[2], *rr, *aux;
    int G;
};

// One PCG iteration on the stream; `par` is the iteration parity (the rz
// partials of this iteration go to rz[par], the previous ones are rz[par ^ 1]).
template <typename T>
void cg_step(hipStream_t st, int64_t n, const Sell &S, const Vecs<T> &v, const Parts &P, int par, double *sc)
{
    const dim3 g(P.G), b(SV_BLOCK);
    if (S.c16)
        hipLaunchKernelGGL((k_cg_spmv<T, int16_t>), g, b, 0, st, n, S.off, (const int16_t *)S.col, sell_vals<T>(S), S.U,
                           v.p, v.q, P.pq);
    else
        hipLaunchKernelGGL((k_cg_spmv<T, int32_t>), g, b, 0, st, n, S.off, (const int32_t *)S.col, sell_vals<T>(S), S.U,
                           v.p, v.q, P.pq);
    if (v.xlo)
        hipLaunchKernelGGL((k_cg_update<T, float>), g, b, 0, st, n, v.dinv, v.xlo, v.r, v.z, v.p, v.q, P.pq,
                           P.rz[par ^ 1], P.rz[par], P.rr);
    else
        hipLaunchKernelGGL((k_cg_update<T, double>), g, b, 0, st, n, v.dinv, v.x, v.r, v.z, v.p, v.q, P.pq,
                           P.rz[par ^ 1], P.rz[par], P.rr);
    hipLaunchKernelGGL((k_cg_dir<T>), g, b, 0, st, n, P.rz[par], P.rz[par ^ 1], P.rr, v.z, v.p, sc);
}

}  // namespace

int pcg_solve(elfel_hip_ctx *c, int64_t nfree, const double *F, double *x, int precision, double tol, int maxit,
              int *iters, int *outer, double *relres)
{
    const int64_t n = nfree, N = c->nrow;
    const int32_t *ptr = result_ptr(c);
    const int32_t *ind = result_ind(c);
    const double *val = (const double *)c->val.p;
    if (n == 0) {
        if (iters) *iters = 0;
        if (outer) *outer = 0;
        if (relres) *relres = 0;
        return 0;
    }
    const bool mixed = precision == ELFEL_MIXED;
    const int64_t nslices = (n + 63) / 64;
    hipStream_t st = c->stream;
    hipDeviceProp_t prop;
    HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
    // the CG grid: 4 workgroups per CU (grid-stride loops), fewer for small
    // systems (8 and 16 per CU measured slower on heat Q4 4096^2: 407 and 412
    // against 381 us per mixed iteration; every workgroup of the next kernel
    // sums G partials)
    const int G = (int)std::max<int64_t>(1, std::min<int64_t>(ceil_div(n, SV_BLOCK), 4 * (int64_t)prop.multiProcessorCount));

    // ---- scratch layout (8-byte words first, then 4-byte) -----------------
    // doubles: xfull[N], b[n], r64[n], dinv64[n], parts[5 G], sc[S_NS],
    //          FP64 vectors r z p q [4n] (FP64 solve only)
    // floats:  r z p q dinv xlo [6n] (mixed solve only; its iterate is the FP64 xfull + xlo)
    // uint32:  off[nslices + 1], len[max(n, nslices)], total[4]
    const size_t nd = (size_t)N + 3 * (size_t)n + 5 * (size_t)G + S_NS + (mixed ? 0 : 4 * (size_t)n);
    const size_t nf = mixed ? 6 * (size_t)n : 0;
    const size_t nu32 = (size_t)nslices + 1 + (size_t)std::max<int64_t>(n, nslices) + (size_t)n + 4;
    int rc = ensure(c, c->pcg, 8 * nd + 4 * nf + 4 * nu32 + 256);
    if (rc) return rc;
    double *dp = (double *)c->pcg.p;
    double *xfull = dp; dp += N;
    double *b = dp; dp += n;
    double *r64 = dp; dp += n;
    double *dinv64 = dp; dp += n;
    Parts P;
    P.G = G;
    P.pq = dp; dp += G; P.rz[0] = dp; dp += G; P.rz[1] = dp; dp += G; P.rr = dp; dp += G; P.aux = dp; dp += G;
    double *sc = dp; dp += S_NS;
    Vecs<double> v64{};
    if (!mixed) {
        v64.x = xfull;  // the free part of the full vector is the iterate
        v64.r = dp; dp += n; v64.z = dp; dp += n; v64.p = dp; dp += n; v64.q = dp; dp += n;
        v64.dinv = dinv64;
    }
    float *fp = (float *)dp;
    Vecs<float> v32{};
    if (mixed) {
        v32.x = xfull;  // the free part of the full vector is the iterate (FP64)
        v32.r = fp; fp += n; v32.z = fp; fp += n; v32.p = fp; fp += n; v32.q = fp; fp += n;
        v32.dinv = fp; fp += n;
        v32.xlo = fp; fp += n;
    }
    uint32_t *up = (uint32_t *)fp;
    uint32_t *off = up; up += nslices + 1;
    uint32_t *len = up; up += std::max<int64_t>(n, nslices);
    uint32_t *far = up; up += n;
    uint32_t *total = up;

    HIPCHK(c, hipMemcpyAsync(xfull, x, 8 * N, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(b, F, 8 * n, hipMemcpyHostToDevice, st));

    // ---- SELL-64 of the free block ------------------------------------------
    hipLaunchKernelGGL(k_sell_rowlen, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, ptr, ind, len, far);
    hipLaunchKernelGGL(k_sell_width, dim3(ceil_div(nslices * 64, 256)), dim3(256), 0, st, n, nslices, len, off);
    if ((rc = exclusive_scan_u32(c, off, nslices, total))) return rc;
    if ((rc = exclusive_scan_u32(c, far, n, total + 1))) return rc;
    uint32_t tot[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(tot, total, 8, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const uint32_t nnzp = tot[0];
    const bool c16 = tot[1] == 0;  // every free column within a 16-bit offset of its row
    HIPCHK(c, hipMemcpyAsync(off + nslices, total, 4, hipMemcpyDeviceToDevice, st));
    // SELL arrays in a buffer of their own (the size is known only now)
    const size_t cbytes = (size_t)nnzp * (c16 ? 2 : 4);
    const size_t nslots = (size_t)nnzp / 64;
    const size_t vbytes = (size_t)nnzp * (8 + (mixed ? 4 : 0));
    const size_t sbytes = ((cbytes + 15) & ~(size_t)15) + vbytes + 4 * (size_t)(nslices + 1) + 4 * nslots + 256;
    if ((rc = ensure(c, c->pcg_sell, sbytes))) return rc;
    void *scol = c->pcg_sell.p;
    double *sv64 = (double *)((char *)c->pcg_sell.p + ((cbytes + 15) & ~(size_t)15));
    float *sv32 = mixed ? (float *)(sv64 + nnzp) : nullptr;
    uint32_t *uflag = (uint32_t *)((char *)sv64 + vbytes);
    int32_t *uoff = (int32_t *)(uflag + nslices + (nslices & 1));
    if (c16)
        hipLaunchKernelGGL((k_sell_fill<int16_t>), dim3(ceil_div(nslices * 64, 256)), dim3(256), 0, st, n, ptr, ind, val,
                           off, (int16_t *)scol, sv64, sv32);
    else
        hipLaunchKernelGGL((k_sell_fill<int32_t>), dim3(ceil_div(nslices * 64, 256)), dim3(256), 0, st, n, ptr, ind, val,
                           off, (int32_t *)scol, sv64, sv32);
    const bool uniform_slots = env_int("ELFEL_HIP_SELL_UNIFORM", 1) != 0;
    if (uniform_slots) {
        if (c16)
            hipLaunchKernelGGL((k_sell_uniform<int16_t>), dim3(ceil_div(nslices * 64, 256)), dim3(256), 0, st, n,
                               nslices, (const uint32_t *)off, (const int16_t *)scol, uflag, uoff);
        else
            hipLaunchKernelGGL((k_sell_uniform<int32_t>), dim3(ceil_div(nslices * 64, 256)), dim3(256), 0, st, n,
                               nslices, (const uint32_t *)off, (const int32_t *)scol, uflag, uoff);
    } else {
        hipLaunchKernelGGL(k_fill<uint32_t>, dim3(ceil_div(nslices, 256)), dim3(256), 0, st, nslices, uflag, 0u);
    }
    hipLaunchKernelGGL(k_diag_inv, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, ptr, ind, val, dinv64,
                       mixed ? v32.dinv : nullptr);
    HIPCHK(c, hipGetLastError());
    const Sell S{off, scol, c16, sv64, sv32, SellU{uflag, uoff}};

    auto read_sc = [&](double *h) -> int {
        HIPCHK(c, hipMemcpyAsync(h, sc, sizeof(double) * S_NS, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        return 0;
    };
    // b = F_f - K_fd x_d, ||b||
    hipLaunchKernelGGL(k_rhs, dim3(G), dim3(SV_BLOCK), 0, st, n, ptr, ind, val, (const double *)xfull,
                       (const double *)b, b, P.aux);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, (const double *)P.aux, G, sc, (int)S_BB,
                       (const double *)nullptr, 0);
    HIPCHK(c, hipGetLastError());
    double h[S_NS];
    if ((rc = read_sc(h))) return rc;
    const double bnorm = h[S_BB] > 0 ? sqrt(h[S_BB]) : 1.0;
    const double target = tol * bnorm;
    // r64 = b - A x (FP64), its norm
    auto true_residual = [&](double *out_norm) -> int {
        if (S.c16)
            hipLaunchKernelGGL((k_sell_resid<int16_t>), dim3(G), dim3(SV_BLOCK), 0, st, n, S.off,
                               (const int16_t *)S.col, S.v64, S.U, (const double *)xfull, (const double *)b, r64,
                               P.aux);
        else
            hipLaunchKernelGGL((k_sell_resid<int32_t>), dim3(G), dim3(SV_BLOCK), 0, st, n, S.off,
                               (const int32_t *)S.col, S.v64, S.U, (const double *)xfull, (const double *)b, r64,
                               P.aux);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, (const double *)P.aux, G, sc, (int)S_RR64,
                           (const double *)nullptr, 0);
        HIPCHK(c, hipGetLastError());
        double hh[S_NS];
        int r2 = read_sc(hh);
        if (r2) return r2;
        *out_norm = sqrt(std::max(hh[S_RR64], 0.0));
        return 0;
    };
    // r = (T) r64, z = M r (p = z on restart); the rz partials go where the next
    // step reads the previous ones (rz[par ^ 1]); sc gets rz and ||r||^2
    int par = 0;
    auto reset = [&](auto &v, int restart) {
        using T = typename std::remove_pointer<decltype(v.r)>::type;
        hipLaunchKernelGGL((k_cg_reset<T>), dim3(G), dim3(SV_BLOCK), 0, st, n, (const double *)r64,
                           (const T *)v.dinv, v.r, v.z, v.p, restart, P.rz[par ^ 1], P.rr);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, st, (const double *)P.rr, G, sc, (int)S_RR,
                           (const double *)P.rz[par ^ 1], (int)S_RZ);
    };

    int it = 0, nupd = 0;
    double rn = 0.0;
    if ((rc = true_residual(&rn))) return rc;
    if (!mixed) {
        reset(v64, 1);
        nupd = 1;
        while (rn > target && it < maxit) {
            // iterate until the recurrence residual reaches the target ...
            double rrec = 0.0;
            for (;;) {
                if ((rc = read_sc(h))) return rc;
                rrec = h[S_RR];
                if (!(rrec > target * target) || it >= maxit) break;
                for (int k = 0; k < CHECK; k++, par ^= 1) cg_step<double>(st, n, S, v64, P, par, sc);
                HIPCHK(c, hipGetLastError());
                it += CHECK;
            }
            // ... then check the true residual and continue from it if needed
            if ((rc = true_residual(&rn))) return rc;
            if (rn > target && it < maxit) {
                reset(v64, !(sqrt(std::max(rrec, 0.0)) >= RU_KEEP * rn));
                nupd++;
            }
        }
    } else {
        hipLaunchKernelGGL(k_fill<float>, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, v32.xlo, 0.f);
        reset(v32, 1);
        double rlast = rn;
        while (rn > target && it < maxit) {
            const double goal = std::max(target, RU_DROP * rlast);
            double rrec = 0.0;  // recurrence residual norm^2 at the last check
            for (;;) {
                if ((rc = read_sc(h))) return rc;
                rrec = h[S_RR];
                if (!(rrec > goal * goal) || it >= maxit) break;
                for (int k = 0; k < CHECK; k++, par ^= 1) cg_step<float>(st, n, S, v32, P, par, sc);
                hipLaunchKernelGGL(k_cg_flush, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, v32.x, v32.xlo);
                HIPCHK(c, hipGetLastError());
                it += CHECK;
            }
            // reliable update: the FP64 residual b - A x replaces the FP32 one
            if ((rc = true_residual(&rn))) return rc;
            nupd++;
            rlast = rn;
            // The search direction is kept while the recurrence still tracks the true residual; once
            // it has run far below it (FP32 round-off, or a tiny system solved exactly) p no longer
            // belongs to the new residual and the iteration restarts from p = z.
            if (rn > target && it < maxit) reset(v32, !(sqrt(std::max(rrec, 0.0)) >= RU_KEEP * rn));
        }
    }
    HIPCHK(c, hipGetLastError());
    HIPCHK(c, hipMemcpyAsync(x, xfull, 8 * n, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (iters) *iters = it;
    if (outer) *outer = nupd;
    if (relres) *relres = rn / bnorm;
    return 0;
}
