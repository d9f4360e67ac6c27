// solve.hip -- Jacobi-preconditioned conjugate gradients on the free block of
// the assembled matrix, replacing the driver's `K[1:nu,1:nu] \ (F - K*T)[1:nu]`
// (test/test_heat.jl:73-76) for SPD problems.
//
// FP64 PCG: textbook PCG (Saad, Iterative Methods for Sparse Linear Systems,
// Alg. 9.1) with every vector and the matrix in FP64; when the recurrence
// residual reaches the tolerance, the true residual b - A x is formed and
// replaces it (the iteration goes on if the true one is not yet small enough).
// Mixed: the same PCG with an FP32 matrix copy and FP32 vectors, FP64 dot
// products, and reliable updates (Sleijpen & van der Vorst, Computing 56,
// 1996): whenever the recurrence residual has fallen by RU_DROP since the last
// update, the FP32 iterate is added into the FP64 solution, the FP64 residual
// b - A x is formed with the FP64 matrix and replaces the FP32 one, and the
// search direction is kept (unless the recurrence has drifted far below the
// true residual: then p restarts from z).  The FP32 iterations keep their
// Krylov space, so they take about as many steps as FP64 PCG (+9% on heat Q4
// 256x256 with lognormal conductivity, in a numpy emulation of this code) at
// about 0.6 of the bytes each.
// Convergence is declared on the FP64 residual: ||b - A x|| / ||b|| <= tol.
//
// Storage: the free-free block in SELL-64 (one slice per 64 consecutive rows,
// each slice as wide as its longest row, entries column-major inside the
// slice), built once per solve from the CSR result: one thread per row, every
// load coalesced, data columns left out (their part of b is formed once with
// the CSR matrix).
// Per iteration three kernels: q = A p with the partial sums of p.q; the
// x / r / z update with the partials of r.z and r.r; p = z + beta p.  Each
// reduction finishes in the last workgroup to arrive (a counter, then a fixed-
// order sum of the partials), so every run gives the same bits; the CG scalars
// stay on the device and the host reads the residual every CHECK iterations.
#include <math.h>

#include <algorithm>

#include "common.h"

#define SV_BLOCK 256
#define CHECK 8
#define RU_DROP 1e-2   // reliable update when the recurrence residual has dropped by this factor

namespace {

// Scalars (double) on the device.
enum { S_RZ = 0, S_PQ = 1, S_RR = 2, S_ALPHA = 3, S_BETA = 4, S_BB = 5, S_RR64 = 6, S_NS = 8 };

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();  // sh may still be read by a previous call
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sh[k];
    return s;  // valid in thread 0
}

// True in every thread of the last workgroup to finish; the partials written
// before the call are then visible to it.
__device__ __forceinline__ bool last_block(unsigned *counter)
{
    __shared__ bool last;
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) last = atomicAdd(counter, 1u) == gridDim.x - 1;
    __syncthreads();
    if (last) __threadfence();  // acquire: the other workgroups' partials
    return last;
}

// Fixed-order sum of n partials by one workgroup (valid in thread 0).
__device__ __forceinline__ double sum_partials(const double *p, int n, double *sh)
{
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += p[i];
    return block_sum(v, sh);
}

// ---- SELL-64 construction ----------------------------------------------------
__global__ void k_sell_rowlen(int64_t nfree, const int32_t *__restrict__ ptr, const int32_t *__restrict__ ind,
                              uint32_t *__restrict__ len)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nfree) return;
    uint32_t k = 0;
    for (int32_t j = ptr[i]; j < ptr[i + 1]; j++) k += ind[j] < nfree ? 1u : 0u;
    len[i] = k;
}

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

__global__ void k_sell_fill(int64_t nfree, const int32_t *__restrict__ ptr, const int32_t *__restrict__ ind,
                            const double *__restrict__ val, const uint32_t *__restrict__ off,
                            int32_t *__restrict__ col, double *__restrict__ v64, float *__restrict__ v32)
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
            col[at] = cj;
            v64[at] = val[j];
            if (v32) v32[at] = (float)val[j];
            k++;
        }
    }
    for (; k < w; k++) {  // padding: a zero on the row's own column (or row 0 beyond nfree)
        const size_t at = (size_t)o + (size_t)k * 64 + lane;
        col[at] = i < nfree ? (int32_t)i : 0;
        v64[at] = 0.0;
        if (v32) v32[at] = 0.f;
    }
}

template <typename T>
__device__ __forceinline__ T sell_row(int64_t i, const uint32_t *__restrict__ off, const int32_t *__restrict__ col,
                                      const T *__restrict__ val, const T *__restrict__ x)
{
    const int64_t s = i >> 6;
    const int lane = (int)(i & 63);
    const uint32_t o = off[s], w = (off[s + 1] - o) >> 6;
    T acc = T(0);
    const int32_t *c = col + o + lane;
    const T *v = val + o + lane;
#pragma unroll 4
    for (uint32_t k = 0; k < w; k++) acc += v[(size_t)k * 64] * x[c[(size_t)k * 64]];
    return acc;
}

// q = A p, then alpha = rz / (p.q) in the last workgroup.
template <typename T>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_spmv(int64_t n, const uint32_t *__restrict__ off,
                                                      const int32_t *__restrict__ col, const T *__restrict__ val,
                                                      const T *__restrict__ p, T *__restrict__ q,
                                                      double *__restrict__ part, unsigned *counter,
                                                      double *__restrict__ sc)
{
    __shared__ double sh[SV_BLOCK / 64];
    const int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x;
    double d = 0.0;
    if (i < n) {
        const T y = sell_row<T>(i, off, col, val, p);
        q[i] = y;
        d = (double)p[i] * (double)y;
    }
    const double s = block_sum(d, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
    if (last_block(counter)) {
        const double pq = sum_partials(part, gridDim.x, sh);
        if (threadIdx.x == 0) {
            sc[S_PQ] = pq;
            sc[S_ALPHA] = pq != 0.0 ? sc[S_RZ] / pq : 0.0;
            *counter = 0u;
        }
    }
}

// x += alpha p; r -= alpha q; z = M r; then rz_new, rr and beta = rz_new / rz.
template <typename T>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_update(int64_t n, const T *__restrict__ dinv, T *__restrict__ x,
                                                        T *__restrict__ r, T *__restrict__ z,
                                                        const T *__restrict__ p, const T *__restrict__ q,
                                                        double *__restrict__ part, unsigned *counter,
                                                        double *__restrict__ sc)
{
    __shared__ double sh[SV_BLOCK / 64];
    const T alpha = (T)sc[S_ALPHA];
    const int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (i < n) {
        x[i] += alpha * p[i];
        const T ri = r[i] - alpha * q[i];
        const T zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        rz = (double)ri * (double)zi;
        rr = (double)ri * (double)ri;
    }
    const double s1 = block_sum(rz, sh);
    const double s2 = block_sum(rr, sh);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s1; part[2 * blockIdx.x + 1] = s2; }
    if (last_block(counter)) {
        double a = 0.0, b = 0.0;
        for (int k = threadIdx.x; k < (int)gridDim.x; k += blockDim.x) { a += part[2 * k]; b += part[2 * k + 1]; }
        a = block_sum(a, sh);
        b = block_sum(b, sh);
        if (threadIdx.x == 0) {
            const double rz_old = sc[S_RZ];
            sc[S_BETA] = rz_old != 0.0 ? a / rz_old : 0.0;
            sc[S_RZ] = a;
            sc[S_RR] = b;
            *counter = 0u;
        }
    }
}

template <typename T>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_dir(int64_t n, const double *__restrict__ sc,
                                                     const T *__restrict__ z, T *__restrict__ p)
{
    const T beta = (T)sc[S_BETA];
    const int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

// r = (T) r64, z = M r, and rz = r.z, rr = r.r; p = z when `restart`.
template <typename T>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_reset(int64_t n, const double *__restrict__ r64,
                                                       const T *__restrict__ dinv, T *__restrict__ r,
                                                       T *__restrict__ z, T *__restrict__ p, int restart,
                                                       double *__restrict__ part, unsigned *counter,
                                                       double *__restrict__ sc)
{
    __shared__ double sh[SV_BLOCK / 64];
    const int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (i < n) {
        const T ri = (T)r64[i];
        const T zi = dinv[i] * ri;
        r[i] = ri;
        z[i] = zi;
        if (restart) p[i] = zi;
        rz = (double)ri * (double)zi;
        rr = (double)ri * (double)ri;
    }
    const double s1 = block_sum(rz, sh);
    const double s2 = block_sum(rr, sh);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = s1; part[2 * blockIdx.x + 1] = s2; }
    if (last_block(counter)) {
        double a = 0.0, b = 0.0;
        for (int k = threadIdx.x; k < (int)gridDim.x; k += blockDim.x) { a += part[2 * k]; b += part[2 * k + 1]; }
        a = block_sum(a, sh);
        b = block_sum(b, sh);
        if (threadIdx.x == 0) {
            sc[S_RZ] = a;
            sc[S_RR] = b;
            *counter = 0u;
        }
    }
}

// x64 += (double) x32, x32 = 0 (FP32 iterate flushed into the FP64 solution);
// r64 = b - A64 x64 and ||r64||^2 follow in k_sell_resid.
__global__ void k_flush(int64_t n, float *__restrict__ x32, double *__restrict__ x64)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) {
        x64[i] += (double)x32[i];
        x32[i] = 0.f;
    }
}

// r = b - A x over the free block (SELL, FP64), ||r||^2 into sc[slot].
__global__ __launch_bounds__(SV_BLOCK) void k_sell_resid(int64_t n, const uint32_t *__restrict__ off,
                                                         const int32_t *__restrict__ col,
                                                         const double *__restrict__ val,
                                                         const double *__restrict__ x, const double *__restrict__ b,
                                                         double *__restrict__ r, double *__restrict__ part,
                                                         unsigned *counter, double *__restrict__ sc, int slot)
{
    __shared__ double sh[SV_BLOCK / 64];
    const int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x;
    double rr = 0.0;
    if (i < n) {
        const double v = b[i] - sell_row<double>(i, off, col, val, x);
        r[i] = v;
        rr = v * v;
    }
    const double s = block_sum(rr, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
    if (last_block(counter)) {
        const double t = sum_partials(part, gridDim.x, sh);
        if (threadIdx.x == 0) {
            sc[slot] = t;
            *counter = 0u;
        }
    }
}

// b = F_f - K_fd x_d with the CSR matrix (data columns only), ||b||^2.
__global__ __launch_bounds__(SV_BLOCK) void k_rhs(int64_t nfree, const int32_t *__restrict__ ptr,
                                                  const int32_t *__restrict__ ind, const double *__restrict__ val,
                                                  const double *__restrict__ xfull, const double *F,
                                                  double *b, double *__restrict__ part, unsigned *counter,
                                                  double *__restrict__ sc)
{
    __shared__ double sh[SV_BLOCK / 64];
    const int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x;
    double bb = 0.0;
    if (i < nfree) {
        double acc = 0.0;
        for (int32_t j = ptr[i]; j < ptr[i + 1]; j++)
            if (ind[j] >= nfree) acc += val[j] * xfull[ind[j]];
        const double v = F[i] - acc;
        b[i] = v;
        bb = v * v;
    }
    const double s = block_sum(bb, sh);
    if (threadIdx.x == 0) part[blockIdx.x] = s;
    if (last_block(counter)) {
        const double t = sum_partials(part, gridDim.x, sh);
        if (threadIdx.x == 0) {
            sc[S_BB] = t;
            *counter = 0u;
        }
    }
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

__global__ void k_fill_u32(int64_t n, uint32_t *__restrict__ a, uint32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

template <typename T>
__global__ void k_fill(int64_t n, T *__restrict__ a, T v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) a[i] = v;
}

template <typename T>
struct Vecs {
    T *x, *r, *z, *p, *q, *dinv;
};

struct Sell {
    const uint32_t *off;
    const int32_t *col;
    const double *v64;
    const float *v32;
};

template <typename T>
const T *sell_vals(const Sell &S);
template <>
const double *sell_vals<double>(const Sell &S) { return S.v64; }
template <>
const float *sell_vals<float>(const Sell &S) { return S.v32; }

// One PCG iteration on the stream.
template <typename T>
void cg_step(hipStream_t st, int64_t n, const Sell &S, const Vecs<T> &v, double *part, unsigned *counter, double *sc)
{
    const dim3 g(ceil_div(n, SV_BLOCK)), b(SV_BLOCK);
    hipLaunchKernelGGL((k_cg_spmv<T>), g, b, 0, st, n, S.off, S.col, sell_vals<T>(S), v.p, v.q, part, counter, sc);
    hipLaunchKernelGGL((k_cg_update<T>), g, b, 0, st, n, v.dinv, v.x, v.r, v.z, v.p, v.q, part, counter, sc);
    hipLaunchKernelGGL((k_cg_dir<T>), g, b, 0, st, n, sc, v.z, v.p);
}

}  // namespace

int pcg_solve(elfel_hip_ctx *c, int64_t nfree, const double *F, double *x, int precision, double tol, int maxit,
              int *iters, int *outer, double *relres)
{
    const int64_t n = nfree, N = c->nrow;
    const int32_t *ptr = (const int32_t *)(c->result_is_pattern ? c->pat.ptr.p : c->ptr.p);
    const int32_t *ind = (const int32_t *)(c->result_is_pattern ? c->pat.ind.p : c->ind.p);
    const double *val = (const double *)c->val.p;
    if (n == 0) {
        if (iters) *iters = 0;
        if (outer) *outer = 0;
        if (relres) *relres = 0;
        return 0;
    }
    const bool mixed = precision == ELFEL_MIXED;
    const int64_t nslices = (n + 63) / 64;
    const int nb = ceil_div(n, SV_BLOCK);
    hipStream_t st = c->stream;

    // ---- scratch layout (8-byte words first, then 4-byte) -----------------
    // doubles: xfull[N], b[n], r64[n], dinv64[n], part[2 nb + 8], sc[S_NS],
    //          fp64 vectors r z p q [4n] (FP64 solve only)
    // uint32:  counter[2], off[nslices + 1], len/size scratch[n]
    // after the solve-independent part: SELL col[nnzp] (int32), v64[nnzp], v32[nnzp]
    const size_t nd = (size_t)N + 3 * (size_t)n + 2 * (size_t)nb + 8 + S_NS + (mixed ? 0 : 4 * (size_t)n);
    const size_t nf = mixed ? 6 * (size_t)n : 0;  // x32 r z p q dinv32
    const size_t nu32 = 4 + (size_t)nslices + 1 + (size_t)std::max<int64_t>(n, nslices);
    size_t bytes = 8 * nd + 4 * nf + 4 * nu32 + 256;
    int rc = ensure(c, c->pcg, bytes);
    if (rc) return rc;
    char *base = (char *)c->pcg.p;
    double *dp = (double *)base;
    double *xfull = dp; dp += N;
    double *b = dp; dp += n;
    double *r64 = dp; dp += n;
    double *dinv64 = dp; dp += n;
    double *part = dp; dp += 2 * (size_t)nb + 8;
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
        v32.x = fp; fp += n; v32.r = fp; fp += n; v32.z = fp; fp += n; v32.p = fp; fp += n; v32.q = fp; fp += n;
        v32.dinv = fp; fp += n;
    }
    uint32_t *up = (uint32_t *)fp;
    unsigned *counter = up; up += 4;
    uint32_t *off = up; up += nslices + 1;
    uint32_t *len = up;

    hipLaunchKernelGGL(k_fill_u32, dim3(1), dim3(64), 0, st, (int64_t)4, counter, 0u);
    HIPCHK(c, hipMemcpyAsync(xfull, x, 8 * N, hipMemcpyHostToDevice, st));
    HIPCHK(c, hipMemcpyAsync(b, F, 8 * n, hipMemcpyHostToDevice, st));

    // ---- SELL-64 of the free block ------------------------------------------
    hipLaunchKernelGGL(k_sell_rowlen, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, ptr, ind, len);
    hipLaunchKernelGGL(k_sell_width, dim3(ceil_div(nslices * 64, 256)), dim3(256), 0, st, n, nslices, len, off);
    uint32_t *total = counter + 2;
    if ((rc = exclusive_scan_u32(c, off, nslices, total))) return rc;
    uint32_t nnzp = 0;
    HIPCHK(c, hipMemcpyAsync(&nnzp, total, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    HIPCHK(c, hipMemcpyAsync(off + nslices, total, 4, hipMemcpyDeviceToDevice, st));
    // SELL arrays in a second buffer (grown separately: the size is known only now)
    const size_t sbytes = (size_t)nnzp * (4 + 8 + (mixed ? 4 : 0)) + 256;
    if ((rc = ensure(c, c->pcg_sell, sbytes))) return rc;
    int32_t *scol = (int32_t *)c->pcg_sell.p;
    double *sv64 = (double *)((char *)c->pcg_sell.p + (((size_t)nnzp * 4 + 15) & ~(size_t)15));
    float *sv32 = mixed ? (float *)(sv64 + nnzp) : nullptr;
    hipLaunchKernelGGL(k_sell_fill, dim3(ceil_div(nslices * 64, 256)), dim3(256), 0, st, n, ptr, ind, val, off, scol,
                       sv64, sv32);
    hipLaunchKernelGGL(k_diag_inv, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, ptr, ind, val, dinv64,
                       mixed ? v32.dinv : nullptr);
    HIPCHK(c, hipGetLastError());
    const Sell S{off, scol, sv64, sv32};

    // b = F_f - K_fd x_d, ||b||
    hipLaunchKernelGGL(k_rhs, dim3(nb), dim3(SV_BLOCK), 0, st, n, ptr, ind, val, xfull, b, b, part, counter, sc);
    HIPCHK(c, hipGetLastError());
    auto true_residual = [&](double *out_norm) -> int {
        hipLaunchKernelGGL(k_sell_resid, dim3(nb), dim3(SV_BLOCK), 0, st, n, S.off, S.col, S.v64,
                           (const double *)xfull, (const double *)b, r64, part, counter, sc, (int)S_RR64);
        double h[S_NS];
        HIPCHK(c, hipMemcpyAsync(h, sc, sizeof(h), hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        *out_norm = sqrt(std::max(h[S_RR64], 0.0));
        return 0;
    };
    double hbb[S_NS];
    HIPCHK(c, hipMemcpyAsync(hbb, sc, sizeof(hbb), hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const double bnorm = hbb[S_BB] > 0 ? sqrt(hbb[S_BB]) : 1.0;
    const double target = tol * bnorm;

    int it = 0, nupd = 0;
    double rn = 0.0;
    if ((rc = true_residual(&rn))) return rc;
    if (!mixed) {
        hipLaunchKernelGGL((k_cg_reset<double>), dim3(nb), dim3(SV_BLOCK), 0, st, n, (const double *)r64,
                           (const double *)v64.dinv, v64.r, v64.z, v64.p, 1, part, counter, sc);
        nupd = 1;
        while (rn > target && it < maxit) {
            // iterate until the recurrence residual reaches the target ...
            double rrec = 0.0;
            for (;;) {
                double h[S_NS];
                HIPCHK(c, hipMemcpyAsync(h, sc, sizeof(h), hipMemcpyDeviceToHost, st));
                HIPCHK(c, hipStreamSynchronize(st));
                rrec = h[S_RR];
                if (!(rrec > target * target) || it >= maxit) break;
                for (int k = 0; k < CHECK; k++) cg_step<double>(st, n, S, v64, part, counter, sc);
                HIPCHK(c, hipGetLastError());
                it += CHECK;
            }
            // ... then check the true residual and continue from it if needed
            if ((rc = true_residual(&rn))) return rc;
            if (rn > target && it < maxit) {
                const int restart = !(sqrt(std::max(rrec, 0.0)) >= 0.5 * rn);
                hipLaunchKernelGGL((k_cg_reset<double>), dim3(nb), dim3(SV_BLOCK), 0, st, n, (const double *)r64,
                                   (const double *)v64.dinv, v64.r, v64.z, v64.p, restart, part, counter, sc);
                nupd++;
            }
        }
    } else {
        hipLaunchKernelGGL(k_fill<float>, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, v32.x, 0.f);
        hipLaunchKernelGGL((k_cg_reset<float>), dim3(nb), dim3(SV_BLOCK), 0, st, n, (const double *)r64,
                           (const float *)v32.dinv, v32.r, v32.z, v32.p, 1, part, counter, sc);
        double rlast = rn;
        while (rn > target && it < maxit) {
            const double goal = std::max(target, RU_DROP * rlast);
            double rrec = 0.0;  // recurrence residual norm^2 at the last check
            for (;;) {
                double h[S_NS];
                HIPCHK(c, hipMemcpyAsync(h, sc, sizeof(h), hipMemcpyDeviceToHost, st));
                HIPCHK(c, hipStreamSynchronize(st));
                rrec = h[S_RR];
                if (!(rrec > goal * goal) || it >= maxit) break;
                for (int k = 0; k < CHECK; k++) cg_step<float>(st, n, S, v32, part, counter, sc);
                HIPCHK(c, hipGetLastError());
                it += CHECK;
            }
            // reliable update: FP32 iterate into the FP64 solution, FP64 residual replaces the FP32 one
            hipLaunchKernelGGL(k_flush, dim3(ceil_div(n, 256)), dim3(256), 0, st, n, v32.x, xfull);
            if ((rc = true_residual(&rn))) return rc;
            nupd++;
            rlast = rn;
            // The search direction is kept while the recurrence still tracks the true residual; once
            // it has run far below it (FP32 round-off, or a tiny system solved exactly) p no longer
            // belongs to the new residual and the iteration restarts from p = z.
            const int restart = !(sqrt(std::max(rrec, 0.0)) >= 0.5 * rn);
            if (rn > target && it < maxit)
                hipLaunchKernelGGL((k_cg_reset<float>), dim3(nb), dim3(SV_BLOCK), 0, st, n, (const double *)r64,
                                   (const float *)v32.dinv, v32.r, v32.z, v32.p, restart, part, counter, sc);
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
