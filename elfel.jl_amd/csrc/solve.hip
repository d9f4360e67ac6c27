// solve.hip -- Jacobi-preconditioned conjugate gradients on the free block of
// the assembled matrix, replacing the driver's `K[1:nu,1:nu] \ (F - K*T)[1:nu]`
// (test/test_heat.jl:73-76) for SPD problems.
//
// FP64 PCG: textbook PCG (Saad, Iterative Methods for Sparse Linear Systems,
// Alg. 9.1) with every vector and the matrix in FP64.
// Mixed: iterative refinement (Wilkinson; Carson & Higham 2018) -- the
// residual r = b - A x is formed in FP64 with the FP64 matrix, the correction
// A d = r is solved by FP32 PCG (FP32 matrix copy and vectors, FP64 dot
// products), and x += d in FP64, until ||b - A x|| / ||b|| <= tol.
//
// SpMV: a group of W lanes per row (W picked from the mean row length) reads
// colind/values with unit stride; y = A p and the dot product p.Ap are fused.
// Reductions are two-stage (per-block partials, then one block), so every run
// is deterministic.  The CG scalars stay on the device; the host reads the
// residual norm every CHECK iterations.
#include <math.h>

#include <algorithm>

#include "common.h"

#define SV_BLOCK 256
#define CHECK 8

namespace {

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    if (threadIdx.x == 0)
        for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sh[k];
    return s;
}

// Scalars (double): 0 rz, 1 pq, 2 rr, 3 rz_new, 4 bb, 5 alpha/beta scratch
enum { S_RZ = 0, S_PQ = 1, S_RR = 2, S_RZN = 3, S_BB = 4, S_RR64 = 5 };

template <typename T, int W>
__global__ __launch_bounds__(SV_BLOCK) void k_spmv(int64_t nfree, const int32_t *__restrict__ ptr,
                                                   const int32_t *__restrict__ ind, const T *__restrict__ val,
                                                   const T *__restrict__ x, T *__restrict__ y,
                                                   double *__restrict__ partial)
{
    __shared__ double sh[SV_BLOCK / 64];
    const int64_t row = ((int64_t)blockIdx.x * SV_BLOCK + threadIdx.x) / W;
    const int lane = threadIdx.x % W;
    T acc = T(0);
    if (row < nfree) {
        const int32_t b = ptr[row], e = ptr[row + 1];
        for (int32_t k = b + lane; k < e; k += W) {
            const int32_t j = ind[k];
            if (j < nfree) acc += val[k] * x[j];
        }
    }
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) acc += __shfl_down(acc, d, W);
    double dp = 0.0;
    if (lane == 0 && row < nfree) {
        y[row] = acc;
        if (partial) dp = (double)acc * (double)x[row];
    }
    if (partial) {
        const double s = block_sum(dp, sh);
        if (threadIdx.x == 0) partial[blockIdx.x] = s;
    }
}

// r = b - A x (FP64, free block), partial ||r||^2.  b = F_f - K_fd x_d when
// `data` is set (x holds the data values beyond nfree).
template <int W>
__global__ __launch_bounds__(SV_BLOCK) void k_residual(int64_t nfree, const int32_t *__restrict__ ptr,
                                                       const int32_t *__restrict__ ind,
                                                       const double *__restrict__ val, const double *__restrict__ x,
                                                       const double *__restrict__ rhs, double *__restrict__ r,
                                                       double *__restrict__ partial, int data_only)
{
    __shared__ double sh[SV_BLOCK / 64];
    const int64_t row = ((int64_t)blockIdx.x * SV_BLOCK + threadIdx.x) / W;
    const int lane = threadIdx.x % W;
    double acc = 0.0;
    if (row < nfree) {
        const int32_t b = ptr[row], e = ptr[row + 1];
        for (int32_t k = b + lane; k < e; k += W) {
            const int32_t j = ind[k];
            if ((j >= nfree) == (data_only != 0)) acc += val[k] * x[j];
        }
    }
#pragma unroll
    for (int d = W / 2; d > 0; d >>= 1) acc += __shfl_down(acc, d, W);
    double rr = 0.0;
    if (lane == 0 && row < nfree) {
        const double v = rhs[row] - acc;
        r[row] = v;
        rr = v * v;
    }
    const double s = block_sum(rr, sh);
    if (threadIdx.x == 0) partial[blockIdx.x] = s;
}

__global__ void k_reduce(const double *__restrict__ partial, int n, double *__restrict__ out)
{
    __shared__ double sh[1024 / 64];
    double v = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) v += partial[i];
    const double s = block_sum(v, sh);
    if (threadIdx.x == 0) *out = s;
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

__global__ void k_to_f32(int64_t n, const double *__restrict__ a, float *__restrict__ b)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) b[i] = (float)a[i];
}

// r, z = M r, p = z, partial r.z and r.r
template <typename T>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_init(int64_t n, const T *__restrict__ dinv, T *__restrict__ r,
                                                      T *__restrict__ z, T *__restrict__ p, T *__restrict__ x,
                                                      const double *__restrict__ rhs64, double *__restrict__ prz,
                                                      double *__restrict__ prr)
{
    __shared__ double sh[SV_BLOCK / 64];
    const int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (i < n) {
        const T ri = rhs64 ? (T)rhs64[i] : r[i];
        const T zi = dinv[i] * ri;
        r[i] = ri; z[i] = zi; p[i] = zi;
        if (rhs64) x[i] = T(0);
        rz = (double)ri * (double)zi;
        rr = (double)ri * (double)ri;
    }
    double s = block_sum(rz, sh);
    if (threadIdx.x == 0) prz[blockIdx.x] = s;
    __syncthreads();
    s = block_sum(rr, sh);
    if (threadIdx.x == 0) prr[blockIdx.x] = s;
}

// alpha = rz / pq; x += alpha p; r -= alpha q; z = M r; partials of r.z, r.r
template <typename T>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_update(int64_t n, const double *__restrict__ sc,
                                                        const T *__restrict__ dinv, T *__restrict__ x,
                                                        T *__restrict__ r, T *__restrict__ z,
                                                        const T *__restrict__ p, const T *__restrict__ q,
                                                        double *__restrict__ prz, double *__restrict__ prr)
{
    __shared__ double sh[SV_BLOCK / 64];
    const T alpha = (T)(sc[S_RZ] / sc[S_PQ]);
    const int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x;
    double rz = 0.0, rr = 0.0;
    if (i < n) {
        x[i] += alpha * p[i];
        const T ri = r[i] - alpha * q[i];
        const T zi = dinv[i] * ri;
        r[i] = ri; z[i] = zi;
        rz = (double)ri * (double)zi;
        rr = (double)ri * (double)ri;
    }
    double s = block_sum(rz, sh);
    if (threadIdx.x == 0) prz[blockIdx.x] = s;
    __syncthreads();
    s = block_sum(rr, sh);
    if (threadIdx.x == 0) prr[blockIdx.x] = s;
}

// beta = rz_new / rz; p = z + beta p; rz = rz_new (the last block updates rz
// after every block has read it: done by a separate tiny kernel)
template <typename T>
__global__ __launch_bounds__(SV_BLOCK) void k_cg_dir(int64_t n, const double *__restrict__ sc,
                                                     const T *__restrict__ z, T *__restrict__ p)
{
    const T beta = (T)(sc[S_RZN] / sc[S_RZ]);
    const int64_t i = (int64_t)blockIdx.x * SV_BLOCK + threadIdx.x;
    if (i < n) p[i] = z[i] + beta * p[i];
}

__global__ void k_roll(double *sc) { sc[S_RZ] = sc[S_RZN]; }

// x64 += d32 (refinement step)
__global__ void k_add_correction(int64_t n, const float *__restrict__ d, double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += (double)d[i];
}

__global__ void k_axpy64(int64_t n, const double *__restrict__ d, double *__restrict__ x)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) x[i] += d[i];
}

template <typename T>
struct Vecs {
    T *r, *z, *p, *q, *x, *dinv;
};

}  // namespace

typedef void (*spmv_fn)(dim3, hipStream_t, int64_t, const int32_t *, const int32_t *, const void *, const void *,
                        void *, double *);

template <typename T, int W>
static void spmv_launch(dim3 g, hipStream_t s, int64_t n, const int32_t *ptr, const int32_t *ind, const void *val,
                        const void *x, void *y, double *partial)
{
    hipLaunchKernelGGL((k_spmv<T, W>), g, dim3(SV_BLOCK), 0, s, n, ptr, ind, (const T *)val, (const T *)x, (T *)y,
                       partial);
}

template <int W>
static void resid_launch(dim3 g, hipStream_t s, int64_t n, const int32_t *ptr, const int32_t *ind, const double *val,
                         const double *x, const double *rhs, double *r, double *partial, int data_only)
{
    hipLaunchKernelGGL((k_residual<W>), g, dim3(SV_BLOCK), 0, s, n, ptr, ind, val, x, rhs, r, partial, data_only);
}

// Runs PCG on A (ptr, ind, val of type T) for rhs (FP64, device); x (T) out.
// Returns iterations; *rr_out = final ||r||^2 of the recurrence.
template <typename T>
static int cg_loop(elfel_hip_ctx *c, int64_t n, const int32_t *ptr, const int32_t *ind, const T *val, int W,
                   const Vecs<T> &v, const double *rhs, double *sc, double *prz, double *prr, double tol_abs,
                   int maxit, int *iters, double *rr_out)
{
    const int nb_v = ceil_div(n, SV_BLOCK);
    const int rows_per_block = SV_BLOCK / W;
    const int nb_s = ceil_div(n, rows_per_block);
    spmv_fn sp = W == 4 ? spmv_launch<T, 4> : W == 8 ? spmv_launch<T, 8> : W == 16 ? spmv_launch<T, 16> : spmv_launch<T, 32>;
    hipLaunchKernelGGL((k_cg_init<T>), dim3(nb_v), dim3(SV_BLOCK), 0, c->stream, n, v.dinv, v.r, v.z, v.p, v.x, rhs,
                       prz, prr);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c->stream, prz, nb_v, sc + S_RZ);
    hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c->stream, prr, nb_v, sc + S_RR);
    HIPCHK(c, hipGetLastError());
    int it = 0;
    double hsc[8];
    for (;;) {
        if (it % CHECK == 0 || it >= maxit) {
            HIPCHK(c, hipMemcpyAsync(hsc, sc, sizeof(hsc), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            if (!(hsc[S_RR] > tol_abs * tol_abs) || it >= maxit) break;
        }
        sp(dim3(nb_s), c->stream, n, ptr, ind, val, v.p, v.q, prz);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c->stream, prz, nb_s, sc + S_PQ);
        hipLaunchKernelGGL((k_cg_update<T>), dim3(nb_v), dim3(SV_BLOCK), 0, c->stream, n, sc, v.dinv, v.x, v.r, v.z,
                           v.p, v.q, prz, prr);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c->stream, prz, nb_v, sc + S_RZN);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c->stream, prr, nb_v, sc + S_RR);
        hipLaunchKernelGGL((k_cg_dir<T>), dim3(nb_v), dim3(SV_BLOCK), 0, c->stream, n, sc, v.z, v.p);
        hipLaunchKernelGGL(k_roll, dim3(1), dim3(1), 0, c->stream, sc);
        HIPCHK(c, hipGetLastError());
        it++;
    }
    *iters = it;
    *rr_out = hsc[S_RR];
    return 0;
}

int pcg_solve(elfel_hip_ctx *c, int64_t nfree, const double *F, double *x, int precision, double tol, int maxit,
              int *iters, int *outer, double *relres)
{
    const int64_t n = nfree, N = c->nrow, nnz = c->nnz;
    const int32_t *ptr = (const int32_t *)(c->result_is_pattern ? c->pat.ptr.p : c->ptr.p);
    const int32_t *ind = (const int32_t *)(c->result_is_pattern ? c->pat.ind.p : c->ind.p);
    const double *val = (const double *)c->val.p;
    if (n == 0) {
        if (iters) *iters = 0;
        if (outer) *outer = 0;
        if (relres) *relres = 0;
        return 0;
    }
    const double mean_row = (double)nnz / (double)std::max<int64_t>(N, 1);
    const int W = mean_row <= 6 ? 4 : mean_row <= 12 ? 8 : mean_row <= 24 ? 16 : 32;
    const int nb_v = ceil_div(n, SV_BLOCK);
    const int nb_s = ceil_div(n, SV_BLOCK / W);
    const int nbp = std::max(nb_v, nb_s) + 8;
    const bool mixed = precision == ELFEL_MIXED;
    // layout of the scratch: doubles then floats
    size_t nd = (size_t)N /*xfull*/ + 3 * n /*b, r64, x64-view unused*/ + (mixed ? n : 5 * n) + 2 * (size_t)nbp + 16;
    size_t nf = mixed ? (size_t)nnz + 6 * n : 0;
    int rc = ensure(c, c->pcg, 8 * nd + 4 * nf + 256);
    if (rc) return rc;
    double *dp = (double *)c->pcg.p;
    double *xfull = dp; dp += N;
    double *b = dp; dp += n;
    double *r64 = dp; dp += n;
    double *dinv64 = dp; dp += n;
    double *prz = dp; dp += nbp;
    double *prr = dp; dp += nbp;
    double *sc = dp; dp += 16;
    Vecs<double> v64{};
    if (!mixed) {
        v64.r = dp; dp += n; v64.z = dp; dp += n; v64.p = dp; dp += n; v64.q = dp; dp += n; v64.x = xfull;
        v64.dinv = dinv64;
    } else {
        dp += n;
    }
    float *fp = (float *)dp;
    float *val32 = nullptr;
    Vecs<float> v32{};
    if (mixed) {
        val32 = fp; fp += nnz;
        v32.r = fp; fp += n; v32.z = fp; fp += n; v32.p = fp; fp += n; v32.q = fp; fp += n; v32.x = fp; fp += n;
        v32.dinv = fp; fp += n;
    }
    HIPCHK(c, hipMemcpyAsync(xfull, x, 8 * N, hipMemcpyHostToDevice, c->stream));
    HIPCHK(c, hipMemcpyAsync(b, F, 8 * n, hipMemcpyHostToDevice, c->stream));
    hipLaunchKernelGGL(k_diag_inv, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, n, ptr, ind, val, dinv64,
                       mixed ? v32.dinv : nullptr);
    if (mixed) hipLaunchKernelGGL(k_to_f32, dim3(ceil_div(nnz, 256)), dim3(256), 0, c->stream, nnz, val, val32);
    HIPCHK(c, hipGetLastError());
    auto resid = [&](const double *xx, const double *rhs, double *r, int data_only, double *out) {
        const dim3 g(nb_s);
        if (W == 4) resid_launch<4>(g, c->stream, n, ptr, ind, val, xx, rhs, r, prr, data_only);
        else if (W == 8) resid_launch<8>(g, c->stream, n, ptr, ind, val, xx, rhs, r, prr, data_only);
        else if (W == 16) resid_launch<16>(g, c->stream, n, ptr, ind, val, xx, rhs, r, prr, data_only);
        else resid_launch<32>(g, c->stream, n, ptr, ind, val, xx, rhs, r, prr, data_only);
        hipLaunchKernelGGL(k_reduce, dim3(1), dim3(1024), 0, c->stream, prr, nb_s, out);
    };
    // b = F_f - K_fd x_d
    resid(xfull, b, b, 1, sc + S_BB);
    double hbb = 0;
    HIPCHK(c, hipMemcpyAsync(&hbb, sc + S_BB, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const double bnorm = hbb > 0 ? sqrt(hbb) : 1.0;
    int total = 0, nouter = 0;
    double rn = 0.0;
    if (!mixed) {
        // r = b - A x (x = initial guess in xfull[0:n])
        resid(xfull, b, v64.r, 0, sc + S_RR64);
        // PCG from x0 in correction form: A d = r0 with d0 = 0, then x = x0 + d.
        Vecs<double> vv = v64;
        vv.x = r64;  // the correction d
        double rr = 0;
        int it = 0;
        rc = cg_loop<double>(c, n, ptr, ind, val, W, vv, v64.r, sc, prz, prr, tol * bnorm, maxit, &it, &rr);
        if (rc) return rc;
        hipLaunchKernelGGL(k_axpy64, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, n, (const double *)r64, xfull);
        HIPCHK(c, hipGetLastError());
        total = it;
        nouter = 1;
        resid(xfull, b, r64, 0, sc + S_RR64);
    } else {
        const double inner_tol = 1e-5;
        for (nouter = 0; nouter < 50; nouter++) {
            resid(xfull, b, r64, 0, sc + S_RR64);
            double h = 0;
            HIPCHK(c, hipMemcpyAsync(&h, sc + S_RR64, 8, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
            rn = sqrt(h);
            if (rn <= tol * bnorm || total >= maxit) break;
            double rr = 0;
            int it = 0;
            rc = cg_loop<float>(c, n, ptr, ind, val32, W, v32, r64, sc, prz, prr, inner_tol * rn, maxit - total, &it,
                                &rr);
            if (rc) return rc;
            total += it;
            hipLaunchKernelGGL(k_add_correction, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, n,
                               (const float *)v32.x, xfull);
            HIPCHK(c, hipGetLastError());
        }
    }
    double h = 0;
    HIPCHK(c, hipMemcpyAsync(&h, sc + S_RR64, 8, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(x, xfull, 8 * n, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    if (iters) *iters = total;
    if (outer) *outer = nouter;
    if (relres) *relres = sqrt(h) / bnorm;
    return 0;
}
