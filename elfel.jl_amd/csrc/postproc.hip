// postproc.hip -- post-processing integrals on the device: the L2 errors of
// the colliding-flow manufactured solution (test/test_stokes.jl:354-356 exact
// fields, :438-496 and :675-731 the norms), on the same gather + quadrature
// structure as the assembly kernels.
//
// Per element, in FP64 and in the reference's operation order (this file is
// compiled with -ffp-contract=off): J = sum_k x_k (x) grad N_k with the first
// term assigned, JxW = det J w_q, location x = sum_k x_k N_k (first term
// assigned), the discrete field sum_j u_j N_j from 0.0, and the integrand
// JxW (u_h - u)^2.  The element terms are summed by a fixed two-stage
// reduction (per workgroup, then in workgroup order), not left to right over
// all elements as the reference does, so the norms agree to rounding.
#include <math.h>

#include "common.h"

namespace {

__device__ __forceinline__ double truep(double x, double y) { return 60 * (x * x) * y - 20 * (y * y * y); }
__device__ __forceinline__ double trueux(double x, double y) { return 20 * x * (y * y * y); }
__device__ __forceinline__ double trueuy(double x, double y) { return 5 * (x * x * x * x) - 5 * (y * y * y * y); }

template <int NEN>
__device__ __forceinline__ void jac_loc(const QTable &t, int q, const double *X, const double *Y, double *jxw,
                                        double *lx, double *ly)
{
    double j11 = X[0] * t.G[q][0][0], j21 = Y[0] * t.G[q][0][0];
    double j12 = X[0] * t.G[q][0][1], j22 = Y[0] * t.G[q][0][1];
    double x = X[0] * t.N[q][0], y = Y[0] * t.N[q][0];
#pragma unroll
    for (int k = 1; k < NEN; k++) {
        j11 += X[k] * t.G[q][k][0];
        j21 += Y[k] * t.G[q][k][0];
        j12 += X[k] * t.G[q][k][1];
        j22 += Y[k] * t.G[q][k][1];
        x += X[k] * t.N[q][k];
        y += Y[k] * t.N[q][k];
    }
    *jxw = (j11 * j22 - j21 * j12) * t.w[q];
    *lx = x;
    *ly = y;
}

__device__ __forceinline__ double block_sum(double v, double *sh)
{
    for (int d = 32; d > 0; d >>= 1) v += __shfl_down(v, d, 64);
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    __syncthreads();
    if (lane == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int k = 0; k < (int)(blockDim.x >> 6); k++) s += sh[k];
    return s;
}

// Workgroup partials of the pressure (T3, field 1) and velocity (T6 vector,
// field 0) error integrands.
__global__ __launch_bounds__(256) void k_colliding(int64_t Ep, const int32_t *__restrict__ conn3,
                                                   const double2 *__restrict__ xy3, const int32_t *__restrict__ dp,
                                                   int64_t Ev, const int32_t *__restrict__ conn6,
                                                   const double2 *__restrict__ xy6, const int32_t *__restrict__ du,
                                                   const double *__restrict__ U, QTable t3, QTable t6,
                                                   double *__restrict__ part)
{
    __shared__ double sh[4];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double errp = 0.0, errv = 0.0;
    if (e < Ep) {
        double X[3], Y[3], dv[3];
#pragma unroll
        for (int k = 0; k < 3; k++) {
            const int32_t nd = conn3[e * 3 + k];
            X[k] = xy3[nd].x; Y[k] = xy3[nd].y;
            dv[k] = U[dp[nd]];
        }
#pragma unroll
        for (int q = 0; q < 3; q++) {
            double jxw, lx, ly;
            jac_loc<3>(t3, q, X, Y, &jxw, &lx, &ly);
            const double pt = truep(lx, ly);
            double pa = 0.0;
#pragma unroll
            for (int j = 0; j < 3; j++) pa += dv[j] * t3.N[q][j];
            errp += jxw * (pa - pt) * (pa - pt);
        }
    }
    if (e < Ev) {
        double X[6], Y[6], vx[6], vy[6];
#pragma unroll
        for (int k = 0; k < 6; k++) {
            const int32_t nd = conn6[e * 6 + k];
            X[k] = xy6[nd].x; Y[k] = xy6[nd].y;
            vx[k] = U[du[2 * nd]];
            vy[k] = U[du[2 * nd + 1]];
        }
#pragma unroll
        for (int q = 0; q < 3; q++) {
            double jxw, lx, ly;
            jac_loc<6>(t6, q, X, Y, &jxw, &lx, &ly);
            double ua = 0.0, va = 0.0;
#pragma unroll
            for (int j = 0; j < 6; j++) { ua += vx[j] * t6.N[q][j]; va += vy[j] * t6.N[q][j]; }
            const double ux = trueux(lx, ly), uy = trueuy(lx, ly);
            errv += jxw * ((ua - ux) * (ua - ux) + (va - uy) * (va - uy));
        }
    }
    const double a = block_sum(errp, sh);
    const double b = block_sum(errv, sh);
    if (threadIdx.x == 0) { part[2 * blockIdx.x] = a; part[2 * blockIdx.x + 1] = b; }
}

__global__ void k_sum_pairs(const double *__restrict__ part, int n, double *__restrict__ out)
{
    __shared__ double sh[16];
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < n; i += blockDim.x) { a += part[2 * i]; b += part[2 * i + 1]; }
    a = block_sum(a, sh);
    b = block_sum(b, sh);
    if (threadIdx.x == 0) { out[0] = a; out[1] = b; }
}

}  // namespace

extern "C" int elfel_hip_colliding_errors(elfel_hip_ctx *c, int64_t ndofs, const double *U, double *ep, double *ev)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        const Field &V = c->f[0], &P = c->f[1];
        if (!V.has_mesh || !V.has_dofs || !P.has_mesh || !P.has_dofs)
            return set_err(c, ELFEL_ESTATE, "colliding_errors: set_mesh/set_dofs of both fields first");
        if (V.nen != 6 || V.ncomp != 2 || P.nen != 3 || P.ncomp != 1 || V.sdim != 2 || P.sdim != 2)
            return set_err(c, ELFEL_EINVAL, "colliding_errors: field 0 must be vector T6, field 1 scalar T3");
        if (!U || ndofs < std::max(V.ndofs, P.ndofs)) return set_err(c, ELFEL_EINVAL, "colliding_errors: bad vector");
        QTable t3, t6;
        if (make_qtable(ELFEL_T3, 3, &t3) || make_qtable(ELFEL_T6, 3, &t6))
            return set_err(c, ELFEL_EINVAL, "no quadrature table");
        const int64_t E = std::max(V.E, P.E);
        const int nb = ceil_div(std::max<int64_t>(E, 1), 256);
        int rc = ensure(c, c->pcg, 8 * ((size_t)ndofs + 2 * (size_t)nb + 2) + 256);
        if (rc) return rc;
        double *d_U = (double *)c->pcg.p, *part = d_U + ndofs, *out = part + 2 * nb;
        hipStream_t st = c->stream;
        HIPCHK(c, hipMemcpyAsync(d_U, U, 8 * (size_t)ndofs, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_colliding, dim3(nb), dim3(256), 0, st, P.E, (const int32_t *)P.conn.p,
                           (const double2 *)P.xy.p, (const int32_t *)P.dof.p, V.E, (const int32_t *)V.conn.p,
                           (const double2 *)V.xy.p, (const int32_t *)V.dof.p, (const double *)d_U, t3, t6, part);
        hipLaunchKernelGGL(k_sum_pairs, dim3(1), dim3(1024), 0, st, (const double *)part, nb, out);
        HIPCHK(c, hipGetLastError());
        double h[2];
        HIPCHK(c, hipMemcpyAsync(h, out, 16, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (ep) *ep = sqrt(h[0]);
        if (ev) *ev = sqrt(h[1]);
        return ELFEL_OK;
    });
}
