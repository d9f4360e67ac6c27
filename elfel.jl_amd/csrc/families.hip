// families.hip -- the element families beyond T3/T6/Q4 and the weak forms
// that use them, on the plain FP64 method (COO in Elfel's stream order, then
// the stable radix sort and in-order segmented reduce of plain.hip) and on the
// REUSE method (that sort done once and kept as a permutation; each assembly
// sums the runs through it in the same order).
// Compiled with -ffp-contract=off: every product and sum rounds as the
// reference's unfused Julia arithmetic does.
//
//   FEH1_L2 (src/FElements.jl:203-225)         heat, mass        1D, Gauss 1-3
//   FEH1_T4 (:356-384)                         heat, mass        3D, 1 or 4 points
//   FEH1_T3_BUBBLE (:322-354)                  Reddy Stokes velocity (P1b/P1)
//   FEL2_Q4 (:386-413), Q4 geometry carrier    Reddy Stokes pressure (Q1/Q0)
//   FEH1_T3 / T6 / Q4                          mass; T6/T3 Reddy Stokes
//
// One work-item per element.  Geometry (src/FElements.jl:113-168): J = sum_k
// x_k (outer) dN_k/dxi over the geometry nodes (first term assigned, then
// added); det J = |J11| in 1D, J11 J22 - J21 J12 in 2D, the unrolled 3x3
// expansion in 3D.  Spatial gradients (src/QPIterators.jl:132-140, g / J):
// g / J11 in 1D, else the row vector g times the adjugate of J over det J,
// each component summed in index order.
#include <string.h>

#include <algorithm>
#include <chrono>

#include "common.h"

namespace {

constexpr int FQ = 4;   // quadrature points
constexpr int FB = 6;   // basis functions of the scalar element

// Tables of one (element, rule) pair, reaching the kernels by value (every
// index is wave-uniform: scalar loads from the argument segment).
struct FamTab {
    int nq;
    double w[FQ];
    double Nu[FQ][FB];       // primary (velocity / scalar) basis
    double Gu[FQ][FB][3];    // its parametric gradients
    double Np[FQ][3];        // pressure basis (Reddy)
    double Gg[FQ][FB][3];    // parametric gradients of the geometry element
};

__host__ __device__ constexpr int fam_nbf(int fe)
{
    return fe == ELFEL_T3 ? 3 : fe == ELFEL_T6 ? 6 : fe == ELFEL_Q4 ? 4 : fe == ELFEL_T3B ? 4
         : fe == ELFEL_Q4L2 ? 1 : fe == ELFEL_T4 ? 4 : fe == ELFEL_L2 ? 2 : -1;
}
__host__ __device__ constexpr int fam_ngeo(int fe)
{
    return fe == ELFEL_T3 || fe == ELFEL_T3B ? 3 : fe == ELFEL_T6 ? 6 : fe == ELFEL_Q4 || fe == ELFEL_Q4L2 ? 4
         : fe == ELFEL_T4 ? 4 : fe == ELFEL_L2 ? 2 : -1;
}
__host__ __device__ constexpr int fam_mdim(int fe) { return fe == ELFEL_L2 ? 1 : fe == ELFEL_T4 ? 3 : 2; }

// Rules: interval Gauss (src/RefShapes.jl:85-110, literal constants),
// triangle 1/3 points (:113-119), square Gauss tensor i outer / j inner
// (:333-366), tetrahedron 1/4 points (:232-245).
int fam_rule(int fe, int rule, double (*pc)[3], double *w)
{
    for (int q = 0; q < FQ; q++) pc[q][0] = pc[q][1] = pc[q][2] = 0.0;
    auto gauss = [](int order, double *p, double *wt) -> int {
        if (order == 1) { p[0] = 0.0; wt[0] = 2.0; return 1; }
        if (order == 2) { p[0] = -0.577350269189626; p[1] = 0.577350269189626; wt[0] = 1.0; wt[1] = 1.0; return 2; }
        if (order == 3) {
            p[0] = -0.774596669241483; p[1] = 0.0; p[2] = 0.774596669241483;
            wt[0] = 0.5555555555555556; wt[1] = 0.8888888888888889; wt[2] = 0.5555555555555556;
            return 3;
        }
        return -1;
    };
    if (fe == ELFEL_L2) {
        double p[3], wt[3];
        const int n = gauss(rule, p, wt);
        for (int q = 0; q < n; q++) { pc[q][0] = p[q]; w[q] = wt[q]; }
        return n;
    }
    if (fe == ELFEL_T4) {
        if (rule == 1) { pc[0][0] = pc[0][1] = pc[0][2] = 0.25; w[0] = 1.0 / 6.0; return 1; }
        if (rule == 4) {
            const double a = 0.13819660, b = 0.58541020;
            const double P[4][3] = {{a, a, a}, {b, a, a}, {a, b, a}, {a, a, b}};
            for (int q = 0; q < 4; q++) {
                for (int d = 0; d < 3; d++) pc[q][d] = P[q][d];
                w[q] = 0.041666666666666666667;
            }
            return 4;
        }
        return -1;
    }
    if (fe == ELFEL_Q4 || fe == ELFEL_Q4L2) {
        double p[3], wt[3];
        const int n1 = rule <= 2 ? gauss(rule, p, wt) : -1;
        if (n1 < 0) return -1;
        int n = 0;
        for (int i = 0; i < n1; i++)
            for (int j = 0; j < n1; j++) { pc[n][0] = p[i]; pc[n][1] = p[j]; w[n] = wt[i] * wt[j]; n++; }
        return n;
    }
    if (rule == 1) { pc[0][0] = 1.0 / 3.; pc[0][1] = 1.0 / 3.; w[0] = 1.0 / 2.0; return 1; }
    if (rule == 3) {
        pc[0][0] = 2.0 / 3; pc[0][1] = 1.0 / 6;
        pc[1][0] = 1.0 / 6; pc[1][1] = 2.0 / 3;
        pc[2][0] = 1.0 / 6; pc[2][1] = 1.0 / 6;
        for (int q = 0; q < 3; q++) w[q] = (1.0 / 3) / 2;
        return 3;
    }
    return -1;
}

// Basis values and parametric gradients g[k][d] (src/FElements.jl; the
// nodal T3 / T6 / Q4 are qtable.h's).
void fam_basis(int fe, const double *pc, double *N, double (*g)[3])
{
    const double r = pc[0], s = pc[1], t = pc[2];
    for (int k = 0; k < FB; k++) g[k][0] = g[k][1] = g[k][2] = 0.0;
    switch (fe) {
    case ELFEL_T3B:
        N[0] = (1 - r - s); N[1] = r; N[2] = s;
        N[3] = (1 - r - s) * r * s;
        g[0][0] = -1.; g[0][1] = -1.; g[1][0] = +1.; g[1][1] = 0.; g[2][0] = 0.; g[2][1] = +1.;
        g[3][0] = (-r * s + (1 - r - s) * s);
        g[3][1] = (-r * s + (1 - r - s) * r);
        return;
    case ELFEL_Q4L2:
        N[0] = 1.0;
        return;
    case ELFEL_T4:
        N[0] = (1 - r - s - t); N[1] = r; N[2] = s; N[3] = t;
        g[0][0] = -1.0; g[0][1] = -1.0; g[0][2] = -1.0;
        g[1][0] = +1.0; g[2][1] = +1.0; g[3][2] = +1.0;
        return;
    case ELFEL_L2:
        N[0] = (1. - r) / 2.0; N[1] = (1. + r) / 2.0;
        g[0][0] = -1.0 / 2.0; g[1][0] = +1.0 / 2.0;
        return;
    default: {
        double g2[ELFEL_MAXBF][2];
        qt_bfun(fe, r, s, N);
        qt_bgrad(fe, r, s, g2);
        for (int k = 0; k < fam_nbf(fe); k++) { g[k][0] = g2[k][0]; g[k][1] = g2[k][1]; }
    }
    }
}

// ufe: the primary element; pfe: the pressure element (Reddy) or -1; gfe:
// the element whose nodes carry the geometry.
int fam_table(int ufe, int pfe, int gfe, int rule, FamTab *t)
{
    double pc[FQ][3], w[FQ];
    const int nq = fam_rule(ufe, rule, pc, w);
    if (nq < 1) return -1;
    *t = FamTab{};
    t->nq = nq;
    for (int q = 0; q < nq; q++) {
        double N[FB] = {}, g[FB][3];
        t->w[q] = w[q];
        fam_basis(ufe, pc[q], t->Nu[q], t->Gu[q]);
        if (pfe >= 0) {
            fam_basis(pfe, pc[q], N, g);
            for (int k = 0; k < 3; k++) t->Np[q][k] = N[k];
        }
        fam_basis(gfe, pc[q], N, t->Gg[q]);
    }
    return 0;
}

// ---- geometry on the device ---------------------------------------------------
template <int M, int NG>
struct Jac {
    double a[M][M];
    double det;

    __device__ __forceinline__ void make(const double *X, const double (*G)[3])
    {
#pragma unroll
        for (int s = 0; s < M; s++)
#pragma unroll
            for (int d = 0; d < M; d++) a[s][d] = X[s] * G[0][d];
#pragma unroll
        for (int k = 1; k < NG; k++)
#pragma unroll
            for (int s = 0; s < M; s++)
#pragma unroll
                for (int d = 0; d < M; d++) a[s][d] += X[k * M + s] * G[k][d];
        if constexpr (M == 1) det = fabs(a[0][0]);
        else if constexpr (M == 2) det = a[0][0] * a[1][1] - a[1][0] * a[0][1];
        else
            det = (+a[0][0] * (a[1][1] * a[2][2] - a[2][1] * a[1][2])
                   - a[0][1] * (a[1][0] * a[2][2] - a[1][2] * a[2][0])
                   + a[0][2] * (a[1][0] * a[2][1] - a[1][1] * a[2][0]));
    }

    // gx[s] = (g / J)_s for one basis function with parametric gradient g
    __device__ __forceinline__ void grad(const double *g, double *gx) const
    {
        if constexpr (M == 1) {
            gx[0] = g[0] / a[0][0];
        } else if constexpr (M == 2) {
            gx[0] = (a[1][1] * g[0] - a[1][0] * g[1]) / det;
            gx[1] = (a[0][0] * g[1] - a[0][1] * g[0]) / det;
        } else {
            double C[3][3];
            C[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
            C[0][1] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
            C[0][2] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
            C[1][0] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
            C[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
            C[1][2] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
            C[2][0] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
            C[2][1] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
            C[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
#pragma unroll
            for (int s = 0; s < 3; s++) gx[s] = (g[0] * C[s][0] + g[1] * C[s][1] + g[2] * C[s][2]) / det;
        }
    }

    // 1 / det J (1 / J11 in 1D), for grad_r
    __device__ __forceinline__ double rdet() const { return 1.0 / (M == 1 ? a[0][0] : det); }

    // grad with the divisions replaced by products with r = rdet() (the
    // mixed kernels: a few FP64 units in the last place, below the FP32
    // rounding of what they feed)
    __device__ __forceinline__ void grad_r(const double *g, double *gx, double r) const
    {
        if constexpr (M == 1) {
            gx[0] = g[0] * r;
        } else if constexpr (M == 2) {
            gx[0] = (a[1][1] * g[0] - a[1][0] * g[1]) * r;
            gx[1] = (a[0][0] * g[1] - a[0][1] * g[0]) * r;
        } else {
            double C[3][3];
            C[0][0] = a[1][1] * a[2][2] - a[1][2] * a[2][1];
            C[0][1] = a[1][2] * a[2][0] - a[1][0] * a[2][2];
            C[0][2] = a[1][0] * a[2][1] - a[1][1] * a[2][0];
            C[1][0] = a[0][2] * a[2][1] - a[0][1] * a[2][2];
            C[1][1] = a[0][0] * a[2][2] - a[0][2] * a[2][0];
            C[1][2] = a[0][1] * a[2][0] - a[0][0] * a[2][1];
            C[2][0] = a[0][1] * a[1][2] - a[0][2] * a[1][1];
            C[2][1] = a[0][2] * a[1][0] - a[0][0] * a[1][2];
            C[2][2] = a[0][0] * a[1][1] - a[0][1] * a[1][0];
#pragma unroll
            for (int s = 0; s < 3; s++) gx[s] = (g[0] * C[s][0] + g[1] * C[s][1] + g[2] * C[s][2]) * r;
        }
    }
};

// An element's N consecutive values: 16-byte stores when they fill whole
// 16-byte words from a 16-byte aligned start (every buffer starts 256-byte
// aligned and the element's offset is a multiple of N), else one per value.
template <int N, typename VT>
__device__ __forceinline__ void store_block(VT *dst, const VT (&v)[N])
{
    constexpr int W = 16 / (int)sizeof(VT);
    if constexpr (N % W == 0) {
#pragma unroll
        for (int k = 0; k < N / W; k++) {
            uint4 q;
            __builtin_memcpy(&q, &v[k * W], 16);
            *(uint4 *)(dst + k * W) = q;
        }
    } else {
#pragma unroll
        for (int k = 0; k < N; k++) dst[k] = v[k];
    }
}

// The workgroup's blocks of N values (elements e0 .. e0 + 255, contiguous in
// dst) written through LDS: each work-item's N values into its LDS row (N + 1
// apart, against bank conflicts), then the range out with consecutive lanes
// on consecutive 16-byte words.  Every work-item of the workgroup calls it.
template <int N, typename VT>
__device__ __forceinline__ void staged_store(VT *dst, int64_t e0, int64_t E, const VT (&v)[N], VT *st)
{
    constexpr int W = 16 / (int)sizeof(VT);
    const int t = threadIdx.x;
#pragma unroll
    for (int k = 0; k < N; k++) st[t * (N + 1) + k] = v[k];
    __syncthreads();
    const int nv = (int)(E - e0 < 256 ? E - e0 : 256) * N;
    VT *d = dst + e0 * N;
    if constexpr (N % W == 0) {
        for (int i = t * W; i < nv; i += 256 * W) {
            VT q[W];
#pragma unroll
            for (int w = 0; w < W; w++) q[w] = st[((i + w) / N) * (N + 1) + (i + w) % N];
            uint4 u;
            __builtin_memcpy(&u, q, 16);
            *(uint4 *)(d + i) = u;
        }
    } else {
        for (int i = t; i < nv; i += 256) d[i] = st[(i / N) * (N + 1) + i % N];
    }
    __syncthreads();
}

__device__ __forceinline__ uint64_t coo_key(int32_t r, int32_t col, int minor_bits, int csc)
{
    const uint64_t mj = (uint64_t)(csc ? col : r), mn = (uint64_t)(csc ? r : col);
    return (mj << minor_bits) | mn;
}

// ---- heat / mass on a nodal family --------------------------------------------
// heat: ke[i,j] += dot(gradN_i, gradN_j) * (kappa JxW), fe[j] += N_j Q JxW
//       (examples/heat/poisson/t4.jl:33-55, the loop of test/test_heat.jl:36-60);
// mass: me[i,j] += (N_i N_j) * (kappa JxW) (examples/klein_gordon/klein_gordon.jl:34-51).
//
// VT double: the FP64 kernel (the reference's operation order).  VT float:
// the REUSE method's MIXED kernel -- FP64 geometry (J, det J, one reciprocal
// of det J in place of the divisions), FP32 gradients, products and sums over
// the points, FP32 stored values.  keys / fkeys null: values only (the REUSE
// method's numeric phase, whose sort order is cached).
template <int KIND, int M, int NB, int NQ, typename VT>
__global__ __launch_bounds__(256) void k_fam_scalar(int64_t E, const int32_t *__restrict__ conn,
                                                    const double *__restrict__ x, const int32_t *__restrict__ dof,
                                                    FamTab tab, double kappa, double Q, int minor_bits, int csc,
                                                    uint64_t *__restrict__ keys, VT *__restrict__ vals,
                                                    uint64_t *__restrict__ fkeys, VT *__restrict__ fvals)
{
    constexpr bool F32 = sizeof(VT) == 4;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    // past the end: a copy of the last element, computed and not stored (the
    // workgroup's stores may go through LDS together)
    const bool valid = e < E;
    const int64_t ec = valid ? e : E - 1;
    double X[NB * M];
    int32_t d[NB];
    const bool want_keys = keys || fkeys;
#pragma unroll
    for (int k = 0; k < NB; k++) {
        const int32_t n = conn[ec * NB + k];
#pragma unroll
        for (int s = 0; s < M; s++) X[k * M + s] = x[(int64_t)n * M + s];
        d[k] = want_keys ? dof[n] : 0;
    }
    VT ke[NB * NB], fe[NB];
#pragma unroll
    for (int i = 0; i < NB * NB; i++) ke[i] = VT(0);
#pragma unroll
    for (int i = 0; i < NB; i++) fe[i] = VT(0);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        Jac<M, NB> J;
        J.make(X, tab.Gg[q]);
        const double JxW = J.det * tab.w[q];
        const VT kj = (VT)(kappa * JxW);
        VT g[NB][M];
        if constexpr (KIND == ELFEL_HEAT) {
            if constexpr (F32) {
                const double r = J.rdet();
#pragma unroll
                for (int k = 0; k < NB; k++) {
                    double gd[M];
                    J.grad_r(tab.Gu[q][k], gd, r);
#pragma unroll
                    for (int c = 0; c < M; c++) g[k][c] = (VT)gd[c];
                }
            } else {
#pragma unroll
                for (int k = 0; k < NB; k++) J.grad(tab.Gu[q][k], g[k]);
            }
        }
#pragma unroll
        for (int j = 0; j < NB; j++) {
#pragma unroll
            for (int i = 0; i < NB; i++) {
                if constexpr (KIND == ELFEL_HEAT) {
                    VT s = g[i][0] * g[j][0];
#pragma unroll
                    for (int c = 1; c < M; c++) s += g[i][c] * g[j][c];
                    ke[i + NB * j] += s * kj;
                } else {
                    ke[i + NB * j] += (VT)(tab.Nu[q][i] * tab.Nu[q][j]) * kj;
                }
            }
            if constexpr (KIND == ELFEL_HEAT) {
                if constexpr (F32) fe[j] += (VT)(tab.Nu[q][j] * Q) * (VT)JxW;
                else fe[j] += tab.Nu[q][j] * Q * JxW;
            }
        }
    }
    if (keys && valid) {
        int64_t o = e * NB * NB;
#pragma unroll
        for (int j = 0; j < NB; j++)
#pragma unroll
            for (int i = 0; i < NB; i++) keys[o++] = coo_key(d[i], d[j], minor_bits, csc);
    }
    if constexpr (KIND == ELFEL_HEAT) {
        if (fkeys && valid) {
#pragma unroll
            for (int j = 0; j < NB; j++) fkeys[e * NB + j] = (uint64_t)d[j];
        }
    }
    // values through LDS where the stage takes at most 40 KiB (heat T4 64^3
    // element kernel 51 -> 40 us); the larger blocks (FP64 T6 mass, 76 KiB)
    // keep direct stores and their occupancy
    constexpr bool STAGE = (size_t)256 * (NB * NB + 1) * sizeof(VT) <= 40 * 1024;
    if constexpr (STAGE) {
        __shared__ VT stage[256 * (NB * NB + 1)];
        const int64_t e0 = (int64_t)blockIdx.x * blockDim.x;
        staged_store<NB * NB>(vals, e0, E, ke, stage);
        if constexpr (KIND == ELFEL_HEAT) staged_store<NB>(fvals, e0, E, fe, stage);
    } else if (valid) {
        store_block<NB * NB>(vals + e * NB * NB, ke);
        if constexpr (KIND == ELFEL_HEAT) store_block<NB>(fvals + e * NB, fe);
    }
}

// ---- Reddy-form Stokes over three scalar spaces ---------------------------------
// test/test_stokes.jl:373-431 (T6, T3), examples/stokes/colliding_flow/
// p1b_p1.jl:50-108 (T3 + bubble, T3), q1_q0.jl:52-110 (Q4, FEL2_Q4):
//   kxx += (mu JxW)(2 gx_i gx_j + gy_i gy_j), kyy += (mu JxW)(gx_i gx_j + 2 gy_i gy_j),
//   kxy += (mu JxW)(gx_i gy_j), kxp += (-JxW)(gx_i Np_j), kyp += (-JxW)(gy_i Np_j).
// The Jacobian comes from the T3 pressure element (jacjac(pel, pqp)) or, for
// the cell-constant pressure, from the velocity Q4 (jacjac(uxel, uxqp)).
// Appended kxx, kxy, kxy', kyy, kxp, kxp', kyp, kyp': column-major, a
// transpose row-major of its parent.  The spatial gradients and JxW of every
// point are formed once; each block is then summed over the points in order
// and written before the next, so that one block at a time is live.
template <int VFE, int PFE, int NQ, typename VT>
__global__ __launch_bounds__(256) void k_fam_reddy(int64_t E, const int32_t *__restrict__ conn_v,
                                                   const double2 *__restrict__ xy_v,
                                                   const int32_t *__restrict__ dof_v,
                                                   const int32_t *__restrict__ edof_v,
                                                   const int32_t *__restrict__ conn_p,
                                                   const double2 *__restrict__ xy_p,
                                                   const int32_t *__restrict__ dof_p,
                                                   const int32_t *__restrict__ edof_p, FamTab tab, double mu,
                                                   int minor_bits, int csc, uint64_t *__restrict__ keys,
                                                   VT *__restrict__ vals)
{
    constexpr bool F32 = sizeof(VT) == 4;  // the MIXED kernel (as k_fam_scalar)
    constexpr int NU = fam_nbf(VFE), NV = fam_ngeo(VFE), NP = fam_nbf(PFE);
    constexpr bool PNODAL = PFE == ELFEL_T3;
    constexpr int NG = PNODAL ? 3 : NV;
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    double X[2 * NG];
#pragma unroll
    for (int k = 0; k < NG; k++) {
        const double2 p = PNODAL ? xy_p[conn_p[e * 3 + k]] : xy_v[conn_v[e * NV + k]];
        X[2 * k] = p.x; X[2 * k + 1] = p.y;
    }
    // per point: the gradients, mu JxW and -JxW in VT (FP64: as formed)
    VT gx[NQ][NU], gy[NQ][NU], mj[NQ], nj[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        Jac<2, NG> J;
        J.make(X, tab.Gg[q]);
        const double JxW = J.det * tab.w[q];
        mj[q] = (VT)(mu * JxW);
        nj[q] = (VT)(-JxW);
        const double r = F32 ? J.rdet() : 0.0;
#pragma unroll
        for (int k = 0; k < NU; k++) {
            double g[2];
            if constexpr (F32) J.grad_r(tab.Gu[q][k], g, r);
            else J.grad(tab.Gu[q][k], g);
            gx[q][k] = (VT)g[0]; gy[q][k] = (VT)g[1];
        }
    }
    // dof numbers (for the keys only)
    int32_t dx[NU], dy[NU], dp[NP];
#pragma unroll
    for (int k = 0; k < NU; k++) dx[k] = dy[k] = 0;
#pragma unroll
    for (int k = 0; k < NP; k++) dp[k] = 0;
    if (keys) {
#pragma unroll
        for (int k = 0; k < NV; k++) {
            const int32_t n = conn_v[e * NV + k];
            dx[k] = dof_v[2 * (int64_t)n];
            dy[k] = dof_v[2 * (int64_t)n + 1];
        }
        if constexpr (NU > NV) {  // the bubble's cell dofs
            dx[NV] = edof_v[2 * e];
            dy[NV] = edof_v[2 * e + 1];
        }
        if constexpr (PNODAL) {
#pragma unroll
            for (int k = 0; k < 3; k++) dp[k] = dof_p[conn_p[e * 3 + k]];
        } else {
            dp[0] = edof_p[e];
        }
    }
    // values in stream order, W at a time through a register word (cnt is a
    // constant at every call once the loops are unrolled)
    constexpr int NOUT = 4 * NU * NU + 4 * NU * NP, W = 16 / (int)sizeof(VT);
    static_assert(NOUT % W == 0, "a Reddy element's values fill whole 16-byte words");
    const int64_t o0 = e * NOUT;
    VT *vout = vals + o0;
    VT wb[W];
    int cnt = 0;
    auto put = [&](int32_t r, int32_t col, VT v) {
        if (keys) keys[o0 + cnt] = coo_key(r, col, minor_bits, csc);
        wb[cnt % W] = v;
        if (cnt % W == W - 1) {
            uint4 q;
            __builtin_memcpy(&q, wb, 16);
            *(uint4 *)(vout + cnt - (W - 1)) = q;
        }
        cnt++;
    };
    VT k[NU * NU];
    // kxx
#pragma unroll
    for (int i = 0; i < NU * NU; i++) k[i] = VT(0);
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
        for (int j = 0; j < NU; j++)
#pragma unroll
            for (int i = 0; i < NU; i++)
                k[i + NU * j] += mj[q] * (2 * gx[q][i] * gx[q][j] + gy[q][i] * gy[q][j]);
#pragma unroll
    for (int j = 0; j < NU; j++)
#pragma unroll
        for (int i = 0; i < NU; i++) put(dx[i], dx[j], k[i + NU * j]);
    // kxy and its transpose
#pragma unroll
    for (int i = 0; i < NU * NU; i++) k[i] = VT(0);
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
        for (int j = 0; j < NU; j++)
#pragma unroll
            for (int i = 0; i < NU; i++) k[i + NU * j] += mj[q] * (gx[q][i] * gy[q][j]);
#pragma unroll
    for (int j = 0; j < NU; j++)
#pragma unroll
        for (int i = 0; i < NU; i++) put(dx[i], dy[j], k[i + NU * j]);
#pragma unroll
    for (int b = 0; b < NU; b++)
#pragma unroll
        for (int a = 0; a < NU; a++) put(dy[a], dx[b], k[b + NU * a]);
    // kyy
#pragma unroll
    for (int i = 0; i < NU * NU; i++) k[i] = VT(0);
#pragma unroll
    for (int q = 0; q < NQ; q++)
#pragma unroll
        for (int j = 0; j < NU; j++)
#pragma unroll
            for (int i = 0; i < NU; i++)
                k[i + NU * j] += mj[q] * (gx[q][i] * gx[q][j] + 2 * gy[q][i] * gy[q][j]);
#pragma unroll
    for (int j = 0; j < NU; j++)
#pragma unroll
        for (int i = 0; i < NU; i++) put(dy[i], dy[j], k[i + NU * j]);
    // kxp, kyp and their transposes
#pragma unroll
    for (int c = 0; c < 2; c++) {
#pragma unroll
        for (int i = 0; i < NU * NP; i++) k[i] = VT(0);
#pragma unroll
        for (int q = 0; q < NQ; q++)
#pragma unroll
            for (int j = 0; j < NP; j++)
#pragma unroll
                for (int i = 0; i < NU; i++)
                    k[i + NU * j] += nj[q] * ((c == 0 ? gx[q][i] : gy[q][i]) * (VT)tab.Np[q][j]);
#pragma unroll
        for (int j = 0; j < NP; j++)
#pragma unroll
            for (int i = 0; i < NU; i++) put(c == 0 ? dx[i] : dy[i], dp[j], k[i + NU * j]);
#pragma unroll
        for (int b = 0; b < NU; b++)
#pragma unroll
            for (int a = 0; a < NP; a++) put(dp[a], c == 0 ? dx[b] : dy[b], k[b + NU * a]);
    }
}

bool reddy_pair(int vfe, int pfe)
{
    return (vfe == ELFEL_T6 && pfe == ELFEL_T3) || (vfe == ELFEL_T3B && pfe == ELFEL_T3) ||
           (vfe == ELFEL_Q4 && pfe == ELFEL_Q4L2);
}

// The element kernel's outputs: keys (or null) and values of the COO stream
// and, for heat, of the load-vector stream.
using FamOut = StreamOut;

template <int KIND, int M, int NB, int NQ, typename VT>
int launch_scalar(elfel_hip_ctx *c, const FamTab &tab, double kappa, double Q, int minor_bits, int csc,
                  const FamOut &o)
{
    const Field &F = c->f[0];
    const bool heat = KIND == ELFEL_HEAT;
    hipEvent_t ev = nullptr;
    tm_begin(c, ELFEL_PH_ELEMENT, &ev);
    if (F.E > 0)
        hipLaunchKernelGGL((k_fam_scalar<KIND, M, NB, NQ, VT>), dim3(ceil_div(F.E, 256)), dim3(256), 0, c->stream, F.E,
                           (const int32_t *)F.conn.p, (const double *)F.xy.p, (const int32_t *)F.dof.p, tab, kappa, Q,
                           minor_bits, csc, o.keys, (VT *)o.vals, heat ? o.fkeys : nullptr,
                           heat ? (VT *)o.fvals : nullptr);
    HIPCHK(c, hipGetLastError());
    tm_end(c, ELFEL_PH_ELEMENT, ev);
    return 0;
}

template <int VFE, int PFE, int NQ, typename VT>
int launch_reddy(elfel_hip_ctx *c, const FamTab &tab, double mu, int minor_bits, int csc, const FamOut &o)
{
    const Field &V = c->f[0], &P = c->f[1];
    hipEvent_t ev = nullptr;
    tm_begin(c, ELFEL_PH_ELEMENT, &ev);
    if (V.E > 0)
        hipLaunchKernelGGL((k_fam_reddy<VFE, PFE, NQ, VT>), dim3(ceil_div(V.E, 256)), dim3(256), 0, c->stream, V.E,
                           (const int32_t *)V.conn.p, (const double2 *)V.xy.p, (const int32_t *)V.dof.p,
                           VFE == ELFEL_T3B ? (const int32_t *)V.edof.p : nullptr,
                           PFE == ELFEL_T3 ? (const int32_t *)P.conn.p : nullptr,
                           PFE == ELFEL_T3 ? (const double2 *)P.xy.p : nullptr,
                           PFE == ELFEL_T3 ? (const int32_t *)P.dof.p : nullptr,
                           PFE == ELFEL_Q4L2 ? (const int32_t *)P.edof.p : nullptr, tab, mu, minor_bits, csc,
                           o.keys, (VT *)o.vals);
    HIPCHK(c, hipGetLastError());
    tm_end(c, ELFEL_PH_ELEMENT, ev);
    return 0;
}

// ---- the REUSE method on the families ------------------------------------------
// The plain method's stable sort orders the COO stream by (major, minor) and
// keeps stream order among equal keys; each matrix entry is then the sum of its
// run, left to right.  The order depends on the mesh and the dof map only, so
// the symbolic phase sorts the stream once with each entry's stream position
// as payload; every assembly after it runs the element kernel (values only, in
// stream order) and sums each run through the cached permutation, in the same
// order: FP64 bit for bit the plain method's result.  (Storing each value at
// its sorted position instead, so that the sums read contiguous runs, measured
// slower: the scattered 4-byte stores took heat T4 64^3's element kernel from
// 0.14 to 0.38 ms.)
__global__ void k_fam_iota(uint64_t *__restrict__ v, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) v[i] = (uint64_t)i;
}

__global__ void k_fam_narrow(const uint64_t *__restrict__ src, int32_t *__restrict__ dst, int64_t n)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) dst[i] = (int32_t)src[i];
}

// run starts of the matrix entries: the segment heads, then n
__global__ void k_fam_seg(const uint32_t *__restrict__ hpos, int64_t nnz, int64_t n, int32_t *__restrict__ seg)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i <= nnz) seg[i] = i < nnz ? (int32_t)hpos[i] : (int32_t)n;
}

// run start of every dof in the sorted load-vector keys (empty runs for dofs
// no element touches): beg[d] = first sorted position with key >= d
__global__ void k_fam_dense_begin(const uint64_t *__restrict__ keys, int64_t nv, int64_t ndofs,
                                  int32_t *__restrict__ beg)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= nv) return;
    const int64_t k = (int64_t)keys[i];
    const int64_t kp = i == 0 ? -1 : (int64_t)keys[i - 1];
    for (int64_t d = kp + 1; d <= k; d++) beg[d] = (int32_t)i;
    if (i == nv - 1)
        for (int64_t d = k + 1; d <= ndofs; d++) beg[d] = (int32_t)nv;
}

__global__ void k_fam_fill_i32(int32_t *__restrict__ p, int64_t n, int32_t v)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

// Work-item s < nnz: matrix entry s, its run summed left to right from the
// first entry (sparse()'s combination of duplicates), in FP64.  Work-item
// nnz + d (d < ndofs, heat): load-vector entry d the same way (start! zeros
// it, assemble! adds: src/Assemblers.jl:194-221).
//
// The run sorted positions i0 .. i1 - 1 (i0 < i1), summed left to right from
// the first: its stream positions, then its values, RUNB at a time, so that a
// run's loads are in flight together instead of one dependent pair per entry.
template <int RUNB, typename VT>
__device__ __forceinline__ double run_sum(const int32_t *__restrict__ perm, const VT *__restrict__ v, int32_t i0,
                                          int32_t i1)
{
    double acc = 0.0;
    for (int32_t b = i0; b < i1; b += RUNB) {
        int32_t p[RUNB];
        VT x[RUNB];
#pragma unroll
        for (int u = 0; u < RUNB; u++) p[u] = b + u < i1 ? perm[b + u] : 0;
#pragma unroll
        for (int u = 0; u < RUNB; u++) x[u] = b + u < i1 ? v[p[u]] : VT(0);
#pragma unroll
        for (int u = 0; u < RUNB; u++)
            if (b + u < i1) acc = (b + u == i0) ? (double)x[u] : acc + (double)x[u];
    }
    return acc;
}

// Workgroup b < nbm: matrix entries 256 b .. 256 b + 255; workgroup nbm + b:
// load-vector entries likewise.  The runs of a workgroup's entries are one
// contiguous range of sorted positions: when it holds at most 16 KiB of
// values, the workgroup first gathers them into LDS, one sorted position per
// work-item (the permutation read in whole lines, every lane's gather useful),
// and each work-item then sums its run from LDS; a longer range gathers run by
// run (run_sum).  Gathering run by run everywhere left most lanes idle in the
// waves that hold a long run (a heat T4 diagonal entry sums 24 values, an
// off-diagonal one 1 to 6): heat T4 64^3 0.20 ms.  (Contiguous eighths of the
// entries per XCD, so that neighbouring rows share one XCD's L2, measured
// slower still: 0.48 ms.)
template <typename VT>
__global__ __launch_bounds__(256) void k_fam_runs(const int32_t *__restrict__ perm, const int32_t *__restrict__ seg,
                                                  const VT *__restrict__ coo, int64_t nnz, double *__restrict__ val,
                                                  const int32_t *__restrict__ fperm, const int32_t *__restrict__ fbeg,
                                                  const VT *__restrict__ fcoo, int64_t ndofs, double *__restrict__ F,
                                                  int64_t nbm)
{
    constexpr int RUNS_CAP = 16384 / (int)sizeof(VT);  // 32 KiB of FP64 values: P1b/P1 FP64 0.72 against 0.59 ms
    __shared__ VT sv[RUNS_CAP];
    const bool mat = blockIdx.x < nbm;
    const int32_t *__restrict__ pm = mat ? perm : fperm;
    const int32_t *__restrict__ sg = mat ? seg : fbeg;
    const VT *__restrict__ v = mat ? coo : fcoo;
    const int64_t cnt = mat ? nnz : ndofs;
    const int64_t e0 = (mat ? (int64_t)blockIdx.x : (int64_t)blockIdx.x - nbm) * 256;
    const int64_t e1 = e0 + 256 < cnt ? e0 + 256 : cnt;
    const int32_t p0 = sg[e0];
    const int32_t np = sg[e1] - p0;
    const bool staged = np <= RUNS_CAP;  // the same in the whole workgroup
    if (staged) {
        for (int32_t k = threadIdx.x; k < np; k += 8 * 256) {
            int32_t q[8];
            VT x[8];
#pragma unroll
            for (int u = 0; u < 8; u++) q[u] = k + u * 256 < np ? pm[p0 + k + u * 256] : 0;
#pragma unroll
            for (int u = 0; u < 8; u++) x[u] = k + u * 256 < np ? v[q[u]] : VT(0);
#pragma unroll
            for (int u = 0; u < 8; u++)
                if (k + u * 256 < np) sv[k + u * 256] = x[u];
        }
        __syncthreads();
    }
    const int64_t s = e0 + threadIdx.x;
    if (s >= e1) return;
    const int32_t i0 = sg[s], i1 = sg[s + 1];
    double acc = 0.0;
    if (i0 < i1) {
        if (staged) {
            acc = (double)sv[i0 - p0];
            for (int32_t i = i0 + 1; i < i1; i++) acc += (double)sv[i - p0];
        } else {
            acc = run_sum<8>(pm, v, i0, i1);
        }
    }
    if (mat) __builtin_nontemporal_store(acc, &val[s]);
    else __builtin_nontemporal_store(i0 == i1 ? 0.0 : 0.0 + acc, &F[s]);  // start! zeros the vector, assemble! adds
}

// The element kernel of the problem (its weak form, element and rule) with
// FP64 or FP32 stored values.
int fam_elements(elfel_hip_ctx *c, const elfel_hip_problem *pb, const FamTab &tab, int minor_bits, int csc,
                 const FamOut &o, bool f32)
{
    const int fe = pb->fe;
    const bool heat = pb->kind == ELFEL_HEAT;
    const bool reddy = pb->kind == ELFEL_STOKES_REDDY;
    const int nq = tab.nq;
    int rc = 1;  // stays 1 when no kernel instance fits
#define S(K_, M_, NB_, NQ_)                                                                          \
    rc = f32 ? launch_scalar<K_, M_, NB_, NQ_, float>(c, tab, pb->kappa, pb->Q, minor_bits, csc, o) \
             : launch_scalar<K_, M_, NB_, NQ_, double>(c, tab, pb->kappa, pb->Q, minor_bits, csc, o)
#define R(V_, P_, NQ_)                                                                 \
    rc = f32 ? launch_reddy<V_, P_, NQ_, float>(c, tab, pb->mu, minor_bits, csc, o) \
             : launch_reddy<V_, P_, NQ_, double>(c, tab, pb->mu, minor_bits, csc, o)
    if (heat || pb->kind == ELFEL_MASS) {
        const bool h = heat;
        if (fe == ELFEL_L2) {
            if (nq == 1) { if (h) S(ELFEL_HEAT, 1, 2, 1); else S(ELFEL_MASS, 1, 2, 1); }
            else if (nq == 2) { if (h) S(ELFEL_HEAT, 1, 2, 2); else S(ELFEL_MASS, 1, 2, 2); }
            else if (nq == 3) { if (h) S(ELFEL_HEAT, 1, 2, 3); else S(ELFEL_MASS, 1, 2, 3); }
        } else if (fe == ELFEL_T4) {
            if (nq == 1) { if (h) S(ELFEL_HEAT, 3, 4, 1); else S(ELFEL_MASS, 3, 4, 1); }
            else if (nq == 4) { if (h) S(ELFEL_HEAT, 3, 4, 4); else S(ELFEL_MASS, 3, 4, 4); }
        } else if (!h && fe == ELFEL_T3) {
            if (nq == 1) S(ELFEL_MASS, 2, 3, 1); else if (nq == 3) S(ELFEL_MASS, 2, 3, 3);
        } else if (!h && fe == ELFEL_T6) {
            if (nq == 1) S(ELFEL_MASS, 2, 6, 1); else if (nq == 3) S(ELFEL_MASS, 2, 6, 3);
        } else if (!h && fe == ELFEL_Q4) {
            if (nq == 1) S(ELFEL_MASS, 2, 4, 1); else if (nq == 4) S(ELFEL_MASS, 2, 4, 4);
        }
    } else if (reddy) {
        if (fe == ELFEL_T6 && nq == 3) R(ELFEL_T6, ELFEL_T3, 3);
        else if (fe == ELFEL_T3B && nq == 3) R(ELFEL_T3B, ELFEL_T3, 3);
        else if (fe == ELFEL_T3B && nq == 1) R(ELFEL_T3B, ELFEL_T3, 1);
        else if (fe == ELFEL_Q4 && nq == 4) R(ELFEL_Q4, ELFEL_Q4L2, 4);
        else if (fe == ELFEL_Q4 && nq == 1) R(ELFEL_Q4, ELFEL_Q4L2, 1);
    }
#undef S
#undef R
    if (rc == 1) return set_err(c, ELFEL_ENOTSUP, "unsupported (weak form, element, rule)");
    return rc;
}

#define FAM_LAUNCH(kern, n, ...)                                                                           \
    do {                                                                                                   \
        if ((n) > 0) {                                                                                     \
            hipLaunchKernelGGL(kern, dim3(ceil_div((n), 256)), dim3(256), 0, c->stream, __VA_ARGS__);      \
            HIPCHK(c, hipGetLastError());                                                                  \
        }                                                                                                  \
    } while (0)

// Symbolic phase of the stream REUSE method (c->fpat), when the cached one
// does not describe this mesh, dof map and weak form.
int stream_symbolic(elfel_hip_ctx *c, const elfel_hip_problem *pb, int64_t n, int64_t nv, int64_t ndofs,
                    int minor_bits, const StreamElements &elements)
{
    FamPattern &P = c->fpat;
    const bool heat = pb->kind == ELFEL_HEAT;
    const Field &F0 = c->f[0], &F1 = c->f[1];
    const bool reddy = pb->kind == ELFEL_STOKES_REDDY;
    const bool two = reddy || pb->kind == ELFEL_STOKES_GEN || pb->kind == ELFEL_STOKES_VECLAP;
    const int pfe = reddy ? pb->fe_p : -1;
    const uint64_t key[6] = {F0.mesh_version, F0.dof_version, F0.edof_version, two ? F1.mesh_version : 0,
                             two ? F1.dof_version : 0, two ? F1.edof_version : 0};
    if (P.valid && P.kind == pb->kind && P.fe == pb->fe && P.fe_p == pfe && P.layout == pb->layout && P.n == n &&
        P.nv == nv && P.ndofs == ndofs && memcmp(P.key, key, sizeof key) == 0)
        return 0;
    const auto t0 = std::chrono::steady_clock::now();
    P.valid = false;
    int rc;
    if ((rc = plain_buffers(c, n, nv, ndofs, heat)) || (rc = ensure(c, P.coo, 8 * std::max<int64_t>(n, 1))) ||
        (rc = ensure(c, P.fcoo, 8 * std::max<int64_t>(nv, 1))) || (rc = ensure(c, P.perm, 4 * std::max<int64_t>(n, 1))) ||
        (rc = ensure(c, P.fperm, 4 * std::max<int64_t>(nv, 1))) || (rc = ensure(c, P.fbeg, 4 * (ndofs + 1))) ||
        (rc = ensure(c, P.ptr, 4 * (ndofs + 1))) || (rc = ensure(c, P.ind, 4 * std::max<int64_t>(n, 1))))
        return rc;
    uint64_t *k0 = (uint64_t *)c->keys[0].p, *k1 = (uint64_t *)c->keys[1].p;
    uint64_t *v0 = (uint64_t *)c->vals[0].p, *v1 = (uint64_t *)c->vals[1].p;
    uint64_t *fk0 = (uint64_t *)c->fkeys[0].p, *fk1 = (uint64_t *)c->fkeys[1].p;
    uint64_t *fv0 = (uint64_t *)c->fvals[0].p, *fv1 = (uint64_t *)c->fvals[1].p;
    // the stream's keys (its values land in coo, unused here)
    const StreamOut o{k0, heat ? fk0 : nullptr, P.coo.p, heat ? P.fcoo.p : nullptr};
    if ((rc = elements(o, false))) return rc;
    hipEvent_t ev = nullptr;
    tm_begin(c, ELFEL_PH_SORT, &ev);
    FAM_LAUNCH(k_fam_iota, n, v0, n);
    if ((rc = radix_sort_pairs(c, k0, v0, k1, v1, n, bits_for(ndofs) + minor_bits))) return rc;
    if (heat) {
        FAM_LAUNCH(k_fam_iota, nv, fv0, nv);
        if ((rc = radix_sort_pairs(c, fk0, fv0, fk1, fv1, nv, bits_for(ndofs)))) return rc;
    }
    tm_end(c, ELFEL_PH_SORT, ev);
    if ((rc = build_csr_from_sorted(c, k0, nullptr, n, minor_bits, ndofs, (int32_t *)P.ptr.p, (int32_t *)P.ind.p,
                                    nullptr, (int32_t *)c->dscal.p, false)))
        return rc;
    int32_t hnnz = 0;
    HIPCHK(c, hipMemcpyAsync(&hnnz, c->dscal.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t nnz = hnnz;
    if ((rc = ensure(c, P.seg, 4 * (nnz + 1)))) return rc;
    FAM_LAUNCH(k_fam_narrow, n, v0, (int32_t *)P.perm.p, n);
    if (n > 0) FAM_LAUNCH(k_fam_seg, nnz + 1, (const uint32_t *)c->segpos.p, nnz, n, (int32_t *)P.seg.p);
    else FAM_LAUNCH(k_fam_fill_i32, (int64_t)1, (int32_t *)P.seg.p, (int64_t)1, 0);
    if (heat) {
        FAM_LAUNCH(k_fam_narrow, nv, fv0, (int32_t *)P.fperm.p, nv);
        if (nv > 0) FAM_LAUNCH(k_fam_dense_begin, nv, fk0, nv, ndofs, (int32_t *)P.fbeg.p);
        else FAM_LAUNCH(k_fam_fill_i32, ndofs + 1, (int32_t *)P.fbeg.p, ndofs + 1, 0);
    }
    HIPCHK(c, hipStreamSynchronize(c->stream));
    P.kind = pb->kind; P.fe = pb->fe; P.fe_p = pfe; P.layout = pb->layout;
    memcpy(P.key, key, sizeof key);
    P.n = n; P.nv = nv; P.nnz = nnz; P.ndofs = ndofs;
    P.valid = true;
    P.symbolic_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

}  // namespace

// Numeric phase: element values only, then every run summed through the
// cached permutation.
int stream_reuse(elfel_hip_ctx *c, const elfel_hip_problem *pb, int64_t n, int64_t nv, int64_t ndofs, int minor_bits,
                 bool f32, const StreamElements &elements)
{
    const bool heat = pb->kind == ELFEL_HEAT;
    int rc = stream_symbolic(c, pb, n, nv, ndofs, minor_bits, elements);
    if (rc) return rc;
    const FamPattern &P = c->fpat;
    if ((rc = ensure(c, c->val, 8 * std::max<int64_t>(P.nnz, 1)))) return rc;
    if (heat && (rc = ensure(c, c->F, 8 * std::max<int64_t>(ndofs, 1)))) return rc;
    const StreamOut o{nullptr, nullptr, P.coo.p, heat ? P.fcoo.p : nullptr};
    if ((rc = elements(o, f32))) return rc;
    hipEvent_t ev = nullptr;
    tm_begin(c, ELFEL_PH_REDUCE, &ev);
    const int64_t nf = heat ? ndofs : 0;
    const int64_t nbm = ceil_div(P.nnz, 256), nbw = nbm + ceil_div(nf, 256);
    if (nbw > 0) {
        if (f32)
            hipLaunchKernelGGL((k_fam_runs<float>), dim3(nbw), dim3(256), 0, c->stream, (const int32_t *)P.perm.p,
                               (const int32_t *)P.seg.p, (const float *)P.coo.p, P.nnz, (double *)c->val.p,
                               (const int32_t *)P.fperm.p, (const int32_t *)P.fbeg.p, (const float *)P.fcoo.p, nf,
                               heat ? (double *)c->F.p : nullptr, nbm);
        else
            hipLaunchKernelGGL((k_fam_runs<double>), dim3(nbw), dim3(256), 0, c->stream, (const int32_t *)P.perm.p,
                               (const int32_t *)P.seg.p, (const double *)P.coo.p, P.nnz, (double *)c->val.p,
                               (const int32_t *)P.fperm.p, (const int32_t *)P.fbeg.p, (const double *)P.fcoo.p, nf,
                               heat ? (double *)c->F.p : nullptr, nbm);
        HIPCHK(c, hipGetLastError());
    }
    tm_end(c, ELFEL_PH_REDUCE, ev);
    c->nnz = P.nnz;
    c->nmaj = c->nmin = ndofs;
    c->layout = pb->layout;
    c->result_src = RES_FAMILY;
    c->has_F = heat;
    c->nF = heat ? ndofs : 0;
    return 0;
}

#undef FAM_LAUNCH

bool family_problem(const elfel_hip_problem *pb)
{
    return pb->kind == ELFEL_STOKES_REDDY || pb->kind == ELFEL_MASS || pb->fe == ELFEL_T3B || pb->fe == ELFEL_Q4L2 ||
           pb->fe == ELFEL_T4 || pb->fe == ELFEL_L2;
}

int family_validate(elfel_hip_ctx *c, const elfel_hip_problem *pb)
{
    if (pb->method != ELFEL_SORT && pb->method != ELFEL_REUSE) return set_err(c, ELFEL_EINVAL, "bad method");
    if (pb->precision != ELFEL_FP64 && pb->precision != ELFEL_MIXED) return set_err(c, ELFEL_EINVAL, "bad precision");
    if (pb->method == ELFEL_SORT && pb->precision != ELFEL_FP64)
        return set_err(c, ELFEL_ENOTSUP, "the sort method is the plain FP64 method");
    if (pb->layout != ELFEL_CSR && pb->layout != ELFEL_CSC) return set_err(c, ELFEL_EINVAL, "bad layout");
    if (c->has_kappa_e || c->has_Q_e || c->has_scale_e)
        return set_err(c, ELFEL_ENOTSUP, "element families take uniform coefficients (clear the per-element ones)");
    const Field &F = c->f[0];
    const int fe = pb->fe;
    if (fam_nbf(fe) < 0) return set_err(c, ELFEL_EINVAL, "bad element");
    if (!F.has_mesh) return set_err(c, ELFEL_ESTATE, "set the mesh of field 0 first");
    if (F.nen != fam_ngeo(fe) || F.sdim != fam_mdim(fe))
        return set_err(c, ELFEL_EINVAL, "element type does not match the mesh (nodes per element, space dimension)");
    int64_t ndofs = 0;
    if (pb->kind == ELFEL_HEAT || pb->kind == ELFEL_MASS) {
        if (fe == ELFEL_T3B || fe == ELFEL_Q4L2)
            return set_err(c, ELFEL_ENOTSUP, "heat and mass on the nodal families (T3, T6, Q4, T4, L2)");
        if (!F.has_dofs || F.ncomp != 1) return set_err(c, ELFEL_ESTATE, "a scalar dof map of field 0 first");
        ndofs = F.ndofs;
    } else if (pb->kind == ELFEL_STOKES_REDDY) {
        const Field &P = c->f[1];
        if (!reddy_pair(fe, pb->fe_p))
            return set_err(c, ELFEL_ENOTSUP, "Reddy Stokes pairs: (T6, T3), (T3B, T3), (Q4, Q4L2)");
        if (!F.has_dofs || F.ncomp != 2)
            return set_err(c, ELFEL_ESTATE, "field 0 needs node dofs with 2 components (the ux and uy spaces)");
        if (fe == ELFEL_T3B && (!F.has_edofs || F.edof_ncomp != 2 || F.edof_E != F.E))
            return set_err(c, ELFEL_ESTATE, "the bubble needs cell dofs of field 0 with 2 components");
        if (pb->fe_p == ELFEL_T3) {
            if (!P.has_mesh || P.nen != 3 || P.sdim != 2 || P.E != F.E || !P.has_dofs || P.ncomp != 1)
                return set_err(c, ELFEL_ESTATE, "the T3 pressure needs field 1's T3 mesh (same elements) and node dofs");
        } else if (!P.has_edofs || P.edof_ncomp != 1 || P.edof_E != F.E) {
            return set_err(c, ELFEL_ESTATE, "the FEL2_Q4 pressure needs cell dofs of field 1");
        }
        // the system size from the dof maps this pair reads (a map left from an
        // earlier problem on the context does not count)
        ndofs = F.ndofs;
        if (fe == ELFEL_T3B) ndofs = std::max(ndofs, F.edof_ndofs);
        ndofs = std::max(ndofs, pb->fe_p == ELFEL_T3 ? P.ndofs : P.edof_ndofs);
    } else {
        return set_err(c, ELFEL_ENOTSUP, "weak form not available for this element");
    }
    FamTab tab;
    const int pfe = pb->kind == ELFEL_STOKES_REDDY ? pb->fe_p : -1;
    const int gfe = pfe == ELFEL_T3 ? ELFEL_T3 : fe;
    if (fam_table(fe, pfe, gfe, pb->rule, &tab)) return set_err(c, ELFEL_ENOTSUP, "unsupported quadrature rule");
    c->nrow = c->ncol = ndofs;
    return 0;
}

int family_assemble(elfel_hip_ctx *c, const elfel_hip_problem *pb)
{
    const Field &F = c->f[0];
    const int fe = pb->fe;
    const bool reddy = pb->kind == ELFEL_STOKES_REDDY;
    const int pfe = reddy ? pb->fe_p : -1;
    FamTab tab;
    if (fam_table(fe, pfe, pfe == ELFEL_T3 ? ELFEL_T3 : fe, pb->rule, &tab))
        return set_err(c, ELFEL_ENOTSUP, "unsupported quadrature rule");
    const int nu = fam_nbf(fe), np = reddy ? fam_nbf(pfe) : 0;
    const int64_t n = F.E * (reddy ? 4 * nu * nu + 4 * nu * np : nu * nu);
    const bool heat = pb->kind == ELFEL_HEAT;
    const int64_t nv = heat ? F.E * nu : 0;
    const int64_t ndofs = c->nrow;
    const int minor_bits = bits_for(ndofs);
    const int csc = pb->layout == ELFEL_CSC ? 1 : 0;
    // the cached permutation and run starts are int32: a longer stream runs the
    // plain FP64 method (its sort carries 64-bit positions)
    if (pb->method == ELFEL_REUSE && n < INT32_MAX && nv < INT32_MAX)
        return stream_reuse(c, pb, n, nv, ndofs, minor_bits, pb->precision == ELFEL_MIXED,
                            [&](const StreamOut &o, bool f32) { return fam_elements(c, pb, tab, minor_bits, csc, o, f32); });
    if (pb->method == ELFEL_REUSE) {
        c->used_method = ELFEL_SORT;
        c->used_precision = ELFEL_FP64;
        c->used_note = "plain FP64 method: a COO stream of 2^31 entries or more";
    }
    int rc = plain_buffers(c, n, nv, ndofs, heat);
    if (rc) return rc;
    const FamOut o{(uint64_t *)c->keys[0].p, heat ? (uint64_t *)c->fkeys[0].p : nullptr, c->vals[0].p,
                   heat ? c->fvals[0].p : nullptr};
    if ((rc = fam_elements(c, pb, tab, minor_bits, csc, o, false))) return rc;
    return plain_finish(c, pb->layout, n, nv, ndofs, minor_bits, heat);
}
