// elem.h -- per-element arithmetic shared by the FP64 and mixed kernels.
//
// R = double reproduces the reference's arithmetic operation by operation
// (the including translation unit is compiled with -ffp-contract=off):
//   jacjac   src/FElements.jl:148-156,165-168 and :120-129
//   bfungrad src/QPIterators.jl:132-140 (2x2 solve J' \ g by Cramer's rule)
//   heat     test/test_heat.jl:49-54
//   elast    test/test_elasticity.jl:150-166 (B(g,1)=(g1,0,g2), B(g,2)=(0,g2,g1))
//   stokes   examples/stokes/colliding_flow/ht_p2_p1_gen.jl:65-74 (gen),
//            test/test_stokes.jl:645-652 (vector Laplacian)
// Each matrix entry is accumulated over quadrature points in point order
// starting from 0.0, which is the order of `ke[i, j] += ...` in the loops.
// The mixed kernels take their geometry from geometry_mixed(): Jacobian,
// determinant and one reciprocal of det per point in FP64 from coordinates
// relative to the element's first node; the inverse Jacobian and JxW rounded
// to FP32, and the spatial gradients formed from them in FP32.
#pragma once
#include "common.h"

template <typename R> __device__ __forceinline__ R tab_w(const QTable &t, int q);
template <> __device__ __forceinline__ double tab_w<double>(const QTable &t, int q) { return t.w[q]; }
template <> __device__ __forceinline__ float tab_w<float>(const QTable &t, int q) { return t.wf[q]; }
template <typename R> __device__ __forceinline__ R tab_g(const QTable &t, int q, int k, int d);
template <> __device__ __forceinline__ double tab_g<double>(const QTable &t, int q, int k, int d) { return t.G[q][k][d]; }
template <> __device__ __forceinline__ float tab_g<float>(const QTable &t, int q, int k, int d) { return t.Gf[q][k][d]; }
template <typename R> __device__ __forceinline__ R tab_n(const QTable &t, int q, int k);
template <> __device__ __forceinline__ double tab_n<double>(const QTable &t, int q, int k) { return t.N[q][k]; }
template <> __device__ __forceinline__ float tab_n<float>(const QTable &t, int q, int k) { return t.Nf[q][k]; }
template <typename R> __device__ __forceinline__ R tab_np(const QTable &t, int q, int k);
template <> __device__ __forceinline__ double tab_np<double>(const QTable &t, int q, int k) { return t.Np[q][k]; }
template <> __device__ __forceinline__ float tab_np<float>(const QTable &t, int q, int k) { return t.Npf[q][k]; }

template <typename R, int NEN, int NQ>
struct Geo {
    R gx[NQ][NEN], gy[NQ][NEN], jxw[NQ], det[NQ];
};

// Jacobian, determinant, spatial gradients and JxW at every quadrature point.
template <typename R, int NEN, int NQ>
__device__ __forceinline__ void geometry(const QTable &t, const R *X, const R *Y, Geo<R, NEN, NQ> &g)
{
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        R j11 = X[0] * tab_g<R>(t, q, 0, 0), j21 = Y[0] * tab_g<R>(t, q, 0, 0);
        R j12 = X[0] * tab_g<R>(t, q, 0, 1), j22 = Y[0] * tab_g<R>(t, q, 0, 1);
#pragma unroll
        for (int k = 1; k < NEN; k++) {
            j11 += X[k] * tab_g<R>(t, q, k, 0);
            j21 += Y[k] * tab_g<R>(t, q, k, 0);
            j12 += X[k] * tab_g<R>(t, q, k, 1);
            j22 += Y[k] * tab_g<R>(t, q, k, 1);
        }
        const R det = j11 * j22 - j21 * j12;
        g.det[q] = det;
        if constexpr (sizeof(R) == 8) {
#pragma unroll
            for (int k = 0; k < NEN; k++) {
                const R a = tab_g<R>(t, q, k, 0), b = tab_g<R>(t, q, k, 1);
                g.gx[q][k] = (j22 * a - j21 * b) / det;
                g.gy[q][k] = (j11 * b - j12 * a) / det;
            }
        } else {
            const R inv = R(1) / det;
#pragma unroll
            for (int k = 0; k < NEN; k++) {
                const R a = tab_g<R>(t, q, k, 0), b = tab_g<R>(t, q, k, 1);
                g.gx[q][k] = (j22 * a - j21 * b) * inv;
                g.gy[q][k] = (j11 * b - j12 * a) * inv;
            }
        }
        g.jxw[q] = det * tab_w<R>(t, q);
    }
}

// Mixed-precision geometry.  An FP32 Jacobian loses up to ~1e-5 of the matrix
// to the cancellation in sum_k x_k grad N_k on curved T6 elements, so J, det
// and 1/det stay in FP64; the inverse and JxW are rounded to FP32, and the
// gradients formed from them in FP32.  X, Y are relative to
// node 0 (X[0] = Y[0] = 0, so node 0 adds nothing to J) and t is a
// compile-time table (qt_fixed), so the zero entries of grad N drop out.
template <int NEN, int NQ>
__device__ __forceinline__ void geometry_mixed(const QTable &t, const double *X, const double *Y,
                                               Geo<float, NEN, NQ> &g)
{
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        double j11 = 0.0, j21 = 0.0, j12 = 0.0, j22 = 0.0;
#pragma unroll
        for (int k = 1; k < NEN; k++) {
            const double a = t.G[q][k][0], b = t.G[q][k][1];
            if (a != 0.0) { j11 = fma(X[k], a, j11); j21 = fma(Y[k], a, j21); }
            if (b != 0.0) { j12 = fma(X[k], b, j12); j22 = fma(Y[k], b, j22); }
        }
        const double det = j11 * j22 - j21 * j12;
        // 1/det: the hardware reciprocal estimate (v_rcp_f64) refined by two
        // Newton-Raphson steps r <- r + r (1 - det r), which square the
        // relative error each time -- far below the FP32 rounding that follows
        double inv = __builtin_amdgcn_rcp(det);
        inv = fma(inv, fma(-det, inv, 1.0), inv);
        inv = fma(inv, fma(-det, inv, 1.0), inv);
        // J^-T rounded to FP32; the spatial gradients J^-T grad N in FP32 (C2:
        // 5.8e-8 against FP64, a little below the 7.5e-8 of FP64 gradients
        // rounded to FP32, in 63 instead of 65 us)
        const float f11 = (float)(j22 * inv), f12 = (float)(-j21 * inv);
        const float f21 = (float)(-j12 * inv), f22 = (float)(j11 * inv);
        g.det[q] = (float)det;
        g.jxw[q] = (float)(det * t.w[q]);
#pragma unroll
        for (int k = 0; k < NEN; k++) {
            const float a = t.Gf[q][k][0], b = t.Gf[q][k][1];
            g.gx[q][k] = a == 0.f ? b * f12 : (b == 0.f ? a * f11 : fmaf(a, f11, b * f12));
            g.gy[q][k] = b == 0.f ? a * f21 : (a == 0.f ? b * f22 : fmaf(b, f22, a * f21));
        }
    }
}

// B(g, c) . (D B(g', c')) with D column-major, evaluated term by term as
// dot(D * B(g', c'), B(g, c)) including the explicit zero entries of B.
template <typename R>
__device__ __forceinline__ void dmul_b(const R *D, R gx, R gy, int c, R *o)
{
    R b0, b1, b2;
    if (c == 0) { b0 = gx; b1 = R(0); b2 = gy; }
    else { b0 = R(0); b1 = gy; b2 = gx; }
#pragma unroll
    for (int r = 0; r < 3; r++) o[r] = D[r] * b0 + D[r + 3] * b1 + D[r + 6] * b2;
}

template <typename R>
__device__ __forceinline__ R dot_b(const R *db, R gx, R gy, int c)
{
    R b0, b1, b2;
    if (c == 0) { b0 = gx; b1 = R(0); b2 = gy; }
    else { b0 = R(0); b1 = gy; b2 = gx; }
    return db[0] * b0 + db[1] * b1 + db[2] * b2;
}

// Kind-specific entry of the velocity/displacement/temperature block.
// i, j: element dof indices (node = idx / NC, component = idx % NC).
template <int KIND, typename R, int NEN, int NC, int NQ>
struct Entry {
    // Column-j precomputation (D*B_j per quadrature point).
    R db[NQ][3];
    __device__ __forceinline__ void column(const Geo<R, NEN, NQ> &g, const R *D, int j)
    {
        if constexpr (KIND == ELFEL_ELAST || KIND == ELFEL_STOKES_GEN) {
#pragma unroll
            for (int q = 0; q < NQ; q++) dmul_b<R>(D, g.gx[q][j / NC], g.gy[q][j / NC], j % NC, db[q]);
        }
    }
    __device__ __forceinline__ R value(const Geo<R, NEN, NQ> &g, R kap, R mu, int i, int j) const
    {
        R acc = R(0);
        if constexpr (KIND == ELFEL_HEAT) {
#pragma unroll
            for (int q = 0; q < NQ; q++)
                acc += (g.gx[q][i] * g.gx[q][j] + g.gy[q][i] * g.gy[q][j]) * (kap * g.jxw[q]);
        } else if constexpr (KIND == ELFEL_ELAST || KIND == ELFEL_STOKES_GEN) {
#pragma unroll
            for (int q = 0; q < NQ; q++) acc += dot_b<R>(db[q], g.gx[q][i / NC], g.gy[q][i / NC], i % NC) * g.jxw[q];
        } else {
            if ((i % NC) == (j % NC)) {
                const int bi = i / NC, bj = j / NC;
#pragma unroll
                for (int q = 0; q < NQ; q++)
                    acc += (mu * g.jxw[q]) * (g.gx[q][bi] * g.gx[q][bj] + g.gy[q][bi] * g.gy[q][bj]);
            }
        }
        return acc;
    }
};

// Stokes velocity-pressure coupling kup[i, jp] (jp = pressure node 0..2).
template <typename R, int NEN, int NC, int NQ>
__device__ __forceinline__ R kup_value(const QTable &t, const Geo<R, NEN, NQ> &g, int i, int jp)
{
    R acc = R(0);
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        const R gi = (i % NC) == 0 ? g.gx[q][i / NC] : g.gy[q][i / NC];
        acc += (-g.jxw[q] * tab_np<R>(t, q, jp)) * gi;
    }
    return acc;
}

// Heat load vector entry fe[j] = sum_q N_j * Q * JxW (test/test_heat.jl:53).
template <typename R, int NEN, int NQ>
__device__ __forceinline__ R fe_value(const QTable &t, const Geo<R, NEN, NQ> &g, R Qe, int j)
{
    R acc = R(0);
#pragma unroll
    for (int q = 0; q < NQ; q++) acc += tab_n<R>(t, q, j) * Qe * g.jxw[q];
    return acc;
}
