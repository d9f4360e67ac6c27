// em.hip -- element matrices of plane elasticity on T6 (3-point rule), mixed
// arithmetic: the ke blocks LocalMatrixAssembler holds for every element
// (src/LocalAssemblers.jl:68-83; the element loop of test/test_elasticity.jl:
// 150-172), written to device memory as ke[e][i + 12 j] (column-major, the
// order of LocalMatrixAssembler.M), by one of two kernels:
//
//   VALU: one work-item per element; FP64 Jacobian, determinant and inverse,
//         FP32 gradients (geometry_mixed), then ke[i][j] = sum_q JxW
//         B_i . (D B_j) with vector FMAs (B's structural zeros skipped).
//   MFMA: the same contraction on the matrix cores.  ke = G^T H with
//         G[kk][i] = B_q[r][i], H[kk][j] = JxW_q (D B_q)[r][j], kk = (q, r),
//         9 terms.  v_mfma_f32_4x4x1f32 computes 16 independent 4x4 outer
//         products per instruction; a wave holds 16 elements, 4 lanes each
//         (lane 4b + s: element b, row/column s of a 4x4 block), and the
//         12x12 block is 3 x 3 such blocks, so 9 blocks x 9 terms = 81
//         instructions per 16 elements.  Each lane forms the geometry of its
//         element itself (4 lanes per element) and its own G and H entries.
//
// The kernels exist to measure the matrix cores on the one dense contraction
// of the path (B^T D B); the assembly kernels use the vector form.
#include <algorithm>

#include "common.h"
#include "elem.h"

namespace {

typedef float floatx4 __attribute__((ext_vector_type(4)));

// B_q[r][dof] for dof = 2 n + c (c: x or y component of node n)
__device__ __forceinline__ float bmat(float gx, float gy, int r, int c)
{
    return r == 0 ? (c == 0 ? gx : 0.f) : (r == 1 ? (c == 1 ? gy : 0.f) : (c == 0 ? gy : gx));
}

// coordinates of element e relative to its first node (FP64)
__device__ __forceinline__ void rel_coords(int64_t e, const int32_t *conn, const double2 *xy, double *X, double *Y)
{
    int32_t nd[6];
#pragma unroll
    for (int k = 0; k < 6; k++) nd[k] = conn[e * 6 + k];
#pragma unroll
    for (int k = 0; k < 6; k++) { const double2 p = xy[nd[k]]; X[k] = p.x; Y[k] = p.y; }
    const double x0 = X[0], y0 = Y[0];
#pragma unroll
    for (int k = 0; k < 6; k++) { X[k] -= x0; Y[k] -= y0; }
}

__global__ __launch_bounds__(256) void k_ke_valu(int64_t E, const int32_t *__restrict__ conn,
                                                 const double2 *__restrict__ xy, Coef cf, float *__restrict__ ke)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= E) return;
    constexpr QTable CT = qt_fixed<6, 3>();
    double X[6], Y[6];
    rel_coords(e, conn, xy, X, Y);
    Geo<float, 6, 3> g;
    geometry_mixed<6, 3>(CT, X, Y, g);
    float D[9];
#pragma unroll
    for (int i = 0; i < 9; i++) D[i] = (float)cf.D[i];
    float k[12][12];
#pragma unroll
    for (int i = 0; i < 12; i++)
#pragma unroll
        for (int j = 0; j < 12; j++) k[i][j] = 0.f;
#pragma unroll
    for (int q = 0; q < 3; q++) {
#pragma unroll
        for (int j = 0; j < 12; j++) {
            // JxW (D B_j): B_j has two nonzeros
            float db[3];
#pragma unroll
            for (int r = 0; r < 3; r++) {
                float v = 0.f;
#pragma unroll
                for (int t = 0; t < 3; t++) {
                    const float b = bmat(g.gx[q][j / 2], g.gy[q][j / 2], t, j % 2);
                    if (!((t == 0 && j % 2 == 1) || (t == 1 && j % 2 == 0))) v = fmaf(D[r + 3 * t], b, v);
                }
                db[r] = v * g.jxw[q];
            }
#pragma unroll
            for (int i = 0; i < 12; i++) {
                const float gx = g.gx[q][i / 2], gy = g.gy[q][i / 2];
                k[i][j] = (i % 2) == 0 ? fmaf(db[0], gx, fmaf(db[2], gy, k[i][j]))
                                       : fmaf(db[1], gy, fmaf(db[2], gx, k[i][j]));
            }
        }
    }
    // SoA: entry m of every element is contiguous across the wave
#pragma unroll
    for (int j = 0; j < 12; j++)
#pragma unroll
        for (int i = 0; i < 12; i++) ke[(int64_t)(i + 12 * j) * E + e] = k[i][j];
}

__global__ __launch_bounds__(256) void k_ke_mfma(int64_t E, const int32_t *__restrict__ conn,
                                                 const double2 *__restrict__ xy, Coef cf, float *__restrict__ ke)
{
    const int lane = threadIdx.x & 63;
    const int b = lane >> 2, s = lane & 3;  // element b of the wave's 16, slot s of a 4x4 block
    const int64_t e0 = ((int64_t)blockIdx.x * blockDim.x + (threadIdx.x & ~63)) / 4;  // the wave's first element
    const int64_t e = e0 + b;
    const int64_t ec = e < E ? e : E - 1;  // every lane takes part in the matrix instructions
    constexpr QTable CT = qt_fixed<6, 3>();
    double X[6], Y[6];
    rel_coords(ec, conn, xy, X, Y);
    Geo<float, 6, 3> g;
    geometry_mixed<6, 3>(CT, X, Y, g);
    float D[9];
#pragma unroll
    for (int i = 0; i < 9; i++) D[i] = (float)cf.D[i];
    // this lane's G (rows) and H (columns) entries: dofs s, 4 + s, 8 + s
    float G[3][9], H[3][9];
#pragma unroll
    for (int I = 0; I < 3; I++) {
        const int d = 4 * I + s, n = d >> 1, c = d & 1;
#pragma unroll
        for (int q = 0; q < 3; q++) {
            // selects on the lane's node index keep the gradients in registers
            float gx = g.gx[q][0], gy = g.gy[q][0];
#pragma unroll
            for (int m = 1; m < 6; m++) { gx = n == m ? g.gx[q][m] : gx; gy = n == m ? g.gy[q][m] : gy; }
            float bq[3];
#pragma unroll
            for (int r = 0; r < 3; r++) bq[r] = bmat(gx, gy, r, c);
#pragma unroll
            for (int r = 0; r < 3; r++) {
                G[I][3 * q + r] = bq[r];
                H[I][3 * q + r] = (D[r] * bq[0] + D[r + 3] * bq[1] + D[r + 6] * bq[2]) * g.jxw[q];
            }
        }
    }
    floatx4 acc[3][3];
#pragma unroll
    for (int I = 0; I < 3; I++)
#pragma unroll
        for (int J = 0; J < 3; J++) acc[I][J] = (floatx4){0.f, 0.f, 0.f, 0.f};
#pragma unroll
    for (int kk = 0; kk < 9; kk++)
#pragma unroll
        for (int I = 0; I < 3; I++)
#pragma unroll
            for (int J = 0; J < 3; J++)
                acc[I][J] = __builtin_amdgcn_mfma_f32_4x4x1f32(G[I][kk], H[J][kk], acc[I][J], 0, 0, 0);
    // lane 4b + s holds column 4J + s, rows 4I .. 4I + 3 of element b
    if (e < E) {
#pragma unroll
        for (int I = 0; I < 3; I++)
#pragma unroll
            for (int J = 0; J < 3; J++)
#pragma unroll
                for (int r = 0; r < 4; r++) ke[(int64_t)((4 * I + r) + 12 * (4 * J + s)) * E + e] = acc[I][J][r];
    }
}

}  // namespace

extern "C" int elfel_hip_element_matrices(elfel_hip_ctx *c, const elfel_hip_problem *pb, int method, float *ke,
                                          int64_t n, double *ms)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        if (!pb || (method != ELFEL_EM_VALU && method != ELFEL_EM_MFMA))
            return set_err(c, ELFEL_EINVAL, "elfel_hip_element_matrices: bad arguments");
        if (pb->kind != ELFEL_ELAST || pb->fe != ELFEL_T6 || pb->rule != 3 || pb->precision != ELFEL_MIXED)
            return set_err(c, ELFEL_ENOTSUP, "element matrices: plane elasticity on T6, 3-point rule, mixed arithmetic");
        const Field &F = c->f[0];
        if (!F.has_mesh || F.nen != 6 || F.sdim != 2) return set_err(c, ELFEL_ESTATE, "a T6 mesh in field 0 first");
        const int64_t E = F.E;
        const bool copy = ke != nullptr;  // ke NULL with n 0: the kernel alone, timed, results left on the device
        if ((copy && n != E * 144) || (!copy && n != 0))
            return set_err(c, ELFEL_EINVAL, "element matrices: ke must hold 144 E values");
        if (E * 144 >= INT32_MAX * 8LL) return set_err(c, ELFEL_ENOTSUP, "element matrices: too many elements");
        int rc = ensure(c, c->geom, sizeof(float) * 144 * (size_t)std::max<int64_t>(E, 1));
        if (rc) return rc;
        Coef cf{};
        for (int i = 0; i < 9; i++) cf.D[i] = pb->D[i];
        hipEvent_t a, b;
        HIPCHK(c, hipEventCreate(&a));
        HIPCHK(c, hipEventCreate(&b));
        HIPCHK(c, hipEventRecord(a, c->stream));
        if (E > 0) {
            if (method == ELFEL_EM_VALU)
                hipLaunchKernelGGL(k_ke_valu, dim3(ceil_div(E, 256)), dim3(256), 0, c->stream, E, (const int32_t *)F.conn.p,
                                   (const double2 *)F.xy.p, cf, (float *)c->geom.p);
            else  // 16 elements per wave, 64 per workgroup
                hipLaunchKernelGGL(k_ke_mfma, dim3(ceil_div(E, 64)), dim3(256), 0, c->stream, E, (const int32_t *)F.conn.p,
                                   (const double2 *)F.xy.p, cf, (float *)c->geom.p);
            HIPCHK(c, hipGetLastError());
        }
        HIPCHK(c, hipEventRecord(b, c->stream));
        HIPCHK(c, hipEventSynchronize(b));
        float t = 0.f;
        hipEventElapsedTime(&t, a, b);
        hipEventDestroy(a);
        hipEventDestroy(b);
        if (ms) *ms = t;
        // element-major out of the device's entry-major layout
        if (copy && E > 0) {
            std::vector<float> h((size_t)E * 144);
            HIPCHK(c, hipMemcpy(h.data(), c->geom.p, sizeof(float) * h.size(), hipMemcpyDeviceToHost));
            for (int64_t m = 0; m < 144; m++)
                for (int64_t e = 0; e < E; e++) ke[e * 144 + m] = h[(size_t)(m * E + e)];
        }
        return ELFEL_OK;
    });
}
