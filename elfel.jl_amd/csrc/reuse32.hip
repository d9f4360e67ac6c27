// reuse32.hip -- mixed-precision instantiations of the reuse kernels: FP64
// Jacobian, determinant and inverse, FP32 shape gradients and element
// products (contracted into FMAs), FP64 sums into the matrix.
#include "reuse_kernel.h"

static reuse_launch_fn pick128(int kind, int fe, int nq, int csc) { REUSE_PICK_BODY(float, 128) }
static reuse_launch_fn pick256(int kind, int fe, int nq, int csc) { REUSE_PICK_BODY(float, 256) }
reuse_launch_fn pick_heat_rows32(int fe, int nq, int csc) { HEAT_ROWS_PICK_BODY(float) }
reuse_launch_fn pick_reuse32(int kind, int fe, int nq, int csc, int block)
{
    return block == 128 ? pick128(kind, fe, nq, csc) : pick256(kind, fe, nq, csc);
}

int gather_geom_rel(elfel_hip_ctx *c, int nen, int64_t E, const int32_t *conn, const double2 *xy, double *geom)
{
    if (E <= 0) return 0;
    const dim3 g(ceil_div(E * nen, 256)), b(256);
    if (nen == 3) hipLaunchKernelGGL((k_gather_geom<3, true>), g, b, 0, c->stream, E, conn, xy, geom);
    else if (nen == 4) hipLaunchKernelGGL((k_gather_geom<4, true>), g, b, 0, c->stream, E, conn, xy, geom);
    else hipLaunchKernelGGL((k_gather_geom<6, true>), g, b, 0, c->stream, E, conn, xy, geom);
    HIPCHK(c, hipGetLastError());
    return 0;
}
