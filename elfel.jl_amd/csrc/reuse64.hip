// reuse64.hip -- FP64 instantiations of the reuse kernels.  Compiled with
// -ffp-contract=off: every entry is computed operation by operation as the
// reference computes it and summed in element order, so the FP64 reuse result
// equals the plain method's bit for bit.
#include "reuse_kernel.h"

static reuse_launch_fn pick128(int kind, int fe, int nq, int csc) { REUSE_PICK_BODY(double, 128) }
static reuse_launch_fn pick256(int kind, int fe, int nq, int csc) { REUSE_PICK_BODY(double, 256) }
reuse_launch_fn pick_heat_rows64(int fe, int nq, int csc) { HEAT_ROWS_PICK_BODY(double) }
reuse_launch_fn pick_reuse64(int kind, int fe, int nq, int csc, int block)
{
    return block == 128 ? pick128(kind, fe, nq, csc) : pick256(kind, fe, nq, csc);
}

int gather_geom64(elfel_hip_ctx *c, int nen, int64_t E, const int32_t *conn, const double2 *xy, double *geom)
{
    if (E <= 0) return 0;
    const dim3 g(ceil_div(E * nen, 256)), b(256);
    if (nen == 3) hipLaunchKernelGGL((k_gather_geom<3, false>), g, b, 0, c->stream, E, conn, xy, geom);
    else if (nen == 4) hipLaunchKernelGGL((k_gather_geom<4, false>), g, b, 0, c->stream, E, conn, xy, geom);
    else hipLaunchKernelGGL((k_gather_geom<6, false>), g, b, 0, c->stream, E, conn, xy, geom);
    HIPCHK(c, hipGetLastError());
    return 0;
}
