// reuse.hip -- cached-pattern fused row-owner kernels (placeholder).
#include "elem.h"

int reuse_assemble(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf)
{
    (void)pb; (void)tab; (void)cf;
    return set_err(c, ELFEL_ENOTSUP, "reuse method not built yet");
}

int geometry_stats_run(elfel_hip_ctx *c, double *min_det, int64_t *nfallback)
{
    (void)min_det; (void)nfallback;
    return set_err(c, ELFEL_ENOTSUP, "geometry statistics not built yet");
}
