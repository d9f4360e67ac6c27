// solve.hip -- Jacobi-PCG (placeholder).
#include "common.h"

int pcg_solve(elfel_hip_ctx *c, int64_t nfree, const double *F, double *x, int precision, double tol, int maxit,
              int *iters, int *outer, double *relres)
{
    (void)nfree; (void)F; (void)x; (void)precision; (void)tol; (void)maxit; (void)iters; (void)outer; (void)relres;
    return set_err(c, ELFEL_ENOTSUP, "pcg not built yet");
}
