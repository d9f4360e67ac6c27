// batch64.hip -- the FP64 instantiation of the batched kernel (k_reuse_batch,
// reuse_kernel.h), built like reuse64.hip without floating-point contraction:
// every entry has the bits of the single-problem FP64 method.
#include "reuse_kernel.h"

int launch_batch64(hipStream_t st, const BatchLaunch &L) { return launch_batch_t<double>(st, L); }
