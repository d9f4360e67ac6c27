// numbering.hip -- dof numbering and system-vector gather/scatter on the device.
//
// numberdofs! (src/FESpaces.jl:250-273, src/FEFields.jl:137-177): over the
// spaces in order, the free dofs of each space are numbered consecutively in
// entity order (components minor), the first space from 1 and every further
// space from the previous space's highest free number + 1; then the data dofs
// the same way, the first space's from the last space's highest free number
// + 1.  "Highest" is 0 for a space without free (data) dofs, as in
// highestfreedofnum / highestdatadofnum, so a later space then restarts at 1;
// the rule is kept as the reference has it.
// Device form: one exclusive scan of the free flags and one of the data flags
// over all spaces; the per-space starting numbers follow from the per-space
// counts on the host; one kernel writes the numbers.
//
// gathersysvec! / scattersysvec! (src/FEFields.jl:247-257 and :264-279, per
// field of src/FESpaces.jl:312-365): v[dofnums[k]] = dofvals[k] for every
// dof, and dofvals[k] = v[dofnums[k]] for the dofs numbered within v.  A dof
// number outside v (gather) or an unnumbered dof (either) is the reference's
// BoundsError: ELFEL_ERANGE.
#include <algorithm>
#include <vector>

#include "common.h"

namespace {

__global__ void k_flags(int64_t n, const uint8_t *__restrict__ isdatum, uint32_t *__restrict__ fr,
                        uint32_t *__restrict__ da)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const uint32_t d = isdatum[i] ? 1u : 0u;
    fr[i] = 1u - d;
    da[i] = d;
}

// dofnums[k] = start of k's space (free or data) + rank of k among that
// space's free (data) entries; spaces are found by binary search of base[].
__global__ void k_number(int64_t n, int nspaces, const int64_t *__restrict__ base,
                         const int64_t *__restrict__ free_start, const int64_t *__restrict__ data_start,
                         const uint8_t *__restrict__ isdatum, const uint32_t *__restrict__ fr,
                         const uint32_t *__restrict__ da, int64_t *__restrict__ dofnums)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    int lo = 0, hi = nspaces - 1;
    while (lo < hi) {
        const int m = (lo + hi + 1) >> 1;
        if (base[m] <= i) lo = m;
        else hi = m - 1;
    }
    const int64_t b = base[lo];
    if (isdatum[i]) dofnums[i] = data_start[lo] + (int64_t)(da[i] - da[b]);
    else dofnums[i] = free_start[lo] + (int64_t)(fr[i] - fr[b]);
}

__global__ void k_set_u32(uint32_t *p, uint32_t v)
{
    if (threadIdx.x == 0) p[0] = v;
}

__global__ void k_gather_sysvec(int64_t n, const int32_t *__restrict__ dof, const double *__restrict__ vals,
                                int64_t ndofs, double *__restrict__ v, uint32_t *__restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t d = dof[i];
    if (d >= 0 && d < ndofs) v[d] = vals[i];
    else *bad = 1u;
}

__global__ void k_scatter_sysvec(int64_t n, const int32_t *__restrict__ dof, const double *__restrict__ v,
                                 int64_t ndofs, double *__restrict__ vals, uint32_t *__restrict__ bad)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const int32_t d = dof[i];
    if (d < 0) *bad = 1u;
    else if (d < ndofs) vals[i] = v[d];
}

}  // namespace

extern "C" int elfel_hip_number_dofs(elfel_hip_ctx *c, int32_t nspaces, const int64_t *nterms, const int32_t *ncomp,
                                     const uint8_t *isdatum, int64_t *dofnums)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        if (nspaces < 1 || !nterms || !ncomp || !isdatum || !dofnums)
            return set_err(c, ELFEL_EINVAL, "elfel_hip_number_dofs: bad arguments");
        std::vector<int64_t> base(nspaces + 1, 0);
        for (int s = 0; s < nspaces; s++) {
            if (nterms[s] < 0 || ncomp[s] < 1) return set_err(c, ELFEL_EINVAL, "elfel_hip_number_dofs: bad space size");
            base[s + 1] = base[s] + nterms[s] * ncomp[s];
        }
        const int64_t n = base[nspaces];
        if (n == 0) return ELFEL_OK;
        if (n >= INT32_MAX) return set_err(c, ELFEL_ENOTSUP, "elfel_hip_number_dofs: more than 2^31 dofs");
        // scratch: isdatum (n bytes), two uint32 flag arrays (n + 1 each), the
        // output (n int64) and the per-space tables
        const size_t a0 = ((size_t)n + 15) & ~(size_t)15;
        const size_t a1 = a0 + 4 * ((size_t)n + 4);
        const size_t a2 = a1 + 4 * ((size_t)n + 4);
        const size_t a3 = (a2 + 8 * (size_t)n + 15) & ~(size_t)15;
        const size_t bytes = a3 + 8 * 3 * ((size_t)nspaces + 1);
        int rc = ensure(c, c->pcg, bytes);
        if (rc) return rc;
        char *p = (char *)c->pcg.p;
        uint8_t *d_isd = (uint8_t *)p;
        uint32_t *fr = (uint32_t *)(p + a0);
        uint32_t *da = (uint32_t *)(p + a1);
        int64_t *d_num = (int64_t *)(p + a2);
        int64_t *d_tab = (int64_t *)(p + a3);
        hipStream_t st = c->stream;
        HIPCHK(c, hipMemcpyAsync(d_isd, isdatum, (size_t)n, hipMemcpyHostToDevice, st));
        const int nb = ceil_div(n, 256);
        hipLaunchKernelGGL(k_flags, dim3(nb), dim3(256), 0, st, n, (const uint8_t *)d_isd, fr, da);
        // exclusive scans over n + 1 entries: position n then holds the total
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, fr + n, 0u);
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, da + n, 0u);
        if ((rc = exclusive_scan_u32(c, fr, n + 1, nullptr))) return rc;
        if ((rc = exclusive_scan_u32(c, da, n + 1, nullptr))) return rc;
        // per-space counts -> starting numbers (host, nspaces entries)
        std::vector<uint32_t> hf(nspaces + 1), hd(nspaces + 1);
        for (int s = 0; s <= nspaces; s++) {
            HIPCHK(c, hipMemcpyAsync(&hf[s], fr + base[s], 4, hipMemcpyDeviceToHost, st));
            HIPCHK(c, hipMemcpyAsync(&hd[s], da + base[s], 4, hipMemcpyDeviceToHost, st));
        }
        HIPCHK(c, hipStreamSynchronize(st));
        std::vector<int64_t> tab(3 * ((size_t)nspaces + 1), 0);
        int64_t *fstart = tab.data(), *dstart = tab.data() + (nspaces + 1), *tbase = tab.data() + 2 * (nspaces + 1);
        int64_t hi_free = 0;
        for (int s = 0; s < nspaces; s++) {
            const int64_t cnt = (int64_t)hf[s + 1] - hf[s];
            fstart[s] = s == 0 ? 1 : hi_free + 1;
            hi_free = cnt > 0 ? fstart[s] + cnt - 1 : 0;
        }
        int64_t hi_data = 0;
        for (int s = 0; s < nspaces; s++) {
            const int64_t cnt = (int64_t)hd[s + 1] - hd[s];
            dstart[s] = s == 0 ? hi_free + 1 : hi_data + 1;
            hi_data = cnt > 0 ? dstart[s] + cnt - 1 : 0;
        }
        for (int s = 0; s <= nspaces; s++) tbase[s] = base[s];
        HIPCHK(c, hipMemcpyAsync(d_tab, tab.data(), 8 * tab.size(), hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_number, dim3(nb), dim3(256), 0, st, n, (int)nspaces, (const int64_t *)(d_tab + 2 * (nspaces + 1)),
                           (const int64_t *)d_tab, (const int64_t *)(d_tab + (nspaces + 1)), (const uint8_t *)d_isd,
                           (const uint32_t *)fr, (const uint32_t *)da, d_num);
        HIPCHK(c, hipGetLastError());
        HIPCHK(c, hipMemcpyAsync(dofnums, d_num, 8 * (size_t)n, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        return ELFEL_OK;
    });
}

static int sysvec_common(elfel_hip_ctx *c, int field, int64_t nvals, int64_t ndofs)
{
    if (field < 0 || field > 1) return set_err(c, ELFEL_EINVAL, "bad field");
    const Field &F = c->f[field];
    if (!F.has_dofs) return set_err(c, ELFEL_ESTATE, "set_dofs of this field first");
    if (nvals != (int64_t)F.h_dof.size() || ndofs < 0) return set_err(c, ELFEL_EINVAL, "dof value array size mismatch");
    return 0;
}

extern "C" int elfel_hip_gather_sysvec(elfel_hip_ctx *c, int field, int64_t nvals, const double *dofvals, int64_t ndofs,
                                       double *v)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        int rc = sysvec_common(c, field, nvals, ndofs);
        if (rc) return rc;
        if (nvals == 0 || ndofs == 0) return ELFEL_OK;
        if (!dofvals || !v) return set_err(c, ELFEL_EINVAL, "null array");
        if ((rc = ensure(c, c->pcg, 8 * ((size_t)nvals + (size_t)ndofs) + 256))) return rc;
        double *d_vals = (double *)c->pcg.p, *d_v = d_vals + nvals;
        uint32_t *bad = (uint32_t *)(d_v + ndofs);
        hipStream_t st = c->stream;
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, bad, 0u);
        HIPCHK(c, hipMemcpyAsync(d_vals, dofvals, 8 * (size_t)nvals, hipMemcpyHostToDevice, st));
        HIPCHK(c, hipMemcpyAsync(d_v, v, 8 * (size_t)ndofs, hipMemcpyHostToDevice, st));  // entries no dof writes keep their values
        hipLaunchKernelGGL(k_gather_sysvec, dim3(ceil_div(nvals, 256)), dim3(256), 0, st, nvals,
                           (const int32_t *)c->f[field].dof.p, (const double *)d_vals, ndofs, d_v, bad);
        HIPCHK(c, hipGetLastError());
        uint32_t hbad = 0;
        HIPCHK(c, hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (hbad) return set_err(c, ELFEL_ERANGE, "gather_sysvec: a dof number lies outside the vector");
        HIPCHK(c, hipMemcpy(v, d_v, 8 * (size_t)ndofs, hipMemcpyDeviceToHost));
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_scatter_sysvec(elfel_hip_ctx *c, int field, int64_t ndofs, const double *v, int64_t nvals,
                                        double *dofvals)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        int rc = sysvec_common(c, field, nvals, ndofs);
        if (rc) return rc;
        if (nvals == 0) return ELFEL_OK;
        if (!dofvals || (!v && ndofs)) return set_err(c, ELFEL_EINVAL, "null array");
        if ((rc = ensure(c, c->pcg, 8 * ((size_t)nvals + (size_t)ndofs) + 256))) return rc;
        double *d_vals = (double *)c->pcg.p, *d_v = d_vals + nvals;
        uint32_t *bad = (uint32_t *)(d_v + ndofs);
        hipStream_t st = c->stream;
        hipLaunchKernelGGL(k_set_u32, dim3(1), dim3(64), 0, st, bad, 0u);
        HIPCHK(c, hipMemcpyAsync(d_vals, dofvals, 8 * (size_t)nvals, hipMemcpyHostToDevice, st));  // dofs past v keep theirs
        if (ndofs) HIPCHK(c, hipMemcpyAsync(d_v, v, 8 * (size_t)ndofs, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_scatter_sysvec, dim3(ceil_div(nvals, 256)), dim3(256), 0, st, nvals,
                           (const int32_t *)c->f[field].dof.p, (const double *)d_v, ndofs, d_vals, bad);
        HIPCHK(c, hipGetLastError());
        uint32_t hbad = 0;
        HIPCHK(c, hipMemcpyAsync(&hbad, bad, 4, hipMemcpyDeviceToHost, st));
        HIPCHK(c, hipStreamSynchronize(st));
        if (hbad) return set_err(c, ELFEL_ERANGE, "scatter_sysvec: unnumbered dof");
        HIPCHK(c, hipMemcpy(dofvals, d_vals, 8 * (size_t)nvals, hipMemcpyDeviceToHost));
        return ELFEL_OK;
    });
}
