// noalloc.hip -- elfel_hip_sparse_noalloc: the COO -> CSR/CSC build of
// elfel_hip_sparse on device-resident arrays, with every scratch buffer taken
// from one file-scope device array, so that the call allocates nothing.
//
// The array is `scratch` below.  In host code its name is only a stand-in for
// the device symbol: the device address of each GPU's copy comes from
// hipGetSymbolAddress, called once per device at elfel_hip_init (with the
// device current) and kept in noalloc_base[device].  Every call on a GPU uses
// that GPU's one copy, so calls that overlap on one GPU (two contexts, two
// host threads) race; the calls of one context are serialised by its stream.
#include <mutex>

#include "common.h"

int plain_coo_keys(elfel_hip_ctx *c, const int64_t *I, const int64_t *J, int64_t n, int64_t nrow, int64_t ncol,
                   int minor_bits, int csc, uint64_t *keys);

namespace {

constexpr int64_t NMAX = ELFEL_NOALLOC_MAX_COO;
constexpr size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }
// the layout of the scratch: keys and values twice (the radix sort's
// ping-pong), the digit histogram, the scan's block sums, the segment-head
// flags and positions, and the entry count
constexpr int64_t RS_TILE_KEYS = 4096;  // keys per radix tile (sort.hip)
constexpr size_t B_KEYS = align256(8 * (size_t)NMAX);
constexpr size_t B_HIST = align256(4 * (size_t)(256 * ((NMAX + RS_TILE_KEYS - 1) / RS_TILE_KEYS)));
constexpr size_t B_SCAN = align256(4 * (size_t)(NMAX / 4095 + 64));
constexpr size_t B_FLAGS = align256(4 * (size_t)NMAX);
constexpr size_t B_SEG = align256(4 * (size_t)(NMAX + 1));
constexpr size_t O_K0 = 0, O_K1 = O_K0 + B_KEYS, O_V0 = O_K1 + B_KEYS, O_V1 = O_V0 + B_KEYS;
constexpr size_t O_HIST = O_V1 + B_KEYS, O_SCAN = O_HIST + B_HIST, O_FLAGS = O_SCAN + B_SCAN;
constexpr size_t O_SEG = O_FLAGS + B_FLAGS, O_NNZ = O_SEG + B_SEG;
constexpr size_t SCRATCH_BYTES = O_NNZ + 256;

}  // namespace

static __device__ __align__(256) unsigned char scratch[SCRATCH_BYTES];

namespace {
constexpr int MAXDEV = 64;
std::mutex noalloc_mu;
void *noalloc_base[MAXDEV] = {};
}  // namespace

// Called by elfel_hip_init with `device` current.
int noalloc_init_device(int device)
{
    if (device < 0 || device >= MAXDEV) return ELFEL_ENOTSUP;
    std::lock_guard<std::mutex> g(noalloc_mu);
    if (noalloc_base[device]) return 0;
    void *p = nullptr;
    const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
    if (e != hipSuccess || !p) {
        (void)hipGetLastError();
        return ELFEL_EHIP;
    }
    noalloc_base[device] = p;
    return 0;
}

static void *noalloc_scratch(int device)
{
    if (device < 0 || device >= MAXDEV) return nullptr;
    std::lock_guard<std::mutex> g(noalloc_mu);
    return noalloc_base[device];
}

extern "C" int elfel_hip_sparse_noalloc(elfel_hip_ctx *c, int64_t nrow, int64_t ncol, int64_t n, const int64_t *row,
                                        const int64_t *col, const double *val, int layout, int32_t *ptr,
                                        int32_t *ind, double *out, int64_t *nnz)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        if (nrow < 0 || ncol < 0 || n < 0 || (n > 0 && (!row || !col || !val || !ind || !out)) || !ptr || !nnz ||
            (layout != ELFEL_CSR && layout != ELFEL_CSC))
            return set_err(c, ELFEL_EINVAL, "elfel_hip_sparse_noalloc: bad arguments");
        if (n > NMAX) {
            char msg[128];
            snprintf(msg, sizeof msg, "elfel_hip_sparse_noalloc: %lld entries exceed the limit of %lld", (long long)n,
                     (long long)NMAX);
            return set_err(c, ELFEL_EINVAL, msg);
        }
        if (nrow >= INT32_MAX || ncol >= INT32_MAX)
            return set_err(c, ELFEL_ENOTSUP, "elfel_hip_sparse_noalloc: dimensions beyond int32");
        unsigned char *base = (unsigned char *)noalloc_scratch(c->device);
        if (!base) return set_err(c, ELFEL_ESTATE, "elfel_hip_sparse_noalloc: no scratch address for this device");
        const int csc = layout == ELFEL_CSC;
        const int64_t nmaj = csc ? ncol : nrow, nmin = csc ? nrow : ncol;
        const int minor_bits = bits_for(nmin), major_bits = bits_for(nmaj);
        uint64_t *k0 = (uint64_t *)(base + O_K0), *k1 = (uint64_t *)(base + O_K1);
        uint64_t *v0 = (uint64_t *)(base + O_V0), *v1 = (uint64_t *)(base + O_V1);
        SortScratch s;
        s.hist = (uint32_t *)(base + O_HIST);
        s.hist_cap = (int64_t)(B_HIST / 4);
        s.scan = (uint32_t *)(base + O_SCAN);
        s.scan_cap = (int64_t)(B_SCAN / 4);
        s.flags = (uint32_t *)(base + O_FLAGS);
        s.segpos = (uint32_t *)(base + O_SEG);
        s.n_cap = NMAX;
        int32_t *nnz_dev = (int32_t *)(base + O_NNZ);
        int rc;
        if (n > 0) HIPCHK(c, hipMemcpyAsync(v0, val, 8 * (size_t)n, hipMemcpyDeviceToDevice, c->stream));
        if ((rc = plain_coo_keys(c, row, col, n, nrow, ncol, minor_bits, csc, k0))) return rc;
        // the range flags live in the context's device scalars (set at init)
        int flags = 0;
        HIPCHK(c, hipMemcpyAsync(&flags, (int *)c->dscal.p + 1, sizeof(int), hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (flags) {
            HIPCHK(c, hipMemsetAsync((int *)c->dscal.p + 1, 0, sizeof(int), c->stream));
            if (flags & DERR_RANGE) return set_err(c, ELFEL_ERANGE, "row or column index out of range");
        }
        if ((rc = radix_sort_pairs(c, k0, v0, k1, v1, n, major_bits + minor_bits, &s))) return rc;
        if ((rc = build_csr_from_sorted(c, k0, (const double *)v0, n, minor_bits, nmaj, ptr, ind, out, nnz_dev, true,
                                        &s)))
            return rc;
        int32_t h = 0;
        HIPCHK(c, hipMemcpyAsync(&h, nnz_dev, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        *nnz = h;
        return ELFEL_OK;
    });
}
