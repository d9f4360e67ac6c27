// sort.hip -- deterministic COO -> CSR build: stable LSD radix sort of 64-bit
// keys with 64-bit payloads, then an in-order segmented reduce.
//
// This replaces SparseArrays.sparse called by finish!(SysmatAssemblerSparse)
// (src/Assemblers.jl:121-123).  Stability of every pass keeps equal keys in
// COO stream order, and each segment is summed sequentially from its first
// entry, so duplicates are combined left to right exactly as the stream
// presents them; explicit zeros stay structural entries.
//
// Radix pass (8-bit digits, tiles of 4096 keys, 256 threads = 4 wave64):
//   k_rs_hist    per-tile digit histogram (per-wave LDS sub-histograms)
//   scan         exclusive scan of the digit-major [256][ntiles] table
//   k_rs_scatter stable rank: per round of 256 keys, lanes holding the same
//                digit are matched with 8 ballots; ranks = popcount of lower
//                peers + counts of earlier waves + earlier rounds of the tile;
//                the tile is staged in LDS in digit order and written back
//                one digit range after another.
#include <algorithm>

#include "common.h"

#define RS_BLOCK 256
#define RS_ITEMS 16
#define RS_TILE (RS_BLOCK * RS_ITEMS)
#define RS_RADIX 256

#define SC_BLOCK 256
#define SC_ITEMS 16
#define SC_CHUNK (SC_BLOCK * SC_ITEMS)

int bits_for(int64_t n)
{
    int b = 0;
    while (b < 63 && ((int64_t)1 << b) < n) b++;
    return b;
}

// ---------------------------------------------------------------------------
// Exclusive scan of uint32 (in place), chunk per block, recursive on sums.
// ---------------------------------------------------------------------------
__device__ __forceinline__ uint32_t wave_incl_scan(uint32_t v, int lane)
{
    for (int d = 1; d < 64; d <<= 1) {
        uint32_t t = __shfl_up(v, d, 64);
        if (lane >= d) v += t;
    }
    return v;
}

__global__ __launch_bounds__(SC_BLOCK) void k_scan_chunks(uint32_t *data, int64_t n, uint32_t *sums)
{
    __shared__ uint32_t s[SC_CHUNK];
    __shared__ uint32_t wsum[SC_BLOCK / 64];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int64_t base = (int64_t)blockIdx.x * SC_CHUNK;
    for (int i = tid; i < SC_CHUNK; i += SC_BLOCK) {
        int64_t g = base + i;
        s[i] = g < n ? data[g] : 0u;
    }
    __syncthreads();
    uint32_t loc[SC_ITEMS];
    uint32_t tsum = 0;
    for (int k = 0; k < SC_ITEMS; k++) {
        loc[k] = tsum;
        tsum += s[tid * SC_ITEMS + k];
    }
    uint32_t inc = wave_incl_scan(tsum, lane);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t woff = 0;
    for (int k = 0; k < w; k++) woff += wsum[k];
    uint32_t excl = woff + inc - tsum;
    for (int k = 0; k < SC_ITEMS; k++) s[tid * SC_ITEMS + k] = excl + loc[k];
    if (tid == SC_BLOCK - 1) sums[blockIdx.x] = excl + tsum;
    __syncthreads();
    for (int i = tid; i < SC_CHUNK; i += SC_BLOCK) {
        int64_t g = base + i;
        if (g < n) data[g] = s[i];
    }
}

__global__ void k_scan_add(uint32_t *data, int64_t n, const uint32_t *sums)
{
    int64_t g = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (g < n) data[g] += sums[g / SC_CHUNK];
}

__global__ void k_copy_u32(const uint32_t *src, uint32_t *dst) { *dst = *src; }

// total_dev (optional) receives the sum of all n inputs.
static int scan_rec(elfel_hip_ctx *c, uint32_t *data, int64_t n, uint32_t *tmp, int64_t tmp_cap,
                    uint32_t *total_dev)
{
    if (n <= 0) return 0;
    int64_t nb = (n + SC_CHUNK - 1) / SC_CHUNK;
    if (nb + 1 > tmp_cap) return set_err(c, ELFEL_EINVAL, "scan scratch too small");
    uint32_t *sums = tmp;
    hipLaunchKernelGGL(k_scan_chunks, dim3((unsigned)nb), dim3(SC_BLOCK), 0, c->stream, data, n, sums);
    HIPCHK(c, hipGetLastError());
    if (nb > 1) {
        // sums[nb] receives the grand total via the recursive call.
        int rc = scan_rec(c, sums, nb, tmp + nb + 1, tmp_cap - nb - 1, total_dev);
        if (rc) return rc;
        hipLaunchKernelGGL(k_scan_add, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, data, n, sums);
        HIPCHK(c, hipGetLastError());
    } else if (total_dev) {
        hipLaunchKernelGGL(k_copy_u32, dim3(1), dim3(1), 0, c->stream, sums, total_dev);
        HIPCHK(c, hipGetLastError());
    }
    return 0;
}

int64_t scan_scratch_entries(int64_t n) { return n / (SC_CHUNK - 1) + 64; }

int exclusive_scan_u32(elfel_hip_ctx *c, uint32_t *data, int64_t n, uint32_t *total_dev, const SortScratch *s)
{
    // Scratch: nb + nb/4096 + ... < n/4000 + 64 entries.
    int64_t cap = scan_scratch_entries(n);
    if (s) {
        if (s->scan_cap < cap) return set_err(c, ELFEL_EINVAL, "internal: scan scratch too small");
        return scan_rec(c, data, n, s->scan, s->scan_cap, total_dev);
    }
    int rc = ensure(c, c->scan_tmp, sizeof(uint32_t) * cap);
    if (rc) return rc;
    return scan_rec(c, data, n, (uint32_t *)c->scan_tmp.p, cap, total_dev);
}

// ---------------------------------------------------------------------------
// Radix sort
// ---------------------------------------------------------------------------
__global__ __launch_bounds__(RS_BLOCK) void k_rs_hist(const uint64_t *__restrict__ keys, int64_t n,
                                                      int shift, uint32_t *__restrict__ hist, int ntiles)
{
    __shared__ uint32_t h[4][RS_RADIX];
    const int tid = threadIdx.x, w = tid >> 6;
    for (int i = tid; i < 4 * RS_RADIX; i += RS_BLOCK) (&h[0][0])[i] = 0;
    __syncthreads();
    const int64_t base = (int64_t)blockIdx.x * RS_TILE;
    for (int r = 0; r < RS_ITEMS; r++) {
        int64_t idx = base + r * RS_BLOCK + tid;
        if (idx < n) {
            uint32_t d = (uint32_t)(keys[idx] >> shift) & (RS_RADIX - 1);
            atomicAdd(&h[w][d], 1u);
        }
    }
    __syncthreads();
    hist[(int64_t)tid * ntiles + blockIdx.x] = h[0][tid] + h[1][tid] + h[2][tid] + h[3][tid];
}

// Stable scatter of one tile.  Ranks come round by round as before (wave64
// ballots within a round, earlier waves and rounds of the tile through LDS
// counts); the tile is then laid out in LDS in digit order and written back
// bucket by bucket, so consecutive lanes store consecutive addresses of one
// digit's range instead of 8 bytes into each of up to 256 buckets per round.
// Every key lands where the unstaged scatter put it: the order is the same.
__global__ __launch_bounds__(RS_BLOCK) void k_rs_scatter(const uint64_t *__restrict__ kin,
                                                         const uint64_t *__restrict__ vin,
                                                         uint64_t *__restrict__ kout,
                                                         uint64_t *__restrict__ vout, int64_t n, int shift,
                                                         const uint32_t *__restrict__ hist, int ntiles)
{
    extern __shared__ uint64_t stage[];  // [RS_TILE] keys, then [RS_TILE] values
    uint64_t *sk = stage, *sv = stage + RS_TILE;
    __shared__ uint32_t gbase[RS_RADIX];  // this tile's first global slot of each digit
    __shared__ uint32_t run[RS_RADIX];    // keys of each digit met so far, then the tile-local starts
    __shared__ uint32_t wc[4][RS_RADIX];
    __shared__ uint32_t wsum[RS_BLOCK / 64];
    const int tid = threadIdx.x, w = tid >> 6, lane = tid & 63;
    gbase[tid] = hist[(int64_t)tid * ntiles + blockIdx.x];
    run[tid] = 0;
    for (int ww = 0; ww < 4; ww++) wc[ww][tid] = 0;
    __syncthreads();
    const uint64_t lt = (lane == 0) ? 0ull : (~0ull >> (64 - lane));
    const int64_t t0 = (int64_t)blockIdx.x * RS_TILE;
    uint64_t kr[RS_ITEMS], vr[RS_ITEMS];
    uint32_t lr[RS_ITEMS];
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const int64_t idx = t0 + r * RS_BLOCK + tid;
        const bool valid = idx < n;
        kr[r] = 0; vr[r] = 0;
        if (valid) { kr[r] = kin[idx]; vr[r] = vin[idx]; }
    }
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        const bool valid = t0 + r * RS_BLOCK + tid < n;
        const uint32_t d = (uint32_t)(kr[r] >> shift) & (RS_RADIX - 1);
        uint64_t peers = __ballot(valid);
#pragma unroll
        for (int b = 0; b < 8; b++) {
            const bool bit = (d >> b) & 1u;
            const uint64_t m = __ballot(bit);
            peers &= bit ? m : ~m;
        }
        const uint32_t rank = (uint32_t)__popcll(peers & lt);
        if (valid && rank == 0) wc[w][d] = (uint32_t)__popcll(peers);
        __syncthreads();
        uint32_t acc = 0;
        for (int ww = 0; ww < 4; ww++) {
            const uint32_t cnt = wc[ww][tid];
            wc[ww][tid] = acc;
            acc += cnt;
        }
        __syncthreads();
        lr[r] = valid ? run[d] + wc[w][d] + rank : 0u;
        __syncthreads();
        run[tid] += acc;
        for (int ww = 0; ww < 4; ww++) wc[ww][tid] = 0;
        __syncthreads();
    }
    // tile-local start of each digit: exclusive scan of the counts
    const uint32_t total = run[tid];
    const uint32_t inc = wave_incl_scan(total, lane);
    if (lane == 63) wsum[w] = inc;
    __syncthreads();
    uint32_t woff = 0;
    for (int k = 0; k < w; k++) woff += wsum[k];
    run[tid] = woff + inc - total;
    __syncthreads();
#pragma unroll
    for (int r = 0; r < RS_ITEMS; r++) {
        if (t0 + r * RS_BLOCK + tid < n) {
            const uint32_t d = (uint32_t)(kr[r] >> shift) & (RS_RADIX - 1);
            const uint32_t p = run[d] + lr[r];
            sk[p] = kr[r];
            sv[p] = vr[r];
        }
    }
    __syncthreads();
    const int m = (int)((n - t0) < RS_TILE ? (n - t0) : RS_TILE);
    for (int i = tid; i < m; i += RS_BLOCK) {
        const uint64_t k = sk[i];
        const uint32_t d = (uint32_t)(k >> shift) & (RS_RADIX - 1);
        const uint32_t dst = gbase[d] + (uint32_t)i - run[d];
        kout[dst] = k;
        vout[dst] = sv[i];
    }
}

int64_t radix_hist_entries(int64_t n) { return (int64_t)RS_RADIX * ceil_div(std::max<int64_t>(n, 1), RS_TILE); }

int radix_sort_pairs(elfel_hip_ctx *c, uint64_t *&keys, uint64_t *&vals, uint64_t *&keys_alt,
                     uint64_t *&vals_alt, int64_t n, int nbits, const SortScratch *s)
{
    if (n <= 1 || nbits <= 0) return 0;
    if (n >= ((int64_t)1 << 32)) return set_err(c, ELFEL_ENOTSUP, "sort of more than 2^32 keys");
    const int ntiles = ceil_div(n, RS_TILE);
    const int64_t nh = (int64_t)RS_RADIX * ntiles;
    uint32_t *hist = nullptr;
    int rc;
    if (s) {
        if (s->hist_cap < nh) return set_err(c, ELFEL_EINVAL, "internal: histogram scratch too small");
        hist = s->hist;
    } else {
        if ((rc = ensure(c, c->hist, sizeof(uint32_t) * nh))) return rc;
        hist = (uint32_t *)c->hist.p;
    }
    for (int shift = 0; shift < nbits; shift += 8) {
        hipLaunchKernelGGL(k_rs_hist, dim3(ntiles), dim3(RS_BLOCK), 0, c->stream, keys, n, shift, hist, ntiles);
        HIPCHK(c, hipGetLastError());
        rc = exclusive_scan_u32(c, hist, nh, nullptr, s);
        if (rc) return rc;
        hipLaunchKernelGGL(k_rs_scatter, dim3(ntiles), dim3(RS_BLOCK), 2 * RS_TILE * sizeof(uint64_t), c->stream,
                           keys, vals, keys_alt, vals_alt, n, shift, hist, ntiles);
        HIPCHK(c, hipGetLastError());
        uint64_t *t = keys; keys = keys_alt; keys_alt = t;
        t = vals; vals = vals_alt; vals_alt = t;
    }
    return 0;
}

// ---------------------------------------------------------------------------
// Segmented reduce of sorted pairs into CSR
// ---------------------------------------------------------------------------
__global__ void k_mark_heads(const uint64_t *__restrict__ keys, int64_t n, uint32_t *__restrict__ flags)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) flags[i] = (i == 0 || keys[i] != keys[i - 1]) ? 1u : 0u;
}

// Head positions: hpos[seg[i]] = i for every segment head i, and the segment
// count (seg is the exclusive scan of the head flags).
__global__ void k_head_pos(const uint64_t *__restrict__ keys, int64_t n, const uint32_t *__restrict__ seg,
                           uint32_t *__restrict__ hpos, uint32_t *__restrict__ nseg)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const bool head = (i == 0) || (keys[i - 1] != keys[i]);
    if (head) hpos[seg[i]] = (uint32_t)i;
    if (i == n - 1) *nseg = seg[i] + (head ? 1u : 0u);
}

// One work-item per segment s (grid-stride): the segment's entries summed left
// to right from its head, its column, and ptr for the majors that end at it.
__global__ __launch_bounds__(256) void k_seg_sum(const uint64_t *__restrict__ keys, const double *__restrict__ vals,
                                                 int64_t n, const uint32_t *__restrict__ hpos,
                                                 const uint32_t *__restrict__ nseg_dev, int minor_bits,
                                                 int64_t nmaj, int32_t *__restrict__ ptr, int32_t *__restrict__ ind,
                                                 double *__restrict__ val, int32_t *__restrict__ nnz_dev,
                                                 int with_values)
{
    const int64_t nseg = *nseg_dev;
    const uint64_t mmask = (minor_bits >= 64) ? ~0ull : ((1ull << minor_bits) - 1);
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i0 = hpos[s];
        const int64_t i1 = s + 1 < nseg ? (int64_t)hpos[s + 1] : n;
        const uint64_t k = keys[i0];
        if (with_values) {
            double acc = vals[i0];
            for (int64_t j = i0 + 1; j < i1; j++) acc += vals[j];
            val[s] = acc;
        }
        ind[s] = (int32_t)(k & mmask);
        const int64_t m = (int64_t)(k >> minor_bits);
        const int64_t mp = (i0 == 0) ? -1 : (int64_t)(keys[i0 - 1] >> minor_bits);
        for (int64_t r = mp + 1; r <= m; r++) ptr[r] = (int32_t)s;
        if (s == nseg - 1) {
            for (int64_t r = m + 1; r <= nmaj; r++) ptr[r] = (int32_t)nseg;
            *nnz_dev = (int32_t)nseg;
        }
    }
}

__global__ void k_fill_i32(int32_t *p, int64_t n, int32_t v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

__global__ void k_fill_f64(double *p, int64_t n, double v)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = v;
}

int build_csr_from_sorted(elfel_hip_ctx *c, const uint64_t *keys, const double *vals, int64_t n,
                          int minor_bits, int64_t nmaj, int32_t *ptr, int32_t *ind, double *val,
                          int32_t *nnz_dev, bool values, const SortScratch *s)
{
    if (n == 0) {
        hipLaunchKernelGGL(k_fill_i32, dim3(ceil_div(nmaj + 1, 256)), dim3(256), 0, c->stream, ptr, nmaj + 1, 0);
        HIPCHK(c, hipGetLastError());
        hipLaunchKernelGGL(k_fill_i32, dim3(1), dim3(1), 0, c->stream, nnz_dev, (int64_t)1, 0);
        HIPCHK(c, hipGetLastError());
        return 0;
    }
    int rc;
    uint32_t *seg = nullptr, *hpos = nullptr;
    if (s) {
        if (s->n_cap < n) return set_err(c, ELFEL_EINVAL, "internal: segment scratch too small");
        seg = s->flags;
        hpos = s->segpos;
    } else {
        if ((rc = ensure(c, c->flags, sizeof(uint32_t) * n))) return rc;
        seg = (uint32_t *)c->flags.p;
    }
    hipLaunchKernelGGL(k_mark_heads, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, keys, n, seg);
    HIPCHK(c, hipGetLastError());
    rc = exclusive_scan_u32(c, seg, n, nullptr, s);
    if (rc) return rc;
    // one work-item per segment over the compacted head positions (the
    // per-entry form idled the lanes of every entry that does not start a
    // segment while the head lanes summed)
    if (!s) {
        if ((rc = ensure(c, c->segpos, sizeof(uint32_t) * (n + 1)))) return rc;
        hpos = (uint32_t *)c->segpos.p;
    }
    uint32_t *nseg = hpos + n;
    hipLaunchKernelGGL(k_head_pos, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, keys, n, seg, hpos, nseg);
    HIPCHK(c, hipGetLastError());
    const int grid = (int)std::min<int64_t>(ceil_div(n, 256), 8192);
    hipLaunchKernelGGL(k_seg_sum, dim3(grid), dim3(256), 0, c->stream, keys, vals, n, hpos, nseg, minor_bits,
                       nmaj, ptr, ind, val, nnz_dev, values ? 1 : 0);
    HIPCHK(c, hipGetLastError());
    return 0;
}

// One work-item per segment of equal keys (compacted heads, as for the
// matrix): the segment summed left to right into out[key].
__global__ __launch_bounds__(256) void k_dense_sum(const uint64_t *__restrict__ keys, const double *__restrict__ vals,
                                                   int64_t n, const uint32_t *__restrict__ hpos,
                                                   const uint32_t *__restrict__ nseg_dev, double *__restrict__ out)
{
    const int64_t nseg = *nseg_dev;
    for (int64_t s = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; s < nseg; s += (int64_t)gridDim.x * blockDim.x) {
        const int64_t i0 = hpos[s];
        const int64_t i1 = s + 1 < nseg ? (int64_t)hpos[s + 1] : n;
        double acc = vals[i0];
        for (int64_t j = i0 + 1; j < i1; j++) acc += vals[j];
        out[keys[i0]] = 0.0 + acc;  // start! zeros the vector, assemble! adds (src/Assemblers.jl:194-221)
    }
}

int build_dense_from_sorted(elfel_hip_ctx *c, const uint64_t *keys, const double *vals, int64_t n,
                            int64_t ndofs, double *out)
{
    hipLaunchKernelGGL(k_fill_f64, dim3(ceil_div(ndofs, 256)), dim3(256), 0, c->stream, out, ndofs, 0.0);
    HIPCHK(c, hipGetLastError());
    if (n == 0) return 0;
    int rc;
    if ((rc = ensure(c, c->flags, sizeof(uint32_t) * n)) || (rc = ensure(c, c->segpos, sizeof(uint32_t) * (n + 1))))
        return rc;
    uint32_t *seg = (uint32_t *)c->flags.p, *hpos = (uint32_t *)c->segpos.p, *nseg = hpos + n;
    hipLaunchKernelGGL(k_mark_heads, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, keys, n, seg);
    HIPCHK(c, hipGetLastError());
    if ((rc = exclusive_scan_u32(c, seg, n, nullptr))) return rc;
    hipLaunchKernelGGL(k_head_pos, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, keys, n, seg, hpos, nseg);
    HIPCHK(c, hipGetLastError());
    const int grid = (int)std::min<int64_t>(ceil_div(n, 256), 8192);
    hipLaunchKernelGGL(k_dense_sum, dim3(grid), dim3(256), 0, c->stream, keys, vals, n, hpos, nseg, out);
    HIPCHK(c, hipGetLastError());
    return 0;
}
