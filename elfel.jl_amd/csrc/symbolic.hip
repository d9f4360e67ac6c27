// symbolic.hip -- the symbolic phase of the REUSE method for task tiles
// (plane elasticity and Stokes: two dofs per velocity node) on the device.
//
// The host builder (reuse.hip: adjacency, the pattern from the plain method's
// sort, greedy tile packing, task tables) spends 100-200 ms on a 400x400 T6
// mesh, 40 times the whole FP64 plain assembly.  Here every step is a kernel
// over nodes, rows, row ranges or tiles, and the host reads three scalars:
//
//   1. node -> element adjacency of each field (counts, scan, atomic fill,
//      then each node's list sorted by element: the rank order of sparse()'s
//      sums, src/Assemblers.jl:121-123);
//   2. dof -> (node, component) inverse map (a dof claimed twice: the sort
//      method instead), and per element the check that its columns are
//      distinct dofs;
//   3. each node's sorted column list: the dofs of every element at the node
//      (plus the element's pressure dofs for Stokes velocity rows), sorted
//      and made unique by one wavefront in LDS; the rows of a node share it.
//      The pattern of sparse() on the element blocks is this union, so ptr is
//      the scan of the row lengths and ind the concatenated lists.  The blocks
//      are symmetric (elasticity K; Stokes K, B and B^T), so the CSC arrays
//      are the CSR arrays;
//   4. greedy tile packing of row groups (a node's two rows together) in dof
//      order, one thread per range of 2048 rows: the same limits as the host
//      builder (tasks <= slots, values <= the LDS cap, rows <= 256), a tile
//      never spans two ranges;
//   5. the task tables, one workgroup per tile: within a tile the tasks are
//      laid out rank by rank (each rank step's threads in as few waves as
//      possible), each task's column positions found by binary search in its
//      row and the first contribution to each entry flagged;
//   6. tiles sorted by their first element (the host builder's order) and
//      placed so that launch slot b (XCD b mod 8) walks a contiguous eighth
//      of the sorted list.
//
// Nothing depends on the tile packing but speed: every value is the sum of
// its contributions in element order, whatever tile owns it, so the FP64
// result equals the plain method's bit for bit.
#include "reuse_kernel.h"

namespace {

constexpr int SY_CMAX = 128;  // longest row (entries): the tasks' 7-bit positions
constexpr int SY_SEG = 2048;  // rows per packing range

enum { SYF_DEG = 1, SYF_LONG = 2, SYF_SHARED = 4, SYF_REPEAT = 8, SYF_PAIR = 16 };
enum { SYS_FLAGS = 0, SYS_MAXV = 1, SYS_MAXRANK = 2, SYS_N = 4 };

struct SymIn {
    const int32_t *conn0, *dof0;  // velocity / displacement field: E x nen0, Nn0 x nc (0-based)
    const int32_t *conn1, *dof1;  // Stokes pressure field: E x 3, Np x 1 (null otherwise)
    int64_t E, Nn0, Np, n;
    int nen0, nc, stokes;
};

// element e's columns for a row of field f (f 0: every dof of its nodes, then
// for Stokes its three pressure dofs; f 1: its velocity dofs)
__device__ __forceinline__ int sy_ncol(const SymIn &S, int field)
{
    return S.nen0 * S.nc + ((S.stokes && field == 0) ? 3 : 0);
}
__device__ __forceinline__ int32_t sy_col(const SymIn &S, int64_t e, int j)
{
    const int nd = S.nen0 * S.nc;
    if (j < nd) return S.dof0[(int64_t)S.conn0[e * S.nen0 + j / S.nc] * S.nc + j % S.nc];
    return S.dof1[S.conn1[e * 3 + (j - nd)]];
}

__global__ void k_sy_adj_count(SymIn S, uint32_t *__restrict__ cnt)
{
    const int64_t m0 = S.E * S.nen0, m1 = S.stokes ? S.E * 3 : 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m0 + m1; i += (int64_t)gridDim.x * blockDim.x) {
        const int64_t node = i < m0 ? S.conn0[i] : S.Nn0 + S.conn1[i - m0];
        atomicAdd(&cnt[node], 1u);
    }
}

__global__ void k_sy_adj_fill(SymIn S, const uint32_t *__restrict__ aptr, uint32_t *__restrict__ fill,
                              int32_t *__restrict__ aelem, uint8_t *__restrict__ aloc)
{
    const int64_t m0 = S.E * S.nen0, m1 = S.stokes ? S.E * 3 : 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m0 + m1; i += (int64_t)gridDim.x * blockDim.x) {
        int64_t node, e;
        int k;
        if (i < m0) { node = S.conn0[i]; e = i / S.nen0; k = (int)(i % S.nen0); }
        else { node = S.Nn0 + S.conn1[i - m0]; e = (i - m0) / 3; k = (int)((i - m0) % 3); }
        const uint32_t p = aptr[node] + atomicAdd(&fill[node], 1u);
        aelem[p] = (int32_t)e;
        aloc[p] = (uint8_t)k;
    }
}

// each node's elements in ascending order (insertion sort; at most 16 --
// more turns the method off for the mesh)
__global__ void k_sy_adj_sort(int64_t N, const uint32_t *__restrict__ aptr, int32_t *__restrict__ aelem,
                              uint8_t *__restrict__ aloc, uint32_t *__restrict__ scal)
{
    const int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= N) return;
    const uint32_t p0 = aptr[nd], deg = aptr[nd + 1] - p0;
    if (deg > 16) { atomicOr(&scal[SYS_FLAGS], (uint32_t)SYF_DEG); return; }
    int32_t e[16];
    uint8_t l[16];
    for (uint32_t i = 0; i < deg; i++) { e[i] = aelem[p0 + i]; l[i] = aloc[p0 + i]; }
    for (uint32_t i = 1; i < deg; i++) {
        const int32_t ke = e[i];
        const uint8_t kl = l[i];
        int j = (int)i - 1;
        while (j >= 0 && (e[j] > ke || (e[j] == ke && l[j] > kl))) { e[j + 1] = e[j]; l[j + 1] = l[j]; j--; }
        e[j + 1] = ke;
        l[j + 1] = kl;
    }
    for (uint32_t i = 0; i < deg; i++) { aelem[p0 + i] = e[i]; aloc[p0 + i] = l[i]; }
}

// dof -> node (field 0 nodes 0..Nn0-1, field 1 nodes Nn0..) and field/component
__global__ void k_sy_inv(SymIn S, int32_t *__restrict__ inv_node, uint8_t *__restrict__ inv_fc,
                         uint32_t *__restrict__ scal)
{
    const int64_t m0 = S.Nn0 * S.nc, m1 = S.stokes ? S.Np : 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < m0 + m1; i += (int64_t)gridDim.x * blockDim.x) {
        int32_t d, node;
        uint8_t fc;
        if (i < m0) { d = S.dof0[i]; node = (int32_t)(i / S.nc); fc = (uint8_t)(i % S.nc); }
        else { d = S.dof1[i - m0]; node = (int32_t)(S.Nn0 + (i - m0)); fc = (uint8_t)(1 << 4); }
        if (d < 0 || d >= S.n || atomicCAS(&inv_node[d], -1, node) != -1) {
            atomicOr(&scal[SYS_FLAGS], (uint32_t)SYF_SHARED);
            continue;
        }
        inv_fc[d] = fc;
    }
}

// an element whose columns repeat a dof (a task reads all of its partial
// sums before it writes any, so its columns must be distinct)
__global__ void k_sy_elem_check(SymIn S, uint32_t *__restrict__ scal)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= S.E) return;
    const int nc0 = sy_ncol(S, 0);
    int32_t c[16];
    for (int j = 0; j < nc0; j++) c[j] = sy_col(S, e, j);
    bool rep = false;
    for (int j = 1; j < nc0; j++)
        for (int i = 0; i < j; i++) rep |= c[i] == c[j];
    if (rep) atomicOr(&scal[SYS_FLAGS], (uint32_t)SYF_REPEAT);
}

// One wavefront (a workgroup of 64) per node: the columns of every element
// at the node, bitonic-sorted in LDS, duplicates dropped; the list goes to
// lists[node * SY_CMAX ..] and its length to nodeL[node].
__global__ __launch_bounds__(64) void k_sy_node_cols(SymIn S, int64_t N, const uint32_t *__restrict__ aptr,
                                                     const int32_t *__restrict__ aelem, int32_t *__restrict__ lists,
                                                     uint32_t *__restrict__ nodeL, uint32_t *__restrict__ scal)
{
    __shared__ int32_t s[256];
    const int64_t nd = blockIdx.x;
    const int lane = threadIdx.x;
    const uint32_t p0 = aptr[nd], deg = aptr[nd + 1] - p0;
    if (deg > 16) {  // flagged by k_sy_adj_sort
        if (lane == 0) nodeL[nd] = 0;
        return;
    }
    const int field = nd >= S.Nn0 ? 1 : 0;
    const int ncol = sy_ncol(S, field);
    const int cnt = (int)deg * ncol;  // <= 16 x 15
    int size = 64;
    while (size < cnt) size <<= 1;
    for (int i = lane; i < size; i += 64) {
        int32_t v = INT32_MAX;
        if (i < cnt) v = sy_col(S, aelem[p0 + i / ncol], i % ncol);
        s[i] = v;
    }
    __syncthreads();
    for (int k = 2; k <= size; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int p = lane; p < (size >> 1); p += 64) {
                const int i = (p / j) * 2 * j + (p % j), q = i + j;
                const int32_t a = s[i], b = s[q];
                const bool up = (i & k) == 0;
                if ((a > b) == up) { s[i] = b; s[q] = a; }
            }
            __syncthreads();
        }
    int32_t *out = lists + nd * SY_CMAX;
    int run = 0;
    for (int base = 0; base < size; base += 64) {
        const int i = base + lane;
        const int32_t v = s[i];
        const bool keep = v != INT32_MAX && (i == 0 || v != s[i - 1]);
        const uint64_t m = __ballot(keep);
        const int pos = run + __popcll(m & ((1ull << lane) - 1ull));
        if (keep && pos < SY_CMAX) out[pos] = v;
        run += __popcll(m);
    }
    if (lane == 0) {
        nodeL[nd] = (uint32_t)run;
        if (run > SY_CMAX) atomicOr(&scal[SYS_FLAGS], (uint32_t)SYF_LONG);
    }
}

// Row r: length of its node's list; rinfo packs what the packing walk reads:
// 0xFFFFFFFF for the second row of a node's pair, else deg (bits 0-4), rows
// of the group (bits 5-6), length (bits 8-15), row code (bits 16-17),
// field (bit 18).
__global__ void k_sy_rows(SymIn S, const int32_t *__restrict__ inv_node, const uint8_t *__restrict__ inv_fc,
                          const uint32_t *__restrict__ aptr, const uint32_t *__restrict__ nodeL,
                          uint32_t *__restrict__ rowL, uint32_t *__restrict__ rinfo, uint32_t *__restrict__ scal)
{
    const int64_t r = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (r > S.n) return;
    if (r == S.n) { rowL[r] = 0; return; }
    const int32_t node = inv_node[r];
    const uint8_t fc = node >= 0 ? inv_fc[r] : 0;
    const uint32_t L = node >= 0 ? min(nodeL[node], (uint32_t)SY_CMAX) : 0u;
    rowL[r] = L;
    auto second = [&](int64_t q) {
        return q > 0 && q < S.n && S.nc == 2 && inv_node[q] >= 0 && inv_fc[q] == 1 && inv_node[q - 1] == inv_node[q] &&
               inv_fc[q - 1] == 0;
    };
    if (second(r)) { rinfo[r] = 0xFFFFFFFFu; return; }
    const int nr = second(r + 1) ? 2 : 1;
    const uint32_t deg = node >= 0 ? aptr[node + 1] - aptr[node] : 0u;
    const int field = fc >> 4, comp = fc & 15;
    const int rowcode = nr == 2 ? 2 : (comp == 1 ? 1 : 0);
    rinfo[r] = min(deg, 31u) | ((uint32_t)nr << 5) | (min(L, 255u) << 8) | ((uint32_t)rowcode << 16) |
               ((uint32_t)field << 18);
}

// first row of packing range k: k * SY_SEG, moved off the second row of a pair
__device__ __forceinline__ int64_t sy_seg_start(int64_t k, int64_t n, const uint32_t *rinfo)
{
    int64_t r = k * SY_SEG;
    if (r >= n) return n;
    if (r > 0 && rinfo[r] == 0xFFFFFFFFu) r++;
    return r;
}

// Greedy packing of each range's row groups (one thread per range).  Pass 0
// counts the range's tiles; pass 1 writes each group's tile and LDS offset and
// each tile's rows and sizes (dof-ordered tile numbers from segbase).
__global__ void k_sy_pack(int64_t n, int64_t nseg, const uint32_t *__restrict__ rinfo,
                          const int32_t *__restrict__ ptr, int cap, int vcap, int rowcap, int pass,
                          uint32_t *__restrict__ segtiles, int32_t *__restrict__ gtile, uint32_t *__restrict__ ginfo,
                          int4 *__restrict__ tdo, int4 *__restrict__ tro, const int32_t *__restrict__ inv_node,
                          const uint32_t *__restrict__ aptr, const int32_t *__restrict__ aelem,
                          uint64_t *__restrict__ tkey, uint64_t *__restrict__ tval, uint32_t *__restrict__ scal)
{
    const int64_t k = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= nseg) return;
    const int64_t r0 = sy_seg_start(k, n, rinfo), r1 = sy_seg_start(k + 1, n, rinfo);
    int64_t t = pass ? (int64_t)segtiles[k] : 0;
    const int64_t tfirst = t;
    bool open = false;
    int ty = 0, tw = 0, trows = 0, nranks = 0;
    int64_t row0 = 0;
    int32_t emin = INT32_MAX;  // the tile's first element: its sort key
    bool bad = false;
    auto close = [&]() {
        if (!open) return;
        if (pass) {
            tdo[t] = make_int4((int)row0, ty, ptr[row0], tw);
            tro[t] = make_int4((int)row0, trows, nranks, 0);
            tkey[t] = ((uint64_t)(uint32_t)emin << 32) | (uint64_t)t;
            tval[t] = (uint64_t)t;
        }
        t++;
        open = false;
    };
    for (int64_t r = r0; r < r1; r++) {
        const uint32_t ri = rinfo[r];
        if (ri == 0xFFFFFFFFu) {
            if (pass) gtile[r] = -1;
            continue;
        }
        const int deg = (int)(ri & 31u), nr = (int)((ri >> 5) & 3u), L = (int)((ri >> 8) & 255u);
        if (deg > cap || nr * L > vcap || deg > 16 || L > SY_CMAX) { bad = true; continue; }
        if (open && (ty + deg > cap || tw + nr * L > vcap || trows + nr > rowcap)) close();
        if (!open) { open = true; row0 = r; ty = tw = trows = nranks = 0; emin = INT32_MAX; }
        if (pass) {
            if (deg > 0) emin = min(emin, aelem[aptr[inv_node[r]]]);  // the group's lowest element
            const int l8 = nr == 2 ? L : trows;
            gtile[r] = (int32_t)t;
            ginfo[r] = (uint32_t)tw | ((uint32_t)l8 << 16) | (((ri >> 16) & 3u) << 24) | (((ri >> 18) & 1u) << 26);
        }
        ty += deg;
        tw += nr * L;
        trows += nr;
        nranks = max(nranks, deg);
    }
    close();
    if (!pass) segtiles[k] = (uint32_t)(t - tfirst);
    if (bad) atomicOr(&scal[SYS_FLAGS], (uint32_t)SYF_DEG);
}

// sorted position of each tile (tiles sorted by their first element, so that
// tiles reading the same elements -- a Stokes element's velocity and pressure
// rows lie far apart in dof order -- run close together)
__global__ void k_sy_tpos(int64_t nt, const uint64_t *__restrict__ sorted_tile, uint32_t *__restrict__ tpos)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < nt) tpos[sorted_tile[i]] = (uint32_t)i;
}

// launch slot of the tile at sorted position t: the sorted list is cut into
// 8 contiguous chunks, chunk x runs on XCD x (slots x, x + 8, ...)
__device__ __forceinline__ int64_t sy_slot(int64_t t, int64_t nt)
{
    const int64_t q = nt / 8, rem = nt % 8;
    int64_t x, off;
    if (t < rem * (q + 1)) { x = t / (q + 1); off = x * (q + 1); }
    else { x = rem + (t - rem * (q + 1)) / q; off = rem * (q + 1) + (x - rem) * q; }
    return (t - off) * 8 + x;
}

// The task tables of one tile (a workgroup of 256 threads; thread i owns
// tile-local row i when that row starts a group).
__global__ __launch_bounds__(256) void k_sy_tasks(SymIn S, int64_t nt, int cap, int ncolp,
                                                  const uint32_t *__restrict__ tpos,
                                                  const int4 *__restrict__ tdo, const int4 *__restrict__ tro,
                                                  const uint32_t *__restrict__ rinfo, const uint32_t *__restrict__ ginfo,
                                                  const int32_t *__restrict__ inv_node,
                                                  const uint32_t *__restrict__ aptr, const int32_t *__restrict__ aelem,
                                                  const uint8_t *__restrict__ aloc, const int32_t *__restrict__ ptr,
                                                  const int32_t *__restrict__ ind, int4 *__restrict__ tiles,
                                                  int4 *__restrict__ tile_rows, int32_t *__restrict__ task_elem,
                                                  uint32_t *__restrict__ task_info, uint32_t *__restrict__ task_pos,
                                                  uint32_t *__restrict__ scal)
{
    __shared__ uint16_t idx[16][256];
    __shared__ int rank_off[17];
    __shared__ int wsum[4];
    const int64_t t = blockIdx.x;
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int4 d = tdo[t], rr = tro[t];
    const int64_t b = sy_slot(tpos[t], nt);
    const int64_t r = (int64_t)rr.x + tid;
    int deg = 0;
    uint32_t ri = 0xFFFFFFFFu;
    if (tid < rr.y) {
        ri = rinfo[r];
        if (ri != 0xFFFFFFFFu) deg = (int)(ri & 31u);
    }
    // rank by rank: the position of this row's rank-s task among the tile's
    // rank-s tasks (block scan of deg > s), and where rank s starts
    if (tid == 0) rank_off[0] = 0;
    for (int s = 0; s < rr.z; s++) {
        const bool f = deg > s;
        const uint64_t m = __ballot(f);
        const int below = __popcll(m & ((1ull << lane) - 1ull));
        if (lane == 0) wsum[w] = __popcll(m);
        __syncthreads();
        int pre = 0, tot = 0;
        for (int k = 0; k < 4; k++) { pre += k < w ? wsum[k] : 0; tot += wsum[k]; }
        idx[s][tid] = (uint16_t)(pre + below);
        if (tid == 0) rank_off[s + 1] = rank_off[s] + tot;
        __syncthreads();
    }
    const int64_t slot0 = b * cap;
    if (tid == 0) {
        tiles[b] = make_int4((int)slot0, d.y, d.z, d.w);
        tile_rows[b] = make_int4(rr.x, rr.y, rr.z, 0);
        atomicMax(&scal[SYS_MAXV], (uint32_t)d.w);
        atomicMax(&scal[SYS_MAXRANK], (uint32_t)rr.z);
    }
    // pad slots: no task
    const int nw = ncolp / 4;
    for (int k = d.y + tid; k < cap; k += 256) {
        task_elem[slot0 + k] = 0;
        task_info[slot0 + k] = TASK_PAD;
        for (int q = 0; q < nw; q++) task_pos[(slot0 + k) * nw + q] = 0u;
    }
    if (deg == 0) return;
    const uint32_t gi = ginfo[r];
    const int vbase = (int)(gi & 0xFFFFu), l8 = (int)((gi >> 16) & 255u), rowcode = (int)((gi >> 24) & 3u);
    const int field = (int)((gi >> 26) & 1u);
    const int ncol = sy_ncol(S, field);
    const int32_t node = inv_node[r];
    const uint32_t a0 = aptr[node];
    const int32_t *row = ind + ptr[r];
    const int L = ptr[r + 1] - ptr[r];
    uint32_t seen[4] = {0u, 0u, 0u, 0u};
    bool bad = false;
    for (int s = 0; s < deg; s++) {
        const int32_t e = aelem[a0 + s];
        const int a = aloc[a0 + s];
        const int64_t slot = slot0 + rank_off[s] + idx[s][tid];
        task_elem[slot] = e;
        task_info[slot] = (uint32_t)a | ((uint32_t)s << 3) | ((uint32_t)rowcode << 7) | ((uint32_t)field << 9) |
                          ((uint32_t)vbase << 10) | ((uint32_t)l8 << 24);
        uint32_t pw[4] = {0u, 0u, 0u, 0u};
        for (int j = 0; j < ncol; j++) {
            const int32_t col = sy_col(S, e, j);
            int lo = 0, hi = L;  // first position with row[pos] >= col
            while (lo < hi) {
                const int mid = (lo + hi) >> 1;
                if (row[mid] < col) lo = mid + 1;
                else hi = mid;
            }
            if (lo >= L || row[lo] != col) { bad = true; continue; }
            uint32_t pv = (uint32_t)lo;
            if (!((seen[lo >> 5] >> (lo & 31)) & 1u)) { seen[lo >> 5] |= 1u << (lo & 31); pv |= 0x80u; }
            pw[j >> 2] |= pv << (8 * (j & 3));
        }
        for (int q = 0; q < nw; q++) task_pos[slot * nw + q] = pw[q];
    }
    if (bad) atomicOr(&scal[SYS_FLAGS], (uint32_t)SYF_PAIR);
}

__global__ void k_sy_fill_ind(int64_t n, const int32_t *__restrict__ inv_node, const int32_t *__restrict__ ptr,
                              const int32_t *__restrict__ lists, int32_t *__restrict__ ind)
{
    const int64_t r = blockIdx.x;
    const int32_t node = inv_node[r];
    if (node < 0) return;
    const int L = ptr[r + 1] - ptr[r];
    const int32_t *src = lists + (int64_t)node * SY_CMAX;
    for (int i = threadIdx.x; i < L; i += 64) ind[ptr[r] + i] = src[i];
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }

}  // namespace

// The pattern and the task tables of a task-tile problem, built on the device
// into c->pat.  Returns 0 with c->pat.fallback set when the mesh does not fit
// the tiles (a node in more than 16 elements, a row of more than 128 entries,
// a dof shared by two nodes, an element that repeats a dof).
int device_symbolic(elfel_hip_ctx *c, const elfel_hip_problem *pb, int block)
{
    Pattern &P = c->pat;
    const bool stokes = pb->kind == ELFEL_STOKES_GEN || pb->kind == ELFEL_STOKES_VECLAP;
    const Field &F0 = c->f[0];
    const Field *F1 = stokes ? &c->f[1] : nullptr;
    SymIn S;
    S.conn0 = (const int32_t *)F0.conn.p;
    S.dof0 = (const int32_t *)F0.dof.p;
    S.conn1 = stokes ? (const int32_t *)F1->conn.p : nullptr;
    S.dof1 = stokes ? (const int32_t *)F1->dof.p : nullptr;
    S.E = F0.E;
    S.Nn0 = F0.Nn;
    S.Np = stokes ? F1->Nn : 0;
    S.n = c->nrow;
    S.nen0 = F0.nen;
    S.nc = F0.ncomp;
    S.stokes = stokes ? 1 : 0;
    if (S.nc != 2 || S.nen0 * S.nc + (stokes ? 3 : 0) > 16 || (stokes && (F1->nen != 3 || F1->E != S.E)))
        return set_err(c, ELFEL_EINVAL, "internal: device symbolic phase for vector task tiles only");
    const int64_t N = S.Nn0 + S.Np, n = S.n;
    const int64_t nent = S.E * S.nen0 + (stokes ? S.E * 3 : 0);
    const int cap = block;  // one task per thread (elasticity, Stokes)
    const int vcap = std::min(std::max(env_int("ELFEL_HIP_VCAP", stokes ? 20 : REUSE_VCAP_PER_THREAD), 8), 24) * cap;
    const int rowcap = std::min(cap, REUSE_MAX_TILE_ROWS);
    const int ncolp = (S.nen0 * S.nc + (stokes ? 3 : 0) + 3) & ~3;
    const int64_t nseg = std::max<int64_t>(1, (n + SY_SEG - 1) / SY_SEG);
    // scratch, one allocation: adjacency, inverse map, node lists, row data,
    // packing ranges, dof-ordered tiles (at most one per row)
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += align256(bytes); return o; };
    const size_t o_scal = take(4 * SYS_N), o_aptr = take(4 * (N + 1)), o_fill = take(4 * N), o_aelem = take(4 * nent),
                 o_aloc = take(nent), o_inv = take(4 * n), o_fc = take(n), o_nodeL = take(4 * N),
                 o_lists = take(4 * (size_t)N * SY_CMAX), o_rinfo = take(4 * n), o_ginfo = take(4 * n),
                 o_gtile = take(4 * n), o_seg = take(4 * (nseg + 1)), o_tdo = take(16 * (size_t)std::max<int64_t>(n, 1)),
                 o_tro = take(16 * (size_t)std::max<int64_t>(n, 1)), o_k0 = take(8 * (size_t)std::max<int64_t>(n, 1)),
                 o_k1 = take(8 * (size_t)std::max<int64_t>(n, 1)), o_v0 = take(8 * (size_t)std::max<int64_t>(n, 1)),
                 o_v1 = take(8 * (size_t)std::max<int64_t>(n, 1)), o_tpos = take(4 * (size_t)std::max<int64_t>(n, 1));
    int rc;
    if ((rc = ensure(c, P.sym, off))) return rc;
    unsigned char *base = (unsigned char *)P.sym.p;
    uint32_t *scal = (uint32_t *)(base + o_scal), *aptr = (uint32_t *)(base + o_aptr), *fill = (uint32_t *)(base + o_fill);
    int32_t *aelem = (int32_t *)(base + o_aelem), *inv_node = (int32_t *)(base + o_inv), *lists = (int32_t *)(base + o_lists);
    uint8_t *aloc = base + o_aloc, *inv_fc = base + o_fc;
    uint32_t *nodeL = (uint32_t *)(base + o_nodeL), *rinfo = (uint32_t *)(base + o_rinfo), *ginfo = (uint32_t *)(base + o_ginfo);
    int32_t *gtile = (int32_t *)(base + o_gtile);
    uint32_t *segtiles = (uint32_t *)(base + o_seg);
    int4 *tdo = (int4 *)(base + o_tdo), *tro = (int4 *)(base + o_tro);
    uint64_t *k0 = (uint64_t *)(base + o_k0), *k1 = (uint64_t *)(base + o_k1);
    uint64_t *v0 = (uint64_t *)(base + o_v0), *v1 = (uint64_t *)(base + o_v1);
    uint32_t *tpos = (uint32_t *)(base + o_tpos);
    hipStream_t st = c->stream;
    const int g = 256;
    const int grid_ent = std::max(1, std::min(ceil_div(nent, g), 8192));
    HIPCHK(c, hipMemsetAsync(scal, 0, 4 * SYS_N, st));
    HIPCHK(c, hipMemsetAsync(aptr, 0, 4 * (N + 1), st));
    HIPCHK(c, hipMemsetAsync(fill, 0, 4 * N, st));
    HIPCHK(c, hipMemsetAsync(inv_node, 0xFF, 4 * n, st));
    // 1. adjacency
    hipLaunchKernelGGL(k_sy_adj_count, dim3(grid_ent), dim3(g), 0, st, S, aptr);
    if ((rc = exclusive_scan_u32(c, aptr, N + 1, nullptr))) return rc;
    hipLaunchKernelGGL(k_sy_adj_fill, dim3(grid_ent), dim3(g), 0, st, S, (const uint32_t *)aptr, fill, aelem, aloc);
    if (N > 0)
        hipLaunchKernelGGL(k_sy_adj_sort, dim3(ceil_div(N, g)), dim3(g), 0, st, N, (const uint32_t *)aptr, aelem, aloc,
                           scal);
    // 2. inverse dof map, element check
    const int64_t ndofref = S.Nn0 * S.nc + S.Np;
    hipLaunchKernelGGL(k_sy_inv, dim3(std::max(1, std::min(ceil_div(ndofref, g), 8192))), dim3(g), 0, st, S, inv_node,
                       inv_fc, scal);
    if (S.E > 0) hipLaunchKernelGGL(k_sy_elem_check, dim3(ceil_div(S.E, g)), dim3(g), 0, st, S, scal);
    // 3. node column lists, row lengths, ptr
    if (N > 0)
        hipLaunchKernelGGL(k_sy_node_cols, dim3((unsigned)N), dim3(64), 0, st, S, N, (const uint32_t *)aptr,
                           (const int32_t *)aelem, lists, nodeL, scal);
    if ((rc = ensure(c, P.ptr, 4 * (n + 1)))) return rc;
    int32_t *ptr = (int32_t *)P.ptr.p;
    hipLaunchKernelGGL(k_sy_rows, dim3(ceil_div(n + 1, g)), dim3(g), 0, st, S, (const int32_t *)inv_node,
                       (const uint8_t *)inv_fc, (const uint32_t *)aptr, (const uint32_t *)nodeL, (uint32_t *)ptr, rinfo,
                       scal);
    if ((rc = exclusive_scan_u32(c, (uint32_t *)ptr, n + 1, nullptr))) return rc;
    uint32_t hv[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(&hv[0], ptr + n, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(&hv[1], scal + SYS_FLAGS, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const int64_t nnz = hv[0];
    P.fallback = false;
    if (hv[1]) {
        P.fallback = true;
        const uint32_t f = hv[1];
        P.fallback_note = std::string("plain FP64 method: ") +
                          ((f & SYF_SHARED) ? "a dof shared by two nodes or components"
                           : (f & SYF_REPEAT) ? "an element with a repeated dof"
                                              : "node too highly connected for the reuse tiles");
        P.nnz = nnz;
        return 0;
    }
    if ((rc = ensure(c, P.ind, 4 * (size_t)std::max<int64_t>(nnz, 1)))) return rc;
    if (n > 0)
        hipLaunchKernelGGL(k_sy_fill_ind, dim3((unsigned)n), dim3(64), 0, st, n, (const int32_t *)inv_node,
                           (const int32_t *)ptr, (const int32_t *)lists, (int32_t *)P.ind.p);
    // 4. packing
    hipLaunchKernelGGL(k_sy_pack, dim3(ceil_div(nseg, 64)), dim3(64), 0, st, n, nseg, (const uint32_t *)rinfo,
                       (const int32_t *)ptr, cap, vcap, rowcap, 0, segtiles, gtile, ginfo, tdo, tro,
                       (const int32_t *)inv_node, (const uint32_t *)aptr, (const int32_t *)aelem, k0, v0, scal);
    HIPCHK(c, hipMemsetAsync(segtiles + nseg, 0, 4, st));
    if ((rc = exclusive_scan_u32(c, segtiles, nseg + 1, nullptr))) return rc;
    hipLaunchKernelGGL(k_sy_pack, dim3(ceil_div(nseg, 64)), dim3(64), 0, st, n, nseg, (const uint32_t *)rinfo,
                       (const int32_t *)ptr, cap, vcap, rowcap, 1, segtiles, gtile, ginfo, tdo, tro,
                       (const int32_t *)inv_node, (const uint32_t *)aptr, (const int32_t *)aelem, k0, v0, scal);
    uint32_t hn[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(&hn[0], segtiles + nseg, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(&hn[1], scal + SYS_FLAGS, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (hn[1]) {
        P.fallback = true;
        P.fallback_note = "plain FP64 method: node too highly connected for the reuse tiles";
        P.nnz = nnz;
        return 0;
    }
    const int64_t nt = hn[0];
    // 6. launch order: tiles sorted by first element (key: element << 32 | tile)
    if (nt > 1) {
        if ((rc = radix_sort_pairs(c, k0, v0, k1, v1, nt, 32 + bits_for(std::max<int64_t>(S.E, 1))))) return rc;
        hipLaunchKernelGGL(k_sy_tpos, dim3(ceil_div(nt, g)), dim3(g), 0, st, nt, (const uint64_t *)v0, tpos);
    } else {
        HIPCHK(c, hipMemsetAsync(tpos, 0, 4, st));
    }
    // 5. task tables in launch order
    const size_t slots = (size_t)nt * cap;
    if ((rc = ensure(c, P.tile, sizeof(int4) * std::max<int64_t>(nt, 1))) ||
        (rc = ensure(c, P.tile_rows, sizeof(int4) * std::max<int64_t>(nt, 1))) ||
        (rc = ensure(c, P.task_elem, 4 * std::max<size_t>(slots, 1))) ||
        (rc = ensure(c, P.task_info, 4 * std::max<size_t>(slots, 1))) ||
        (rc = ensure(c, P.task_pos, std::max<size_t>(slots * ncolp, 4))))
        return rc;
    if (nt > 0)
        hipLaunchKernelGGL(k_sy_tasks, dim3((unsigned)nt), dim3(256), 0, st, S, nt, cap, ncolp, (const uint32_t *)tpos,
                           (const int4 *)tdo,
                           (const int4 *)tro, (const uint32_t *)rinfo, (const uint32_t *)ginfo, (const int32_t *)inv_node,
                           (const uint32_t *)aptr, (const int32_t *)aelem, (const uint8_t *)aloc, (const int32_t *)ptr,
                           (const int32_t *)P.ind.p, (int4 *)P.tile.p, (int4 *)P.tile_rows.p, (int32_t *)P.task_elem.p,
                           (uint32_t *)P.task_info.p, (uint32_t *)P.task_pos.p, scal);
    HIPCHK(c, hipGetLastError());
    uint32_t hs[SYS_N] = {0, 0, 0, 0};
    HIPCHK(c, hipMemcpyAsync(hs, scal, 4 * SYS_N, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (hs[SYS_FLAGS])
        return set_err(c, ELFEL_EINVAL, "internal: an element column is missing from the device pattern");
    release(P.rrec); release(P.tconn); release(P.txy); release(P.truns); release(P.runs); release(P.tile_elem);
    P.rw = 0;
    P.tstride = cap;
    P.tgeom = false;
    P.has_runs = false;
    P.elem8 = false;
    P.maxu = 0;
    P.ntiles = nt;
    P.ntasks = (int64_t)slots;
    P.ncolmax = ncolp;
    P.maxrank = (int)hs[SYS_MAXRANK];
    P.lds_vals = (int)hs[SYS_MAXV];
    P.nnz = nnz;
    return 0;
}
