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
//      and made unique by one wavefront in registers (a bitonic network over
//      shuffles); the rows of a node share it.  The pattern of sparse() on
//      the element blocks is this union, so ptr is the scan of the row
//      lengths; a second sort writes the lists into ind and, for every
//      (element, column) at the node, its position in the list.  The blocks
//      are symmetric (elasticity K; Stokes K, B and B^T), so the CSC arrays
//      are the CSR arrays;
//   4. greedy tile packing of row groups (a node's two rows together) in dof
//      order, one wavefront per range of 2048 rows (wave prefix sums find
//      where each tile overflows): the same limits as the host builder
//      (tasks <= slots, values <= the LDS cap, rows <= 256), a tile never
//      spans two ranges;
//   5. tiles sorted by their first element (the host builder's order) and
//      placed so that launch slot b (XCD b mod 8) walks a contiguous eighth
//      of the sorted list;
//   6. the task tables, one workgroup per tile: within a tile the tasks are
//      laid out rank by rank (each rank step's threads in as few waves as
//      possible), with their column positions from step 3 and the first
//      contribution to each entry flagged.
//
// Nothing depends on the tile packing but speed: every value is the sum of
// its contributions in element order, whatever tile owns it, so the FP64
// result equals the plain method's bit for bit.
#include "reuse_kernel.h"

namespace {

constexpr int SY_CMAX = 128;  // longest row (entries): the tasks' 7-bit positions
constexpr int SY_SEG = 2048;  // rows per packing range

enum { SYF_DEG = 1, SYF_LONG = 2, SYF_SHARED = 4, SYF_REPEAT = 8 };
enum { SYS_FLAGS = 0, SYS_MAXV = 1, SYS_MAXRANK = 2, SYS_MAXU = 3, SYS_MAXDEG = 4, SYS_MAXL = 5, SYS_N = 8 };

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
// sums before it writes any, so its columns must be distinct); edof keeps
// every element's columns (field 0 order) for the node lists
__global__ void k_sy_elem_check(SymIn S, int pstride, int32_t *__restrict__ edof, uint32_t *__restrict__ scal)
{
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= S.E) return;
    const int nc0 = sy_ncol(S, 0);
    int32_t c[16];
    for (int j = 0; j < nc0; j++) {
        c[j] = sy_col(S, e, j);
        edof[e * pstride + j] = c[j];  // the element's columns, for the node lists
    }
    bool rep = false;
    for (int j = 1; j < nc0; j++)
        for (int i = 0; i < j; i++) rep |= c[i] == c[j];
    if (rep) atomicOr(&scal[SYS_FLAGS], (uint32_t)SYF_REPEAT);
}

// Bitonic sort of 64 R values across one wavefront, value i = r * 64 + lane
// in register r of lane i mod 64: exchanges between lanes by shuffles,
// between registers in place; no LDS, no barrier.
__device__ __forceinline__ uint64_t sy_shfl_xor(uint64_t v, int m)
{
    const uint32_t lo = (uint32_t)__shfl_xor((int)(uint32_t)v, m, 64);
    const uint32_t hi = (uint32_t)__shfl_xor((int)(uint32_t)(v >> 32), m, 64);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t sy_shfl(uint64_t v, int src)
{
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, src, 64);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), src, 64);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint64_t sy_shfl_up1(uint64_t v)
{
    const uint32_t lo = (uint32_t)__shfl_up((int)(uint32_t)v, 1, 64);
    const uint32_t hi = (uint32_t)__shfl_up((int)(uint32_t)(v >> 32), 1, 64);
    return ((uint64_t)hi << 32) | lo;
}
__device__ __forceinline__ uint32_t sy_shfl_xor(uint32_t v, int m) { return (uint32_t)__shfl_xor((int)v, m, 64); }
__device__ __forceinline__ uint32_t sy_shfl(uint32_t v, int src) { return (uint32_t)__shfl((int)v, src, 64); }
__device__ __forceinline__ uint32_t sy_shfl_up1(uint32_t v) { return (uint32_t)__shfl_up((int)v, 1, 64); }

template <int R, typename K, int SIZE = 64 * R>
__device__ __forceinline__ void sy_wave_sort(K (&v)[R], int lane)
{
    // SIZE < 64 (R = 1): each group of SIZE lanes sorts its own values
    // (partners lane ^ j, j < SIZE, stay inside the group; the direction
    // comes from the lane's index in its group)
    const int dl = SIZE < 64 ? (lane & (SIZE - 1)) : lane;
#pragma unroll
    for (int k = 2; k <= SIZE; k <<= 1) {
#pragma unroll
        for (int j = k >> 1; j > 0; j >>= 1) {
            if (j >= 64) {
                const int rj = j >> 6;
#pragma unroll
                for (int r = 0; r < R; r++) {
                    if ((r & rj) != 0) continue;
                    const bool up = ((r * 64 + dl) & k) == 0;
                    const K a = v[r], b = v[r | rj];
                    const bool sw = (a > b) == up;
                    v[r] = sw ? b : a;
                    v[r | rj] = sw ? a : b;
                }
            } else {
#pragma unroll
                for (int r = 0; r < R; r++) {
                    const K p = sy_shfl_xor(v[r], j);
                    const bool up = ((r * 64 + dl) & k) == 0;
                    const bool takemin = ((dl & j) == 0) == up;
                    v[r] = takemin ? (v[r] < p ? v[r] : p) : (v[r] < p ? p : v[r]);
                }
            }
        }
    }
}

// One wavefront per node.  Pass 0: the length of the node's column list
// (nodeL).  Pass 1: the columns of every element at the node (key =
// dof << 8 | candidate index, candidate = rank * ncol + column), sorted,
// duplicates of a dof dropped; the list goes into ind at each of the node's
// rows, and each candidate's position in it into npos (one byte, at
// (aptr[node] + rank) * pstride + column).
template <int R, typename K, int W = 64>
__device__ __forceinline__ void sy_node_cols_body(const SymIn &S, int64_t nd, int lane, uint32_t p0, int cnt, int ncol,
                                                  const int32_t *__restrict__ aelem, const int32_t *__restrict__ edof,
                                                  int pass, uint32_t *__restrict__ nodeL,
                                                  const int32_t *__restrict__ ptr, int32_t *__restrict__ ind,
                                                  uint8_t *__restrict__ npos, int pstride, uint32_t *__restrict__ scal)
{
    // W < 64: the node is one of 64 / W in the wavefront, on lanes
    // [g W, g W + W); its candidate i sits on lane g W + i (R = 1)
    static_assert(W == 64 || R == 1, "groups of fewer than 64 lanes hold one value a lane");
    const int glane = lane % W;
    const uint64_t gmask = W == 64 ? ~0ull : (((1ull << (W & 63)) - 1ull) << (lane - glane));
    // keys: pass 0 node numbers, pass 1 dof << 8 | candidate (32-bit keys
    // when every dof is below 2^23)
    constexpr K PAD = (K)~(K)0;
    K v[R];
    // pass 0 counts the distinct nodes of the elements at the node (no dof is
    // shared -- that case falls back -- so each velocity node brings nc dofs
    // and each pressure node one); pass 1 sorts the dofs themselves
    const int nnv = S.nen0 + ((S.stokes && nd < S.Nn0) ? 3 : 0);
#pragma unroll
    for (int r = 0; r < R; r++) {
        const int i = r * 64 + glane;
        v[r] = PAD;
        if (pass == 0) {
            if (i < cnt) {
                const int64_t e = aelem[p0 + i / nnv];
                const int k = i % nnv;
                v[r] = k < S.nen0 ? (K)S.conn0[e * S.nen0 + k] : (K)(S.Nn0 + S.conn1[e * 3 + (k - S.nen0)]);
            }
        } else if (i < cnt) {
            v[r] = ((K)(uint32_t)edof[(int64_t)aelem[p0 + i / ncol] * pstride + i % ncol] << 8) | (K)i;
        }
    }
    sy_wave_sort<R, K, W * R>(v, lane);
    // a value is kept when it differs from the one before; every candidate's
    // position is the number of kept values up to it, less one
    int run = 0;
    int32_t rows[2] = {-1, -1};
    int nrows = 0;
    if (pass == 1) {
        if (nd < S.Nn0) {
            nrows = S.nc;
            for (int c = 0; c < S.nc; c++) rows[c] = S.dof0[nd * S.nc + c];
        } else {
            nrows = 1;
            rows[0] = S.dof1[nd - S.Nn0];
        }
    }
    const int sh = pass == 0 ? 0 : 8;
    int nl = 0;  // pass 0: the list length
#pragma unroll
    for (int r = 0; r < R; r++) {
        const K up = sy_shfl_up1(v[r]);
        const K last = r > 0 ? sy_shfl(v[r > 0 ? r - 1 : 0], 63) : PAD;
        const K prev = lane == 0 ? last : up;
        const bool real = v[r] != PAD;
        const bool keep = real && (r * 64 + glane == 0 || (v[r] >> sh) != (prev >> sh));
        if (pass == 0) {
            const uint64_t mv = __ballot(keep && v[r] < (K)S.Nn0) & gmask,
                           mp = __ballot(keep && v[r] >= (K)S.Nn0) & gmask;
            nl += S.nc * __popcll(mv) + __popcll(mp);
            continue;
        }
        const uint64_t m = __ballot(keep) & gmask;
        const int incl = run + __popcll(m & ((2ull << lane) - 1ull));  // lanes of the group up to this one
        if (pass == 1 && real) {
            const int pos = incl - 1;
            if (keep && pos < SY_CMAX)
                for (int k = 0; k < nrows; k++) ind[ptr[rows[k]] + pos] = (int32_t)(v[r] >> 8);
            const int cand = (int)(v[r] & 255u);
            npos[((int64_t)p0 + cand / ncol) * pstride + cand % ncol] = (uint8_t)min(pos, 255);
        }
        run += __popcll(m);
    }
    if (pass == 0 && glane == 0) {
        nodeL[nd] = (uint32_t)nl;
        if (nl > SY_CMAX) atomicOr(&scal[SYS_FLAGS], (uint32_t)SYF_LONG);
    }
}

__global__ __launch_bounds__(256) void k_sy_node_cols(SymIn S, int64_t N, const uint32_t *__restrict__ aptr,
                                                      const int32_t *__restrict__ aelem,
                                                      const int32_t *__restrict__ edof, int pass, int npw,
                                                      uint32_t *__restrict__ nodeL, const int32_t *__restrict__ ptr,
                                                      int32_t *__restrict__ ind, uint8_t *__restrict__ npos,
                                                      int pstride, uint32_t *__restrict__ scal)
{
    // four wavefronts per workgroup (they share nothing), npw nodes per
    // wavefront: 1, or 2 / 4 when every node has at most 32 / 16 candidates
    const int lane = threadIdx.x & 63;
    const int64_t nd = ((int64_t)blockIdx.x * 4 + (threadIdx.x >> 6)) * npw + lane / (64 / npw);
    if (nd >= N) return;
    const uint32_t p0 = aptr[nd], deg = aptr[nd + 1] - p0;
    if (deg > 16) {  // flagged by k_sy_adj_sort
        if (lane % (64 / npw) == 0 && pass == 0) nodeL[nd] = 0;
        return;
    }
    const int ncol = sy_ncol(S, nd >= S.Nn0 ? 1 : 0);
    // candidates: pass 0 the nodes, pass 1 the columns of each element (<= 16 x 15)
    const int cnt = (int)deg * (pass == 0 ? S.nen0 + ((S.stokes && nd < S.Nn0) ? 3 : 0) : ncol);
    const bool k32 = pass == 0 || S.n <= (1 << 23);
    if (npw == 4) {
        if (k32) sy_node_cols_body<1, uint32_t, 16>(S, nd, lane, p0, cnt, ncol, aelem, edof, pass, nodeL, ptr, ind, npos, pstride, scal);
        else sy_node_cols_body<1, uint64_t, 16>(S, nd, lane, p0, cnt, ncol, aelem, edof, pass, nodeL, ptr, ind, npos, pstride, scal);
        return;
    }
    if (npw == 2) {
        if (k32) sy_node_cols_body<1, uint32_t, 32>(S, nd, lane, p0, cnt, ncol, aelem, edof, pass, nodeL, ptr, ind, npos, pstride, scal);
        else sy_node_cols_body<1, uint64_t, 32>(S, nd, lane, p0, cnt, ncol, aelem, edof, pass, nodeL, ptr, ind, npos, pstride, scal);
        return;
    }
#define SY_BODY(RR, KK) \
    sy_node_cols_body<RR, KK>(S, nd, lane, p0, cnt, ncol, aelem, edof, pass, nodeL, ptr, ind, npos, pstride, scal)
    if (k32) {
        if (cnt <= 64) SY_BODY(1, uint32_t);
        else if (cnt <= 128) SY_BODY(2, uint32_t);
        else SY_BODY(4, uint32_t);
    } else {
        if (cnt <= 64) SY_BODY(1, uint64_t);
        else if (cnt <= 128) SY_BODY(2, uint64_t);
        else SY_BODY(4, uint64_t);
    }
#undef SY_BODY
}

// Row r: length of its node's list; rinfo packs what the packing walk reads
// (with gmin, the group's lowest element):
// 0xFFFFFFFF for the second row of a node's pair, else deg (bits 0-4), rows
// of the group (bits 5-6), length (bits 8-15), row code (bits 16-17),
// field (bit 18).
__global__ void k_sy_rows(SymIn S, const int32_t *__restrict__ inv_node, const uint8_t *__restrict__ inv_fc,
                          const uint32_t *__restrict__ aptr, const int32_t *__restrict__ aelem,
                          const uint32_t *__restrict__ nodeL, uint32_t *__restrict__ rowL, uint32_t *__restrict__ rinfo,
                          int32_t *__restrict__ gmin)
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
    gmin[r] = deg > 0 ? aelem[aptr[node]] : INT32_MAX;  // the group's lowest element
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

// inclusive sums over the wavefront
__device__ __forceinline__ int sy_wave_incl(int v, int lane)
{
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int u = __shfl_up(v, d, 64);
        if (lane >= d) v += u;
    }
    return v;
}
__device__ __forceinline__ int sy_wave_min(int v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = min(v, __shfl_xor(v, d, 64));
    return v;
}
__device__ __forceinline__ int sy_wave_max(int v)
{
#pragma unroll
    for (int d = 32; d > 0; d >>= 1) v = max(v, __shfl_xor(v, d, 64));
    return v;
}

// Greedy packing of each range's row groups, one wavefront per range: the
// rows are read 64 at a time (one per lane); prefix sums of the groups'
// tasks, values and rows say where the open tile overflows, the tile is
// closed before that group and the rest of the chunk goes on from it, so the
// result is the row-by-row greedy walk's.  Each group gets its LDS offset and
// tile-local row (ginfo); each tile's rows, sizes and first element go to
// sparse index k * SY_SEG + l (a range of SY_SEG rows has at most SY_SEG
// tiles); segtiles[k] receives the range's tile count.
__global__ __launch_bounds__(64) void k_sy_pack(int64_t n, int64_t nseg, const uint32_t *__restrict__ rinfo,
                                                const int32_t *__restrict__ gmin, const int32_t *__restrict__ ptr,
                                                int cap, int vcap, int rowcap, uint32_t *__restrict__ segtiles,
                                                uint32_t *__restrict__ ginfo, int4 *__restrict__ tds,
                                                int4 *__restrict__ trs, int32_t *__restrict__ tmin,
                                                uint32_t *__restrict__ scal)
{
    const int64_t k = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t r0 = sy_seg_start(k, n, rinfo), r1 = sy_seg_start(k + 1, n, rinfo);
    int64_t t = k * SY_SEG;
    const int64_t tfirst = t;
    // the open tile (wave-uniform)
    bool open = false;
    int ty = 0, tw = 0, trows = 0, nranks = 0, emin = INT32_MAX;
    int64_t row0 = 0;
    bool bad = false;
    for (int64_t c0 = r0; c0 < r1; c0 += 64) {
        const int64_t r = c0 + lane;
        const uint32_t ri = r < r1 ? rinfo[r] : 0xFFFFFFFFu;
        const int32_t gm = r < r1 ? gmin[r] : INT32_MAX;
        int deg = (int)(ri & 31u), nr = (int)((ri >> 5) & 3u), L = (int)((ri >> 8) & 255u);
        bool isg = ri != 0xFFFFFFFFu;
        if (isg && (deg > cap || nr * L > vcap || deg > 16 || L > SY_CMAX)) { bad = true; isg = false; }
        if (!isg) deg = nr = L = 0;
        int pos = 0;  // first lane of the chunk not yet placed
        for (;;) {
            const bool mine = isg && lane >= pos;
            const int ad = mine ? deg : 0, av = mine ? nr * L : 0, ar = mine ? nr : 0;
            const int sd = sy_wave_incl(ad, lane), sv = sy_wave_incl(av, lane), sr = sy_wave_incl(ar, lane);
            const bool fail = mine && (ty + sd > cap || tw + sv > vcap || trows + sr > rowcap);
            const uint64_t fm = __ballot(fail);
            // the group that overflows an open tile closes it; a group alone
            // always fits (checked above), so a fresh tile's first group never fails
            const int j = fm ? __ffsll((long long)fm) - 1 : 64;  // lanes pos .. j-1 join the tile
            const bool join = mine && lane < j;
            const uint64_t jm = __ballot(join);
            if (jm) {
                if (!open) {
                    open = true;
                    row0 = c0 + (__ffsll((long long)jm) - 1);
                }
                if (join) {
                    const int l8 = nr == 2 ? L : trows + (sr - ar);
                    ginfo[r] = (uint32_t)(tw + (sv - av)) | ((uint32_t)l8 << 16) | (((ri >> 16) & 3u) << 24) |
                               (((ri >> 18) & 1u) << 26);
                }
                const int last = 63 - __clzll((long long)jm);  // the last joining lane
                ty += __shfl(sd, last, 64);
                tw += __shfl(sv, last, 64);
                trows += __shfl(sr, last, 64);
                nranks = max(nranks, sy_wave_max(join ? deg : 0));
                emin = min(emin, sy_wave_min(join ? gm : INT32_MAX));
            }
            if (!fm) break;
            // close the tile before lane j's group; lane j's group opens the next
            if (open) {
                if (lane == 0) {
                    tds[t] = make_int4((int)row0, ty, ptr[row0], tw);
                    trs[t] = make_int4((int)row0, trows, nranks, 0);
                    tmin[t] = emin;
                }
                t++;
            }
            open = false;
            ty = tw = trows = nranks = 0;
            emin = INT32_MAX;
            pos = j;
        }
    }
    if (open) {
        if (lane == 0) {
            tds[t] = make_int4((int)row0, ty, ptr[row0], tw);
            trs[t] = make_int4((int)row0, trows, nranks, 0);
            tmin[t] = emin;
        }
        t++;
    }
    if (lane == 0) segtiles[k] = (uint32_t)(t - tfirst);
    if (__ballot(bad) && lane == 0) atomicOr(&scal[SYS_FLAGS], (uint32_t)SYF_DEG);
}

// Dense tile numbers: range k's tiles follow base[k] (the scanned counts);
// the sort key is the first element, then the tile number.
__global__ void k_sy_dense(int64_t nseg, const uint32_t *__restrict__ base, const int4 *__restrict__ tds,
                           const int4 *__restrict__ trs, const int32_t *__restrict__ tmin, int4 *__restrict__ tdo,
                           int4 *__restrict__ tro, uint64_t *__restrict__ tkey, uint64_t *__restrict__ tval)
{
    const int64_t k = blockIdx.x;
    const uint32_t b0 = base[k], cnt = base[k + 1] - b0;
    for (uint32_t l = threadIdx.x; l < cnt; l += blockDim.x) {
        const int64_t src = k * SY_SEG + l, t = (int64_t)b0 + l;
        tdo[t] = tds[src];
        tro[t] = trs[src];
        tkey[t] = ((uint64_t)(uint32_t)tmin[src] << 32) | (uint64_t)t;
        tval[t] = (uint64_t)t;
    }
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
                                                  const uint8_t *__restrict__ aloc, const uint8_t *__restrict__ npos,
                                                  int pstride, int4 *__restrict__ tiles,
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
    // positions of each task's columns in the row (from the node's sort),
    // flagged where the task is the first (lowest rank) to add into the entry
    uint32_t seen[4] = {0u, 0u, 0u, 0u};
    for (int s = 0; s < deg; s++) {
        const int32_t e = aelem[a0 + s];
        const int a = aloc[a0 + s];
        const int64_t slot = slot0 + rank_off[s] + idx[s][tid];
        task_elem[slot] = e;
        task_info[slot] = (uint32_t)a | ((uint32_t)s << 3) | ((uint32_t)rowcode << 7) | ((uint32_t)field << 9) |
                          ((uint32_t)vbase << 10) | ((uint32_t)l8 << 24);
        uint32_t pw[4] = {0u, 0u, 0u, 0u};
        const uint8_t *pp = npos + (int64_t)(a0 + s) * pstride;
        for (int j = 0; j < ncol; j++) {
            const uint32_t lo = pp[j] & 127u;
            uint32_t pv = lo;
            if (!((seen[lo >> 5] >> (lo & 31)) & 1u)) { seen[lo >> 5] |= 1u << (lo & 31); pv |= 0x80u; }
            pw[j >> 2] |= pv << (8 * (j & 3));
        }
        for (int q = 0; q < nw; q++) task_pos[slot * nw + q] = pw[q];
    }
}

size_t align256(size_t b) { return (b + 255) & ~(size_t)255; }




// ---- heat row-owner tiles ------------------------------------------------------
// The tiles of heat_rows_tile (reuse_kernel.h) for scalar T3 / Q4 problems:
// rows packed greedily in a locality order of their nodes (bands of about
// eight node spacings, in x then y; the host builder's band_order), at most
// 256 rows and 256 distinct elements a tile; inside a tile the rows' values
// lie in dof order, written out in runs of consecutive rows; one record per
// row; the tile's elements and the coordinates of their nodes in tile-local
// form.  Distinct elements of a candidate tile come from prev: for each
// (node, element) pair, the latest earlier position in the row order of
// another node of the element -- the element is new to a tile starting at s
// exactly when prev < s.

// order-preserving integer image of a double (for min / max atomics)
__device__ __forceinline__ unsigned long long sy_ord(double d)
{
    const unsigned long long u = (unsigned long long)__double_as_longlong(d);
    return (u >> 63) ? ~u : (u | (1ull << 63));
}
__device__ __forceinline__ double sy_unord(unsigned long long u)
{
    u = (u >> 63) ? (u & ~(1ull << 63)) : ~u;
    return __longlong_as_double((long long)u);
}

__global__ void k_sy_bbox(int64_t Nn, const double2 *__restrict__ xy, unsigned long long *__restrict__ bb)
{
    unsigned long long mnx = ~0ull, mxx = 0, mny = ~0ull, mxy = 0;
    for (int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; i < Nn; i += (int64_t)gridDim.x * blockDim.x) {
        const double2 p = xy[i];
        const unsigned long long ox = sy_ord(p.x), oy = sy_ord(p.y);
        mnx = min(mnx, ox); mxx = max(mxx, ox); mny = min(mny, oy); mxy = max(mxy, oy);
    }
    for (int d = 32; d > 0; d >>= 1) {
        mnx = min(mnx, (unsigned long long)__shfl_xor((long long)mnx, d, 64));
        mxx = max(mxx, (unsigned long long)__shfl_xor((long long)mxx, d, 64));
        mny = min(mny, (unsigned long long)__shfl_xor((long long)mny, d, 64));
        mxy = max(mxy, (unsigned long long)__shfl_xor((long long)mxy, d, 64));
    }
    __shared__ unsigned long long sm[4][4];
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sm[0][w] = mnx; sm[1][w] = mxx; sm[2][w] = mny; sm[3][w] = mxy; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < (int)(blockDim.x >> 6); q++) {
            mnx = min(mnx, sm[0][q]); mxx = max(mxx, sm[1][q]); mny = min(mny, sm[2][q]); mxy = max(mxy, sm[3][q]);
        }
        atomicMin(&bb[0], mnx); atomicMax(&bb[1], mxx); atomicMin(&bb[2], mny); atomicMax(&bb[3], mxy);
    }
}

// key = (band, x, y): x and y quantised to 16 bits in the bounding box, bands
// band_rows node spacings high (spacing from the box and the node count)
__global__ void k_sy_band_keys(int64_t Nn, const double2 *__restrict__ xy, const unsigned long long *__restrict__ bb,
                               int band_rows, uint64_t *__restrict__ key, uint64_t *__restrict__ val)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= Nn) return;
    const double x0 = sy_unord(bb[0]), x1 = sy_unord(bb[1]), y0 = sy_unord(bb[2]), y1 = sy_unord(bb[3]);
    const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
    const double spacing = sqrt(fmax((x1 - x0) * (y1 - y0), 1e-300) / fmax((double)Nn, 1.0));
    const double bh = spacing * (double)band_rows;
    const double2 p = xy[i];
    const uint32_t qx = (uint32_t)((p.x - x0) * sx), qy = (uint32_t)((p.y - y0) * sy);
    const uint64_t band = (uint64_t)((p.y - y0 + 0.5 * spacing) / bh);
    key[i] = (band << 32) | ((uint64_t)qx << 16) | qy;
    val[i] = (uint64_t)i;
}

// row order: the rows of the nodes in the given node order, then the rows
// without a node in dof order (slot Nn + their rank, from the scan of noder);
// without a node order, dof order
__global__ void k_sy_seq(int64_t n, int64_t Nn, const uint64_t *__restrict__ node_order, const int32_t *__restrict__ dof,
                         const int32_t *__restrict__ inv_node, const uint32_t *__restrict__ noder_rank,
                         int32_t *__restrict__ seq, int32_t *__restrict__ P)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (!node_order) {
        if (i < n) { seq[i] = (int32_t)i; P[i] = (int32_t)i; }
        return;
    }
    if (i < Nn) {
        const int32_t r = dof[node_order ? (int64_t)node_order[i] : i];
        seq[i] = r;
        P[r] = (int32_t)i;
    }
    if (i < n && inv_node[i] < 0) {
        const int64_t q = Nn + noder_rank[i];
        seq[q] = (int32_t)i;
        P[i] = (int32_t)q;
    }
}
__global__ void k_sy_noder(int64_t n, const int32_t *__restrict__ inv_node, uint32_t *__restrict__ f)
{
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) f[i] = inv_node[i] < 0 ? 1u : 0u;
    if (i == n) f[i] = 0u;
}

// prev of each (node, element) pair: the latest position before the node's
// own of another node of the element (-1: none)
__global__ void k_sy_prev(int64_t Nn, int nen, const int32_t *__restrict__ conn, const int32_t *__restrict__ dof,
                          const uint32_t *__restrict__ aptr, const int32_t *__restrict__ aelem,
                          const int32_t *__restrict__ P, int32_t *__restrict__ prev)
{
    const int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (nd >= Nn) return;
    const int32_t p = P[dof[nd]];
    for (uint32_t k = aptr[nd]; k < aptr[nd + 1]; k++) {
        const int64_t e = aelem[k];
        int32_t best = -1;
        for (int j = 0; j < nen; j++) {
            const int32_t m = conn[e * nen + j];
            if (m == nd) continue;
            const int32_t q = P[dof[m]];
            if (q < p) best = max(best, q);
        }
        prev[k] = best;
    }
}

// The most elements at a node (scal[SYS_MAXDEG]) and, given nodeL, the
// longest node list (scal[SYS_MAXL]): a grid-stride loop, wave and
// workgroup reductions, one atomic per workgroup.
__global__ __launch_bounds__(256) void k_sy_max(int64_t N, const uint32_t *__restrict__ aptr,
                                                const uint32_t *__restrict__ nodeL, uint32_t *__restrict__ scal)
{
    __shared__ uint32_t sm[2][4];
    uint32_t dg = 0, l = 0;
    for (int64_t nd = (int64_t)blockIdx.x * blockDim.x + threadIdx.x; nd < N; nd += (int64_t)gridDim.x * blockDim.x) {
        dg = max(dg, aptr[nd + 1] - aptr[nd]);
        if (nodeL) l = max(l, nodeL[nd]);
    }
    for (int d = 32; d > 0; d >>= 1) {
        dg = max(dg, (uint32_t)__shfl_xor((int)dg, d, 64));
        l = max(l, (uint32_t)__shfl_xor((int)l, d, 64));
    }
    const int w = threadIdx.x >> 6;
    if ((threadIdx.x & 63) == 0) { sm[0][w] = dg; sm[1][w] = l; }
    __syncthreads();
    if (threadIdx.x == 0) {
        for (int q = 1; q < 4; q++) { dg = max(dg, sm[0][q]); l = max(l, sm[1][q]); }
        atomicMax(&scal[SYS_MAXDEG], dg);
        if (nodeL) atomicMax(&scal[SYS_MAXL], l);
    }
}

// Greedy packing of the row order, one wavefront per range of `seg`
// positions (as k_sy_pack), with the distinct-element limit: a row's new
// elements for a tile starting at s are those with prev < s.  Tile info at
// sparse index k * seg + l: {start position, rows, tasks, values} and
// {0, 0, most elements of a row, distinct elements}.
template <int KMAX>
__global__ __launch_bounds__(64) void k_sy_hpack(int64_t n, int64_t nseg, int seg, const int32_t *__restrict__ seq,
                                                 const int32_t *__restrict__ inv_node, const uint32_t *__restrict__ aptr,
                                                 const int32_t *__restrict__ prev, const int32_t *__restrict__ ptr,
                                                 int vcap, uint32_t *__restrict__ segtiles, int4 *__restrict__ tds,
                                                 int4 *__restrict__ trs)
{
    const int64_t k = blockIdx.x;
    const int lane = threadIdx.x;
    const int64_t r0 = k * seg, r1 = min((int64_t)(k + 1) * seg, n);
    int64_t t = k * seg;
    const int64_t tfirst = t;
    bool open = false;
    int ty = 0, tw = 0, trows = 0, tu = 0, nranks = 0;
    int64_t s0 = 0;  // the open tile's first position
    for (int64_t c0 = r0; c0 < r1; c0 += 64) {
        const int64_t q = c0 + lane;
        const bool valid = q < r1;
        int deg = 0, L = 0;
        int32_t pv[KMAX];
#pragma unroll
        for (int j = 0; j < KMAX; j++) pv[j] = INT32_MAX;
        if (valid) {
            const int32_t r = seq[q];
            const int32_t nd = inv_node[r];
            L = ptr[r + 1] - ptr[r];
            if (nd >= 0) {
                const uint32_t a0 = aptr[nd];
                deg = (int)min(aptr[nd + 1] - a0, (uint32_t)KMAX);
#pragma unroll
                for (int j = 0; j < KMAX; j++)
                    if (j < deg) pv[j] = prev[a0 + j];
            }
        }
        int pos = 0;
        for (;;) {
            const bool mine = valid && lane >= pos;
            const int64_t s = open ? s0 : c0 + pos;  // the start of the tile the lanes from pos join
            int nu = 0;
#pragma unroll
            for (int j = 0; j < KMAX; j++) nu += (j < deg && pv[j] < s) ? 1 : 0;
            const int ad = mine ? deg : 0, av = mine ? L : 0, ar = mine ? 1 : 0, au = mine ? nu : 0;
            const int sd = sy_wave_incl(ad, lane), sv = sy_wave_incl(av, lane), sr = sy_wave_incl(ar, lane),
                      su = sy_wave_incl(au, lane);
            const bool fail = mine && (ty + sd > 256 * KMAX || tw + sv > vcap || trows + sr > 256 || tu + su > 256);
            const uint64_t fm = __ballot(fail);
            const int j = fm ? __ffsll((long long)fm) - 1 : 64;
            const bool join = mine && lane < j;
            const uint64_t jm = __ballot(join);
            if (jm) {
                if (!open) { open = true; s0 = c0 + pos; }
                const int last = 63 - __clzll((long long)jm);
                ty += __shfl(sd, last, 64);
                tw += __shfl(sv, last, 64);
                trows += __shfl(sr, last, 64);
                tu += __shfl(su, last, 64);
                nranks = max(nranks, sy_wave_max(join ? deg : 0));
            }
            if (!fm) break;
            if (open) {
                if (lane == 0) {
                    tds[t] = make_int4((int)s0, trows, ty, tw);
                    trs[t] = make_int4(0, 0, nranks, tu);
                }
                t++;
            }
            open = false;
            ty = tw = trows = tu = nranks = 0;
            pos = j;
        }
    }
    if (open) {
        if (lane == 0) {
            tds[t] = make_int4((int)s0, trows, ty, tw);
            trs[t] = make_int4(0, 0, nranks, tu);
        }
        t++;
    }
    if (lane == 0) segtiles[k] = (uint32_t)(t - tfirst);
}

__global__ void k_sy_hdense(int64_t nseg, int seg, const uint32_t *__restrict__ base, const int4 *__restrict__ tds,
                            const int4 *__restrict__ trs, int4 *__restrict__ tdo, int4 *__restrict__ tro)
{
    const int64_t k = blockIdx.x;
    const uint32_t b0 = base[k], cnt = base[k + 1] - b0;
    for (uint32_t l = threadIdx.x; l < cnt; l += blockDim.x) {
        tdo[b0 + l] = tds[k * seg + l];
        tro[b0 + l] = trs[k * seg + l];
    }
}

// bitonic sort of `size` (a power of two, <= 4 x 256) keys in LDS by a
// workgroup of 256 threads
template <typename T>
__device__ __forceinline__ void sy_block_sort(T *a, int size)
{
    for (int k = 2; k <= size; k <<= 1)
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int p = threadIdx.x; p < (size >> 1); p += 256) {
                const int i = (p / j) * 2 * j + (p % j), q = i + j;
                const T x = a[i], y = a[q];
                const bool up = (i & k) == 0;
                if ((x > y) == up) { a[i] = y; a[q] = x; }
            }
            __syncthreads();
        }
}
__device__ __forceinline__ int sy_block_excl(int v, int *wsum, int *total)
{
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int incl = sy_wave_incl(v, lane);
    __syncthreads();
    if (lane == 63) wsum[w] = incl;
    __syncthreads();
    int pre = 0, tot = 0;
    for (int q = 0; q < 4; q++) { pre += q < w ? wsum[q] : 0; tot += wsum[q]; }
    if (total) *total = tot;
    return pre + incl - v;
}
template <typename T>
__device__ __forceinline__ int sy_lower_bound(const T *a, int n, T v)
{
    int lo = 0, hi = n;
    while (lo < hi) {
        const int m = (lo + hi) >> 1;
        if (a[m] < v) lo = m + 1;
        else hi = m;
    }
    return lo;
}

// One tile per workgroup of 256.  Pass 0 counts its runs of consecutive rows
// and the distinct nodes of its elements; pass 1 writes, at launch slot b,
// the tile descriptors, runs, element list, row records, tile-local
// coordinates and element nodes.
template <int NEN>
__global__ __launch_bounds__(256) void k_sy_htile(int64_t nt, int pass, int patch, const int32_t *__restrict__ seq,
                                                  const int32_t *__restrict__ inv_node, const uint32_t *__restrict__ aptr,
                                                  const int32_t *__restrict__ aelem, const uint8_t *__restrict__ aloc,
                                                  const int32_t *__restrict__ prev, const uint8_t *__restrict__ npos,
                                                  const int32_t *__restrict__ ptr, const int32_t *__restrict__ conn,
                                                  const double2 *__restrict__ xy, const int4 *__restrict__ tdo,
                                                  const int4 *__restrict__ tro, uint32_t *__restrict__ nruns,
                                                  uint32_t *__restrict__ nnodes, int4 *__restrict__ tiles,
                                                  int4 *__restrict__ tile_rows, int2 *__restrict__ truns,
                                                  int4 *__restrict__ runs, int32_t *__restrict__ telem,
                                                  uint32_t *__restrict__ rrec, double2 *__restrict__ txy,
                                                  ushort4 *__restrict__ tconn, uint32_t *__restrict__ scal)
{
    constexpr int KMAX = hrow_kmax<NEN>();
    constexpr int RW = hrow_words<NEN, KMAX>();
    constexpr int FB = hrow_field_bits<NEN>();
    __shared__ uint64_t rk[256];        // (row << 8 | tid), sorted: the tile's rows in dof order
    __shared__ int vb[256];             // LDS value offset of sorted row q
    __shared__ int32_t el[256];         // the tile's elements, sorted
    __shared__ int32_t nds[1024];       // their nodes, sorted, then unique
    __shared__ int32_t un[1024];
    __shared__ int runend[256];
    __shared__ int wsum[4];
    const int64_t t = blockIdx.x;
    const int tid = threadIdx.x;
    const int4 d = tdo[t], rr = tro[t];
    const int s = d.x, nrows = d.y, nU = rr.w;
    // this thread's row (row order position s + tid)
    int32_t row = -1, nd = -1;
    int deg = 0;
    uint32_t a0 = 0;
    if (tid < nrows) {
        row = seq[s + tid];
        nd = inv_node[row];
        if (nd >= 0) { a0 = aptr[nd]; deg = (int)(aptr[nd + 1] - a0); }
    }
    // rows in dof order
    rk[tid] = tid < nrows ? (((uint64_t)(uint32_t)row << 8) | (uint64_t)tid) : ~0ull;
    __syncthreads();
    sy_block_sort<uint64_t>(rk, 256);
    // sorted position q = tid: its row's length -> value offsets
    const int qtid = tid < nrows ? (int)(rk[tid] & 255u) : 0;  // the thread (row) at sorted position tid
    const int32_t qrow = tid < nrows ? (int32_t)(rk[tid] >> 8) : -1;
    const int qL = tid < nrows ? ptr[qrow + 1] - ptr[qrow] : 0;
    int nvals = 0;
    const int qvb = sy_block_excl(qL, wsum, &nvals);
    vb[tid] = qvb;
    // runs of consecutive dofs (patch tiles; otherwise the rows are one run)
    const bool rstart = tid < nrows && (tid == 0 || qrow != (int32_t)(rk[tid - 1] >> 8) + 1);
    int nr = 0;
    const int ridx = sy_block_excl(rstart ? 1 : 0, wsum, &nr);
    // the tile's elements: each row's elements new to the tile (prev < s)
    int cnew = 0;
    for (int k = 0; k < deg; k++) cnew += prev[a0 + k] < s ? 1 : 0;
    int nu2 = 0;
    const int eoff = sy_block_excl(cnew, wsum, &nu2);
    el[tid] = INT32_MAX;
    __syncthreads();
    for (int k = 0, o = eoff; k < deg; k++)
        if (prev[a0 + k] < s) el[o++] = aelem[a0 + k];
    __syncthreads();
    sy_block_sort<int32_t>(el, 256);
    // their nodes, sorted and unique
    for (int i = tid; i < 1024; i += 256) {
        const int u = i / NEN, j = i % NEN;
        nds[i] = (u < nU && i < NEN * 256) ? conn[(int64_t)el[u] * NEN + j] : INT32_MAX;
    }
    __syncthreads();
    sy_block_sort<int32_t>(nds, 1024);
    int fl[4], cu = 0;
    for (int m = 0; m < 4; m++) {
        const int i = tid * 4 + m;
        fl[m] = nds[i] != INT32_MAX && (i == 0 || nds[i] != nds[i - 1]);
        cu += fl[m];
    }
    int nn = 0;
    const int uoff = sy_block_excl(cu, wsum, &nn);
    if (pass == 0) {
        if (tid == 0) {
            nruns[t] = (uint32_t)(patch ? nr : 0);
            nnodes[t] = (uint32_t)nn;
            atomicMax(&scal[SYS_MAXV], (uint32_t)nvals);
            atomicMax(&scal[SYS_MAXRANK], (uint32_t)rr.z);
            atomicMax(&scal[SYS_MAXU], (uint32_t)nU);
        }
        return;
    }
    for (int m = 0, o = uoff; m < 4; m++)
        if (fl[m]) un[o++] = nds[tid * 4 + m];
    __syncthreads();
    const int64_t b = sy_slot(t, nt);
    const int xb = (int)nnodes[t];  // scanned: this tile's first tile-local coordinate
    // descriptors
    if (tid == 0) {
        const int32_t row0 = (int32_t)(rk[0] >> 8);
        tiles[b] = make_int4(xb, d.z, ptr[row0], nvals);
        tile_rows[b] = make_int4(row0, nrows, rr.z, nU);
        if (patch) truns[b] = make_int2((int)nruns[t], nr);
    }
    // runs: the last row of each run records where it ends (the run of sorted
    // row tid is the number of run starts up to it, less one)
    if (patch) {
        const bool rend = tid < nrows && (tid == nrows - 1 || (int32_t)(rk[tid + 1] >> 8) != qrow + 1);
        if (rend) runend[ridx + (rstart ? 1 : 0) - 1] = tid;
    }
    __syncthreads();
    if (patch && rstart) {
        // this run: sorted rows tid .. runend[ridx]
        const int qe = runend[ridx];
        const int nv = vb[qe] + (ptr[(int32_t)(rk[qe] >> 8) + 1] - ptr[(int32_t)(rk[qe] >> 8)]) - vb[tid];
        runs[nruns[t] + ridx] = make_int4(ptr[qrow], qrow, vb[tid] | (nv << 16), tid | ((qe - tid + 1) << 16));
    }
    // element list (padded with the first element)
    telem[b * 256 + tid] = tid < nU ? el[tid] : el[0];
    // tile-local coordinates and element nodes
    for (int i = tid; i < nn; i += 256) txy[xb + i] = xy[un[i]];
    {
        uint16_t l[4] = {0, 0, 0, 0};
        if (tid < nU)
            for (int j = 0; j < NEN; j++)
                l[j] = (uint16_t)sy_lower_bound<int32_t>(un, nn, conn[(int64_t)el[tid] * NEN + j]);
        tconn[b * 256 + tid] = make_ushort4(l[0], l[1], l[2], l[3]);
    }
    // row records, at the row's sorted (dof-order) position
    uint32_t w[RW];
    for (int i = 0; i < RW; i++) w[i] = 0u;
    int myq = 0;
    if (tid < nrows) {
        // my sorted position: the q with rk[q] & 255 == tid
        myq = sy_lower_bound<uint64_t>(rk, nrows, ((uint64_t)(uint32_t)row << 8) | (uint64_t)tid);
        auto put = [&](int o, uint32_t v, int nb) {
            for (int bb = 0; bb < nb; bb++)
                if ((v >> bb) & 1u) w[(o + bb) >> 5] |= 1u << ((o + bb) & 31);
        };
        put(0, (uint32_t)deg, 4);
        put(4, (uint32_t)vb[myq], 13);
        for (int k = 0; k < deg && k < KMAX; k++) {
            const int o = 17 + k * FB;
            put(o, (uint32_t)sy_lower_bound<int32_t>(el, nU, aelem[a0 + k]), 8);
            put(o + 8, (uint32_t)aloc[a0 + k], 2);
            for (int j = 0; j < NEN; j++) put(o + 10 + 4 * j, (uint32_t)(npos[(int64_t)(a0 + k) * NEN + j] & 15u), 4);
        }
    }
    __syncthreads();
    // position q gets the record of its row; positions past the rows get zeros
    __shared__ uint32_t recs[256 * RW];
    if (tid < nrows)
        for (int i = 0; i < RW; i++) recs[myq * RW + i] = w[i];
    __syncthreads();
    for (int i = 0; i < RW; i++) rrec[(b * 256 + tid) * RW + i] = tid < nrows ? recs[tid * RW + i] : 0u;
    (void)qtid;
}

// nodes per wavefront of k_sy_node_cols for at most c candidates a node
int sy_npw(int64_t c) { return c <= 16 ? 4 : (c <= 32 ? 2 : 1); }

// the most elements at a node, read back (one synchronisation)
int sy_maxdeg(elfel_hip_ctx *c, int64_t N, const uint32_t *aptr, uint32_t *scal, uint32_t *out)
{
    *out = 0;
    if (N <= 0) return 0;
    hipLaunchKernelGGL(k_sy_max, dim3(std::max(1, std::min(ceil_div(N, 256), 1024))), dim3(256), 0, c->stream, N, aptr,
                       (const uint32_t *)nullptr, scal);
    HIPCHK(c, hipMemcpyAsync(out, scal + SYS_MAXDEG, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    return 0;
}

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
    const int pstride = S.nen0 * S.nc + (stokes ? 3 : 0);  // npos bytes per adjacency entry
    const int64_t nseg = std::max<int64_t>(1, (n + SY_SEG - 1) / SY_SEG);
    // scratch, one allocation: adjacency, inverse map, per-candidate
    // positions, row data, packing ranges, dof-ordered tiles (at most one per
    // row), the tile sort
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += align256(bytes); return o; };
    const size_t o_scal = take(4 * SYS_N), o_aptr = take(4 * (N + 1)), o_fill = take(4 * N), o_aelem = take(4 * nent),
                 o_aloc = take(nent), o_inv = take(4 * n), o_fc = take(n), o_nodeL = take(4 * N),
                 o_npos = take((size_t)nent * pstride), o_edof = take(4 * (size_t)S.E * pstride), o_gmin = take(4 * n),
                 o_rinfo = take(4 * n), o_ginfo = take(4 * n), o_seg = take(4 * (nseg + 1)),
                 o_tds = take(16 * (size_t)nseg * SY_SEG), o_trs = take(16 * (size_t)nseg * SY_SEG),
                 o_tmin = take(4 * (size_t)nseg * SY_SEG),
                 o_scan = take(4 * (size_t)scan_scratch_entries(std::max(N, n) + 1)),
                 o_hist = take(4 * (size_t)radix_hist_entries((int64_t)nseg * SY_SEG));
    int rc;
    if ((rc = ensure(c, P.sym, off))) return rc;
    unsigned char *base = (unsigned char *)P.sym.p;
    uint32_t *scal = (uint32_t *)(base + o_scal), *aptr = (uint32_t *)(base + o_aptr), *fill = (uint32_t *)(base + o_fill);
    int32_t *aelem = (int32_t *)(base + o_aelem), *inv_node = (int32_t *)(base + o_inv), *gmin = (int32_t *)(base + o_gmin);
    uint8_t *npos = base + o_npos;
    uint8_t *aloc = base + o_aloc, *inv_fc = base + o_fc;
    uint32_t *nodeL = (uint32_t *)(base + o_nodeL), *rinfo = (uint32_t *)(base + o_rinfo), *ginfo = (uint32_t *)(base + o_ginfo);
    int32_t *edof = (int32_t *)(base + o_edof), *tmin = (int32_t *)(base + o_tmin);
    uint32_t *segtiles = (uint32_t *)(base + o_seg);
    int4 *tds = (int4 *)(base + o_tds), *trs = (int4 *)(base + o_trs);
    // the scans' and the tile sort's scratch come from the same allocation
    SortScratch ss;
    ss.scan = (uint32_t *)(base + o_scan);
    ss.scan_cap = scan_scratch_entries(std::max(N, n) + 1);
    ss.hist = (uint32_t *)(base + o_hist);
    ss.hist_cap = radix_hist_entries((int64_t)nseg * SY_SEG);
    hipStream_t st = c->stream;
    const int g = 256;
    const int grid_ent = std::max(1, std::min(ceil_div(nent, g), 8192));
    HIPCHK(c, hipMemsetAsync(scal, 0, 4 * SYS_N, st));
    HIPCHK(c, hipMemsetAsync(aptr, 0, 4 * (N + 1), st));
    HIPCHK(c, hipMemsetAsync(fill, 0, 4 * N, st));
    HIPCHK(c, hipMemsetAsync(inv_node, 0xFF, 4 * n, st));
    // 1. adjacency
    hipLaunchKernelGGL(k_sy_adj_count, dim3(grid_ent), dim3(g), 0, st, S, aptr);
    if ((rc = exclusive_scan_u32(c, aptr, N + 1, nullptr, &ss))) return rc;
    hipLaunchKernelGGL(k_sy_adj_fill, dim3(grid_ent), dim3(g), 0, st, S, (const uint32_t *)aptr, fill, aelem, aloc);
    if (N > 0)
        hipLaunchKernelGGL(k_sy_adj_sort, dim3(ceil_div(N, g)), dim3(g), 0, st, N, (const uint32_t *)aptr, aelem, aloc,
                           scal);
    // 2. inverse dof map, element check
    const int64_t ndofref = S.Nn0 * S.nc + S.Np;
    hipLaunchKernelGGL(k_sy_inv, dim3(std::max(1, std::min(ceil_div(ndofref, g), 8192))), dim3(g), 0, st, S, inv_node,
                       inv_fc, scal);
    if (S.E > 0) hipLaunchKernelGGL(k_sy_elem_check, dim3(ceil_div(S.E, g)), dim3(g), 0, st, S, pstride, edof, scal);
    // 3. node column lists: lengths, row lengths, ptr; then the lists into ind
    uint32_t maxdeg = 0;
    if ((rc = sy_maxdeg(c, N, aptr, scal, &maxdeg))) return rc;
    const int npw0 = sy_npw((int64_t)maxdeg * (S.nen0 + (stokes ? 3 : 0))), npw1 = sy_npw((int64_t)maxdeg * pstride);
    if (N > 0)
        hipLaunchKernelGGL(k_sy_node_cols, dim3(ceil_div(N, 4 * npw0)), dim3(256), 0, st, S, N, (const uint32_t *)aptr,
                           (const int32_t *)aelem, (const int32_t *)edof, 0, npw0, nodeL, (const int32_t *)nullptr,
                           (int32_t *)nullptr, (uint8_t *)nullptr, pstride, scal);
    if ((rc = ensure(c, P.ptr, 4 * (n + 1)))) return rc;
    int32_t *ptr = (int32_t *)P.ptr.p;
    hipLaunchKernelGGL(k_sy_rows, dim3(ceil_div(n + 1, g)), dim3(g), 0, st, S, (const int32_t *)inv_node,
                       (const uint8_t *)inv_fc, (const uint32_t *)aptr, (const int32_t *)aelem, (const uint32_t *)nodeL,
                       (uint32_t *)ptr, rinfo, gmin);
    if ((rc = exclusive_scan_u32(c, (uint32_t *)ptr, n + 1, nullptr, &ss))) return rc;
    uint32_t hv[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(&hv[0], ptr + n, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(&hv[1], scal + SYS_FLAGS, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const int64_t nnz = hv[0];
    P.fallback = false;
    if (hv[1]) {
        P.fallback = true;
        const uint32_t f = hv[1];
        P.fallback_note = std::string("FP64 runs of the cached sort order: ") +
                          ((f & SYF_SHARED) ? "a dof shared by two nodes or components"
                           : (f & SYF_REPEAT) ? "an element with a repeated dof"
                                              : "node too highly connected for the reuse tiles");
        P.nnz = nnz;
        return 0;
    }
    if ((rc = ensure(c, P.ind, 4 * (size_t)std::max<int64_t>(nnz, 1)))) return rc;
    if (N > 0)
        hipLaunchKernelGGL(k_sy_node_cols, dim3(ceil_div(N, 4 * npw1)), dim3(256), 0, st, S, N, (const uint32_t *)aptr,
                           (const int32_t *)aelem, (const int32_t *)edof, 1, npw1, nodeL, (const int32_t *)ptr,
                           (int32_t *)P.ind.p, npos, pstride, scal);
    // 4. packing (sparse tile numbers), counts scanned, dense tiles and keys
    hipLaunchKernelGGL(k_sy_pack, dim3((unsigned)nseg), dim3(64), 0, st, n, nseg, (const uint32_t *)rinfo,
                       (const int32_t *)gmin, (const int32_t *)ptr, cap, vcap, rowcap, segtiles, ginfo, tds, trs, tmin,
                       scal);
    HIPCHK(c, hipMemsetAsync(segtiles + nseg, 0, 4, st));
    if ((rc = exclusive_scan_u32(c, segtiles, nseg + 1, nullptr, &ss))) return rc;
    uint32_t hn[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(&hn[0], segtiles + nseg, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(&hn[1], scal + SYS_FLAGS, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (hn[1]) {
        P.fallback = true;
        P.fallback_note = "FP64 runs of the cached sort order: node too highly connected for the reuse tiles";
        P.nnz = nnz;
        return 0;
    }
    const int64_t nt = hn[0];
    // the dense tiles, their sort keys and order: nt-sized, a second allocation
    size_t off2 = 0;
    auto take2 = [&](size_t bytes) { const size_t o = off2; off2 += align256(bytes); return o; };
    const int64_t ntc = std::max<int64_t>(nt, 1);
    const size_t o_tdo = take2(16 * ntc), o_tro = take2(16 * ntc), o_k0 = take2(8 * ntc), o_k1 = take2(8 * ntc),
                 o_v0 = take2(8 * ntc), o_v1 = take2(8 * ntc), o_tpos = take2(4 * ntc);
    if ((rc = ensure(c, P.sym2, off2))) return rc;
    unsigned char *base2 = (unsigned char *)P.sym2.p;
    int4 *tdo = (int4 *)(base2 + o_tdo), *tro = (int4 *)(base2 + o_tro);
    uint64_t *k0 = (uint64_t *)(base2 + o_k0), *k1 = (uint64_t *)(base2 + o_k1);
    uint64_t *v0 = (uint64_t *)(base2 + o_v0), *v1 = (uint64_t *)(base2 + o_v1);
    uint32_t *tpos = (uint32_t *)(base2 + o_tpos);
    hipLaunchKernelGGL(k_sy_dense, dim3((unsigned)nseg), dim3(64), 0, st, nseg, (const uint32_t *)segtiles,
                       (const int4 *)tds, (const int4 *)trs, (const int32_t *)tmin, tdo, tro, k0, v0);
    // 5. launch order: tiles sorted by first element (key: element << 32 | tile)
    if (nt > 1) {
        if ((rc = radix_sort_pairs(c, k0, v0, k1, v1, nt, 32 + bits_for(std::max<int64_t>(S.E, 1)), &ss))) return rc;
        hipLaunchKernelGGL(k_sy_tpos, dim3(ceil_div(nt, g)), dim3(g), 0, st, nt, (const uint64_t *)v0, tpos);
    } else {
        HIPCHK(c, hipMemsetAsync(tpos, 0, 4, st));
    }
    // 6. task tables in launch order
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
                           (const uint32_t *)aptr, (const int32_t *)aelem, (const uint8_t *)aloc,
                           (const uint8_t *)npos, pstride, (int4 *)P.tile.p, (int4 *)P.tile_rows.p, (int32_t *)P.task_elem.p,
                           (uint32_t *)P.task_info.p, (uint32_t *)P.task_pos.p, scal);
    HIPCHK(c, hipGetLastError());
    uint32_t hs[SYS_N] = {};
    HIPCHK(c, hipMemcpyAsync(hs, scal, 4 * SYS_N, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    if (hs[SYS_FLAGS]) return set_err(c, ELFEL_EINVAL, "internal: device symbolic phase flagged late");
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

// The row-owner heat tiles of a scalar T3 / Q4 problem built on the device
// into c->pat.  Returns 1 (nothing built) when the mesh does not fit the row
// records (a node in more than kmax elements, a row of more than 16 entries):
// the host builder then makes task tiles.  Returns 0 with c->pat.fallback set
// when the reuse method cannot take the mesh at all (a dof shared by two
// nodes, an element that repeats a dof).
int device_symbolic_heat(elfel_hip_ctx *c, const elfel_hip_problem *pb, bool patch_allowed)
{
    Pattern &P = c->pat;
    const Field &F0 = c->f[0];
    const int nen = F0.nen;
    if (F0.ncomp != 1 || (nen != 3 && nen != 4) || F0.sdim != 2) return 1;
    SymIn S;
    S.conn0 = (const int32_t *)F0.conn.p;
    S.dof0 = (const int32_t *)F0.dof.p;
    S.conn1 = S.dof1 = nullptr;
    S.E = F0.E;
    S.Nn0 = F0.Nn;
    S.Np = 0;
    S.n = c->nrow;
    S.nen0 = nen;
    S.nc = 1;
    S.stokes = 0;
    const int64_t N = S.Nn0, n = S.n, nent = S.E * nen;
    const int kmax = nen == 4 ? hrow_kmax<4>() : hrow_kmax<3>();
    const int RW = nen == 4 ? hrow_words<4, hrow_kmax<4>()>() : hrow_words<3, hrow_kmax<3>()>();
    const int vcap = REUSE_VCAP_PER_THREAD * HEAT_ROWS_BLOCK;
    const int pstride = nen;
    // packing ranges of 16384 positions: a tile never spans two, so longer
    // ranges leave fewer part-filled tiles (heat tiles take about 220 rows)
    const int seg = 16384;
    const int64_t nseg = std::max<int64_t>(1, (n + seg - 1) / seg);
    const bool patch = patch_allowed && n >= (int64_t(1) << 18);
    size_t off = 0;
    auto take = [&](size_t bytes) { const size_t o = off; off += align256(bytes); return o; };
    const int64_t n1 = std::max<int64_t>(n, 1), N1 = std::max<int64_t>(N, 1);
    const size_t o_scal = take(4 * SYS_N), o_aptr = take(4 * (N + 1)), o_fill = take(4 * N1), o_aelem = take(4 * nent),
                 o_aloc = take(nent), o_inv = take(4 * n1), o_fc = take(n1), o_nodeL = take(4 * N1),
                 o_npos = take((size_t)nent * pstride), o_edof = take(4 * (size_t)S.E * pstride),
                 o_rinfo = take(4 * n1), o_gmin = take(4 * n1), o_noder = take(4 * (n + 1)), o_seq = take(4 * n1),
                 o_P = take(4 * n1), o_prev = take(4 * nent), o_bb = take(32), o_k0 = take(8 * N1), o_k1 = take(8 * N1),
                 o_v0 = take(8 * N1), o_v1 = take(8 * N1), o_seg = take(4 * (nseg + 1)),
                 o_tds = take(16 * (size_t)nseg * seg), o_trs = take(16 * (size_t)nseg * seg),
                 o_scan = take(4 * (size_t)scan_scratch_entries(std::max(N, n) + 1)),
                 o_hist = take(4 * (size_t)radix_hist_entries(std::max<int64_t>(N, (int64_t)nseg * seg)));
    int rc;
    if ((rc = ensure(c, P.sym, off))) return rc;
    unsigned char *base = (unsigned char *)P.sym.p;
    uint32_t *scal = (uint32_t *)(base + o_scal), *aptr = (uint32_t *)(base + o_aptr), *fill = (uint32_t *)(base + o_fill);
    int32_t *aelem = (int32_t *)(base + o_aelem), *inv_node = (int32_t *)(base + o_inv);
    uint8_t *aloc = base + o_aloc, *inv_fc = base + o_fc, *npos = base + o_npos;
    uint32_t *nodeL = (uint32_t *)(base + o_nodeL), *rinfo = (uint32_t *)(base + o_rinfo);
    int32_t *edof = (int32_t *)(base + o_edof), *gmin = (int32_t *)(base + o_gmin);
    uint32_t *noder = (uint32_t *)(base + o_noder);
    int32_t *seq = (int32_t *)(base + o_seq), *Ppos = (int32_t *)(base + o_P), *prev = (int32_t *)(base + o_prev);
    unsigned long long *bb = (unsigned long long *)(base + o_bb);
    uint64_t *k0 = (uint64_t *)(base + o_k0), *k1 = (uint64_t *)(base + o_k1);
    uint64_t *v0 = (uint64_t *)(base + o_v0), *v1 = (uint64_t *)(base + o_v1);
    uint32_t *segtiles = (uint32_t *)(base + o_seg);
    int4 *tds = (int4 *)(base + o_tds), *trs = (int4 *)(base + o_trs);
    SortScratch ss;
    ss.scan = (uint32_t *)(base + o_scan);
    ss.scan_cap = scan_scratch_entries(std::max(N, n) + 1);
    ss.hist = (uint32_t *)(base + o_hist);
    ss.hist_cap = radix_hist_entries(std::max<int64_t>(N, (int64_t)nseg * seg));
    hipStream_t st = c->stream;
    const int g = 256;
    const int grid_ent = std::max(1, std::min(ceil_div(nent, g), 8192));
    HIPCHK(c, hipMemsetAsync(scal, 0, 4 * SYS_N, st));
    HIPCHK(c, hipMemsetAsync(aptr, 0, 4 * (N + 1), st));
    HIPCHK(c, hipMemsetAsync(fill, 0, 4 * N1, st));
    HIPCHK(c, hipMemsetAsync(inv_node, 0xFF, 4 * n1, st));
    // adjacency, inverse map, element check, node list lengths, ptr
    hipLaunchKernelGGL(k_sy_adj_count, dim3(grid_ent), dim3(g), 0, st, S, aptr);
    if ((rc = exclusive_scan_u32(c, aptr, N + 1, nullptr, &ss))) return rc;
    hipLaunchKernelGGL(k_sy_adj_fill, dim3(grid_ent), dim3(g), 0, st, S, (const uint32_t *)aptr, fill, aelem, aloc);
    if (N > 0)
        hipLaunchKernelGGL(k_sy_adj_sort, dim3(ceil_div(N, g)), dim3(g), 0, st, N, (const uint32_t *)aptr, aelem, aloc,
                           scal);
    hipLaunchKernelGGL(k_sy_inv, dim3(std::max(1, std::min(ceil_div(N, g), 8192))), dim3(g), 0, st, S, inv_node, inv_fc,
                       scal);
    if (S.E > 0) hipLaunchKernelGGL(k_sy_elem_check, dim3(ceil_div(S.E, g)), dim3(g), 0, st, S, pstride, edof, scal);
    uint32_t maxdeg = 0;
    if ((rc = sy_maxdeg(c, N, aptr, scal, &maxdeg))) return rc;
    if (maxdeg > (uint32_t)kmax) return 1;  // task tiles (host)
    const int npw = sy_npw((int64_t)maxdeg * nen);  // both passes: nen candidates per element
    if (N > 0) {
        hipLaunchKernelGGL(k_sy_node_cols, dim3(ceil_div(N, 4 * npw)), dim3(256), 0, st, S, N, (const uint32_t *)aptr,
                           (const int32_t *)aelem, (const int32_t *)edof, 0, npw, nodeL, (const int32_t *)nullptr,
                           (int32_t *)nullptr, (uint8_t *)nullptr, pstride, scal);
        hipLaunchKernelGGL(k_sy_max, dim3(std::max(1, std::min(ceil_div(N, g), 1024))), dim3(g), 0, st, N,
                           (const uint32_t *)aptr, (const uint32_t *)nodeL, scal);
    }
    if ((rc = ensure(c, P.ptr, 4 * (n + 1)))) return rc;
    int32_t *ptr = (int32_t *)P.ptr.p;
    hipLaunchKernelGGL(k_sy_rows, dim3(ceil_div(n + 1, g)), dim3(g), 0, st, S, (const int32_t *)inv_node,
                       (const uint8_t *)inv_fc, (const uint32_t *)aptr, (const int32_t *)aelem, (const uint32_t *)nodeL,
                       (uint32_t *)ptr, rinfo, gmin);
    if ((rc = exclusive_scan_u32(c, (uint32_t *)ptr, n + 1, nullptr, &ss))) return rc;
    uint32_t hv[SYS_N + 1] = {};
    HIPCHK(c, hipMemcpyAsync(&hv[0], ptr + n, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(&hv[1], scal, 4 * SYS_N, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const int64_t nnz = hv[0];
    const uint32_t flags = hv[1 + SYS_FLAGS];
    P.fallback = false;
    if (flags & (SYF_SHARED | SYF_REPEAT)) {
        P.fallback = true;
        P.fallback_note = std::string("FP64 runs of the cached sort order: ") +
                          ((flags & SYF_SHARED) ? "a dof shared by two nodes or components" : "an element with a repeated dof");
        P.nnz = nnz;
        return 0;
    }
    if (flags || hv[1 + SYS_MAXDEG] > (uint32_t)kmax || hv[1 + SYS_MAXL] > 16u) return 1;  // task tiles (host)
    HIPCHK(c, hipMemsetAsync(scal + SYS_MAXV, 0, 4 * (SYS_N - SYS_MAXV), st));
    if ((rc = ensure(c, P.ind, 4 * (size_t)std::max<int64_t>(nnz, 1)))) return rc;
    if (N > 0)
        hipLaunchKernelGGL(k_sy_node_cols, dim3(ceil_div(N, 4 * npw)), dim3(256), 0, st, S, N, (const uint32_t *)aptr,
                           (const int32_t *)aelem, (const int32_t *)edof, 1, npw, nodeL, (const int32_t *)ptr,
                           (int32_t *)P.ind.p, npos, pstride, scal);
    // row order: nodes in bands (patch tiles) or in dof order, then the rows without a node
    const uint64_t *order = nullptr;
    if (patch && N > 1) {
        const unsigned long long init[4] = {~0ull, 0ull, ~0ull, 0ull};
        HIPCHK(c, hipMemcpyAsync(bb, init, sizeof init, hipMemcpyHostToDevice, st));
        hipLaunchKernelGGL(k_sy_bbox, dim3(std::max(1, std::min(ceil_div(N, g), 1024))), dim3(g), 0, st, N,
                           (const double2 *)F0.xy.p, bb);
        hipLaunchKernelGGL(k_sy_band_keys, dim3(ceil_div(N, g)), dim3(g), 0, st, N, (const double2 *)F0.xy.p,
                           (const unsigned long long *)bb, std::max(1, env_int("ELFEL_HIP_BAND_ROWS", 8)), k0, v0);
        if ((rc = radix_sort_pairs(c, k0, v0, k1, v1, N, 64, &ss))) return rc;
        order = v0;
    }
    hipLaunchKernelGGL(k_sy_noder, dim3(ceil_div(n + 1, g)), dim3(g), 0, st, n, (const int32_t *)inv_node, noder);
    if ((rc = exclusive_scan_u32(c, noder, n + 1, nullptr, &ss))) return rc;
    hipLaunchKernelGGL(k_sy_seq, dim3(ceil_div(std::max(n, N), g)), dim3(g), 0, st, n, N, order,
                       (const int32_t *)S.dof0, (const int32_t *)inv_node, (const uint32_t *)noder, seq, Ppos);
    if (N > 0)
        hipLaunchKernelGGL(k_sy_prev, dim3(ceil_div(N, g)), dim3(g), 0, st, N, nen, S.conn0, S.dof0,
                           (const uint32_t *)aptr, (const int32_t *)aelem, (const int32_t *)Ppos, prev);
    // packing, dense tiles
    if (nen == 4)
        hipLaunchKernelGGL((k_sy_hpack<hrow_kmax<4>()>), dim3((unsigned)nseg), dim3(64), 0, st, n, nseg, seg,
                           (const int32_t *)seq, (const int32_t *)inv_node, (const uint32_t *)aptr,
                           (const int32_t *)prev, (const int32_t *)ptr, vcap, segtiles, tds, trs);
    else
        hipLaunchKernelGGL((k_sy_hpack<hrow_kmax<3>()>), dim3((unsigned)nseg), dim3(64), 0, st, n, nseg, seg,
                           (const int32_t *)seq, (const int32_t *)inv_node, (const uint32_t *)aptr,
                           (const int32_t *)prev, (const int32_t *)ptr, vcap, segtiles, tds, trs);
    HIPCHK(c, hipMemsetAsync(segtiles + nseg, 0, 4, st));
    if ((rc = exclusive_scan_u32(c, segtiles, nseg + 1, nullptr, &ss))) return rc;
    uint32_t hn = 0;
    HIPCHK(c, hipMemcpyAsync(&hn, segtiles + nseg, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const int64_t nt = hn, ntc = std::max<int64_t>(nt, 1);
    size_t off2 = 0;
    auto take2 = [&](size_t bytes) { const size_t o = off2; off2 += align256(bytes); return o; };
    const size_t o_tdo = take2(16 * ntc), o_tro = take2(16 * ntc), o_nr = take2(4 * (ntc + 1)), o_nn = take2(4 * (ntc + 1));
    if ((rc = ensure(c, P.sym2, off2))) return rc;
    unsigned char *base2 = (unsigned char *)P.sym2.p;
    int4 *tdo = (int4 *)(base2 + o_tdo), *tro = (int4 *)(base2 + o_tro);
    uint32_t *nruns = (uint32_t *)(base2 + o_nr), *nnodes = (uint32_t *)(base2 + o_nn);
    hipLaunchKernelGGL(k_sy_hdense, dim3((unsigned)nseg), dim3(64), 0, st, nseg, seg, (const uint32_t *)segtiles,
                       (const int4 *)tds, (const int4 *)trs, tdo, tro);
    // tile pass 0: runs and tile-local nodes per tile, scanned
#define HTILE(NN, PASS)                                                                                                \
    hipLaunchKernelGGL((k_sy_htile<NN>), dim3((unsigned)nt), dim3(256), 0, st, nt, PASS, patch ? 1 : 0,                \
                       (const int32_t *)seq, (const int32_t *)inv_node, (const uint32_t *)aptr, (const int32_t *)aelem, \
                       (const uint8_t *)aloc, (const int32_t *)prev, (const uint8_t *)npos, (const int32_t *)ptr,     \
                       S.conn0, (const double2 *)F0.xy.p, (const int4 *)tdo, (const int4 *)tro, nruns, nnodes,        \
                       (int4 *)P.tile.p, (int4 *)P.tile_rows.p, (int2 *)P.truns.p, (int4 *)P.runs.p,                   \
                       (int32_t *)P.tile_elem.p, (uint32_t *)P.rrec.p, (double2 *)P.txy.p, (ushort4 *)P.tconn.p, scal)
    if (nt > 0) {
        if (nen == 4) HTILE(4, 0);
        else HTILE(3, 0);
    }
    HIPCHK(c, hipMemsetAsync(nruns + nt, 0, 4, st));
    HIPCHK(c, hipMemsetAsync(nnodes + nt, 0, 4, st));
    if ((rc = exclusive_scan_u32(c, nruns, nt + 1, nullptr, &ss)) || (rc = exclusive_scan_u32(c, nnodes, nt + 1, nullptr, &ss)))
        return rc;
    uint32_t ht[2] = {0, 0}, hs[SYS_N] = {};
    HIPCHK(c, hipMemcpyAsync(&ht[0], nruns + nt, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(&ht[1], nnodes + nt, 4, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipMemcpyAsync(hs, scal, 4 * SYS_N, hipMemcpyDeviceToHost, st));
    HIPCHK(c, hipStreamSynchronize(st));
    const int64_t totruns = ht[0], totnodes = ht[1];
    if ((rc = ensure(c, P.tile, sizeof(int4) * ntc)) || (rc = ensure(c, P.tile_rows, sizeof(int4) * ntc)) ||
        (rc = ensure(c, P.tile_elem, 4 * (size_t)ntc * HEAT_ROWS_BLOCK)) ||
        (rc = ensure(c, P.rrec, 4 * (size_t)ntc * HEAT_ROWS_BLOCK * RW)) ||
        (rc = ensure(c, P.tconn, sizeof(ushort4) * (size_t)ntc * HEAT_ROWS_BLOCK)) ||
        (rc = ensure(c, P.txy, sizeof(double2) * (size_t)std::max<int64_t>(totnodes, 1))))
        return rc;
    if (patch) {
        if ((rc = ensure(c, P.truns, sizeof(int2) * ntc)) || (rc = ensure(c, P.runs, sizeof(int4) * std::max<int64_t>(totruns, 1))))
            return rc;
    } else {
        release(P.truns); release(P.runs);
    }
    if (nt > 0) {
        if (nen == 4) HTILE(4, 1);
        else HTILE(3, 1);
    }
#undef HTILE
    HIPCHK(c, hipGetLastError());
    release(P.task_elem); release(P.task_info); release(P.task_pos);
    P.rw = RW;
    P.tstride = HEAT_ROWS_BLOCK;
    P.tgeom = true;
    P.has_runs = patch;
    P.elem8 = false;
    P.maxu = (int)hs[SYS_MAXU];
    P.ntiles = nt;
    P.ntasks = 0;
    P.ncolmax = (nen + 3) & ~3;
    P.maxrank = (int)hs[SYS_MAXRANK];
    P.lds_vals = (int)hs[SYS_MAXV];
    P.nnz = nnz;
    return 0;
}
