// reuse_kernel.h -- fused owner-computes assembly with a cached pattern.
//
// Work unit ("task") = (owner row group, element): a row group is the run of
// consecutive dofs of one node (its components), all of which share one
// sorted column list.  The task computes the owner dofs' entries of that
// element's block -- rows of ke for CSR, columns of ke for CSC (Elfel's
// finish! returns CSC, src/Assemblers.jl:121-123) -- and adds them into an LDS
// copy of the tile's CSR/CSC value range.  Tasks of a group are ordered by
// element, and the adds run rank by rank with a barrier in between, so every
// entry is summed in element order from 0.0: the order in which
// SparseArrays.sparse combines the duplicates of the COO stream (within a rank
// step no two tasks touch the same entry, so Stokes tiles may use LDS atomic
// adds without changing a bit).  The tile's values are then written to HBM
// once, coalesced, with non-temporal stores; no global atomics, no sort.
//
// Task descriptor (uint32 info):
//   bits 0-2   local node a of the owner in the element
//   bits 3-6   rank s of the task in its group (element order)
//   bits 7-8   rows: 0 -> comp 0 only, 1 -> comp 1 only, 2 -> comps 0 and 1
//   bit  9     owner is a pressure dof (Taylor-Hood)
//   bits 10-23 LDS offset of the group's first row in the tile
//   bits 24-31 row length L (two-row groups) or tile-local row (one-row groups)
// A tile holds exactly BLOCK * TPT task slots (tile t's tasks start at
// t * BLOCK * TPT; TPT tasks per thread); unused slots carry TASK_PAD (rows
// code 3) and element 0, and add nothing.  Within a tile the tasks are laid out
// rank by rank.
// task_pos[task][j]: position of element column j in the owner's sorted
// column list (bits 0-6, rows of at most 128 entries); bit 7 says that this
// task is the first to add into that entry (it stores instead of adding, so
// the tile's sums need no zero fill).
#pragma once
#include "elem.h"

#define REUSE_BLOCK 256
#define TASK_PAD (3u << 7)
#ifndef REUSE_WAVES_PER_EU
#define REUSE_WAVES_PER_EU 1   // occupancy hint for the mixed kernels (1 = none)
#endif
#define REUSE_VCAP_PER_THREAD 24   // LDS doubles for tile values per thread of the block (48 KiB at 256)
#define REUSE_VCAP (REUSE_VCAP_PER_THREAD * REUSE_BLOCK)
#define REUSE_COND_MAX 1024.0  // sum_k |grad N_k|^2 * det J: above it an element counts as badly shaped

struct ReuseArgs {
    const int4 *tiles;        // task0, ntasks, v0, nvals
    const int4 *tile_rows;    // row0, nrows, nranks, 0
    const int32_t *task_elem;  // element of the task (not read for heat)
    const int32_t *tile_elem;  // heat: distinct elements of each tile, [ntiles][slots]
    int maxu;                  // heat: most distinct elements in a tile
    const uint32_t *task_info;
    const uint8_t *task_pos;  // [ntasks][ncolp]
    int ncolp;
    int ntiles;
    const void *geom;         // per element [NEN][2] double from a gather kernel, or null:
    const int32_t *conn;      //   then the kernel reads conn and xy itself,
    const double2 *xy;        //   stored padded with gaps ngap and xgap (0: plain)
    int ngap, xgap;
    double *vals;
    double *F;
    // Patch tiles (element-staged heat): a tile's rows are not one dof range
    // but runs of consecutive rows, listed per tile (tile_runs: first run,
    // run count).  A run is {first value, first row, LDS value start | values
    // << 16, tile-local row start | rows << 16}.  Null: contiguous tiles.
    const int2 *tile_runs = nullptr;
    const int4 *runs = nullptr;
    // heat (element-staged tiles): the task's element as its index in the
    // tile's list (< 256), one byte per task
    const uint8_t *task_elem8 = nullptr;
};

// The owner's entries of one element block.  The owner node a and its
// components are run-time values, so the owner's gradients are first pulled
// out of the register arrays with unrolled selects (no dynamic register
// indexing, hence no scratch); every entry is then formed with the same
// operations, in the same order, as the reference loop that produces it.
template <int KIND, typename R, int NEN, int NC, int NQ, bool CSC>
__device__ __forceinline__ void owner_values(const QTable &tab, const Geo<R, NEN, NQ> &g, const R *D, R kap, R mu,
                                             R Qe, int a, int comp0, int nr, bool powner, R (*out)[NEN * NC + 3],
                                             R *fo)
{
    constexpr int ND = NEN * NC;
    constexpr bool STOKES = (KIND == ELFEL_STOKES_GEN || KIND == ELFEL_STOKES_VECLAP);
    constexpr bool BDB = (KIND == ELFEL_ELAST || KIND == ELFEL_STOKES_GEN);
    R ox[NQ], oy[NQ], on[NQ], onp[NQ];
#pragma unroll
    for (int q = 0; q < NQ; q++) {
        ox[q] = g.gx[q][0]; oy[q] = g.gy[q][0]; on[q] = tab_n<R>(tab, q, 0); onp[q] = tab_np<R>(tab, q, 0);
#pragma unroll
        for (int k = 1; k < NEN; k++) {
            const bool m = (k == a);
            ox[q] = m ? g.gx[q][k] : ox[q];
            oy[q] = m ? g.gy[q][k] : oy[q];
            on[q] = m ? tab_n<R>(tab, q, k) : on[q];
            if (k < 3) onp[q] = m ? tab_np<R>(tab, q, k) : onp[q];
        }
    }
    if constexpr (STOKES) {
        if (powner) {
            // kup[:, a]: row a of transpose(kup) (CSR) or column a of kup (CSC)
#pragma unroll
            for (int i = 0; i < ND; i++) {
                R acc = R(0);
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    const R gi = (i % NC) == 0 ? g.gx[q][i / NC] : g.gy[q][i / NC];
                    acc += (-g.jxw[q] * onp[q]) * gi;
                }
                out[0][i] = acc;
            }
            return;
        }
    }
    if constexpr (sizeof(R) == 4) {
        // Mixed form: the owner's factor (D B_owner for a column owner, D^T
        // B_owner for a row owner, or the gradient) is scaled by the
        // coefficient and JxW once per point, and the structural zeros of B
        // are skipped: two products per entry and point.
        if constexpr (KIND == ELFEL_HEAT) {
            R wx[NQ], wy[NQ];
#pragma unroll
            for (int q = 0; q < NQ; q++) { wx[q] = ox[q] * (kap * g.jxw[q]); wy[q] = oy[q] * (kap * g.jxw[q]); }
#pragma unroll
            for (int j = 0; j < ND; j++) {
                R acc = R(0);
#pragma unroll
                for (int q = 0; q < NQ; q++) acc = fmaf(g.gx[q][j], wx[q], fmaf(g.gy[q][j], wy[q], acc));
                out[0][j] = acc;
            }
            R f = R(0);
#pragma unroll
            for (int q = 0; q < NQ; q++) f += on[q] * Qe * g.jxw[q];
            *fo = f;
            return;
        } else {
            // both owner rows are formed even for one-row tasks (rare: a node
            // with one free and one data dof), whose second row is never read:
            // skipping it made the compiler keep two of the entries in scratch
            (void)nr;
#pragma unroll
            for (int c = 0; c < 2; c++) {
                const int oc = comp0 + c;
                if constexpr (BDB) {
                    R t[NQ][3];
#pragma unroll
                    for (int q = 0; q < NQ; q++) {
                        const R b0 = oc == 0 ? ox[q] : R(0), b1 = oc == 0 ? R(0) : oy[q], b2 = oc == 0 ? oy[q] : ox[q];
#pragma unroll
                        for (int r = 0; r < 3; r++) {
                            const R v = CSC ? D[r] * b0 + D[r + 3] * b1 + D[r + 6] * b2
                                            : D[3 * r] * b0 + D[3 * r + 1] * b1 + D[3 * r + 2] * b2;
                            t[q][r] = v * g.jxw[q];
                        }
                    }
#pragma unroll
                    for (int i = 0; i < ND; i++) {
                        R acc = R(0);
#pragma unroll
                        for (int q = 0; q < NQ; q++) {
                            const R gx = g.gx[q][i / NC], gy = g.gy[q][i / NC];
                            acc = (i % NC) == 0 ? fmaf(t[q][0], gx, fmaf(t[q][2], gy, acc))
                                                : fmaf(t[q][1], gy, fmaf(t[q][2], gx, acc));
                        }
                        out[c][i] = acc;
                    }
                } else {  // vector Laplacian
                    R wx[NQ], wy[NQ];
#pragma unroll
                    for (int q = 0; q < NQ; q++) { wx[q] = ox[q] * (mu * g.jxw[q]); wy[q] = oy[q] * (mu * g.jxw[q]); }
#pragma unroll
                    for (int j = 0; j < ND; j++) {
                        R acc = R(0);
                        if ((j % NC) == oc) {
#pragma unroll
                            for (int q = 0; q < NQ; q++)
                                acc = fmaf(g.gx[q][j / NC], wx[q], fmaf(g.gy[q][j / NC], wy[q], acc));
                        }
                        out[c][j] = acc;
                    }
                }
                if constexpr (STOKES) {
#pragma unroll
                    for (int jp = 0; jp < 3; jp++) {
                        R acc = R(0);
#pragma unroll
                        for (int q = 0; q < NQ; q++) acc += (-g.jxw[q] * tab_np<R>(tab, q, jp)) * (oc == 0 ? ox[q] : oy[q]);
                        out[c][ND + jp] = acc;
                    }
                }
            }
            return;
        }
    }
#pragma unroll
    for (int c = 0; c < 2; c++) {
        if (c >= nr) continue;
        const int oc = comp0 + c;  // owner component
        if constexpr (KIND == ELFEL_HEAT) {
#pragma unroll
            for (int j = 0; j < ND; j++) {
                R acc = R(0);
#pragma unroll
                for (int q = 0; q < NQ; q++) {
                    if constexpr (CSC) acc += (g.gx[q][j] * ox[q] + g.gy[q][j] * oy[q]) * (kap * g.jxw[q]);
                    else acc += (ox[q] * g.gx[q][j] + oy[q] * g.gy[q][j]) * (kap * g.jxw[q]);
                }
                out[c][j] = acc;
            }
        } else if constexpr (BDB) {
            if constexpr (CSC) {
                R dbo[NQ][3];
#pragma unroll
                for (int q = 0; q < NQ; q++) dmul_b<R>(D, ox[q], oy[q], oc, dbo[q]);
#pragma unroll
                for (int i = 0; i < ND; i++) {
                    R acc = R(0);
#pragma unroll
                    for (int q = 0; q < NQ; q++) acc += dot_b<R>(dbo[q], g.gx[q][i / NC], g.gy[q][i / NC], i % NC) * g.jxw[q];
                    out[c][i] = acc;
                }
            } else {
#pragma unroll
                for (int j = 0; j < ND; j++) {
                    R acc = R(0);
#pragma unroll
                    for (int q = 0; q < NQ; q++) {
                        R db[3];
                        dmul_b<R>(D, g.gx[q][j / NC], g.gy[q][j / NC], j % NC, db);
                        acc += dot_b<R>(db, ox[q], oy[q], oc) * g.jxw[q];
                    }
                    out[c][j] = acc;
                }
            }
        } else {  // vector Laplacian: only equal components couple
#pragma unroll
            for (int j = 0; j < ND; j++) {
                R acc = R(0);
                if ((j % NC) == oc) {
#pragma unroll
                    for (int q = 0; q < NQ; q++) {
                        if constexpr (CSC) acc += (mu * g.jxw[q]) * (g.gx[q][j / NC] * ox[q] + g.gy[q][j / NC] * oy[q]);
                        else acc += (mu * g.jxw[q]) * (ox[q] * g.gx[q][j / NC] + oy[q] * g.gy[q][j / NC]);
                    }
                }
                out[c][j] = acc;
            }
        }
        if constexpr (STOKES) {
#pragma unroll
            for (int jp = 0; jp < 3; jp++) {
                R acc = R(0);
#pragma unroll
                for (int q = 0; q < NQ; q++) acc += (-g.jxw[q] * tab_np<R>(tab, q, jp)) * (oc == 0 ? ox[q] : oy[q]);
                out[c][ND + jp] = acc;
            }
        }
    }
    if constexpr (KIND == ELFEL_HEAT) {
        R acc = R(0);
#pragma unroll
        for (int q = 0; q < NQ; q++) acc += on[q] * Qe * g.jxw[q];
        *fo = acc;
    }
}

#ifndef REUSE_LDS_SKEW
#define REUSE_LDS_SKEW 0
#endif
// Stokes tasks add 15 columns into one or two rows: there one LDS atomic add
// per entry (ds_add_f64, FP64, no return) beats the read-all / write-all
// round trip (C3 kernel 189 against 206 us); for the other kinds it does not.
template <int KIND>
__host__ __device__ constexpr bool reuse_lds_add() { return KIND == ELFEL_STOKES_GEN || KIND == ELFEL_STOKES_VECLAP; }
// LDS slot of tile value v: with the skew, one spare slot after every 32
// values spreads the strided accesses of a rank step over more banks.
__device__ __forceinline__ int lds_slot(int v) { return REUSE_LDS_SKEW ? v + (v >> 5) : v; }
__host__ __device__ constexpr int lds_span(int nvals) { return REUSE_LDS_SKEW ? nvals + (nvals >> 5) + 1 : nvals; }

// Record length of the element-staged heat tiles, odd: the lanes of a wave
// read records of different elements, and an odd stride spreads them over the
// LDS banks.
template <int NEN>
__host__ __device__ constexpr int heat_rec() { return (NEN * NEN + NEN) | 1; }

// Element matrix ke[i][j] and load vector fe[j] of one heat element into its
// LDS record r (NEN*NEN + NEN values), from node coordinates X, Y (relative to
// node 0 for the mixed kernel), conductivity kap and source Qv.
template <typename R, int NEN, int NQ>
__device__ __forceinline__ void heat_record(const QTable &tab, double kap, double Qv, const double *X, const double *Y,
                                            R *r)
{
    constexpr bool MIXED = sizeof(R) == 4;
    constexpr QTable CT = qt_fixed<NEN, NQ>();
    Geo<R, NEN, NQ> g;
    if constexpr (MIXED) {
        geometry_mixed<NEN, NQ>(CT, X, Y, g);
        float wx[NQ][NEN], wy[NQ][NEN];
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            const float w = (float)kap * g.jxw[q];
#pragma unroll
            for (int j = 0; j < NEN; j++) { wx[q][j] = g.gx[q][j] * w; wy[q][j] = g.gy[q][j] * w; }
        }
#pragma unroll
        for (int i = 0; i < NEN; i++)
#pragma unroll
            for (int j = 0; j < NEN; j++) {
                float acc = 0.f;
#pragma unroll
                for (int q = 0; q < NQ; q++) acc = fmaf(g.gx[q][i], wx[q][j], fmaf(g.gy[q][i], wy[q][j], acc));
                r[i * NEN + j] = acc;
            }
#pragma unroll
        for (int j = 0; j < NEN; j++) {
            float acc = 0.f;
#pragma unroll
            for (int q = 0; q < NQ; q++) acc += CT.Nf[q][j] * (float)Qv * g.jxw[q];
            r[NEN * NEN + j] = acc;
        }
    } else {
        geometry<double, NEN, NQ>(tab, X, Y, g);
#pragma unroll
        for (int i = 0; i < NEN; i++)
#pragma unroll
            for (int j = 0; j < NEN; j++) {
                double acc = 0.0;
#pragma unroll
                for (int q = 0; q < NQ; q++)
                    acc += (g.gx[q][i] * g.gx[q][j] + g.gy[q][i] * g.gy[q][j]) * (kap * g.jxw[q]);
                r[i * NEN + j] = acc;
            }
#pragma unroll
        for (int j = 0; j < NEN; j++) {
            double acc = 0.0;
#pragma unroll
            for (int q = 0; q < NQ; q++) acc += tab.N[q][j] * Qv * g.jxw[q];
            r[NEN * NEN + j] = acc;
        }
    }
}

// Element-staged heat tiles.  A scalar task owns one row (CSR) or column (CSC)
// of NEN entries, so per-task geometry and loads would dominate; instead the
// tile's distinct elements (tile_elem, ~half the task count) are computed once
// each: the whole element matrix ke[i][j] and load vector fe[j], into LDS
// records that alias the tile's value sums.  Each task then reads its row or
// column of its element's record.  The FP64 records repeat the reference's
// operations (ke[i,j] += (grad N_i . grad N_j) kappa JxW, fe[j] += N_j Q JxW, in
// point order), so every entry has the same bits as the per-task form.
// The tile holds at most BLOCK distinct elements (the tile builder closes a
// tile before a row would exceed that), so thread t computes element t of the
// list, eg: its number, loaded by the caller before the task records.
template <typename R, int NEN, int NQ, bool CSC, int BLOCK, int TPT>
__device__ __forceinline__ void heat_tile(const ReuseArgs &A, const Coef &cf, const QTable &tab, double *lds,
                                          int32_t eg, int nU, const int32_t *e, const uint32_t *info,
                                          const bool *active, int *rank, int *vbase, int *L8, R (*out)[2][NEN + 3],
                                          R *fo)
{
    constexpr bool MIXED = sizeof(R) == 4;
    constexpr int RECN = heat_rec<NEN>();
    R *rec = (R *)lds;
    const int le = threadIdx.x;
    if (le < nU) {
        // the coefficients first: issued with the node numbers, they arrive
        // with them instead of adding a load round trip after the coordinates
        double kap = cf.kappa, Qv = cf.Q;
        if (cf.kappa_e) kap = cf.kappa_e[padded(eg, cf.cgap)];
        if (cf.Q_e) Qv = cf.Q_e[padded(eg, cf.cgap)];
        double X[NEN], Y[NEN];
        if (A.geom) {
            const double *gp = (const double *)A.geom + (int64_t)eg * (2 * NEN);
#pragma unroll
            for (int k = 0; k < NEN; k++) { X[k] = gp[2 * k]; Y[k] = gp[2 * k + 1]; }
        } else {
            int32_t nd[NEN];
#pragma unroll
            for (int k = 0; k < NEN; k++) nd[k] = A.conn[padded(eg, A.ngap) * NEN + k];
#pragma unroll
            for (int k = 0; k < NEN; k++) { const double2 p = A.xy[padded(nd[k], A.xgap)]; X[k] = p.x; Y[k] = p.y; }
            if constexpr (MIXED) {
                const double x0 = X[0], y0 = Y[0];
#pragma unroll
                for (int k = 0; k < NEN; k++) { X[k] -= x0; Y[k] -= y0; }
            }
        }
        heat_record<R, NEN, NQ>(tab, kap, Qv, X, Y, rec + (size_t)le * RECN);
    }
    __syncthreads();
#pragma unroll
    for (int u = 0; u < TPT; u++) {
        rank[u] = -1; vbase[u] = 0; L8[u] = 0; fo[u] = R(0);
        if (!active[u]) continue;
        const uint32_t in = info[u];
        const int a = in & 7;
        rank[u] = (in >> 3) & 15;
        vbase[u] = (in >> 10) & 0x3FFF;
        L8[u] = in >> 24;
        const R *r = rec + (size_t)e[u] * RECN;
#pragma unroll
        for (int j = 0; j < NEN; j++) out[u][0][j] = CSC ? r[j * NEN + a] : r[a * NEN + j];
        fo[u] = r[NEN * NEN + a];
    }
    // the caller's barrier separates these reads from the sums that overwrite the records
}

// One tile: a workgroup of BLOCK threads, TPT tasks per thread (task slot
// u * BLOCK + tid).  The loads of all of a thread's tasks are issued stage by
// stage (task records, then node numbers and coefficients, then coordinates)
// before any arithmetic, so light tasks (heat: one row, NEN columns) overlap
// their memory latency.
template <int KIND, typename R, int NEN, int NC, int NQ, bool CSC, int BLOCK, int TPT>
__device__ __forceinline__ void reuse_tile(const ReuseArgs &A, const Coef &cf, const QTable &tab, double *lds,
                                           int lds_vals, int tile)
{
    constexpr int ND = NEN * NC;
    constexpr bool STOKES = (KIND == ELFEL_STOKES_GEN || KIND == ELFEL_STOKES_VECLAP);
    constexpr int NCOL = ND + (STOKES ? 3 : 0);
    constexpr bool MIXED = sizeof(R) == 4;
    constexpr int SLOTS = BLOCK * TPT;
    double *acc = lds;
    double *accF = lds + lds_vals;
    const int tid = threadIdx.x;
    const int4 td = A.tiles[tile];
    const int4 tr = A.tile_rows[tile];
    if constexpr (KIND == ELFEL_HEAT)
        for (int v = tid; v < tr.y; v += BLOCK) accF[v] = 0.0;
    if constexpr (reuse_lds_add<KIND>()) {  // the adds below need the tile's sums zeroed
        double2 *a2 = (double2 *)acc;
        const int n2 = (lds_span(td.w) + 1) >> 1;
#pragma unroll 4
        for (int v = tid; v < n2; v += BLOCK) a2[v] = make_double2(0.0, 0.0);
    }

    // every slot holds a task record; pad slots skip the arithmetic and add nothing
    int64_t task[TPT];
    uint32_t info[TPT];
    int32_t e[TPT];
    bool active[TPT];
    // heat: this thread's element of the tile's list, first in the dependent
    // chain element -> nodes -> coordinates, so issued before everything else
    // (the list is padded to SLOTS entries per tile)
    int32_t eg = 0;
    if constexpr (KIND == ELFEL_HEAT) eg = A.tile_elem[(int64_t)tile * SLOTS + tid];
#pragma unroll
    for (int u = 0; u < TPT; u++) {
        task[u] = (int64_t)tile * SLOTS + u * BLOCK + tid;
        if constexpr (KIND == ELFEL_HEAT) e[u] = A.task_elem8[task[u]];
        else e[u] = A.task_elem[task[u]];
        info[u] = A.task_info[task[u]];
    }
    // column positions and first-contribution flags, prefetched as 32-bit
    // words before the barriered adds
    constexpr int NPW = (NCOL + 3) / 4;
    uint32_t posw[TPT][NPW];
#pragma unroll
    for (int u = 0; u < TPT; u++) {
        active[u] = ((info[u] >> 7) & 3) != 3;
        const uint32_t *pw = (const uint32_t *)(A.task_pos + task[u] * A.ncolp);
#pragma unroll
        for (int w = 0; w < NPW; w++) posw[u][w] = pw[w];
    }
    int rank[TPT], rowcode[TPT], vbase[TPT], L8[TPT];
    bool pown[TPT];
    R out[TPT][2][ND + 3];
    R fo[TPT];
#pragma unroll
    for (int u = 0; u < TPT; u++) { rowcode[u] = 0; pown[u] = false; }
    if constexpr (KIND == ELFEL_HEAT) {
        heat_tile<R, NEN, NQ, CSC, BLOCK, TPT>(A, cf, tab, lds, eg, tr.w, e, info, active, rank, vbase, L8, out, fo);
    } else {
    // node numbers (gather fused in) and per-element coefficients
    int32_t nd[TPT][NEN];
    double sc[TPT], kapd[TPT], Qd[TPT];
#pragma unroll
    for (int u = 0; u < TPT; u++) {
        sc[u] = 1.0; kapd[u] = cf.kappa; Qd[u] = cf.Q;
        if (active[u]) {
            if (!A.geom) {
#pragma unroll
                for (int k = 0; k < NEN; k++) nd[u][k] = A.conn[padded(e[u], A.ngap) * NEN + k];
            }
            if (cf.scale_e) sc[u] = cf.scale_e[padded(e[u], cf.cgap)];
            if (cf.kappa_e) kapd[u] = cf.kappa_e[padded(e[u], cf.cgap)];
            if (cf.Q_e) Qd[u] = cf.Q_e[padded(e[u], cf.cgap)];
        }
    }
    // coordinates: FP64, absolute for the FP64 kernel, relative to the
    // element's first node for the mixed kernel
    double X[TPT][NEN], Y[TPT][NEN];
#pragma unroll
    for (int u = 0; u < TPT; u++) {
        if (!active[u]) continue;
        if (A.geom) {
            const double *gp = (const double *)A.geom + (int64_t)e[u] * (2 * NEN);
#pragma unroll
            for (int k = 0; k < NEN; k++) { X[u][k] = gp[2 * k]; Y[u][k] = gp[2 * k + 1]; }
        } else {
#pragma unroll
            for (int k = 0; k < NEN; k++) { const double2 p = A.xy[padded(nd[u][k], A.xgap)]; X[u][k] = p.x; Y[u][k] = p.y; }
        }
    }
#pragma unroll
    for (int u = 0; u < TPT; u++) {
        rank[u] = -1; rowcode[u] = 0; vbase[u] = 0; L8[u] = 0; pown[u] = false; fo[u] = R(0);
        if (!active[u]) continue;
        const uint32_t in = info[u];
        const int a = in & 7;
        rank[u] = (in >> 3) & 15;
        rowcode[u] = (in >> 7) & 3;
        pown[u] = STOKES && ((in >> 9) & 1);
        vbase[u] = (in >> 10) & 0x3FFF;
        L8[u] = in >> 24;
        const int comp0 = rowcode[u] == 1 ? 1 : 0;
        const int nr = rowcode[u] == 2 ? 2 : 1;
        if constexpr (MIXED) {
            const double x0 = X[u][0], y0 = Y[u][0];
            if (!A.geom) {
#pragma unroll
                for (int k = 0; k < NEN; k++) { X[u][k] -= x0; Y[u][k] -= y0; }
            }
        }
        Geo<R, NEN, NQ> g;
        constexpr QTable CT = qt_fixed<NEN, NQ>();
        if constexpr (MIXED) geometry_mixed<NEN, NQ>(CT, X[u], Y[u], g);
        else geometry<double, NEN, NQ>(tab, X[u], Y[u], g);
        R Dr[9];
#pragma unroll
        for (int i = 0; i < 9; i++) Dr[i] = (R)(cf.scale_e ? sc[u] * cf.D[i] : cf.D[i]);
        owner_values<KIND, R, NEN, NC, NQ, CSC>(MIXED ? CT : tab, g, Dr, (R)kapd[u], (R)cf.mu, (R)Qd[u], a, comp0,
                                                nr, pown[u], out[u], &fo[u]);
    }
    }  // not heat
    __syncthreads();
    const int nranks = tr.z;
    for (int s = 0; s < nranks; s++) {
#pragma unroll
        for (int u = 0; u < TPT; u++) {
            if (!(active[u] && rank[u] == s)) continue;
            // A task's positions are distinct (distinct dofs of one element), so
            // all of its partial sums are read before any is written back: one
            // LDS round trip per rank instead of one per entry.
            const int ncols = (STOKES && pown[u]) ? ND : NCOL;
            const int nr = rowcode[u] == 2 ? 2 : 1;
            if constexpr (reuse_lds_add<KIND>()) {
                // one LDS add per entry; within a rank step no two lanes share an entry
#pragma unroll
                for (int c = 0; c < 2; c++) {
                    if (c >= nr) continue;
                    const int rbase = vbase[u] + c * L8[u];
#pragma unroll
                    for (int j = 0; j < NCOL; j++) {
                        const uint32_t pb = (posw[u][j >> 2] >> (8 * (j & 3))) & 0x7Fu;
                        if (j < ncols) atomicAdd(&acc[lds_slot(rbase + (int)pb)], (double)out[u][c][j]);
                    }
                }
                if constexpr (KIND == ELFEL_HEAT) accF[L8[u]] += (double)fo[u];
                continue;
            }
#pragma unroll
            for (int c = 0; c < 2; c++) {
                if (c >= nr) continue;
                const int rbase = vbase[u] + c * L8[u];
                double cur[NCOL];
#pragma unroll
                for (int j = 0; j < NCOL; j++) {
                    const uint32_t pb = (posw[u][j >> 2] >> (8 * (j & 3))) & 0xFFu;
                    if (j < ncols) cur[j] = (pb & 0x80u) ? 0.0 : acc[lds_slot(rbase + (int)(pb & 0x7Fu))];
                }
#pragma unroll
                for (int j = 0; j < NCOL; j++) {
                    const uint32_t pb = (posw[u][j >> 2] >> (8 * (j & 3))) & 0xFFu;
                    if (j < ncols) acc[lds_slot(rbase + (int)(pb & 0x7Fu))] = cur[j] + (double)out[u][c][j];
                }
            }
            if constexpr (KIND == ELFEL_HEAT) accF[L8[u]] += (double)fo[u];
        }
        __syncthreads();
    }
    // Coalesced write of the tile's values, 16 bytes per lane (the pair of
    // non-temporal stores merges into one): an odd first (or last) value of the
    // range goes out alone.  Non-temporal stores stream the values past the
    // caches -- nothing in this launch reads them again -- and cut the kernel
    // time by a quarter on MI355X.
    if (A.runs) {
        // patch tile: the LDS values are the tile's rows in dof order; each run
        // of consecutive rows goes out as one contiguous range, a wave per run
        const int2 rp = A.tile_runs[tile];
        const int wave = tid >> 6, lane = tid & 63;
        for (int k = wave; k < rp.y; k += BLOCK / 64) {
            const int4 run = A.runs[rp.x + k];
            const int lv0 = run.z & 0xFFFF, nv = run.z >> 16;
            double *rd = A.vals + run.x;
            const int rhead = run.x & 1;
            if (lane == 0) {
                if (rhead && nv > 0) __builtin_nontemporal_store(acc[lds_slot(lv0)], &rd[0]);
                if (((nv - rhead) & 1) && nv > rhead) __builtin_nontemporal_store(acc[lds_slot(lv0 + nv - 1)], &rd[nv - 1]);
            }
            const int np = (nv - rhead) >> 1;
            for (int p = lane; p < np; p += 64) {
                const int v = rhead + 2 * p;
                __builtin_nontemporal_store(acc[lds_slot(lv0 + v)], &rd[v]);
                __builtin_nontemporal_store(acc[lds_slot(lv0 + v + 1)], &rd[v + 1]);
            }
            if constexpr (KIND == ELFEL_HEAT) {
                const int lr0 = run.w & 0xFFFF, nr = run.w >> 16;
                for (int i = lane; i < nr; i += 64) A.F[run.y + i] = accF[lr0 + i];
            }
        }
        return;
    }
    double *dst = A.vals + td.z;
    const int head = td.z & 1;
    const int npair = (td.w - head) >> 1;
    if (tid == 0) {
        if (head && td.w > 0) __builtin_nontemporal_store(acc[lds_slot(0)], &dst[0]);
        if (((td.w - head) & 1) && td.w > head) __builtin_nontemporal_store(acc[lds_slot(td.w - 1)], &dst[td.w - 1]);
    }
#pragma unroll 4
    for (int p = tid; p < npair; p += BLOCK) {
        const int v = head + 2 * p;
        __builtin_nontemporal_store(acc[lds_slot(v)], &dst[v]);
        __builtin_nontemporal_store(acc[lds_slot(v + 1)], &dst[v + 1]);
    }
    if constexpr (KIND == ELFEL_HEAT)
        for (int v = tid; v < tr.y; v += BLOCK) A.F[tr.x + v] = accF[v];
}

// Tasks per thread of the single-problem kernels: heat tasks are light (one
// row of NEN columns), three per thread overlap their loads and make tiles of
// up to 192 rows, whose elements fewer other tiles share (heat Q4 4096^2 in
// band tiles: 2 / 3 / 4 tasks per thread 0.72 / 0.62 / 0.66 ms; heat T3 100^2
// 4.7 / 4.9 / 5.1 us).
__host__ __device__ constexpr int reuse_tpt(int kind) { return kind == ELFEL_HEAT ? 3 : 1; }
#define REUSE_MAX_TILE_ROWS 256  // tile-local row numbers of one-row groups are 8-bit

template <int KIND, typename R, int NEN, int NC, int NQ, bool CSC, int BLOCK>
__global__ __launch_bounds__(BLOCK) __attribute__((amdgpu_waves_per_eu(sizeof(R) == 4 ? REUSE_WAVES_PER_EU : 1)))
void k_reuse(ReuseArgs A, Coef cf, QTable tab, int lds_vals)
{
    extern __shared__ double lds[];
    reuse_tile<KIND, R, NEN, NC, NQ, CSC, BLOCK, reuse_tpt(KIND)>(A, cf, tab, lds, lds_vals, blockIdx.x);
}

// Arguments of the batched kernel, kept in device memory: passed by value, the
// two kinds' arguments (about 120 dwords) were loaded into scalar registers
// at once, and 67 of them were spilled into vector-register lanes.
struct BatchArgs {
    ReuseArgs H, Ev;
    Coef ch, ce;
};

// Batched problems of uneven size in one persistent launch: heat-T3 (1-point)
// tiles first, then elasticity-T6 (3-point) tiles; each workgroup walks the
// tile list with stride gridDim.x.  The tile kind is uniform per workgroup;
// heat tiles hold two tasks per thread, elasticity tiles one.  The mixed
// kernels take their quadrature tables from qt_fixed, so no table is passed.
template <bool CSC, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_reuse_batch(const BatchArgs *__restrict__ B, int nH, int nt,
                                                       int lds_vals)
{
    extern __shared__ double lds[];
    constexpr QTable th = qt_fixed<3, 1>(), te = qt_fixed<6, 3>();
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        if (t < nH) reuse_tile<ELFEL_HEAT, float, 3, 1, 1, CSC, BLOCK, 2>(B->H, B->ch, th, lds, lds_vals, t);
        else reuse_tile<ELFEL_ELAST, float, 6, 2, 3, CSC, BLOCK, 1>(B->Ev, B->ce, te, lds, lds_vals, t - nH);
        __syncthreads();
    }
}

// Geometry gather for the reuse kernels: per element, FP64 coordinates of its
// nodes -- absolute (FP64 kernel: the reference's operands exactly) or
// relative to the element's first node (mixed kernel).
template <int NEN, bool RELATIVE>
__global__ __launch_bounds__(256) void k_gather_geom(int64_t E, const int32_t *__restrict__ conn,
                                                     const double2 *__restrict__ xy, double *__restrict__ geom)
{
    // one work-item per (element, node) slot: conn is read and the records are
    // written with unit stride across the wavefront
    const int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= E * NEN) return;
    const double2 p = xy[conn[i]];
    double2 *g = (double2 *)geom;
    if constexpr (RELATIVE) {
        const double2 p0 = xy[conn[i - i % NEN]];
        g[i] = make_double2(p.x - p0.x, p.y - p0.y);
    } else {
        g[i] = p;
    }
}

// Host-side task tables of the reuse method (reuse.hip).
struct ReuseTables {
    std::vector<int4> tiles, trows;
    std::vector<int2> truns;     // patch tiles: per tile {first run, run count}
    std::vector<int4> runs;      // patch tiles: runs of consecutive rows (ReuseArgs::runs)
    std::vector<int32_t> elem;
    std::vector<uint32_t> info;
    std::vector<uint8_t> pos;
    std::vector<int32_t> telem;  // heat: distinct elements per tile (padded to the slot count)
    int ncolp = 0, maxv = 0, maxrank = 0, maxu = 0;
};
int pattern_to_host(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf,
                    std::vector<int32_t> &hptr, std::vector<int32_t> &hind);
int tile_tables(elfel_hip_ctx *c, bool stokes, const Field &F0, const Field *F1, int64_t n,
                const std::vector<int32_t> &hptr, const std::vector<int32_t> &hind, int cap, int ucap,
                ReuseTables &T, bool patch = false);
int upload_tables(elfel_hip_ctx *c, const ReuseTables &T, Pattern &P);
// Sorts the tiles by their first element and lays them out so that workgroup b
// (dispatched to XCD b mod nxcd) walks a contiguous chunk of the sorted list
// (run = 0), or runs of `run` tiles dealt to the XCDs in turn.
void order_tiles(ReuseTables &T, int nxcd, int cap, int run);
int env_int(const char *name, int dflt);

typedef int (*reuse_launch_fn)(elfel_hip_ctx *, const ReuseArgs &, const Coef &, const QTable &, int);
reuse_launch_fn pick_reuse64(int kind, int fe, int nq, int csc, int block);
reuse_launch_fn pick_reuse32(int kind, int fe, int nq, int csc, int block);
int gather_geom64(elfel_hip_ctx *c, int nen, int64_t E, const int32_t *conn, const double2 *xy, double *geom);
int gather_geom_rel(elfel_hip_ctx *c, int nen, int64_t E, const int32_t *conn, const double2 *xy, double *geom);

// Instantiation helpers for reuse64.hip / reuse32.hip.
template <int KIND, typename R, int NEN, int NC, int NQ, bool CSC, int BLOCK>
int launch_reuse(elfel_hip_ctx *c, const ReuseArgs &A, const Coef &cf, const QTable &tab, int lds_vals)
{
    if (A.ntiles <= 0) return 0;
    if (KIND == ELFEL_HEAT && A.maxu > BLOCK) return set_err(c, ELFEL_EINVAL, "internal: heat tile exceeds one element per thread");
    lds_vals = lds_span(lds_vals);
    if constexpr (KIND == ELFEL_HEAT)  // element records alias the value sums
        lds_vals = std::max(lds_vals, (int)(((size_t)A.maxu * heat_rec<NEN>() * sizeof(R) + 7) / 8));
    lds_vals = (lds_vals + 1) & ~1;  // even: the load-vector sums stay 16-byte aligned
    const size_t lds = sizeof(double) * ((size_t)lds_vals + (KIND == ELFEL_HEAT ? REUSE_MAX_TILE_ROWS : 0));
    hipLaunchKernelGGL((k_reuse<KIND, R, NEN, NC, NQ, CSC, BLOCK>), dim3(A.ntiles), dim3(BLOCK), lds, c->stream, A,
                       cf, tab, lds_vals);
    HIPCHK(c, hipGetLastError());
    return 0;
}

#define REUSE_PICK_BODY(R, B)                                                                           \
    if (kind == ELFEL_HEAT) {                                                                        \
        if (fe == ELFEL_T3 && nq == 1) return csc ? launch_reuse<ELFEL_HEAT, R, 3, 1, 1, true, B> : launch_reuse<ELFEL_HEAT, R, 3, 1, 1, false, B>; \
        if (fe == ELFEL_T3 && nq == 3) return csc ? launch_reuse<ELFEL_HEAT, R, 3, 1, 3, true, B> : launch_reuse<ELFEL_HEAT, R, 3, 1, 3, false, B>; \
        if (fe == ELFEL_T6 && nq == 3) return csc ? launch_reuse<ELFEL_HEAT, R, 6, 1, 3, true, B> : launch_reuse<ELFEL_HEAT, R, 6, 1, 3, false, B>; \
        if (fe == ELFEL_Q4 && nq == 1) return csc ? launch_reuse<ELFEL_HEAT, R, 4, 1, 1, true, B> : launch_reuse<ELFEL_HEAT, R, 4, 1, 1, false, B>; \
        if (fe == ELFEL_Q4 && nq == 4) return csc ? launch_reuse<ELFEL_HEAT, R, 4, 1, 4, true, B> : launch_reuse<ELFEL_HEAT, R, 4, 1, 4, false, B>; \
    } else if (kind == ELFEL_ELAST) {                                                                \
        if (fe == ELFEL_T3 && nq == 1) return csc ? launch_reuse<ELFEL_ELAST, R, 3, 2, 1, true, B> : launch_reuse<ELFEL_ELAST, R, 3, 2, 1, false, B>; \
        if (fe == ELFEL_T3 && nq == 3) return csc ? launch_reuse<ELFEL_ELAST, R, 3, 2, 3, true, B> : launch_reuse<ELFEL_ELAST, R, 3, 2, 3, false, B>; \
        if (fe == ELFEL_T6 && nq == 3) return csc ? launch_reuse<ELFEL_ELAST, R, 6, 2, 3, true, B> : launch_reuse<ELFEL_ELAST, R, 6, 2, 3, false, B>; \
        if (fe == ELFEL_Q4 && nq == 4) return csc ? launch_reuse<ELFEL_ELAST, R, 4, 2, 4, true, B> : launch_reuse<ELFEL_ELAST, R, 4, 2, 4, false, B>; \
    } else if (kind == ELFEL_STOKES_GEN) {                                                           \
        if (fe == ELFEL_T6 && nq == 3) return csc ? launch_reuse<ELFEL_STOKES_GEN, R, 6, 2, 3, true, B> : launch_reuse<ELFEL_STOKES_GEN, R, 6, 2, 3, false, B>; \
    } else if (kind == ELFEL_STOKES_VECLAP) {                                                        \
        if (fe == ELFEL_T6 && nq == 3) return csc ? launch_reuse<ELFEL_STOKES_VECLAP, R, 6, 2, 3, true, B> : launch_reuse<ELFEL_STOKES_VECLAP, R, 6, 2, 3, false, B>; \
    }                                                                                                \
    return nullptr;
