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
    // row-owner heat tiles (heat_rows_tile): one record of RW words per row,
    // BLOCK records per tile; tile_elem holds tstride entries per tile
    const uint32_t *rrec = nullptr;
    int tstride = 0;
    // row-owner heat tiles with tile-local geometry: the coordinates of the
    // nodes of each tile's elements, copied once per mesh into one contiguous
    // range per tile (txy, from tiles[t].x), and each element slot's nodes as
    // 16-bit indices into that range (tconn, tstride entries per tile; null:
    // the kernel gathers through conn and xy)
    const ushort4 *tconn = nullptr;
    const double2 *txy = nullptr;
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

// Coalesced write of a tile's values, 16 bytes per lane (the pair of
// non-temporal stores merges into one): an odd first (or last) value of the
// range goes out alone.  Non-temporal stores stream the values past the
// caches -- nothing in this launch reads them again -- and cut the kernel
// time by a quarter on MI355X.  HEAT: also the tile's load-vector sums accF.
template <bool HEAT, int BLOCK>
__device__ __forceinline__ void tile_writeout(const ReuseArgs &A, const double *acc, const double *accF, int tile,
                                              int4 td, int4 tr)
{
    const int tid = threadIdx.x;
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
            if constexpr (HEAT) {
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
    if constexpr (HEAT)
        for (int v = tid; v < tr.y; v += BLOCK) A.F[tr.x + v] = accF[v];
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

    // Elasticity: the waves issuing a tile's loads (task records, node numbers,
    // coordinates) run at raised priority, so that they get their requests
    // out ahead of waves that are computing (the 1131^2 mesh 0.607 -> 0.596
    // ms, the batch's elasticity problems 0.951 -> 0.934 ms; no change at
    // 400^2, Stokes slightly slower)
    if constexpr (KIND == ELFEL_ELAST) __builtin_amdgcn_s_setprio(3);
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
    if constexpr (KIND == ELFEL_ELAST) __builtin_amdgcn_s_setprio(0);
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
    tile_writeout<KIND == ELFEL_HEAT, BLOCK>(A, acc, accF, tile, td, tr);
}

// ---------------------------------------------------------------------------
// Row-owner heat tiles.  The tile's distinct elements are computed as in
// heat_tile (thread t: element t of the tile's list, its whole record into
// LDS); then thread t owns tile-local row t and adds that row's
// contributions -- its elements in element order, within an element its NEN
// columns -- into the row's LDS values.  A row's sums are one thread's, so no
// rank steps and no barriers between contributions: each entry starts from
// its first contribution and adds the rest in element order, the order of
// sparse().  Per row one record of RW 32-bit words replaces the per-task
// element, descriptor and position bytes:
//   bits 0-3    count: elements of the row (<= KMAX)
//   bits 4-16   LDS offset of the row's values in the tile
//   then per element k < count, at bit 17 + k (10 + 4 NEN):
//     8 bits slot (index in the tile's element list), 2 bits local node a of
//     the row in the element, NEN x 4 bits the position of each element
//     column in the row's sorted column list (rows of <= 16 entries)
// ---------------------------------------------------------------------------
template <int NEN>
__host__ __device__ constexpr int hrow_field_bits() { return 10 + 4 * NEN; }
template <int NEN, int KMAX>
__host__ __device__ constexpr int hrow_words() { return (17 + KMAX * hrow_field_bits<NEN>() + 31) / 32; }
// The largest element count per row of the record formats built here: Q4
// rows (4 elements at interior nodes) fit 4 words, T3 rows 7 elements in 6.
template <int NEN>
__host__ __device__ constexpr int hrow_kmax() { return NEN == 4 ? 4 : 7; }

// bits [o, o + w) of the record (o, w known at compile time after unrolling)
template <int RW>
__device__ __forceinline__ uint32_t hrow_bits(const uint32_t (&w)[RW], int o, int nb)
{
    const int i = o >> 5, sh = o & 31;
    uint32_t v = w[i] >> sh;
    if (sh + nb > 32 && i + 1 < RW) v |= w[i + 1] << (32 - sh);
    return v & ((nb >= 32) ? 0xFFFFFFFFu : ((1u << nb) - 1u));
}

template <typename R, int NEN, int NQ, bool CSC, int BLOCK>
__device__ __forceinline__ void heat_rows_tile(const ReuseArgs &A, const Coef &cf, const QTable &tab, double *lds,
                                               int lds_vals, int tile)
{
    constexpr int KMAX = hrow_kmax<NEN>();
    constexpr int RW = hrow_words<NEN, KMAX>();
    constexpr int FB = hrow_field_bits<NEN>();
    constexpr int RECN = heat_rec<NEN>();
    double *acc = lds;
    double *accF = lds + lds_vals;
    R *rec = (R *)lds;
    const int tid = threadIdx.x;
    // the element number first (head of the chain element -> nodes ->
    // coordinates), with the tile descriptors
    const int32_t eg = A.tile_elem[(int64_t)tile * A.tstride + tid];
    const int4 td = A.tiles[tile];
    const int4 tr = A.tile_rows[tile];
    const int nrows = tr.y, nU = tr.w;
    uint32_t w[RW];
#pragma unroll
    for (int i = 0; i < RW; i++) w[i] = 0u;
    if (tid < nrows) {
        const uint32_t *rp = A.rrec + ((int64_t)tile * BLOCK + tid) * RW;
        if constexpr (RW % 4 == 0) {
#pragma unroll
            for (int i = 0; i < RW; i += 4) {
                const uint4 q = *(const uint4 *)(rp + i);
                w[i] = q.x; w[i + 1] = q.y; w[i + 2] = q.z; w[i + 3] = q.w;
            }
        } else {
#pragma unroll
            for (int i = 0; i < RW; i += 2) {
                const uint2 q = *(const uint2 *)(rp + i);
                w[i] = q.x; w[i + 1] = q.y;
            }
        }
    }
    if (tid < nU) {
        // 32-bit index arithmetic: element and node counts stay below 2^31 / NEN
        double kap = cf.kappa, Qv = cf.Q;
        const uint32_t ec = (uint32_t)padded32(eg, cf.cgap);
        if (cf.kappa_e) kap = cf.kappa_e[ec];
        if (cf.Q_e) Qv = cf.Q_e[ec];
        double X[NEN], Y[NEN];
        if (A.tconn) {
            // tile-local geometry: no dependent load through the element number
            const ushort4 ln = A.tconn[(int64_t)tile * A.tstride + tid];
            const double2 *lxy = A.txy + td.x;
            const uint32_t l[4] = {ln.x, ln.y, ln.z, ln.w};
#pragma unroll
            for (int k = 0; k < NEN; k++) {
                const double2 p = lxy[l[k]];
                X[k] = p.x; Y[k] = p.y;
            }
        } else {
            int32_t nd[NEN];
            const uint32_t en = (uint32_t)padded32(eg, A.ngap) * NEN;
#pragma unroll
            for (int k = 0; k < NEN; k++) nd[k] = A.conn[en + k];
#pragma unroll
            for (int k = 0; k < NEN; k++) {
                const double2 p = A.xy[(uint32_t)padded32(nd[k], A.xgap)];
                X[k] = p.x; Y[k] = p.y;
            }
        }
        if constexpr (sizeof(R) == 4) {
            const double x0 = X[0], y0 = Y[0];
#pragma unroll
            for (int k = 0; k < NEN; k++) { X[k] -= x0; Y[k] -= y0; }
        }
        heat_record<R, NEN, NQ>(tab, kap, Qv, X, Y, rec + (size_t)tid * RECN);
    }
    __syncthreads();
    // the row's contributions into registers (the records are overwritten next)
    const int cnt = (int)(w[0] & 15u);
    const int vbase = (int)((w[0] >> 4) & 0x1FFFu);
    R v[KMAX][NEN];
    uint32_t pos[KMAX];  // NEN x 4 bits
    double fsum = 0.0;   // start! zeros the load vector; the elements add in order
#pragma unroll
    for (int k = 0; k < KMAX; k++) {
        pos[k] = 0u;
#pragma unroll
        for (int j = 0; j < NEN; j++) v[k][j] = R(0);
        if (k < cnt) {
            const int o = 17 + k * FB;
            const int slot = (int)hrow_bits<RW>(w, o, 8);
            const int a = (int)hrow_bits<RW>(w, o + 8, 2);
            pos[k] = hrow_bits<RW>(w, o + 10, 4 * NEN);
            const R *r = rec + (size_t)slot * RECN;
#pragma unroll
            for (int j = 0; j < NEN; j++) v[k][j] = CSC ? r[j * NEN + a] : r[a * NEN + j];
            fsum += (double)r[NEN * NEN + a];
        }
    }
    __syncthreads();
    if (tid < nrows) {
        uint32_t seen = 0u;
#pragma unroll
        for (int k = 0; k < KMAX; k++) {
            if (k >= cnt) break;
#pragma unroll
            for (int j = 0; j < NEN; j++) {
                const int p = (int)((pos[k] >> (4 * j)) & 15u);
                double *dst = &acc[lds_slot(vbase + p)];
                if ((seen >> p) & 1u) *dst = *dst + (double)v[k][j];
                else *dst = (double)v[k][j];
                seen |= 1u << p;
            }
        }
        accF[tid] = fsum;
    }
    __syncthreads();
    tile_writeout<true, BLOCK>(A, acc, accF, tile, td, tr);
}

template <typename R, int NEN, int NQ, bool CSC, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_heat_rows(ReuseArgs A, Coef cf, QTable tab, int lds_vals)
{
    extern __shared__ double lds[];
    heat_rows_tile<R, NEN, NQ, CSC, BLOCK>(A, cf, tab, lds, lds_vals, blockIdx.x);
}

#define HEAT_ROWS_BLOCK 256

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

// Batched problems of uneven size in one persistent launch.  Three tile
// kinds: heat T3 1-point (tiles 0 .. n0-1), heat Q4 2x2 (n0 .. n1-1) and
// plane elasticity T6 3-point (n1 .. nt-1); each workgroup walks the tile list
// with stride gridDim.x, and the tile kind is uniform per workgroup.  Heat
// tiles are row-owner tiles in workgroups of 256 (two tasks per thread in
// workgroups of 128), elasticity tiles hold one task per thread.  R = float: the
// mixed arithmetic; R = double: the FP64 arithmetic (built without
// contraction, equal to the single-problem FP64 kernels bit for bit).  The
// arguments are passed by value: kept in device memory and loaded at each
// tile, they added a dependent load to every tile's start (256 problems:
// 1.47 against 1.24 ms).  The quadrature tables are the compile-time ones.
// KINDS: the tile kinds compiled in (1 heat T3, 2 heat Q4, 4 elasticity T6);
// a kind left out takes no registers (with all three the mixed kernel needs
// 129 VGPRs and drops to 3 waves per SIMD, heat T3 + elasticity T6 alone 127).
// A heat tile of the batch: row-owner records (heat_rows_tile) when the
// kind's tables carry them (workgroups of HEAT_ROWS_BLOCK threads only), else
// two tasks per thread.
template <typename R, int NEN, int NQ, bool CSC, int BLOCK>
__device__ __forceinline__ void batch_heat_tile(const ReuseArgs &A, const Coef &cf, const QTable &tab, double *lds,
                                                int lds_vals, int tile)
{
    if constexpr (BLOCK == HEAT_ROWS_BLOCK) {
        if (A.rrec) {
            heat_rows_tile<R, NEN, NQ, CSC, BLOCK>(A, cf, tab, lds, lds_vals, tile);
            return;
        }
    }
    reuse_tile<ELFEL_HEAT, R, NEN, 1, NQ, CSC, BLOCK, 2>(A, cf, tab, lds, lds_vals, tile);
}

template <typename R, bool CSC, int BLOCK, int KINDS>
__global__ __launch_bounds__(BLOCK) void k_reuse_batch(ReuseArgs H3, ReuseArgs Ev, ReuseArgs H4, Coef ch3, Coef ce,
                                                       Coef ch4, int n0, int n1, int nt, int lds_vals)
{
    extern __shared__ double lds[];
    constexpr QTable t3 = qt_fixed<3, 1>(), t6 = qt_fixed<6, 3>(), t4 = qt_fixed<4, 4>();
    for (int t = blockIdx.x; t < nt; t += gridDim.x) {
        if (t < n0) {
            if constexpr ((KINDS & 1) != 0) batch_heat_tile<R, 3, 1, CSC, BLOCK>(H3, ch3, t3, lds, lds_vals, t);
        } else if (t < n1) {
            if constexpr ((KINDS & 2) != 0) batch_heat_tile<R, 4, 4, CSC, BLOCK>(H4, ch4, t4, lds, lds_vals, t - n0);
        } else {
            if constexpr ((KINDS & 4) != 0)
                reuse_tile<ELFEL_ELAST, R, 6, 2, 3, CSC, BLOCK, 1>(Ev, ce, t6, lds, lds_vals, t - n1);
        }
        __syncthreads();
    }
}

// The same tiles, one workgroup per tile (the launch shape the autotuner
// picks on 256 problems): without the grid-stride loop each branch loads only
// its own kind's arguments.
template <typename R, bool CSC, int BLOCK, int KINDS>
__global__ __launch_bounds__(BLOCK) void k_reuse_batch1(ReuseArgs H3, ReuseArgs Ev, ReuseArgs H4, Coef ch3, Coef ce,
                                                        Coef ch4, int n0, int n1, int lds_vals)
{
    extern __shared__ double lds[];
    constexpr QTable t3 = qt_fixed<3, 1>(), t6 = qt_fixed<6, 3>(), t4 = qt_fixed<4, 4>();
    const int t = blockIdx.x;
    if (t < n0) {
        if constexpr ((KINDS & 1) != 0) batch_heat_tile<R, 3, 1, CSC, BLOCK>(H3, ch3, t3, lds, lds_vals, t);
    } else if (t < n1) {
        if constexpr ((KINDS & 2) != 0) batch_heat_tile<R, 4, 4, CSC, BLOCK>(H4, ch4, t4, lds, lds_vals, t - n0);
    } else {
        if constexpr ((KINDS & 4) != 0)
            reuse_tile<ELFEL_ELAST, R, 6, 2, 3, CSC, BLOCK, 1>(Ev, ce, t6, lds, lds_vals, t - n1);
    }
}

// Launch of the batched kernel (batch.hip: mixed, batch64.hip: FP64).
struct BatchLaunch {
    ReuseArgs A[3];  // heat T3, elasticity T6, heat Q4
    Coef cf[3];
    int n0, n1, nt, lds_vals, block, grid;
    size_t lds;
    bool csc;
};
int launch_batch32(hipStream_t st, const BatchLaunch &L);
int launch_batch64(hipStream_t st, const BatchLaunch &L);

// Instantiations: heat T3 + elasticity T6 (KINDS 5; also batches of one of
// them), and all three kinds (7) for batches with heat Q4.
template <typename R>
int launch_batch_t(hipStream_t st, const BatchLaunch &L)
{
    const bool q4 = L.n1 > L.n0;
#define KB(C, BL, K)                                                                                           \
    do {                                                                                                       \
        if (L.grid >= L.nt)                                                                                    \
            hipLaunchKernelGGL((k_reuse_batch1<R, C, BL, K>), dim3(L.nt), dim3(BL), L.lds, st, L.A[0], L.A[1], \
                               L.A[2], L.cf[0], L.cf[1], L.cf[2], L.n0, L.n1, L.lds_vals);                     \
        else                                                                                                   \
            hipLaunchKernelGGL((k_reuse_batch<R, C, BL, K>), dim3(L.grid), dim3(BL), L.lds, st, L.A[0], L.A[1], \
                               L.A[2], L.cf[0], L.cf[1], L.cf[2], L.n0, L.n1, L.nt, L.lds_vals);               \
    } while (0)
#define KB2(C, BL) do { if (q4) KB(C, BL, 7); else KB(C, BL, 5); } while (0)
    if (L.block == 128) { if (L.csc) KB2(true, 128); else KB2(false, 128); }
    else { if (L.csc) KB2(true, 256); else KB2(false, 256); }
#undef KB2
#undef KB
    return hipGetLastError() == hipSuccess ? 0 : -1;
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
    std::vector<uint32_t> rrec;  // row-owner heat tiles: HEAT_ROWS_BLOCK records of rw words per tile
    std::vector<ushort4> tconn;  // row-owner heat tiles: tile-local nodes per element slot (tstride per tile)
    std::vector<double2> txy;    // row-owner heat tiles: tile-local coordinates (tile t's from tiles[t].x)
    int rw = 0;                  // words per row record (0: task tables)
    int tstride = 0;             // entries of telem per tile
    int ncolp = 0, maxv = 0, maxrank = 0, maxu = 0;
};
int pattern_to_host(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf,
                    std::vector<int32_t> &hptr, std::vector<int32_t> &hind);
int tile_tables(elfel_hip_ctx *c, bool stokes, const Field &F0, const Field *F1, int64_t n,
                const std::vector<int32_t> &hptr, const std::vector<int32_t> &hind, int cap, int ucap,
                ReuseTables &T, bool patch = false, bool rows = false, const double2 *host_xy = nullptr,
                std::string *errp = nullptr);
int upload_tables(elfel_hip_ctx *c, const ReuseTables &T, Pattern &P);
// symbolic.hip: the pattern and task tables of an elasticity or Stokes problem
// built on the device (c->pat; fallback set when the mesh does not fit)
int device_symbolic(elfel_hip_ctx *c, const elfel_hip_problem *pb, int block);
// the same for the row-owner heat tiles of scalar T3 / Q4 problems; 1: the
// mesh does not fit the row records (the host builder makes task tiles)
int device_symbolic_heat(elfel_hip_ctx *c, const elfel_hip_problem *pb, bool patch_allowed);
// Sorts the tiles by their first element and lays them out so that workgroup b
// (dispatched to XCD b mod nxcd) walks a contiguous chunk of the sorted list
// (run = 0), or runs of `run` tiles dealt to the XCDs in turn.
void order_tiles(ReuseTables &T, int nxcd, int cap, int run);
int env_int(const char *name, int dflt);

typedef int (*reuse_launch_fn)(elfel_hip_ctx *, const ReuseArgs &, const Coef &, const QTable &, int);
reuse_launch_fn pick_reuse64(int kind, int fe, int nq, int csc, int block);
reuse_launch_fn pick_reuse32(int kind, int fe, int nq, int csc, int block);
// row-owner heat kernels (T3 and Q4, workgroups of HEAT_ROWS_BLOCK threads)
reuse_launch_fn pick_heat_rows64(int fe, int nq, int csc);
reuse_launch_fn pick_heat_rows32(int fe, int nq, int csc);
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

template <typename R, int NEN, int NQ, bool CSC>
int launch_heat_rows(elfel_hip_ctx *c, const ReuseArgs &A, const Coef &cf, const QTable &tab, int lds_vals)
{
    if (A.ntiles <= 0) return 0;
    if (A.maxu > HEAT_ROWS_BLOCK) return set_err(c, ELFEL_EINVAL, "internal: heat tile exceeds one element per thread");
    lds_vals = lds_span(lds_vals);
    // element records alias the value sums
    lds_vals = std::max(lds_vals, (int)(((size_t)A.maxu * heat_rec<NEN>() * sizeof(R) + 7) / 8));
    lds_vals = (lds_vals + 1) & ~1;
    const size_t lds = sizeof(double) * ((size_t)lds_vals + REUSE_MAX_TILE_ROWS);
    hipLaunchKernelGGL((k_heat_rows<R, NEN, NQ, CSC, HEAT_ROWS_BLOCK>), dim3(A.ntiles), dim3(HEAT_ROWS_BLOCK), lds,
                       c->stream, A, cf, tab, lds_vals);
    HIPCHK(c, hipGetLastError());
    return 0;
}

#define HEAT_ROWS_PICK_BODY(R)                                                                        \
    if (fe == ELFEL_T3 && nq == 1) return csc ? launch_heat_rows<R, 3, 1, true> : launch_heat_rows<R, 3, 1, false>; \
    if (fe == ELFEL_T3 && nq == 3) return csc ? launch_heat_rows<R, 3, 3, true> : launch_heat_rows<R, 3, 3, false>; \
    if (fe == ELFEL_Q4 && nq == 1) return csc ? launch_heat_rows<R, 4, 1, true> : launch_heat_rows<R, 4, 1, false>; \
    if (fe == ELFEL_Q4 && nq == 4) return csc ? launch_heat_rows<R, 4, 4, true> : launch_heat_rows<R, 4, 4, false>; \
    return nullptr;

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
