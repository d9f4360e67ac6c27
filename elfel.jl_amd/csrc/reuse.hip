// reuse.hip -- the REUSE method: symbolic phase (pattern, tiles, task tables)
// built once per mesh + dof map, then one gather + one fused owner-computes
// kernel per assembly (reuse_kernel.h).
//
// Symbolic phase:
//   1. the pattern (ptr/ind) comes from the plain method on the device: the
//      union of the element blocks including structural zeros, exactly the
//      pattern sparse() gives the reference's COO stream;
//   2. node -> element adjacency, in element order (a stable counting sort);
//   3. row groups (consecutive dofs of one node) in dof order, packed into
//      tiles of <= REUSE_BLOCK tasks and <= REUSE_VCAP values;
//   4. per task, the position of every element column in the owner's sorted
//      column list (binary search).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <cmath>
#include <chrono>
#include <mutex>
#include <thread>

#include "reuse_kernel.h"

namespace {

struct Adj {
    std::vector<int32_t> ptr, elem;
    std::vector<uint8_t> loc;
};

void adjacency(const Field &f, Adj &A)
{
    const int64_t Nn = f.Nn, E = f.E;
    const int nen = f.nen;
    A.ptr.assign(Nn + 1, 0);
    for (int64_t i = 0; i < E * nen; i++) A.ptr[f.h_conn[i] + 1]++;
    for (int64_t n = 0; n < Nn; n++) A.ptr[n + 1] += A.ptr[n];
    std::vector<int32_t> fill(A.ptr.begin(), A.ptr.end() - 1);
    A.elem.resize(E * nen);
    A.loc.resize(E * nen);
    for (int64_t e = 0; e < E; e++)
        for (int k = 0; k < nen; k++) {
            const int32_t n = f.h_conn[e * nen + k];
            const int32_t p = fill[n]++;
            A.elem[p] = (int32_t)e;
            A.loc[p] = (uint8_t)k;
        }
}

bool rank_order_enabled();

}  // namespace

int env_int(const char *name, int dflt)
{
    const char *v = getenv(name);
    return (v && *v) ? atoi(v) : dflt;
}

namespace {
bool rank_order_enabled() { return env_int("ELFEL_HIP_RANK_ORDER", 1) != 0; }
}  // namespace

// Pattern of the problem described by the context's fields (plain method, FP64
// values unused), copied to the host; the device copy stays in c->ptr/c->ind.
int pattern_to_host(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf,
                    std::vector<int32_t> &hptr, std::vector<int32_t> &hind)
{
    elfel_hip_problem p2 = *pb;
    p2.method = ELFEL_SORT;
    p2.precision = ELFEL_FP64;
    int rc = plain_assemble(c, &p2, tab, cf);
    if (rc) return rc;
    int32_t hnnz = 0;
    HIPCHK(c, hipMemcpyAsync(&hnnz, c->dscal.p, 4, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    const int64_t n = c->nrow;
    hptr.resize(n + 1);
    hind.resize(hnnz);
    HIPCHK(c, hipMemcpy(hptr.data(), c->ptr.p, 4 * (n + 1), hipMemcpyDeviceToHost));
    if (hnnz) HIPCHK(c, hipMemcpy(hind.data(), c->ind.p, 4 * (size_t)hnnz, hipMemcpyDeviceToHost));
    return 0;
}

namespace {

// What every row range of the tile builder reads.
struct TileInput {
    bool stokes, staged, by_rank;
    const int32_t *seq;  // patch tiles: the rows in packing order (null: dof order)
    int nen, nc, ncol, ncolp, cap, ucap, vcap;
    int rw, kmax;  // row-owner heat tiles: words per row record, most elements per row (rw 0: task tables)
    const double2 *hxy;  // row-owner heat tiles: host coordinates for the tile-local geometry (null: none)
    int64_t n;
    const Field *F0, *F1;
    const Adj *A0, *A1;
    const std::vector<int32_t> *hptr, *hind, *inv_node;
    const std::vector<uint8_t> *inv_fc;
};

// Rows r0..r1-1 (whole row groups) packed into tiles; tile task offsets are
// local to T.  stamp / tile_id belong to the calling worker (a running tile
// counter, so stamp never needs clearing).
int tile_range(const TileInput &I, int64_t r0, int64_t r1, ReuseTables &T, std::vector<int32_t> &stamp,
               int &tile_id, std::string &err)
{
    const int nen = I.nen, nc = I.nc, ncolp = I.ncolp, cap = I.cap, ucap = I.ucap, vcap = I.vcap;
    const bool stokes = I.stokes, staged = I.staged, by_rank = I.by_rank;
    const Field &F0 = *I.F0;
    const Field *F1 = I.F1;
    const std::vector<int32_t> &hptr = *I.hptr, &hind = *I.hind, &inv_node = *I.inv_node;
    const std::vector<uint8_t> &inv_fc = *I.inv_fc;
    const int64_t n = I.n;
    T = ReuseTables{};
    T.ncolp = ncolp;
    T.rw = I.rw;
    T.tstride = I.rw ? HEAT_ROWS_BLOCK : cap;
    int4 cur = {0, 0, 0, 0}, cr = {0, 0, 0, 0};
    bool open = false;
    int nu_open = 0;
    std::vector<int32_t> te;
    std::vector<uint32_t> ti;
    std::vector<uint8_t> tp;
    // patch tiles: the open tile's rows in insertion order (row, values)
    std::vector<std::pair<int32_t, int>> orows;
    // row-owner records of the open tile: task of (row, rank), counts, LDS offsets
    std::vector<int> rrow, rcnt, rvb;
    bool row_err = false;
    std::vector<int> ord, newidx, newvb;
    std::vector<int32_t> lnodes;
    auto close = [&]() {
        if (!open) return;
        if (I.seq) {
            // The rows were taken in locality order; lay them out in the tile's
            // LDS in dof order, re-address the tasks (value base, tile-local
            // row) and list the runs of consecutive rows for the write-out.
            const int nrw = (int)orows.size();
            ord.resize(nrw);
            for (int k = 0; k < nrw; k++) ord[k] = k;
            std::sort(ord.begin(), ord.end(), [&](int a, int b) { return orows[a].first < orows[b].first; });
            newidx.assign(nrw, 0);
            newvb.assign(nrw, 0);
            int vb = 0;
            for (int k = 0; k < nrw; k++) { newidx[ord[k]] = k; newvb[ord[k]] = vb; vb += orows[ord[k]].second; }
            for (size_t k = (size_t)cur.x; k < T.info.size(); k++) {
                const uint32_t in = T.info[k];
                const int old = (int)(in >> 24);
                T.info[k] = (in & 0x3FFu) | ((uint32_t)newvb[old] << 10) | ((uint32_t)newidx[old] << 24);
            }
            const int first = (int)T.runs.size();
            for (int k = 0; k < nrw;) {
                int m = k + 1;
                while (m < nrw && orows[ord[m]].first == orows[ord[m - 1]].first + 1) m++;
                const int32_t r0 = orows[ord[k]].first;
                const int nv = newvb[ord[m - 1]] + orows[ord[m - 1]].second - newvb[ord[k]];
                T.runs.push_back(int4{hptr[r0], r0, newvb[ord[k]] | (nv << 16), k | ((m - k) << 16)});
                k = m;
            }
            T.truns.push_back(int2{first, (int)T.runs.size() - first});
            cur.z = hptr[orows[ord[0]].first];
            cr.x = orows[ord[0]].first;
            orows.clear();
        }
        // Within a tile the tasks are laid out rank by rank (stable), so the
        // threads that add in one rank step sit in as few waves as possible and
        // every other wave skips that step.
        if (by_rank && cur.y > 1) {
            const size_t t0 = (size_t)cur.x, nt = (size_t)cur.y;
            te.assign(T.elem.begin() + t0, T.elem.end());
            ti.assign(T.info.begin() + t0, T.info.end());
            tp.assign(T.pos.begin() + t0 * ncolp, T.pos.end());
            size_t o = t0;
            for (int s = 0; s < cr.z; s++)
                for (size_t k = 0; k < nt; k++)
                    if ((int)((ti[k] >> 3) & 15) == s) {
                        T.elem[o] = te[k];
                        T.info[o] = ti[k];
                        std::copy(tp.begin() + k * ncolp, tp.begin() + (k + 1) * ncolp, T.pos.begin() + o * ncolp);
                        o++;
                    }
        }
        if (staged) {
            // the tile's distinct elements (ascending); tasks then name their
            // element by its index in this list
            const size_t t0 = (size_t)cur.x, nt = (size_t)cur.y;
            te.assign(T.elem.begin() + t0, T.elem.end());
            std::sort(te.begin(), te.end());
            te.erase(std::unique(te.begin(), te.end()), te.end());
            for (size_t k = 0; k < nt; k++)
                T.elem[t0 + k] = (int32_t)(std::lower_bound(te.begin(), te.end(), T.elem[t0 + k]) - te.begin());
            cr.w = (int)te.size();
            T.maxu = std::max(T.maxu, cr.w);
            T.telem.insert(T.telem.end(), te.begin(), te.end());
            T.telem.insert(T.telem.end(), (size_t)T.tstride - te.size(), te.empty() ? 0 : te[0]);
        }
        if (I.rw) {
            // one record per tile-local row: its tasks in rank (element) order
            const size_t t0 = (size_t)cur.x, nt = (size_t)cur.y;
            const int nrw = cr.y;
            rrow.assign((size_t)nrw * I.kmax, -1);
            rcnt.assign((size_t)nrw, 0);
            rvb.assign((size_t)nrw, 0);
            for (size_t k = 0; k < nt; k++) {
                const uint32_t in = T.info[t0 + k];
                const int row = (int)(in >> 24), rank = (int)((in >> 3) & 15);
                if (row >= nrw || rank >= I.kmax) { err = "internal: row record"; row_err = true; return; }
                rrow[(size_t)row * I.kmax + rank] = (int)k;
                rcnt[row] = std::max(rcnt[row], rank + 1);
                rvb[row] = (int)((in >> 10) & 0x3FFF);
            }
            const size_t base = T.rrec.size();
            T.rrec.resize(base + (size_t)HEAT_ROWS_BLOCK * I.rw, 0u);
            uint32_t *rec = T.rrec.data() + base;
            const int fb = 10 + 4 * I.nen;
            for (int r = 0; r < nrw; r++) {
                uint32_t *w = rec + (size_t)r * I.rw;
                auto put = [&](int o, uint32_t v, int nb) {
                    for (int b = 0; b < nb; b++)
                        if ((v >> b) & 1u) w[(o + b) >> 5] |= 1u << ((o + b) & 31);
                };
                put(0, (uint32_t)rcnt[r], 4);
                put(4, (uint32_t)rvb[r], 13);
                for (int k = 0; k < rcnt[r]; k++) {
                    const int q = rrow[(size_t)r * I.kmax + k];
                    if (q < 0) { err = "internal: row record gap"; row_err = true; return; }
                    const uint32_t in = T.info[t0 + q];
                    const int o = 17 + k * fb;
                    put(o, (uint32_t)T.elem[t0 + q], 8);
                    put(o + 8, in & 7u, 2);
                    for (int j = 0; j < I.nen; j++) put(o + 10 + 4 * j, T.pos[(t0 + q) * ncolp + j] & 0x7Fu, 4);
                }
            }
        }
        if (I.rw) {
            // row-owner tiles: the records hold everything; the tasks are dropped
            T.elem.resize((size_t)cur.x);
            T.info.resize((size_t)cur.x);
            T.pos.resize((size_t)cur.x * ncolp);
            if (I.hxy) {
                // tile-local geometry: the distinct nodes of the tile's elements
                // (ascending), their coordinates from txy[cur.x], and per element
                // slot its nodes' indices in that list
                const size_t e0 = T.telem.size() - (size_t)T.tstride;
                lnodes.clear();
                for (int u = 0; u < cr.w; u++)
                    for (int k = 0; k < I.nen; k++)
                        lnodes.push_back(F0.h_conn[(int64_t)T.telem[e0 + u] * I.nen + k]);
                std::sort(lnodes.begin(), lnodes.end());
                lnodes.erase(std::unique(lnodes.begin(), lnodes.end()), lnodes.end());
                cur.x = (int)T.txy.size();
                for (int32_t nd : lnodes) T.txy.push_back(I.hxy[nd]);
                for (int u = 0; u < T.tstride; u++) {
                    uint16_t l[4] = {0, 0, 0, 0};
                    if (u < cr.w)
                        for (int k = 0; k < I.nen; k++)
                            l[k] = (uint16_t)(std::lower_bound(lnodes.begin(), lnodes.end(),
                                                               F0.h_conn[(int64_t)T.telem[e0 + u] * I.nen + k]) -
                                              lnodes.begin());
                    T.tconn.push_back(make_ushort4(l[0], l[1], l[2], l[3]));
                }
            }
        } else {
            // pad to cap tasks, so that tile t's tasks start at t * cap (the kernel
            // computes a thread's task index without reading the tile first)
            for (int k = cur.y; k < cap; k++) {
                T.elem.push_back(0);
                T.info.push_back(TASK_PAD);
                T.pos.insert(T.pos.end(), ncolp, (uint8_t)0);
            }
        }
        T.tiles.push_back(cur);
        T.trows.push_back(cr);
        T.maxv = std::max(T.maxv, cur.w);
        open = false;
        tile_id++;
        nu_open = 0;
    };
    int32_t colbuf[16];
    std::vector<uint8_t> seen;
    for (int64_t k = r0; k < r1;) {
        const int64_t r = I.seq ? I.seq[k] : k;
        const int32_t node = inv_node[r];
        const int field = inv_fc[r] >> 4, comp = inv_fc[r] & 15;
        int nr = 1;
        if (node >= 0 && field == 0 && nc == 2 && comp == 0 && r + 1 < n && inv_node[r + 1] == node &&
            inv_fc[r + 1] == 1)
            nr = 2;
        const int L = hptr[r + 1] - hptr[r];
        if (nr == 2 && hptr[r + 2] - hptr[r + 1] != L) { err = "rows of one node differ in pattern"; return ELFEL_ENOTSUP; }
        if (I.seq && nr != 1) { err = "internal: patch tiles of a vector problem"; return ELFEL_EINVAL; }
        int deg = 0;
        const Adj *A = nullptr;
        if (node >= 0) {
            A = field == 0 ? I.A0 : I.A1;
            deg = A->ptr[node + 1] - A->ptr[node];
        }
        if (deg > cap || nr * L > vcap || deg > 16 || L > 128 || (staged && deg > ucap)) {
            err = "node too highly connected for the reuse tiles";
            return ELFEL_ENOTSUP;
        }
        int newu = 0;
        if (staged)
            for (int s = 0; s < deg; s++) newu += stamp[A->elem[A->ptr[node] + s]] != tile_id;
        if (open && (cur.y + deg > cap || cur.w + nr * L > vcap || cr.y + nr > std::min(cap, REUSE_MAX_TILE_ROWS) ||
                     nu_open + newu > ucap)) {
            close();
            if (row_err) return ELFEL_EINVAL;
            newu = deg;  // a node's elements are distinct
        }
        if (!open) {
            cur = int4{(int)T.elem.size(), 0, hptr[r], 0};
            cr = int4{(int)r, 0, 0, 0};
            open = true;
        }
        const int vbase = cur.w;
        const int rowcode = nr == 2 ? 2 : (comp == 1 ? 1 : 0);
        const int l8 = nr == 2 ? L : cr.y;
        const int32_t *list = hind.data() + hptr[r];
        seen.assign(L, 0);
        for (int s = 0; s < deg; s++) {
            const int32_t e = A->elem[A->ptr[node] + s];
            const int a = A->loc[A->ptr[node] + s];
            int nj = 0;
            for (int k = 0; k < nen; k++)
                for (int cc = 0; cc < nc; cc++)
                    colbuf[nj++] = F0.h_dof[(int64_t)F0.h_conn[(int64_t)e * nen + k] * nc + cc];
            if (stokes && field == 0)
                for (int k = 0; k < 3; k++) colbuf[nj++] = F1->h_dof[F1->h_conn[(int64_t)e * 3 + k]];
            // a task reads all of its partial sums before it writes any back,
            // so its columns must be distinct dofs (an element that repeats a
            // node or a dof goes to the sort method)
            for (int j = 1; j < nj; j++)
                for (int i = 0; i < j; i++)
                    if (colbuf[i] == colbuf[j]) {
                        err = "an element with a repeated dof";
                        return ELFEL_ENOTSUP;
                    }
            for (int j = 0; j < ncolp; j++) {
                uint8_t pv = 0;
                if (j < nj) {
                    const int32_t *it = std::lower_bound(list, list + L, colbuf[j]);
                    if (it == list + L || *it != colbuf[j]) {
                        err = "internal: element column missing from the pattern";
                        return ELFEL_EINVAL;
                    }
                    pv = (uint8_t)(it - list);
                    // bit 7: this task is the first (lowest rank) to add into the entry
                    if (!seen[pv]) { seen[pv] = 1; pv |= 0x80; }
                }
                T.pos.push_back(pv);
            }
            if (staged) stamp[e] = tile_id;
            T.elem.push_back(e);
            T.info.push_back((uint32_t)a | ((uint32_t)s << 3) | ((uint32_t)rowcode << 7) |
                             ((uint32_t)(field == 1) << 9) | ((uint32_t)vbase << 10) | ((uint32_t)l8 << 24));
            T.maxrank = std::max(T.maxrank, s + 1);
            cr.z = std::max(cr.z, s + 1);
        }
        if (I.seq) orows.push_back({(int32_t)r, L});
        nu_open += newu;
        cur.y += deg;
        cur.w += nr * L;
        cr.y += nr;
        k += nr;
    }
    close();
    if (row_err) return ELFEL_EINVAL;
    return 0;
}

// Locality order of the nodes of a 2D mesh: bands about eight node spacings
// high, bottom to top, each traversed in x (then y); ties keep node order.
// With row-owner tiles (at most 256 rows and 256 elements each) on heat Q4
// 4096^2, band heights 3 / 4 / 6 / 8 / 10 / 12 / 16 measured 0.599 / 0.573 /
// 0.563 / 0.545 / 0.558 / 0.559 / 0.582 ms: an 8 x 27 node tile holds 216
// rows in 252 elements (1.17 elements per row, against 1.28 for 4 x 48).
// Packed greedily in this order, a tile's rows cover a few short stretches of
// neighbouring grid lines instead of one stretch of one line: each element of
// a heat Q4 4096^2 mesh is integrated by 1.3 tiles instead of 2, while a tile
// still writes its values in a few long runs (two tasks per thread: band
// height in node rows 2 / 4 / 8 measured 0.78 / 0.74 / 0.82 ms, strips 0.86 ms,
// Z-order patches with 16 runs per tile 0.85 ms, the coordinates stored in this
// order as well no change; three tasks per thread: 3 / 4 / 5 / 6 rows 0.64 /
// 0.62 / 0.62 / 0.64 ms).  Small meshes keep strips: on heat T3 100^2 (10,201 rows, a
// launch-latency-bound case) bands measured 6.1 against 4.8 us.
void band_order(const double2 *xy, size_t nn, std::vector<int32_t> &order)
{
    double x0 = 0, x1 = 0, y0 = 0, y1 = 0;
    for (size_t i = 0; i < nn; i++) {
        if (i == 0 || xy[i].x < x0) x0 = xy[i].x;
        if (i == 0 || xy[i].x > x1) x1 = xy[i].x;
        if (i == 0 || xy[i].y < y0) y0 = xy[i].y;
        if (i == 0 || xy[i].y > y1) y1 = xy[i].y;
    }
    const double sx = x1 > x0 ? 65535.0 / (x1 - x0) : 0.0, sy = y1 > y0 ? 65535.0 / (y1 - y0) : 0.0;
    std::vector<uint64_t> key(nn), k2(nn);
    std::vector<int32_t> o2(nn);
    order.resize(nn);
    // key = (band, x, y), x and y quantised to 16 bits; the node spacing is
    // estimated from the bounding box and the node count
    const double spacing = std::sqrt(std::max((x1 - x0) * (y1 - y0), 1e-300) / std::max<double>((double)nn, 1.0));
    const double bh = spacing * std::max(1, env_int("ELFEL_HIP_BAND_ROWS", 8));
    for (size_t i = 0; i < nn; i++) {
        order[i] = (int32_t)i;
        const uint32_t qx = (uint32_t)((xy[i].x - x0) * sx), qy = (uint32_t)((xy[i].y - y0) * sy);
        const uint64_t band = (uint64_t)((xy[i].y - y0 + 0.5 * spacing) / bh);
        key[i] = (band << 32) | ((uint64_t)qx << 16) | qy;
    }
    // stable LSD radix sort by key, 8 bits a pass
    for (int shift = 0; shift < 64; shift += 8) {
        size_t cnt[257] = {0};
        for (size_t i = 0; i < nn; i++) cnt[((key[i] >> shift) & 255) + 1]++;
        if (cnt[1] == nn) continue;  // this digit is 0 everywhere
        for (int b = 0; b < 256; b++) cnt[b + 1] += cnt[b];
        for (size_t i = 0; i < nn; i++) {
            const size_t d = cnt[(key[i] >> shift) & 255]++;
            k2[d] = key[i];
            o2[d] = order[i];
        }
        key.swap(k2);
        order.swap(o2);
    }
}

}  // namespace

// Row groups in dof order packed into tiles of <= cap tasks and <= 24 cap
// values.  The rows are cut into ranges of about 2^18 rows at row-group
// boundaries -- a fixed cut, whatever the number of host threads, so the
// tiles are the same on every machine -- and the ranges are packed in
// parallel; a tile never spans two ranges.
//
// With errp set the builder touches neither the device nor the context (the
// batch runs one per problem on host threads): the coordinates come from
// host_xy and an error goes to *errp.
int tile_tables(elfel_hip_ctx *c, bool stokes, const Field &F0, const Field *F1, int64_t n,
                const std::vector<int32_t> &hptr, const std::vector<int32_t> &hind, int cap, int ucap,
                ReuseTables &T, bool patch, bool rows, const double2 *host_xy, std::string *errp)
{
    auto fail = [&](int code, const std::string &m) -> int {
        if (errp) { *errp = m; return code; }
        return set_err(c, code, m);
    };
    const int nen = F0.nen, nc = F0.ncomp, nd = nen * nc;
    const int ncol = nd + (stokes ? 3 : 0);
    TileInput I;
    I.stokes = stokes;
    // scalar problems: the kernel computes each distinct element of a tile once
    // (element-staged tiles, reuse_kernel.h); for elasticity the staging was
    // measured slower (C2 83 against 68 us) and tasks compute their own geometry.
    // Staged tiles hold at most ucap distinct elements, one per thread of the
    // workgroup: a second element on one thread would add a whole dependent
    // load chain (element -> nodes -> coordinates) to the tile's critical path.
    I.staged = !stokes && nc == 1;
    I.by_rank = rank_order_enabled();
    I.nen = nen; I.nc = nc; I.ncol = ncol;
    I.ncolp = (ncol + 3) & ~3;  // one byte per column: position (bits 0-6), first contribution (bit 7)
    I.cap = cap; I.ucap = ucap;
    // LDS doubles for a tile's values per task slot.  Stokes tiles take 20:
    // the largest tile sets every workgroup's LDS, and at 24 (45 KiB) only 3
    // Stokes workgroups fit on a CU against 4 at 20 (C3 157 against 137 us);
    // the elasticity kernel is held at 4 by its registers and keeps 24 (at 16:
    // 66 against 64 us).
    I.vcap = std::min(std::max(env_int("ELFEL_HIP_VCAP", stokes ? 20 : REUSE_VCAP_PER_THREAD), 8), 24) * cap;
    I.n = n;
    I.F0 = &F0; I.F1 = F1;
    Adj A0, A1;
    adjacency(F0, A0);
    if (stokes) adjacency(*F1, A1);
    I.A0 = &A0; I.A1 = &A1;
    // Each row must belong to one (node, component): a dof shared by two
    // nodes (or two components) gathers contributions from elements of both,
    // which one row group's element list does not hold.  Such dof maps are
    // valid input (sparse() sums whatever the COO stream holds) and go to the
    // sort method instead.
    std::vector<int32_t> inv_node(n, -1);
    std::vector<uint8_t> inv_fc(n, 0);  // field << 4 | comp
    for (int64_t nd0 = 0; nd0 < F0.Nn; nd0++)
        for (int cc = 0; cc < nc; cc++) {
            const int32_t d = F0.h_dof[nd0 * nc + cc];
            if (inv_node[d] >= 0) return fail(ELFEL_ENOTSUP, "a dof shared by two nodes or components");
            inv_node[d] = (int32_t)nd0;
            inv_fc[d] = (uint8_t)cc;
        }
    if (stokes)
        for (int64_t m = 0; m < F1->Nn; m++) {
            const int32_t d = F1->h_dof[m];
            if (inv_node[d] >= 0) return fail(ELFEL_ENOTSUP, "a dof shared by two nodes or components");
            inv_node[d] = (int32_t)m;
            inv_fc[d] = (uint8_t)(1 << 4);
        }
    I.hptr = &hptr; I.hind = &hind; I.inv_node = &inv_node; I.inv_fc = &inv_fc;
    // row-owner heat tiles (T3 and Q4): when every node lies in at most kmax
    // elements and every row has at most 16 entries, each row gets one record
    // (heat_rows_tile) in place of its tasks' tables; the tiles hold up to
    // HEAT_ROWS_BLOCK rows and as many distinct elements
    I.rw = I.kmax = 0;
    I.hxy = nullptr;
    // host coordinates (tile-local geometry, locality order): the caller's, or
    // a copy of the device array made on first use
    std::vector<double2> hcopy;
    const double2 *hxy = errp ? host_xy : nullptr;
    auto fetch_xy = [&]() -> int {
        if (hxy || F0.Nn == 0) return 0;
        if (errp) return fail(ELFEL_EINVAL, "internal: host coordinates missing");
        hcopy.resize((size_t)F0.Nn);
        HIPCHK(c, hipMemcpy(hcopy.data(), F0.xy.p, 16 * (size_t)F0.Nn, hipMemcpyDeviceToHost));
        hxy = hcopy.data();
        return 0;
    };
    if (rows && I.staged && (nen == 3 || nen == 4)) {
        const int kmax = nen == 4 ? hrow_kmax<4>() : hrow_kmax<3>();
        int maxdeg = 0;
        for (int64_t m = 0; m < F0.Nn; m++) maxdeg = std::max(maxdeg, A0.ptr[m + 1] - A0.ptr[m]);
        int maxl = 0;
        for (int64_t r = 0; r < n; r++) maxl = std::max(maxl, hptr[r + 1] - hptr[r]);
        if (maxdeg <= kmax && maxl <= 16) {
            I.kmax = kmax;
            I.rw = nen == 4 ? hrow_words<4, hrow_kmax<4>()>() : hrow_words<3, hrow_kmax<3>()>();
            I.cap = HEAT_ROWS_BLOCK * kmax;
            I.ucap = HEAT_ROWS_BLOCK;
            I.vcap = REUSE_VCAP_PER_THREAD * HEAT_ROWS_BLOCK;
            // ELFEL_HIP_TILE_GEOM=0: the kernel gathers through conn and xy
            if (F0.sdim == 2 && env_int("ELFEL_HIP_TILE_GEOM", 1) != 0) {
                int rc = fetch_xy();
                if (rc) return rc;
                I.hxy = hxy;
            }
        }
    }
    // patch tiles for the element-staged (scalar) problems of 2^18 rows or
    // more: rows packed in a locality order of their nodes
    std::vector<int32_t> seq;
    I.seq = nullptr;
    if (patch && I.staged && nc == 1 && F0.sdim == 2 && n >= (int64_t(1) << 18)) {
        int rc = fetch_xy();
        if (rc) return rc;
        std::vector<int32_t> node_order;
        band_order(hxy, (size_t)F0.Nn, node_order);
        // the rows in the order of their nodes; rows without a node last
        seq.reserve((size_t)n);
        for (int32_t nd : node_order) {
            const int32_t r = F0.h_dof[nd];
            if (r >= 0 && r < n && inv_node[r] == nd) seq.push_back(r);
        }
        for (int64_t r = 0; r < n; r++)
            if (inv_node[r] < 0) seq.push_back((int32_t)r);
        I.seq = seq.data();
    }
    // range cuts, moved forward off the second row of a two-row group
    const int64_t nranges = std::max<int64_t>(1, std::min<int64_t>(1024, n >> 18));
    std::vector<int64_t> cut(nranges + 1);
    for (int64_t k = 0; k <= nranges; k++) {
        int64_t r = n * k / nranges;
        while (!I.seq && r > 0 && r < n && inv_node[r] >= 0 && inv_node[r] == inv_node[r - 1] && (inv_fc[r] >> 4) == 0 &&
               (inv_fc[r - 1] >> 4) == 0 && nc == 2 && (inv_fc[r] & 15) == 1 && (inv_fc[r - 1] & 15) == 0)
            r++;
        cut[k] = std::max(r, k ? cut[k - 1] : 0);
    }
    std::vector<ReuseTables> part(nranges);
    std::vector<int> prc(nranges, 0);
    std::vector<std::string> perr(nranges);
    const int nthreads = (int)std::max<int64_t>(1, std::min<int64_t>(
        nranges, std::min<int64_t>(16, std::max(1u, std::thread::hardware_concurrency()))));
    std::atomic<int64_t> next{0};
    // an exception on a thread (a failed host allocation) fails the build;
    // the other threads still drain their ranges
    std::mutex exc_mu;
    int exc_rc = 0;
    std::string exc_err;
    auto worker = [&]() {
        try {
            // the last tile each element joined (element-staged tiles)
            std::vector<int32_t> stamp(I.staged ? F0.E : 0, -1);
            int tile_id = 0;
            for (;;) {
                const int64_t k = next.fetch_add(1);
                if (k >= nranges) break;
                prc[k] = tile_range(I, cut[k], cut[k + 1], part[k], stamp, tile_id, perr[k]);
            }
        } catch (...) {
            std::string m;
            const int code = exception_code(&m);
            std::lock_guard<std::mutex> lk(exc_mu);
            if (!exc_rc) { exc_rc = code; exc_err = m; }
        }
    };
    std::vector<std::thread> pool;
    {
        auto joiner = join_on_exit(pool, [] {});
        for (int t = 1; t < nthreads; t++) pool.emplace_back(worker);
        worker();
    }
    if (exc_rc) return fail(exc_rc, exc_err);
    for (int64_t k = 0; k < nranges; k++)
        if (prc[k]) return fail(prc[k], perr[k]);
    // concatenation: tile task offsets become global
    T = ReuseTables{};
    T.ncolp = I.ncolp;
    T.rw = I.rw;
    T.tstride = I.rw ? HEAT_ROWS_BLOCK : I.cap;
    size_t ntiles = 0, nslots = 0, ntel = 0, nruns = 0;
    for (const ReuseTables &p : part) {
        ntiles += p.tiles.size(); nslots += p.elem.size(); ntel += p.telem.size(); nruns += p.runs.size();
    }
    if (I.seq) { T.truns.reserve(ntiles); T.runs.reserve(nruns); }
    T.tiles.reserve(ntiles); T.trows.reserve(ntiles);
    T.elem.reserve(nslots); T.info.reserve(nslots); T.pos.reserve(nslots * I.ncolp); T.telem.reserve(ntel);
    for (const ReuseTables &p : part) {
        // tile .x: the first task, or in row-owner tiles the first tile-local node
        const int off = (int)(T.rw ? T.txy.size() : T.elem.size());
        for (int4 td : p.tiles) { td.x += off; T.tiles.push_back(td); }
        T.tconn.insert(T.tconn.end(), p.tconn.begin(), p.tconn.end());
        T.txy.insert(T.txy.end(), p.txy.begin(), p.txy.end());
        T.trows.insert(T.trows.end(), p.trows.begin(), p.trows.end());
        T.elem.insert(T.elem.end(), p.elem.begin(), p.elem.end());
        T.info.insert(T.info.end(), p.info.begin(), p.info.end());
        T.pos.insert(T.pos.end(), p.pos.begin(), p.pos.end());
        T.telem.insert(T.telem.end(), p.telem.begin(), p.telem.end());
        T.rrec.insert(T.rrec.end(), p.rrec.begin(), p.rrec.end());
        const int roff = (int)T.runs.size();
        for (int2 tr : p.truns) { tr.x += roff; T.truns.push_back(tr); }
        T.runs.insert(T.runs.end(), p.runs.begin(), p.runs.end());
        T.maxv = std::max(T.maxv, p.maxv);
        T.maxrank = std::max(T.maxrank, p.maxrank);
        T.maxu = std::max(T.maxu, p.maxu);
    }
    return 0;
}

// Tile order: tiles are independent (each owns its value range), so their
// launch order is free.  Sorting by the first element a tile touches puts the
// tiles that read the same element records close together; laying the sorted
// list out so that workgroup b (dispatched to XCD b mod 8) walks a contiguous
// chunk keeps those re-reads inside one XCD's L2.  With run > 0 the sorted list
// is cut into runs of that many tiles, dealt to the XCDs in turn: the re-reads
// stay inside runs, and every XCD gets a sample of the whole list, which
// balances the XCDs when tile costs vary along it (batched problems of uneven
// size).
void order_tiles(ReuseTables &T, int nxcd, int cap, int run)
{
    const size_t nt = T.tiles.size();
    if (nt < 2) return;
    std::vector<std::pair<int32_t, int32_t>> key(nt);
    for (size_t t = 0; t < nt; t++) {
        // element-staged tiles hold tile-local element indices: their key is
        // 0 and they keep the row order (measured faster for heat Q4 4096^2
        // than sorting them by their first element: 0.91 against 1.00 ms)
        int32_t m = INT32_MAX;
        if (!T.telem.empty()) m = 0;  // (row-owner tiles keep no tasks at all)
        else
            for (int k = 0; k < T.tiles[t].y; k++) m = std::min(m, T.elem[(size_t)T.tiles[t].x + k]);
        key[t] = {m, (int32_t)t};
    }
    std::stable_sort(key.begin(), key.end());
    // src[b]: the sorted position of the tile that launch slot b runs
    std::vector<size_t> src(nt);
    if (run <= 0) {
        std::vector<int64_t> off(nxcd + 1, 0);
        for (int x = 0; x < nxcd; x++) off[x + 1] = off[x] + (int64_t)(nt / nxcd + ((size_t)x < nt % nxcd ? 1 : 0));
        for (size_t b = 0; b < nt; b++) src[b] = (size_t)(off[b % nxcd] + b / nxcd);
    } else {
        std::vector<std::vector<size_t>> lists(nxcd);
        for (size_t i = 0; i < nt; i++) lists[(i / (size_t)run) % nxcd].push_back(i);
        std::vector<size_t> next(nxcd, 0);
        for (size_t b = 0; b < nt; b++) {
            int x = (int)(b % nxcd);
            if (next[x] == lists[x].size()) {  // this XCD's runs are used up: take from the fullest list
                for (int y = 0; y < nxcd; y++)
                    if (lists[y].size() - next[y] > lists[x].size() - next[x]) x = y;
            }
            src[b] = lists[x][next[x]++];
        }
    }
    std::vector<int4> tiles(nt), trows(nt);
    std::vector<int32_t> elem(T.elem.size());
    std::vector<uint32_t> info(T.info.size());
    std::vector<uint8_t> pos(T.pos.size());
    std::vector<int32_t> telem(T.telem.size());
    std::vector<uint32_t> rrec(T.rrec.size());
    std::vector<ushort4> tconn(T.tconn.size());
    std::vector<int2> truns(T.truns.size());
    const size_t ncolp = (size_t)T.ncolp, ts = (size_t)T.tstride, rs = (size_t)HEAT_ROWS_BLOCK * (size_t)T.rw;
    for (size_t b = 0; b < nt; b++) {
        const size_t o = (size_t)key[src[b]].second;
        tiles[b] = T.tiles[o];
        trows[b] = T.trows[o];
        if (!T.rw) tiles[b].x = (int)(b * cap);  // the tasks move with their tile (row-owner: .x stays)
        if (!T.truns.empty()) truns[b] = T.truns[o];  // the runs stay, their ranges move
        if (!T.rw) {  // row-owner tiles keep no tasks
            std::copy(T.elem.begin() + o * cap, T.elem.begin() + (o + 1) * cap, elem.begin() + b * cap);
            std::copy(T.info.begin() + o * cap, T.info.begin() + (o + 1) * cap, info.begin() + b * cap);
            std::copy(T.pos.begin() + o * cap * ncolp, T.pos.begin() + (o + 1) * cap * ncolp, pos.begin() + b * cap * ncolp);
        }
        if (!T.telem.empty())
            std::copy(T.telem.begin() + o * ts, T.telem.begin() + (o + 1) * ts, telem.begin() + b * ts);
        if (!T.rrec.empty())
            std::copy(T.rrec.begin() + o * rs, T.rrec.begin() + (o + 1) * rs, rrec.begin() + b * rs);
        if (!T.tconn.empty())
            std::copy(T.tconn.begin() + o * ts, T.tconn.begin() + (o + 1) * ts, tconn.begin() + b * ts);
    }
    T.tiles.swap(tiles);
    T.trows.swap(trows);
    T.elem.swap(elem);
    T.info.swap(info);
    T.pos.swap(pos);
    T.telem.swap(telem);
    T.rrec.swap(rrec);
    T.tconn.swap(tconn);
    T.truns.swap(truns);
}

int upload_tables(elfel_hip_ctx *c, const ReuseTables &T, Pattern &P)
{
    P.rw = T.rw;
    P.tstride = T.tstride;
    if (T.rw) {
        // row-owner heat tiles: the row records and the tiles' element lists
        // replace the task tables
        int rc;
        if ((rc = ensure(c, P.tile, sizeof(int4) * std::max<size_t>(T.tiles.size(), 1))) ||
            (rc = ensure(c, P.tile_rows, sizeof(int4) * std::max<size_t>(T.trows.size(), 1))) ||
            (rc = ensure(c, P.rrec, 4 * std::max<size_t>(T.rrec.size(), 1))) ||
            (rc = ensure(c, P.tile_elem, 4 * std::max<size_t>(T.telem.size(), 1))))
            return rc;
        if (!T.tconn.empty()) {
            if ((rc = ensure(c, P.tconn, sizeof(ushort4) * T.tconn.size())) ||
                (rc = ensure(c, P.txy, sizeof(double2) * std::max<size_t>(T.txy.size(), 1))))
                return rc;
            HIPCHK(c, hipMemcpy(P.tconn.p, T.tconn.data(), sizeof(ushort4) * T.tconn.size(), hipMemcpyHostToDevice));
            if (!T.txy.empty())
                HIPCHK(c, hipMemcpy(P.txy.p, T.txy.data(), sizeof(double2) * T.txy.size(), hipMemcpyHostToDevice));
            P.tgeom = true;
        } else {
            release(P.tconn); release(P.txy);
            P.tgeom = false;
        }
        P.has_runs = !T.truns.empty();
        if (P.has_runs) {
            if ((rc = ensure(c, P.truns, sizeof(int2) * T.truns.size())) ||
                (rc = ensure(c, P.runs, sizeof(int4) * std::max<size_t>(T.runs.size(), 1))))
                return rc;
            HIPCHK(c, hipMemcpy(P.truns.p, T.truns.data(), sizeof(int2) * T.truns.size(), hipMemcpyHostToDevice));
            if (!T.runs.empty())
                HIPCHK(c, hipMemcpy(P.runs.p, T.runs.data(), sizeof(int4) * T.runs.size(), hipMemcpyHostToDevice));
        }
        if (!T.tiles.empty()) {
            HIPCHK(c, hipMemcpy(P.tile.p, T.tiles.data(), sizeof(int4) * T.tiles.size(), hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(P.tile_rows.p, T.trows.data(), sizeof(int4) * T.trows.size(), hipMemcpyHostToDevice));
        }
        if (!T.rrec.empty())
            HIPCHK(c, hipMemcpy(P.rrec.p, T.rrec.data(), 4 * T.rrec.size(), hipMemcpyHostToDevice));
        if (!T.telem.empty())
            HIPCHK(c, hipMemcpy(P.tile_elem.p, T.telem.data(), 4 * T.telem.size(), hipMemcpyHostToDevice));
        release(P.task_elem); release(P.task_info); release(P.task_pos);
        P.elem8 = false;
        P.ntiles = (int64_t)T.tiles.size();
        P.ntasks = (int64_t)T.elem.size();
        P.ncolmax = T.ncolp;
        P.maxrank = T.maxrank;
        P.lds_vals = T.maxv;
        P.maxu = T.maxu;
        return 0;
    }
    release(P.rrec); release(P.tconn); release(P.txy);
    P.tgeom = false;
    const int64_t ntasks = (int64_t)T.elem.size();
    int rc;
    if ((rc = ensure(c, P.tile, sizeof(int4) * std::max<size_t>(T.tiles.size(), 1))) ||
        (rc = ensure(c, P.tile_rows, sizeof(int4) * std::max<size_t>(T.trows.size(), 1))) ||
        (rc = ensure(c, P.task_elem, 4 * std::max<int64_t>(ntasks, 1))) ||
        (rc = ensure(c, P.task_info, 4 * std::max<int64_t>(ntasks, 1))) ||
        (rc = ensure(c, P.task_pos, std::max<size_t>(T.pos.size(), 1))) ||
        (rc = ensure(c, P.tile_elem, 4 * std::max<size_t>(T.telem.size(), 1))))
        return rc;
    P.has_runs = !T.truns.empty();
    if (P.has_runs) {
        if ((rc = ensure(c, P.truns, sizeof(int2) * T.truns.size())) ||
            (rc = ensure(c, P.runs, sizeof(int4) * std::max<size_t>(T.runs.size(), 1))))
            return rc;
        HIPCHK(c, hipMemcpy(P.truns.p, T.truns.data(), sizeof(int2) * T.truns.size(), hipMemcpyHostToDevice));
        if (!T.runs.empty())
            HIPCHK(c, hipMemcpy(P.runs.p, T.runs.data(), sizeof(int4) * T.runs.size(), hipMemcpyHostToDevice));
    }
    if (!T.tiles.empty()) {
        HIPCHK(c, hipMemcpy(P.tile.p, T.tiles.data(), sizeof(int4) * T.tiles.size(), hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(P.tile_rows.p, T.trows.data(), sizeof(int4) * T.trows.size(), hipMemcpyHostToDevice));
    }
    // staged tiles name a task's element by its index in the tile's list:
    // one byte per task (the builder keeps at most 256 elements per tile)
    P.elem8 = !T.telem.empty();
    if (ntasks && P.elem8) {
        std::vector<uint8_t> e8((size_t)ntasks);
        for (int64_t k = 0; k < ntasks; k++) {
            if (T.elem[k] < 0 || T.elem[k] > 255) return set_err(c, ELFEL_EINVAL, "internal: tile element index above 255");
            e8[k] = (uint8_t)T.elem[k];
        }
        HIPCHK(c, hipMemcpy(P.task_elem.p, e8.data(), (size_t)ntasks, hipMemcpyHostToDevice));
    } else if (ntasks) {
        HIPCHK(c, hipMemcpy(P.task_elem.p, T.elem.data(), 4 * ntasks, hipMemcpyHostToDevice));
    }
    if (ntasks) {
        HIPCHK(c, hipMemcpy(P.task_info.p, T.info.data(), 4 * ntasks, hipMemcpyHostToDevice));
        HIPCHK(c, hipMemcpy(P.task_pos.p, T.pos.data(), T.pos.size(), hipMemcpyHostToDevice));
    }
    if (!T.telem.empty())
        HIPCHK(c, hipMemcpy(P.tile_elem.p, T.telem.data(), 4 * T.telem.size(), hipMemcpyHostToDevice));
    P.ntiles = (int64_t)T.tiles.size();
    P.ntasks = ntasks;
    P.ncolmax = T.ncolp;
    P.maxrank = T.maxrank;
    P.lds_vals = T.maxv;
    P.maxu = T.maxu;
    return 0;
}

static int build_symbolic(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf,
                          int block)
{
    auto t0 = std::chrono::steady_clock::now();
    Pattern &P = c->pat;
    P.valid = false;
    const bool stokes = pb->kind == ELFEL_STOKES_GEN || pb->kind == ELFEL_STOKES_VECLAP;
    const Field &F0 = c->f[0];
    const Field *F1 = stokes ? &c->f[1] : nullptr;
    const int64_t n = c->nrow;
    // ELFEL_HIP_SYMBOLIC_TIMES=1: the phase's stages on stderr
    static const bool say = env_int("ELFEL_HIP_SYMBOLIC_TIMES", 0) != 0;
    auto lap = [&](const char *what) {
        if (!say) return;
        static thread_local std::chrono::steady_clock::time_point last;
        const auto now = std::chrono::steady_clock::now();
        if (what) fprintf(stderr, "[symbolic] %-28s %8.1f ms\n", what, std::chrono::duration<double, std::milli>(now - last).count());
        last = now;
    };
    lap(nullptr);
    // vector problems (elasticity, Stokes): the whole phase on the device
    // (ELFEL_HIP_DEVICE_SYMBOLIC=0: the host builder below)
    if ((pb->kind == ELFEL_ELAST || stokes) && F0.ncomp == 2 && env_int("ELFEL_HIP_DEVICE_SYMBOLIC", 1) != 0) {
        int rc = device_symbolic(c, pb, block);
        if (rc) return rc;
        lap("device symbolic");
        if (!P.fallback) {
            if ((rc = ensure(c, P.conn_pad, 4 * (size_t)F0.nen * padded_len(F0.E, CONN_GAP))) ||
                (rc = ensure(c, P.xy_pad, 16 * (size_t)padded_len(F0.Nn, XY_GAP))))
                return rc;
            if ((rc = copy_padded(c, P.conn_pad.p, F0.conn.p, F0.E, 4 * (size_t)F0.nen, CONN_GAP)) ||
                (rc = copy_padded(c, P.xy_pad.p, F0.xy.p, F0.Nn, 16, XY_GAP)))
                return rc;
        }
        P.block = block * reuse_tpt(pb->kind);
        P.kind = pb->kind; P.fe = pb->fe; P.layout = pb->layout;
        P.key_mesh0 = F0.mesh_version; P.key_dof0 = F0.dof_version;
        P.key_mesh1 = stokes ? F1->mesh_version : 0; P.key_dof1 = stokes ? F1->dof_version : 0;
        P.nrow = P.ncol = n;
        P.valid = true;
        P.symbolic_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
        return 0;
    }
    // scalar heat problems: row-owner tiles built on the device when the mesh
    // fits the row records (else the host builder below makes task tiles)
    if (pb->kind == ELFEL_HEAT && F0.ncomp == 1 && env_int("ELFEL_HIP_HEAT_ROWS", 1) != 0 &&
        env_int("ELFEL_HIP_DEVICE_SYMBOLIC", 1) != 0 && pick_heat_rows32(pb->fe, tab.nq, pb->layout == ELFEL_CSC)) {
        int rc = device_symbolic_heat(c, pb, env_int("ELFEL_HIP_PATCH_TILES", 1) != 0);
        if (rc < 0) return rc;
        if (rc == 0) {
            lap("device symbolic (heat)");
            P.block = block * reuse_tpt(pb->kind);
            P.kind = pb->kind; P.fe = pb->fe; P.layout = pb->layout;
            P.key_mesh0 = F0.mesh_version; P.key_dof0 = F0.dof_version;
            P.key_mesh1 = P.key_dof1 = 0;
            P.nrow = P.ncol = n;
            P.valid = true;
            P.symbolic_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
            return 0;
        }
    }
    std::vector<int32_t> hptr, hind;
    int rc = pattern_to_host(c, pb, tab, cf, hptr, hind);
    if (rc) return rc;
    lap("pattern (device) + copy out");
    const int64_t nnz = (int64_t)hind.size();
    if ((rc = ensure(c, P.ptr, 4 * (n + 1))) || (rc = ensure(c, P.ind, 4 * std::max<int64_t>(nnz, 1)))) return rc;
    HIPCHK(c, hipMemcpyAsync(P.ptr.p, c->ptr.p, 4 * (n + 1), hipMemcpyDeviceToDevice, c->stream));
    if (nnz) HIPCHK(c, hipMemcpyAsync(P.ind.p, c->ind.p, 4 * nnz, hipMemcpyDeviceToDevice, c->stream));
    ReuseTables T;
    const int cap = block * reuse_tpt(pb->kind);  // task slots per tile
    P.fallback = false;
    rc = tile_tables(c, stokes, F0, F1, n, hptr, hind, cap, block, T, env_int("ELFEL_HIP_PATCH_TILES", 1) != 0,
                     pb->kind == ELFEL_HEAT && env_int("ELFEL_HIP_HEAT_ROWS", 1) != 0);
    if (rc == ELFEL_ENOTSUP) {
        // the mesh does not fit the tiles: this pattern's assemblies sum the
        // runs of the plain method's stream in its cached sort order (FP64)
        P.fallback = true;
        P.fallback_note = "FP64 runs of the cached sort order: " + c->err;
        c->err.clear();
    } else if (rc) {
        return rc;
    }
    lap("tiles and tables (host)");
    if (!P.fallback) {
        if (env_int("ELFEL_HIP_TILE_ORDER", 1)) order_tiles(T, 8, cap, env_int("ELFEL_HIP_XCD_RUN", 0));
        lap("tile order (host)");
        if ((rc = upload_tables(c, T, P))) return rc;
        lap("upload");
        // the numeric kernel gathers from padded copies of the mesh arrays
        if ((rc = ensure(c, P.conn_pad, 4 * (size_t)F0.nen * padded_len(F0.E, CONN_GAP))) ||
            (rc = ensure(c, P.xy_pad, 16 * (size_t)padded_len(F0.Nn, XY_GAP))))
            return rc;
        if ((rc = copy_padded(c, P.conn_pad.p, F0.conn.p, F0.E, 4 * (size_t)F0.nen, CONN_GAP)) ||
            (rc = copy_padded(c, P.xy_pad.p, F0.xy.p, F0.Nn, 16, XY_GAP)))
            return rc;
    }
    P.block = cap;
    P.kind = pb->kind; P.fe = pb->fe; P.layout = pb->layout;
    P.key_mesh0 = F0.mesh_version; P.key_dof0 = F0.dof_version;
    P.key_mesh1 = stokes ? F1->mesh_version : 0; P.key_dof1 = stokes ? F1->dof_version : 0;
    P.nrow = P.ncol = n;
    P.nnz = nnz;
    P.valid = true;
    P.symbolic_ms = std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
    return 0;
}

int reuse_assemble(elfel_hip_ctx *c, const elfel_hip_problem *pb, const QTable &tab, const Coef &cf)
{
    const bool stokes = pb->kind == ELFEL_STOKES_GEN || pb->kind == ELFEL_STOKES_VECLAP;
    const Field &F0 = c->f[0];
    Pattern &P = c->pat;
    const bool fresh = P.valid && P.kind == pb->kind && P.fe == pb->fe && P.layout == pb->layout &&
                       P.key_mesh0 == F0.mesh_version && P.key_dof0 == F0.dof_version &&
                       (!stokes || (P.key_mesh1 == c->f[1].mesh_version && P.key_dof1 == c->f[1].dof_version)) &&
                       P.nrow == c->nrow;
    const int block = env_int("ELFEL_HIP_REUSE_BLOCK", 256) == 128 ? 128 : 256;
    const bool mixed = pb->precision == ELFEL_MIXED;
    const int csc = pb->layout == ELFEL_CSC;
    if (!(mixed ? pick_reuse32(pb->kind, pb->fe, tab.nq, csc, block) : pick_reuse64(pb->kind, pb->fe, tab.nq, csc, block)))
        return set_err(c, ELFEL_ENOTSUP, "unsupported (kind, element, rule) for the reuse method");
    int rc;
    if (!fresh || P.block != block * reuse_tpt(pb->kind)) {
        hipEvent_t ev = nullptr;
        tm_begin(c, ELFEL_PH_SYMBOLIC, &ev);
        rc = build_symbolic(c, pb, tab, cf, block);
        tm_end(c, ELFEL_PH_SYMBOLIC, ev);
        if (rc) return rc;
    }
    if (P.fallback) {
        // the mesh does not fit the tiles: its stream is summed run by run in
        // the cached sort order (FP64, the plain method's bits)
        c->used_method = ELFEL_REUSE;
        c->used_precision = ELFEL_FP64;
        c->used_note = P.fallback_note;
        return plain_stream_reuse(c, pb, tab, cf);
    }
    reuse_launch_fn fn = P.rw ? (mixed ? pick_heat_rows32(pb->fe, tab.nq, csc) : pick_heat_rows64(pb->fe, tab.nq, csc))
                              : (mixed ? pick_reuse32(pb->kind, pb->fe, tab.nq, csc, block)
                                       : pick_reuse64(pb->kind, pb->fe, tab.nq, csc, block));
    if (!fn) return set_err(c, ELFEL_ENOTSUP, "unsupported (kind, element, rule) for the reuse method");
    const int64_t E = F0.E;
    const int nen = F0.nen;
    if ((rc = ensure(c, c->geom, (size_t)E * nen * 2 * 8)) ||
        (rc = ensure(c, c->val, 8 * std::max<int64_t>(P.nnz, 1))))
        return rc;
    const bool heat = pb->kind == ELFEL_HEAT;
    if (heat && (rc = ensure(c, c->F, 8 * std::max<int64_t>(c->nrow, 1)))) return rc;
    hipEvent_t ev = nullptr;
    const bool fused = env_int("ELFEL_HIP_FUSED_GATHER", 1) != 0;
    if (!fused) {
        tm_begin(c, ELFEL_PH_GATHER, &ev);
        if (mixed)
            rc = gather_geom_rel(c, nen, E, (const int32_t *)F0.conn.p, (const double2 *)F0.xy.p, (double *)c->geom.p);
        else
            rc = gather_geom64(c, nen, E, (const int32_t *)F0.conn.p, (const double2 *)F0.xy.p, (double *)c->geom.p);
        if (rc) return rc;
        tm_end(c, ELFEL_PH_GATHER, ev);
    }
    ReuseArgs A;
    A.tiles = (const int4 *)P.tile.p;
    A.tile_rows = (const int4 *)P.tile_rows.p;
    A.task_elem = (const int32_t *)P.task_elem.p;
    A.task_elem8 = P.elem8 ? (const uint8_t *)P.task_elem.p : nullptr;
    A.tile_elem = (const int32_t *)P.tile_elem.p;
    A.maxu = P.maxu;
    A.task_info = (const uint32_t *)P.task_info.p;
    A.task_pos = (const uint8_t *)P.task_pos.p;
    A.ncolp = P.ncolmax;
    A.ntiles = (int)P.ntiles;
    A.geom = fused ? nullptr : c->geom.p;
    A.conn = (const int32_t *)P.conn_pad.p;
    A.xy = (const double2 *)P.xy_pad.p;
    A.ngap = CONN_GAP;
    A.xgap = XY_GAP;
    A.vals = (double *)c->val.p;
    A.F = heat ? (double *)c->F.p : nullptr;
    if (P.has_runs) {
        A.tile_runs = (const int2 *)P.truns.p;
        A.runs = (const int4 *)P.runs.p;
    }
    A.rrec = P.rw ? (const uint32_t *)P.rrec.p : nullptr;
    A.tconn = P.rw && P.tgeom ? (const ushort4 *)P.tconn.p : nullptr;
    A.txy = P.rw && P.tgeom ? (const double2 *)P.txy.p : nullptr;
    A.tstride = P.tstride;
    tm_begin(c, ELFEL_PH_NUMERIC, &ev);
    rc = fn(c, A, cf, tab, P.lds_vals);
    if (rc) return rc;
    tm_end(c, ELFEL_PH_NUMERIC, ev);
    c->nnz = P.nnz;
    c->nmaj = c->nmin = c->nrow;
    c->layout = pb->layout;
    c->result_src = RES_PATTERN;
    c->has_F = heat;
    c->nF = heat ? c->nrow : 0;
    return 0;
}

// ---------------------------------------------------------------------------
// Geometry statistics: smallest det J over all elements and quadrature points
// and the number of badly shaped elements (det J <= 0, or
// sum_k |grad N_k|^2 det J >= REUSE_COND_MAX at some point), all in FP64.
// ---------------------------------------------------------------------------
template <int NEN, int NQ>
__global__ __launch_bounds__(256) void k_geom_stats(int64_t E, const int32_t *__restrict__ conn,
                                                    const double2 *__restrict__ xy, QTable tab,
                                                    double *__restrict__ block_min,
                                                    unsigned long long *__restrict__ nbad,
                                                    unsigned long long *__restrict__ nneg)
{
    __shared__ double smin[256];
    const int64_t e = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    double mn = 1e300;
    if (e < E) {
        double X[NEN], Y[NEN];
#pragma unroll
        for (int k = 0; k < NEN; k++) {
            const double2 p = xy[conn[e * NEN + k]];
            X[k] = p.x; Y[k] = p.y;
        }
        Geo<double, NEN, NQ> g;
        geometry<double, NEN, NQ>(tab, X, Y, g);
        bool good = true, pos = true;
#pragma unroll
        for (int q = 0; q < NQ; q++) {
            mn = fmin(mn, g.det[q]);
            double j2 = 0.0;
#pragma unroll
            for (int k = 0; k < NEN; k++) j2 += g.gx[q][k] * g.gx[q][k] + g.gy[q][k] * g.gy[q][k];
            pos = pos && (g.det[q] > 0.0);
            good = good && (g.det[q] > 0.0) && (j2 * g.det[q] < REUSE_COND_MAX);
        }
        if (!good) atomicAdd(nbad, 1ull);
        if (!pos) atomicAdd(nneg, 1ull);
    }
    smin[threadIdx.x] = mn;
    __syncthreads();
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) smin[threadIdx.x] = fmin(smin[threadIdx.x], smin[threadIdx.x + s]);
        __syncthreads();
    }
    if (threadIdx.x == 0) block_min[blockIdx.x] = smin[0];
}

// Geometry statistics of field 0 at the points of rule `rule` (0: the
// element's common rule: T3 1 point, T6 3 points, Q4 2x2): smallest det J,
// elements over the shape-quality threshold or with det J <= 0 somewhere
// (nbad), and elements with det J <= 0 at some point (nneg).
int geometry_check(elfel_hip_ctx *c, int rule, double *min_det, int64_t *nbad, int64_t *nneg)
{
    const Field &F = c->f[0];
    if (!F.has_mesh) return set_err(c, ELFEL_ESTATE, "no mesh");
    if (F.sdim != 2 || (F.nen != 3 && F.nen != 4 && F.nen != 6))
        return set_err(c, ELFEL_ENOTSUP, "geometry statistics of T3, Q4 and T6 meshes only");
    const int fe = F.nen == 3 ? ELFEL_T3 : (F.nen == 6 ? ELFEL_T6 : ELFEL_Q4);
    if (rule <= 0) rule = fe == ELFEL_T3 ? 1 : (fe == ELFEL_T6 ? 3 : 2);
    QTable tab;
    if (make_qtable(fe, rule, &tab)) return set_err(c, ELFEL_EINVAL, "no quadrature table");
    const int64_t E = F.E;
    const int nb = ceil_div(std::max<int64_t>(E, 1), 256);
    int rc = ensure(c, c->scal, 8 * (size_t)(nb + 1));
    if (rc) return rc;
    double *bm = (double *)c->scal.p;
    unsigned long long *cnt = (unsigned long long *)((int64_t *)c->dscal.p + 3);  // [0] bad shape, [1] det <= 0
    HIPCHK(c, hipMemsetAsync(cnt, 0, 16, c->stream));
#define GS(N, Q) hipLaunchKernelGGL((k_geom_stats<N, Q>), dim3(nb), dim3(256), 0, c->stream, E, \
                                    (const int32_t *)F.conn.p, (const double2 *)F.xy.p, tab, bm, cnt, cnt + 1)
    if (fe == ELFEL_T3 && tab.nq == 1) GS(3, 1);
    else if (fe == ELFEL_T3 && tab.nq == 3) GS(3, 3);
    else if (fe == ELFEL_T6 && tab.nq == 3) GS(6, 3);
    else if (fe == ELFEL_Q4 && tab.nq == 1) GS(4, 1);
    else if (fe == ELFEL_Q4 && tab.nq == 4) GS(4, 4);
    else return set_err(c, ELFEL_ENOTSUP, "geometry statistics: unsupported rule");
#undef GS
    HIPCHK(c, hipGetLastError());
    std::vector<double> h(nb);
    unsigned long long hc[2] = {0, 0};
    HIPCHK(c, hipMemcpyAsync(h.data(), bm, 8 * nb, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipMemcpyAsync(hc, cnt, 16, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    double m = 1e300;
    for (double v : h) m = std::min(m, v);
    if (min_det) *min_det = m;
    if (nbad) *nbad = (int64_t)hc[0];
    if (nneg) *nneg = (int64_t)hc[1];
    return 0;
}

int geometry_stats_run(elfel_hip_ctx *c, double *min_det, int64_t *nbad)
{
    if (!c->has_result && c->pat.fe < 0) return set_err(c, ELFEL_ESTATE, "assemble first");
    return geometry_check(c, 0, min_det, nbad, nullptr);
}
