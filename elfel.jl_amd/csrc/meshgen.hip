// meshgen.hip -- structured block meshes, the T6 -> T3 corner mesh and the
// boundary-node selection, generated on the device straight into a field of
// the context (the MeshSteward generators the reference's drivers call:
// T3block test/test_heat.jl:29, Q4block :144, T6block and T6toT3
// test/test_stokes.jl:32,37; connectedv(boundary(mesh)) test/test_heat.jl:31).
//
// Orderings are those of the generators:
// * nodes x fastest: node j (nA + 1) + i at (A i / nA, B j / nB);
// * cells (i, j) with i outer, j inner; a T3 cell gives the triangles
//   (f, f+1, f+W+1) and (f, f+W+1, f+W), f = j W + i, W = nA + 1; a Q4 cell
//   gives (f, f+1, f+W+1, f+W);
// * T6: the T3 block plus one node per edge, numbered from Nv in order of
//   first appearance over (element, local edge (v0 v1), (v1 v2), (v2 v0))
//   and placed at the mean of the edge's end points.  On the block the first
//   appearances have a closed form: cell (i, j) creates, in this order, the
//   bottom edge if j = 0, the right edge, the diagonal, the top edge, and the
//   left edge if i = 0, so the number of nodes created before it is
//   3 c + (i + [j > 0]) + (i = 0 ? j : nB), c = i nB + j; every cell computes
//   its nodes without looking at any other;
// * T6 -> T3: the corner nodes renumbered compactly in increasing T6 order.
#include <stdio.h>
#include <stdlib.h>

#include "common.h"

namespace {

__device__ __forceinline__ double lin_coord(double L, int64_t n, int64_t i) { return L * (double)i / (double)n; }

__global__ void k_block_nodes(int64_t nA, int64_t nB, double A, double B, double2 *__restrict__ xy)
{
    const int64_t f = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    const int64_t W = nA + 1;
    if (f >= W * (nB + 1)) return;
    const int64_t i = f % W, j = f / W;
    xy[f] = make_double2(lin_coord(A, nA, i), lin_coord(B, nB, j));
}

// nodes created before cell (i, j) (T6 block), see the file comment
__device__ __forceinline__ int64_t t6_created_before(int64_t i, int64_t j, int64_t nB)
{
    return 3 * (i * nB + j) + (i + (j > 0 ? 1 : 0)) + (i == 0 ? j : nB);
}

template <int FE>
__global__ void k_block_cells(int64_t nA, int64_t nB, double A, double B, int32_t *__restrict__ conn,
                              double2 *__restrict__ xy)
{
    const int64_t c = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= nA * nB) return;
    const int64_t i = c / nB, j = c % nB;
    const int64_t W = nA + 1;
    const int32_t f = (int32_t)(j * W + i);
    if constexpr (FE == ELFEL_Q4) {
        int32_t *q = conn + 4 * c;
        q[0] = f; q[1] = f + 1; q[2] = f + (int32_t)W + 1; q[3] = f + (int32_t)W;
    } else if constexpr (FE == ELFEL_T3) {
        int32_t *t = conn + 6 * c;
        t[0] = f; t[1] = f + 1; t[2] = f + (int32_t)W + 1;
        t[3] = f; t[4] = f + (int32_t)W + 1; t[5] = f + (int32_t)W;
    } else {  // T6
        const int64_t Nv = W * (nB + 1);
        int64_t idx = Nv + t6_created_before(i, j, nB);
        // bottom edge H(j, i): created here when j = 0, else as the top edge of cell (i, j-1)
        const int64_t hb = j == 0 ? idx++ : Nv + t6_created_before(i, j - 1, nB) + (j - 1 == 0 ? 1 : 0) + 2;
        const int64_t vr = idx++;  // right edge V(i+1, j)
        const int64_t dg = idx++;  // diagonal
        const int64_t ht = idx++;  // top edge H(j+1, i)
        // left edge V(i, j): created here when i = 0, else as the right edge of cell (i-1, j)
        const int64_t vl = i == 0 ? idx++ : Nv + t6_created_before(i - 1, j, nB) + (j == 0 ? 1 : 0);
        const int32_t v0 = f, v1 = f + 1, v2 = f + (int32_t)W + 1, v3 = f + (int32_t)W;
        int32_t *t = conn + 12 * c;
        t[0] = v0; t[1] = v1; t[2] = v2; t[3] = (int32_t)hb; t[4] = (int32_t)vr; t[5] = (int32_t)dg;
        t[6] = v0; t[7] = v2; t[8] = v3; t[9] = (int32_t)dg; t[10] = (int32_t)ht; t[11] = (int32_t)vl;
        // coordinates of the nodes this cell creates: the mean of the end points
        const double x0 = lin_coord(A, nA, i), x1 = lin_coord(A, nA, i + 1);
        const double y0 = lin_coord(B, nB, j), y1 = lin_coord(B, nB, j + 1);
        if (j == 0) xy[hb] = make_double2((x0 + x1) / 2, (y0 + y0) / 2);
        xy[vr] = make_double2((x1 + x1) / 2, (y0 + y1) / 2);
        xy[dg] = make_double2((x1 + x0) / 2, (y1 + y0) / 2);
        xy[ht] = make_double2((x1 + x0) / 2, (y1 + y1) / 2);
        if (i == 0) xy[vl] = make_double2((x0 + x0) / 2, (y1 + y0) / 2);
    }
}

__global__ void k_mark_corners(int64_t E, const int32_t *__restrict__ conn6, uint32_t *__restrict__ flag)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= E * 3) return;
    flag[conn6[(t / 3) * 6 + t % 3]] = 1u;  // every writer stores the same value
}

// after the exclusive scan, num[n] is the new index of corner n
__global__ void k_corner_xy(int64_t Nn6, const uint32_t *__restrict__ num, const uint32_t *__restrict__ is_corner,
                            const double2 *__restrict__ xy6, double2 *__restrict__ xy3)
{
    const int64_t n = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (n >= Nn6) return;
    if (is_corner[n]) xy3[num[n]] = xy6[n];
}

__global__ void k_corner_conn(int64_t E, const int32_t *__restrict__ conn6, const uint32_t *__restrict__ num,
                              int32_t *__restrict__ conn3)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= E * 3) return;
    conn3[t] = (int32_t)num[conn6[(t / 3) * 6 + t % 3]];
}

__global__ void k_copy_u32_n(const uint32_t *__restrict__ a, uint32_t *__restrict__ b, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) b[t] = a[t];
}

__global__ void k_zero_u32(uint32_t *__restrict__ p, int64_t n)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = 0u;
}

// one key per element edge between consecutive corners: lo * Nn + hi; the
// value names the edge (element * ncorner + local edge)
__global__ void k_edge_keys(int64_t E, int nen, int ncorner, int64_t Nn, const int32_t *__restrict__ conn,
                            uint64_t *__restrict__ keys, uint64_t *__restrict__ vals)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= E * ncorner) return;
    const int64_t e = t / ncorner;
    const int k = (int)(t % ncorner);
    const int64_t a = conn[e * nen + k], b = conn[e * nen + (k + 1) % ncorner];
    const int64_t lo = a < b ? a : b, hi = a < b ? b : a;
    keys[t] = (uint64_t)(lo * Nn + hi);
    vals[t] = (uint64_t)t;
}

// an edge whose key occurs once belongs to one element: its end points (and
// its T6 midside node, local node ncorner + k) are boundary nodes
__global__ void k_mark_boundary(int64_t n, int nen, int ncorner, const int32_t *__restrict__ conn,
                                const uint64_t *__restrict__ keys, const uint64_t *__restrict__ vals,
                                uint8_t *__restrict__ flag)
{
    const int64_t t = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= n) return;
    const uint64_t k = keys[t];
    if ((t > 0 && keys[t - 1] == k) || (t + 1 < n && keys[t + 1] == k)) return;
    const int64_t id = (int64_t)vals[t];
    const int64_t e = id / ncorner;
    const int le = (int)(id % ncorner);
    flag[conn[e * nen + le]] = 1;
    flag[conn[e * nen + (le + 1) % ncorner]] = 1;
    if (nen == 6) flag[conn[e * nen + 3 + le]] = 1;
}

int finish_field_mesh(elfel_hip_ctx *c, Field &f, int nen, int64_t E, int64_t Nn)
{
    field_mesh_changed(f, E, Nn);
    f.nen = nen;
    f.sdim = 2;
    f.E = E;
    f.Nn = Nn;
    f.h_conn.resize((size_t)(E * nen));
    if (E > 0)
        HIPCHK(c, hipMemcpyAsync(f.h_conn.data(), f.conn.p, 4 * (size_t)(E * nen), hipMemcpyDeviceToHost, c->stream));
    HIPCHK(c, hipStreamSynchronize(c->stream));
    f.has_mesh = true;
    f.mesh_version++;
    c->pat.valid = false;
    return ELFEL_OK;
}

}  // namespace

extern "C" int elfel_hip_block_mesh(elfel_hip_ctx *c, int field, int fe, double A, double B, int64_t nA, int64_t nB,
                                    int64_t *E_out, int64_t *Nn_out)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        if (field < 0 || field > 1 || nA < 1 || nB < 1 || (fe != ELFEL_T3 && fe != ELFEL_Q4 && fe != ELFEL_T6))
            return set_err(c, ELFEL_EINVAL, "elfel_hip_block_mesh: bad arguments");
        const int nen = fe == ELFEL_T3 ? 3 : fe == ELFEL_Q4 ? 4 : 6;
        const int64_t cells = nA * nB;
        const int64_t E = fe == ELFEL_Q4 ? cells : 2 * cells;
        const int64_t Nv = (nA + 1) * (nB + 1);
        const int64_t Nn = fe == ELFEL_T6 ? Nv + nA * (nB + 1) + (nA + 1) * nB + cells : Nv;
        if (E * nen >= INT32_MAX || Nn >= INT32_MAX) return set_err(c, ELFEL_ENOTSUP, "mesh beyond int32 indexing");
        Field &f = c->f[field];
        f.has_mesh = false;
        int rc;
        if ((rc = ensure(c, f.conn, 4 * (size_t)(E * nen))) || (rc = ensure(c, f.xy, 16 * (size_t)Nn))) return rc;
        hipLaunchKernelGGL(k_block_nodes, dim3(ceil_div(Nv, 256)), dim3(256), 0, c->stream, nA, nB, A, B,
                           (double2 *)f.xy.p);
        const dim3 g(ceil_div(cells, 256)), b(256);
        if (fe == ELFEL_T3)
            hipLaunchKernelGGL(k_block_cells<ELFEL_T3>, g, b, 0, c->stream, nA, nB, A, B, (int32_t *)f.conn.p, (double2 *)f.xy.p);
        else if (fe == ELFEL_Q4)
            hipLaunchKernelGGL(k_block_cells<ELFEL_Q4>, g, b, 0, c->stream, nA, nB, A, B, (int32_t *)f.conn.p, (double2 *)f.xy.p);
        else
            hipLaunchKernelGGL(k_block_cells<ELFEL_T6>, g, b, 0, c->stream, nA, nB, A, B, (int32_t *)f.conn.p, (double2 *)f.xy.p);
        HIPCHK(c, hipGetLastError());
        if ((rc = finish_field_mesh(c, f, nen, E, Nn))) return rc;
        if (E_out) *E_out = E;
        if (Nn_out) *Nn_out = Nn;
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_t6_to_t3(elfel_hip_ctx *c, int src, int dst, int64_t *Nn_out)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        if (src < 0 || src > 1 || dst < 0 || dst > 1 || src == dst)
            return set_err(c, ELFEL_EINVAL, "elfel_hip_t6_to_t3: bad fields");
        Field &s = c->f[src], &d = c->f[dst];
        if (!s.has_mesh || s.nen != 6 || s.sdim != 2) return set_err(c, ELFEL_ESTATE, "elfel_hip_t6_to_t3: source field has no T6 mesh");
        const int64_t E = s.E, Nn6 = s.Nn;
        int rc;
        // flags (scanned in place), a copy of the flags, the total
        if ((rc = ensure(c, c->flags, 4 * (size_t)(2 * Nn6 + 2)))) return rc;
        uint32_t *num = (uint32_t *)c->flags.p, *isc = num + Nn6 + 1, *total = isc + Nn6;
        hipLaunchKernelGGL(k_zero_u32, dim3(ceil_div(Nn6 + 1, 256)), dim3(256), 0, c->stream, num, Nn6 + 1);
        hipLaunchKernelGGL(k_mark_corners, dim3(ceil_div(std::max<int64_t>(E * 3, 1), 256)), dim3(256), 0, c->stream, E,
                           (const int32_t *)s.conn.p, num);
        hipLaunchKernelGGL(k_copy_u32_n, dim3(ceil_div(Nn6, 256)), dim3(256), 0, c->stream, (const uint32_t *)num, isc,
                           Nn6);
        HIPCHK(c, hipGetLastError());
        if ((rc = exclusive_scan_u32(c, num, Nn6, total))) return rc;
        uint32_t h_total = 0;
        HIPCHK(c, hipMemcpyAsync(&h_total, total, 4, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        const int64_t Nn3 = h_total;
        d.has_mesh = false;
        if ((rc = ensure(c, d.conn, 4 * (size_t)(E * 3))) || (rc = ensure(c, d.xy, 16 * (size_t)std::max<int64_t>(Nn3, 1))))
            return rc;
        hipLaunchKernelGGL(k_corner_xy, dim3(ceil_div(Nn6, 256)), dim3(256), 0, c->stream, Nn6, (const uint32_t *)num,
                           (const uint32_t *)isc, (const double2 *)s.xy.p, (double2 *)d.xy.p);
        hipLaunchKernelGGL(k_corner_conn, dim3(ceil_div(std::max<int64_t>(E * 3, 1), 256)), dim3(256), 0, c->stream, E,
                           (const int32_t *)s.conn.p, (const uint32_t *)num, (int32_t *)d.conn.p);
        HIPCHK(c, hipGetLastError());
        if ((rc = finish_field_mesh(c, d, 3, E, Nn3))) return rc;
        if (Nn_out) *Nn_out = Nn3;
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_boundary_nodes(elfel_hip_ctx *c, int field, int64_t Nn, uint8_t *flags, int64_t *count)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        if (field < 0 || field > 1) return set_err(c, ELFEL_EINVAL, "elfel_hip_boundary_nodes: bad field");
        const Field &f = c->f[field];
        if (!f.has_mesh) return set_err(c, ELFEL_ESTATE, "elfel_hip_boundary_nodes: no mesh");
        if (f.sdim != 2) return set_err(c, ELFEL_ENOTSUP, "elfel_hip_boundary_nodes: 2D meshes only");
        if (!flags || Nn != f.Nn) return set_err(c, ELFEL_EINVAL, "elfel_hip_boundary_nodes: flags must hold Nn entries");
        const int ncorner = f.nen == 6 ? 3 : f.nen;  // T3/T6: 3 corner edges, Q4: 4
        if (f.nen != 3 && f.nen != 4 && f.nen != 6) return set_err(c, ELFEL_ENOTSUP, "boundary of this element shape");
        const int64_t n = f.E * ncorner;
        int rc;
        if ((rc = ensure(c, c->keys[0], 8 * (size_t)std::max<int64_t>(n, 1))) ||
            (rc = ensure(c, c->keys[1], 8 * (size_t)std::max<int64_t>(n, 1))) ||
            (rc = ensure(c, c->vals[0], 8 * (size_t)std::max<int64_t>(n, 1))) ||
            (rc = ensure(c, c->vals[1], 8 * (size_t)std::max<int64_t>(n, 1))) ||
            (rc = ensure(c, c->stage64, (size_t)std::max<int64_t>(f.Nn, 1))))
            return rc;
        uint64_t *k0 = (uint64_t *)c->keys[0].p, *k1 = (uint64_t *)c->keys[1].p;
        uint64_t *v0 = (uint64_t *)c->vals[0].p, *v1 = (uint64_t *)c->vals[1].p;
        uint8_t *dflag = (uint8_t *)c->stage64.p;
        HIPCHK(c, hipMemsetAsync(dflag, 0, (size_t)f.Nn, c->stream));
        if (n > 0) {
            hipLaunchKernelGGL(k_edge_keys, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, f.E, f.nen, ncorner, f.Nn,
                               (const int32_t *)f.conn.p, k0, v0);
            HIPCHK(c, hipGetLastError());
            const int bits = bits_for(f.Nn * f.Nn);
            if ((rc = radix_sort_pairs(c, k0, v0, k1, v1, n, bits))) return rc;
            hipLaunchKernelGGL(k_mark_boundary, dim3(ceil_div(n, 256)), dim3(256), 0, c->stream, n, f.nen, ncorner,
                               (const int32_t *)f.conn.p, (const uint64_t *)k0, (const uint64_t *)v0, dflag);
            HIPCHK(c, hipGetLastError());
        }
        HIPCHK(c, hipMemcpyAsync(flags, dflag, (size_t)f.Nn, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        int64_t cnt = 0;
        for (int64_t i = 0; i < f.Nn; i++) cnt += flags[i] != 0;
        if (count) *count = cnt;
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_copy_mesh(elfel_hip_ctx *c, int field, int64_t E, int64_t *conn, int64_t Nn, double *xy)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        if (field < 0 || field > 1) return set_err(c, ELFEL_EINVAL, "elfel_hip_copy_mesh: bad field");
        const Field &f = c->f[field];
        if (!f.has_mesh) return set_err(c, ELFEL_ESTATE, "elfel_hip_copy_mesh: no mesh");
        if (E != f.E || Nn != f.Nn) return set_err(c, ELFEL_EINVAL, "elfel_hip_copy_mesh: sizes differ from the mesh");
        if (xy && Nn > 0)
            HIPCHK(c, hipMemcpyAsync(xy, f.xy.p, 8 * (size_t)f.sdim * Nn, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        if (conn)
            for (int64_t i = 0; i < E * f.nen; i++) conn[i] = (int64_t)f.h_conn[i] + 1;
        return ELFEL_OK;
    });
}

// ---- mesh files ----------------------------------------------------------------
// The reference's on-disk mesh (MeshSteward's load(Mesh(), "name.mesh"), as the
// tests read test/qmesh.mesh and test/mt3gen3.mesh): a three-line manifest --
// coordinate file, shape name, connectivity file -- next to two whitespace-
// separated text tables, one row per node (x y [z]) and one row per element
// (1-based node numbers).
namespace {

bool read_table(const std::string &path, int want_cols_min, std::vector<std::vector<double>> &rows)
{
    FILE *fh = fopen(path.c_str(), "r");
    if (!fh) return false;
    char line[4096];
    while (fgets(line, sizeof line, fh)) {
        std::vector<double> r;
        char *p = line, *end = nullptr;
        for (;;) {
            const double v = strtod(p, &end);
            if (end == p) break;
            r.push_back(v);
            p = end;
        }
        if (r.empty()) continue;
        if ((int)r.size() < want_cols_min) { fclose(fh); return false; }
        rows.push_back(std::move(r));
    }
    fclose(fh);
    return true;
}

std::string trim(const char *s)
{
    std::string t(s);
    const size_t a = t.find_first_not_of(" \t\r\n"), b = t.find_last_not_of(" \t\r\n");
    return a == std::string::npos ? std::string() : t.substr(a, b - a + 1);
}

}  // namespace

extern "C" int elfel_hip_load_mesh(elfel_hip_ctx *c, int field, const char *path, int32_t *nen_out, int64_t *E_out,
                                   int64_t *Nn_out)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        if (!path || field < 0 || field > 1) return set_err(c, ELFEL_EINVAL, "elfel_hip_load_mesh: bad arguments");
        FILE *fh = fopen(path, "r");
        if (!fh) return set_err(c, ELFEL_EINVAL, std::string("cannot open ") + path);
        std::vector<std::string> lines;
        char buf[4096];
        while (fgets(buf, sizeof buf, fh)) {
            const std::string t = trim(buf);
            if (!t.empty()) lines.push_back(t);
        }
        fclose(fh);
        if (lines.size() < 3) return set_err(c, ELFEL_EINVAL, "mesh manifest needs three lines: xyz file, shape, conn file");
        const std::string p(path);
        const size_t slash = p.find_last_of('/');
        const std::string dir = slash == std::string::npos ? std::string() : p.substr(0, slash + 1);
        const std::string &shape = lines[1];
        const int nen = shape == "T3" ? 3 : shape == "Q4" ? 4 : shape == "T6" ? 6 : 0;
        if (!nen) return set_err(c, ELFEL_ENOTSUP, "mesh shape " + shape + " (T3, Q4 and T6 are supported)");
        std::vector<std::vector<double>> xyz, cn;
        if (!read_table(dir + lines[0], 2, xyz)) return set_err(c, ELFEL_EINVAL, "cannot read " + dir + lines[0]);
        if (!read_table(dir + lines[2], nen, cn)) return set_err(c, ELFEL_EINVAL, "cannot read " + dir + lines[2]);
        const int64_t Nn = (int64_t)xyz.size(), E = (int64_t)cn.size();
        std::vector<double> xy(2 * (size_t)Nn);
        for (int64_t n = 0; n < Nn; n++) { xy[2 * n] = xyz[n][0]; xy[2 * n + 1] = xyz[n][1]; }
        std::vector<int64_t> conn((size_t)(E * nen));
        for (int64_t e = 0; e < E; e++) {
            if ((int)cn[e].size() != nen) return set_err(c, ELFEL_EINVAL, "connectivity row with a wrong node count");
            for (int k = 0; k < nen; k++) conn[e * nen + k] = (int64_t)cn[e][k];
        }
        const int rc = elfel_hip_set_mesh(c, field, nen, E, conn.data(), Nn, xy.data());
        if (rc) return rc;
        if (nen_out) *nen_out = nen;
        if (E_out) *E_out = E;
        if (Nn_out) *Nn_out = Nn;
        return ELFEL_OK;
    });
}
