// batch.hip -- many independent problems of uneven size assembled by one
// persistent fused launch (heat T3 1-point, heat Q4 2x2 and elasticity T6
// 3-point; all mixed or all FP64), replayed from a hipGraph.
//
// build: per problem, the pattern comes from the plain method and the task
// tables from the reuse tiler; the tables of all problems of one kind are
// concatenated (element, node, value and row offsets folded in), so a tile
// needs no problem id.  The launch shape -- tasks per workgroup (block of 128
// or 256 threads, the tile capacity) and the persistent grid size -- is chosen
// by timing every variant on the batch itself (autotune) and cached on disk,
// keyed by device and batch signature.
// assemble: one graph launch of k_reuse_batch (it gathers the coordinates itself).
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <deque>
#include <memory>
#include <mutex>
#include <string>
#include <thread>

#include "reuse_kernel.h"

struct BatchProb {
    int kind = 0;
    int64_t E = 0, Nn = 0, ndofs = 0, nnz = 0;
    int64_t elem_base = 0, node_base = 0, val_base = 0, ptr_base = 0, row_base = 0;
    std::vector<int64_t> conn, dofs;
    std::vector<double> xy, coef;
    bool has_coef = false;
};

struct KindData {
    int64_t E = 0, Nn = 0, ntasks = 0, ntiles = 0;
    Buf conn, xy, coef;
    Pattern tab[2];  // tables for block 128 and 256
};

// The problem kinds of a batch (index into Batch::kd): the element, rule and
// dof layout each stands for.
enum { BK_HEAT_T3 = 0, BK_ELAST_T6 = 1, BK_HEAT_Q4 = 2, BK_N = 3 };
struct KindSpec {
    int kind, fe, rule, nen, nc;
    bool heat;
    const char *name;
};
static const KindSpec KS[BK_N] = {{ELFEL_HEAT, ELFEL_T3, 1, 3, 1, true, "heatT3r1"},
                                  {ELFEL_ELAST, ELFEL_T6, 3, 6, 2, false, "elastT6r3"},
                                  {ELFEL_HEAT, ELFEL_Q4, 2, 4, 1, true, "heatQ4r2"}};

struct Batch {
    std::vector<BatchProb> probs;
    elfel_hip_problem pb[BK_N]{};  // the uniform coefficients of each kind
    bool have[BK_N] = {false, false, false}, built = false;
    int layout = ELFEL_CSC, precision = -1;
    KindData kd[BK_N];
    Buf ptr, ind, val, F;
    int64_t total_nnz = 0, total_ptr = 0, total_rows = 0;
    int block = 256, grid = 0, lds_vals = 0;
    double tuned_ms = 0;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
};

static Batch *get_batch(elfel_hip_ctx *c)
{
    if (!c->batch) c->batch = new Batch();
    return (Batch *)c->batch;
}

static void batch_destroy_graph(Batch *b)
{
    if (b->exec) hipGraphExecDestroy(b->exec);
    if (b->graph) hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
}

void batch_free(elfel_hip_ctx *c)
{
    Batch *b = (Batch *)c->batch;
    if (!b) return;
    batch_destroy_graph(b);
    for (auto &k : b->kd) {
        release(k.conn); release(k.xy); release(k.coef);
        for (auto &t : k.tab) {
            release(t.tile); release(t.tile_rows); release(t.task_elem); release(t.task_info); release(t.task_pos);
            release(t.tile_elem); release(t.rrec); release(t.tconn); release(t.txy);
        }
    }
    release(b->ptr); release(b->ind); release(b->val); release(b->F);
    delete b;
    c->batch = nullptr;
}

extern "C" int elfel_hip_batch_begin(elfel_hip_ctx *c)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        hipStreamSynchronize(c->stream);
        batch_free(c);
        get_batch(c);
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_batch_add(elfel_hip_ctx *c, const elfel_hip_problem *pb, int32_t nen, int64_t E,
                                   const int64_t *conn, int64_t Nn, const double *xy, int32_t ncomp,
                                   const int64_t *dofnums, int64_t ndofs, const double *coef_e)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        Batch *b = get_batch(c);
        if (b->built) return set_err(c, ELFEL_ESTATE, "batch already built; call elfel_hip_batch_begin");
        if (!pb || E <= 0 || Nn <= 0 || !conn || !xy || !dofnums || ndofs <= 0)
            return set_err(c, ELFEL_EINVAL, "elfel_hip_batch_add: bad arguments");
        int kind = -1;
        for (int k = 0; k < BK_N; k++)
            if (pb->kind == KS[k].kind && pb->fe == KS[k].fe && pb->rule == KS[k].rule && nen == KS[k].nen &&
                ncomp == KS[k].nc)
                kind = k;
        if (kind < 0)
            return set_err(c, ELFEL_ENOTSUP, "batch problems: heat T3 1-point, heat Q4 2x2 or elasticity T6 3-point");
        if (pb->precision != ELFEL_MIXED && pb->precision != ELFEL_FP64)
            return set_err(c, ELFEL_EINVAL, "elfel_hip_batch_add: bad precision");
        if (b->precision >= 0 && pb->precision != b->precision)
            return set_err(c, ELFEL_ENOTSUP, "the problems of one batch share their precision");
        if (pb->flags) return set_err(c, ELFEL_ENOTSUP, "batch problems take no flags");
        elfel_hip_problem &ref = b->pb[kind];
        if (!b->have[kind]) { ref = *pb; b->have[kind] = true; }
        else if (memcmp(&ref, pb, sizeof(*pb)) != 0)
            return set_err(c, ELFEL_ENOTSUP, "problems of one kind must share their uniform coefficients");
        if (!b->probs.empty() && pb->layout != b->layout) return set_err(c, ELFEL_EINVAL, "mixed layouts in one batch");
        b->layout = pb->layout;
        b->precision = pb->precision;
        for (int64_t i = 0; i < E * nen; i++)
            if (conn[i] < 1 || conn[i] > Nn) return set_err(c, ELFEL_ERANGE, "connectivity outside 1..Nn");
        for (int64_t i = 0; i < Nn * ncomp; i++)
            if (dofnums[i] < 1 || dofnums[i] > ndofs) return set_err(c, ELFEL_ERANGE, "dof number outside 1..ndofs");
        BatchProb p;
        p.kind = kind;
        p.E = E; p.Nn = Nn; p.ndofs = ndofs;
        p.conn.assign(conn, conn + E * nen);
        p.dofs.assign(dofnums, dofnums + Nn * ncomp);
        p.xy.assign(xy, xy + 2 * Nn);
        if (coef_e) { p.coef.assign(coef_e, coef_e + E); p.has_coef = true; }
        b->probs.push_back(std::move(p));
        return ELFEL_OK;
    });
}

static std::string tune_cache_path()
{
    const char *e = getenv("ELFEL_HIP_TUNE_CACHE");
    if (e && *e) return e;
    const char *h = getenv("HOME");
    return std::string(h ? h : "/tmp") + "/.elfel_hip_tune.txt";
}

static bool tune_lookup(const std::string &key, int *block, int *gmult)
{
    FILE *f = fopen(tune_cache_path().c_str(), "r");
    if (!f) return false;
    char k[512];
    int bl, gm;
    bool found = false;
    // only the launch shapes the autotuner times are taken from the file
    while (fscanf(f, "%511s %d %d", k, &bl, &gm) == 3)
        if (key == k && (bl == 128 || bl == 256) && (gm == 0 || gm == 2 || gm == 4 || gm == 8 || gm == 16)) {
            *block = bl;
            *gmult = gm;
            found = true;
        }
    fclose(f);
    return found;
}

static void tune_store(const std::string &key, int block, int gmult)
{
    FILE *f = fopen(tune_cache_path().c_str(), "a");
    if (!f) return;
    fprintf(f, "%s %d %d\n", key.c_str(), block, gmult);
    fclose(f);
}

// Row-owner heat tiles in the batch (ELFEL_HIP_BATCH_ROWS=0: task tiles).
static bool batch_rows_enabled() { return env_int("ELFEL_HIP_BATCH_ROWS", 1) != 0; }

// Task slots per tile: heat tiles hold two light tasks per thread,
// elasticity tiles one.
static int batch_slots(int kind, int block) { return KS[kind].heat ? 2 * block : block; }

// One batched assembly on the stream (no synchronisation).  Tile ranges:
// heat T3, then heat Q4, then elasticity (k_reuse_batch).
static int batch_launch(elfel_hip_ctx *c, Batch *b, int block, int grid)
{
    const int bi = block == 128 ? 0 : 1;
    BatchLaunch L{};
    int nk[BK_N];
    int lds_vals = 0;
    for (int k = 0; k < BK_N; k++) {
        KindData &K = b->kd[k];
        const Pattern &P = K.tab[bi];
        nk[k] = K.E ? (int)P.ntiles : 0;
        ReuseArgs &A = L.A[k];
        A = ReuseArgs{};
        A.tiles = (const int4 *)P.tile.p;
        A.tile_rows = (const int4 *)P.tile_rows.p;
        A.task_elem = (const int32_t *)P.task_elem.p;
        A.task_elem8 = P.elem8 ? (const uint8_t *)P.task_elem.p : nullptr;
        A.tile_elem = (const int32_t *)P.tile_elem.p;
        A.maxu = P.maxu;
        A.task_info = (const uint32_t *)P.task_info.p;
        A.task_pos = (const uint8_t *)P.task_pos.p;
        A.rrec = P.rw ? (const uint32_t *)P.rrec.p : nullptr;  // row-owner heat tiles (block 256)
        A.tstride = P.tstride;
        A.tconn = P.rw && P.tgeom ? (const ushort4 *)P.tconn.p : nullptr;
        A.txy = P.rw && P.tgeom ? (const double2 *)P.txy.p : nullptr;
        A.ncolp = P.ncolmax;
        A.ntiles = nk[k];
        A.geom = nullptr;  // the kernel gathers conn and xy itself
        A.conn = (const int32_t *)K.conn.p;
        A.xy = (const double2 *)K.xy.p;
        A.ngap = A.xgap = 0;  // the concatenated arrays are stored plain
        A.vals = (double *)b->val.p;
        A.F = KS[k].heat ? (double *)b->F.p : nullptr;
        Coef &cf = L.cf[k];
        cf = Coef{};
        cf.kappa = b->pb[k].kappa; cf.Q = b->pb[k].Q; cf.mu = b->pb[k].mu;
        for (int i = 0; i < 9; i++) cf.D[i] = b->pb[k].D[i];
        const double *ce = K.coef.p ? (const double *)K.coef.p : nullptr;
        if (KS[k].heat) cf.kappa_e = ce;
        else cf.scale_e = ce;
        if (!nk[k]) continue;
        lds_vals = std::max(lds_vals, lds_span((int)P.lds_vals));
        if (KS[k].heat) {  // element records alias the value sums
            const size_t rb = (size_t)P.maxu * (k == BK_HEAT_T3 ? heat_rec<3>() : heat_rec<4>()) *
                              (b->precision == ELFEL_FP64 ? 8 : 4);
            lds_vals = std::max(lds_vals, (int)((rb + 7) / 8));
        }
    }
    L.n0 = nk[BK_HEAT_T3];
    L.n1 = L.n0 + nk[BK_HEAT_Q4];
    L.nt = L.n1 + nk[BK_ELAST_T6];
    if (L.nt == 0) return 0;
    lds_vals = (lds_vals + 1) & ~1;  // even: the load-vector sums stay 16-byte aligned
    L.lds_vals = lds_vals;
    // + the heat tiles' load-vector sums (up to REUSE_MAX_TILE_ROWS rows)
    L.lds = sizeof(double) * ((size_t)lds_vals + REUSE_MAX_TILE_ROWS);
    L.grid = grid > 0 ? std::min(grid, L.nt) : L.nt;
    L.block = block;
    L.csc = b->layout == ELFEL_CSC;
    const int rc = b->precision == ELFEL_FP64 ? launch_batch64(c->stream, L) : launch_batch32(c->stream, L);
    if (rc) return set_err(c, ELFEL_EHIP, std::string("batched launch: ") + hipGetErrorString(hipGetLastError()));
    return 0;
}

int launch_batch32(hipStream_t st, const BatchLaunch &L) { return launch_batch_t<float>(st, L); }

static int time_variant(elfel_hip_ctx *c, Batch *b, int block, int grid, double *ms)
{
    int rc = batch_launch(c, b, block, grid);  // warm-up
    if (rc) return rc;
    hipEvent_t e0, e1;
    HIPCHK(c, hipEventCreate(&e0));
    HIPCHK(c, hipEventCreate(&e1));
    const int reps = 5;
    HIPCHK(c, hipEventRecord(e0, c->stream));
    for (int r = 0; r < reps && !rc; r++) rc = batch_launch(c, b, block, grid);
    HIPCHK(c, hipEventRecord(e1, c->stream));
    HIPCHK(c, hipEventSynchronize(e1));
    float t = 0;
    hipEventElapsedTime(&t, e0, e1);
    hipEventDestroy(e0);
    hipEventDestroy(e1);
    *ms = t / reps;
    return rc;
}

extern "C" int elfel_hip_batch_build(elfel_hip_ctx *c, int autotune)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        Batch *b = get_batch(c);
        if (b->probs.empty()) return set_err(c, ELFEL_ESTATE, "empty batch");
        batch_destroy_graph(b);
        // pass 1: the patterns on the device, for chunks of problems of one kind
        // at a time (each chunk one block-diagonal system: its problems' meshes
        // side by side, their dofs offset, so one plain assembly gives every
        // problem's pattern as a diagonal block), on this thread; each problem's
        // pattern slice, host copies and tile tables (both workgroup sizes) on host
        // threads as soon as its chunk is out
        static const bool say = env_int("ELFEL_HIP_SYMBOLIC_TIMES", 0) != 0;
        auto t_lap = std::chrono::steady_clock::now();
        auto lap = [&](const char *what) {
            const auto now = std::chrono::steady_clock::now();
            if (say) fprintf(stderr, "[batch build] %-28s %8.1f ms\n", what, std::chrono::duration<double, std::milli>(now - t_lap).count());
            t_lap = now;
        };
        const size_t np = b->probs.size();
        std::vector<std::vector<int32_t>> hptrs(np), hinds(np);
        std::vector<ReuseTables> tabs[2];
        tabs[0].resize(np);
        tabs[1].resize(np);
        std::vector<Field> hf(np);  // host copies of each problem's mesh and dof map (0-based)
        std::vector<int> jrc(2 * np, 0);
        std::vector<std::string> jerr(2 * np);
        struct Chunk { std::vector<int32_t> ptr, ind; };
        struct Job { size_t k; std::shared_ptr<const Chunk> ch; int64_t d0; };  // problem k: chunk rows d0..
        std::deque<Job> queue;
        std::mutex mu;
        std::condition_variable cv;
        bool fed = false, stop = false;
        // an exception on a host thread (a failed host allocation) fails the build
        std::mutex exc_mu;
        int exc_rc = 0;
        std::string exc_err;
        auto note_exception = [&]() {
            std::string m;
            const int code = exception_code(&m);
            std::lock_guard<std::mutex> lk(exc_mu);
            if (!exc_rc) { exc_rc = code; exc_err = m; }
        };
        auto worker_jobs = [&]() {
            for (;;) {
                Job jb;
                {
                    std::unique_lock<std::mutex> lk(mu);
                    cv.wait(lk, [&] { return stop || fed || !queue.empty(); });
                    if (stop || queue.empty()) return;
                    jb = std::move(queue.front());
                    queue.pop_front();
                }
                const size_t k = jb.k;
                const BatchProb &p = b->probs[k];
                const int nen = KS[p.kind].nen, nc = KS[p.kind].nc;
                const int32_t *cp = jb.ch->ptr.data() + jb.d0;
                const int32_t v0 = cp[0], d0 = (int32_t)jb.d0;
                std::vector<int32_t> &hp = hptrs[k], &hi = hinds[k];
                hp.resize((size_t)p.ndofs + 1);
                for (int64_t i = 0; i <= p.ndofs; i++) hp[i] = cp[i] - v0;
                hi.resize((size_t)hp[p.ndofs]);
                const int32_t *ci = jb.ch->ind.data() + v0;
                for (size_t i = 0; i < hi.size(); i++) hi[i] = ci[i] - d0;
                jb.ch.reset();
                Field &h = hf[k];
                h.nen = nen; h.ncomp = nc; h.sdim = 2;
                h.E = p.E; h.Nn = p.Nn; h.ndofs = p.ndofs;
                h.h_conn.resize(p.conn.size());
                for (size_t i = 0; i < p.conn.size(); i++) h.h_conn[i] = (int32_t)(p.conn[i] - 1);
                h.h_dof.resize(p.dofs.size());
                for (size_t i = 0; i < p.dofs.size(); i++) h.h_dof[i] = (int32_t)(p.dofs[i] - 1);
                // heat problems get row-owner tiles for workgroups of 256 where their
                // mesh allows (then every problem of the kind must: below)
                for (int bi = 0; bi < 2; bi++)
                    jrc[2 * k + bi] = tile_tables(c, false, h, nullptr, p.ndofs, hp, hi, batch_slots(p.kind, bi ? 256 : 128),
                                                  bi ? 256 : 128, tabs[bi][k], false,
                                                  bi == 1 && KS[p.kind].heat && batch_rows_enabled(),
                                                  (const double2 *)p.xy.data(), &jerr[2 * k + bi]);
            }
        };
        auto worker = [&]() {
            try {
                worker_jobs();
            } catch (...) {
                note_exception();
            }
        };
        const int nth = (int)std::max<size_t>(1, std::min<size_t>(
            np, std::min<size_t>(16, std::max(1u, std::thread::hardware_concurrency()))));
        std::vector<std::thread> pool;
        // on every way out (an error below, an exception) the waiting threads
        // are told to stop and joined
        auto joiner = join_on_exit(pool, [&] {
            {
                std::lock_guard<std::mutex> lk(mu);
                stop = true;
            }
            cv.notify_all();
        });
        for (int t = 0; t < nth; t++) pool.emplace_back(worker);
        auto halt = [&](int rc) {
            {
                std::lock_guard<std::mutex> lk(mu);
                stop = true;
            }
            cv.notify_all();
            for (auto &th : pool) th.join();
            return rc;
        };
        // a chunk holds up to 2^26 element-block entries (the COO stream of a
        // 466,000-element elasticity mesh), or one larger problem
        const int64_t chunk_entries = int64_t(1) << 26;
        std::vector<int64_t> nnzk(np, 0);
        for (int kind = 0; kind < BK_N; kind++) {
            const int nen = KS[kind].nen, nc = KS[kind].nc;
            const int64_t per_el = (int64_t)nen * nen * nc * nc;
            std::vector<size_t> idx;
            for (size_t k = 0; k < np; k++)
                if (b->probs[k].kind == kind) idx.push_back(k);
            const elfel_hip_problem &pb = b->pb[kind];
            QTable tab;
            make_qtable(pb.fe, pb.rule, &tab);
            Coef cf{};
            for (int i = 0; i < 9; i++) cf.D[i] = pb.D[i];
            for (size_t i = 0; i < idx.size();) {
                size_t j = i;
                int64_t Ec = 0, Nc = 0, Dc = 0;
                do {
                    const BatchProb &p = b->probs[idx[j]];
                    Ec += p.E; Nc += p.Nn; Dc += p.ndofs;
                    j++;
                } while (j < idx.size() && (Ec + b->probs[idx[j]].E) * per_el <= chunk_entries);
                std::vector<int64_t> conn((size_t)(Ec * nen)), dofs((size_t)(Nc * nc));
                std::vector<double> xy((size_t)(2 * Nc));
                int64_t eo = 0, no = 0, dofo = 0;
                for (size_t m = i; m < j; m++) {
                    const BatchProb &p = b->probs[idx[m]];
                    for (size_t q = 0; q < p.conn.size(); q++) conn[(size_t)(eo * nen) + q] = p.conn[q] + no;
                    for (size_t q = 0; q < p.dofs.size(); q++) dofs[(size_t)(no * nc) + q] = p.dofs[q] + dofo;
                    std::copy(p.xy.begin(), p.xy.end(), xy.begin() + 2 * no);
                    eo += p.E; no += p.Nn; dofo += p.ndofs;
                }
                int rc = elfel_hip_set_mesh(c, 0, nen, Ec, conn.data(), Nc, xy.data());
                if (!rc) rc = elfel_hip_set_dofs(c, 0, nc, Nc, dofs.data(), Dc);
                if (rc) return halt(rc);
                c->nrow = c->ncol = Dc;
                auto ch = std::make_shared<Chunk>();
                if ((rc = pattern_to_host(c, &pb, tab, cf, ch->ptr, ch->ind))) return halt(rc);
                {
                    std::lock_guard<std::mutex> lk(mu);
                    dofo = 0;
                    for (size_t m = i; m < j; m++) {
                        const BatchProb &p = b->probs[idx[m]];
                        nnzk[idx[m]] = ch->ptr[dofo + p.ndofs] - ch->ptr[dofo];
                        queue.push_back(Job{idx[m], ch, dofo});
                        dofo += p.ndofs;
                    }
                }
                cv.notify_all();
                i = j;
            }
        }
        {
            std::lock_guard<std::mutex> lk(mu);
            fed = true;
        }
        cv.notify_all();
        lap("patterns (device)");
        worker();  // this thread drains the queue with the others
        for (auto &th : pool) th.join();
        if (exc_rc) return set_err(c, exc_rc, exc_err);
        for (size_t j = 0; j < 2 * np; j++)
            if (jrc[j]) return set_err(c, jrc[j], jerr[j]);
        int64_t koff_E[BK_N] = {0, 0, 0}, koff_N[BK_N] = {0, 0, 0};
        int64_t val_base = 0, ptr_base = 0, row_base = 0;
        for (size_t k = 0; k < np; k++) {
            BatchProb &p = b->probs[k];
            p.nnz = nnzk[k];
            p.elem_base = koff_E[p.kind];
            p.node_base = koff_N[p.kind];
            koff_E[p.kind] += p.E;
            koff_N[p.kind] += p.Nn;
            p.val_base = val_base;
            p.ptr_base = ptr_base;
            p.row_base = row_base;
            val_base += p.nnz;
            ptr_base += p.ndofs + 1;
            if (KS[p.kind].heat) row_base += p.ndofs;
        }
        lap("tile tables (host threads)");
        // a kind's tiles are all row-owner tiles or all task tiles: a problem whose
        // mesh the row records do not hold (a node in more elements, a longer row)
        // puts every problem of its kind back on task tiles
        for (int kind = 0; kind < BK_N; kind++) {
            bool any_task = false;
            for (size_t k = 0; k < np; k++)
                if (b->probs[k].kind == kind && !tabs[1][k].rw) any_task = true;
            if (!any_task) continue;
            for (size_t k = 0; k < np; k++) {
                const BatchProb &p = b->probs[k];
                if (p.kind != kind || !tabs[1][k].rw) continue;
                std::string err;
                const int rc = tile_tables(c, false, hf[k], nullptr, p.ndofs, hptrs[k], hinds[k], batch_slots(p.kind, 256),
                                           256, tabs[1][k], false, false, (const double2 *)p.xy.data(), &err);
                if (rc) return set_err(c, rc, err);
            }
        }
        b->total_nnz = val_base;
        b->total_ptr = ptr_base;
        b->total_rows = row_base;
        // pass 2: per kind the concatenated mesh arrays, and per kind and
        // workgroup size the concatenated tables in launch order, on host threads;
        // then the uploads
        struct KindHost {
            std::vector<int32_t> conn;
            std::vector<double> xy, coef;
            bool any_coef = false;
        };
        KindHost kh[BK_N];
        ReuseTables cats[BK_N][2];
        auto mesh_arrays = [&](int kind) {
            KindHost &H = kh[kind];
            const int nen = KS[kind].nen;
            H.conn.resize((size_t)(koff_E[kind] * nen));
            H.xy.resize((size_t)(2 * koff_N[kind]));
            for (const BatchProb &p : b->probs) if (p.kind == kind && p.has_coef) H.any_coef = true;
            if (H.any_coef) H.coef.assign((size_t)koff_E[kind], 1.0);
            for (const BatchProb &p : b->probs) {
                if (p.kind != kind) continue;
                for (int64_t i = 0; i < p.E * nen; i++) H.conn[p.elem_base * nen + i] = (int32_t)(p.conn[i] - 1 + p.node_base);
                std::copy(p.xy.begin(), p.xy.end(), H.xy.begin() + 2 * p.node_base);
                if (H.any_coef) {
                    const double u = KS[kind].heat ? b->pb[kind].kappa : 1.0;
                    for (int64_t e = 0; e < p.E; e++) H.coef[p.elem_base + e] = p.has_coef ? p.coef[e] : u;
                }
            }
        };
        auto concat = [&](int kind, int bi) {
            ReuseTables &cat = cats[kind][bi];
            size_t ntl = 0, nel = 0, ntel = 0, npos = 0, nrec = 0, ntc = 0, ntxy = 0;
            for (size_t k = 0; k < np; k++) {
                if (b->probs[k].kind != kind) continue;
                const ReuseTables &t = tabs[bi][k];
                ntl += t.tiles.size(); nel += t.elem.size(); ntel += t.telem.size(); npos += t.pos.size();
                nrec += t.rrec.size(); ntc += t.tconn.size(); ntxy += t.txy.size();
            }
            cat.tiles.reserve(ntl); cat.trows.reserve(ntl);
            cat.elem.reserve(nel); cat.info.reserve(nel); cat.telem.reserve(ntel); cat.pos.reserve(npos);
            cat.rrec.reserve(nrec); cat.tconn.reserve(ntc); cat.txy.reserve(ntxy);
            for (size_t k = 0; k < np; k++) {
                const BatchProb &p = b->probs[k];
                if (p.kind != kind) continue;
                const ReuseTables &t = tabs[bi][k];
                // tile .x: the first task, or in row-owner tiles the first tile-local node
                const int task0 = (int)(t.rw ? cat.txy.size() : cat.elem.size());
                cat.ncolp = t.ncolp;
                cat.tstride = t.tstride;
                cat.rw = t.rw;
                cat.rrec.insert(cat.rrec.end(), t.rrec.begin(), t.rrec.end());
                cat.tconn.insert(cat.tconn.end(), t.tconn.begin(), t.tconn.end());
                cat.txy.insert(cat.txy.end(), t.txy.begin(), t.txy.end());
                cat.maxv = std::max(cat.maxv, t.maxv);
                cat.maxrank = std::max(cat.maxrank, t.maxrank);
                for (size_t i = 0; i < t.tiles.size(); i++) {
                    int4 td = t.tiles[i], tr = t.trows[i];
                    td.x += task0;
                    td.z += (int)p.val_base;
                    tr.x += (int)p.row_base;
                    cat.tiles.push_back(td);
                    cat.trows.push_back(tr);
                }
                if (t.telem.empty()) {
                    for (int32_t e : t.elem) cat.elem.push_back(e + (int32_t)p.elem_base);
                } else {  // element-staged tiles: tasks hold tile-local indices
                    cat.elem.insert(cat.elem.end(), t.elem.begin(), t.elem.end());
                    for (int32_t e : t.telem) cat.telem.push_back(e + (int32_t)p.elem_base);
                    cat.maxu = std::max(cat.maxu, t.maxu);
                }
                cat.info.insert(cat.info.end(), t.info.begin(), t.info.end());
                cat.pos.insert(cat.pos.end(), t.pos.begin(), t.pos.end());
            }
            // tile t runs on XCD t mod 8 (one workgroup per tile, or a persistent
            // grid of a multiple of 8 workgroups striding the list): runs of 64
            // neighbouring tiles, which read the same elements, go to one XCD's
            // L2, and each XCD gets runs from every part of the list.  This cuts
            // the fetched bytes from 3.2 to 2.1 GB per launch at about the same
            // time (the launch is not fetch-bound); whole chunks per XCD left
            // the XCDs unevenly loaded (1.35 against 1.31 ms)
            order_tiles(cat, 8, batch_slots(kind, bi ? 256 : 128), env_int("ELFEL_HIP_XCD_RUN_BATCH", 64));
        };
        {
            std::vector<std::thread> p2;
            auto joiner2 = join_on_exit(p2, [] {});
            for (int kind = 0; kind < BK_N; kind++) {
                if (!koff_E[kind]) continue;
                p2.emplace_back([&, kind] {
                    try {
                        mesh_arrays(kind);
                    } catch (...) {
                        note_exception();
                    }
                });
                for (int bi = 0; bi < 2; bi++)
                    p2.emplace_back([&, kind, bi] {
                        try {
                            concat(kind, bi);
                        } catch (...) {
                            note_exception();
                        }
                    });
            }
        }
        if (exc_rc) return set_err(c, exc_rc, exc_err);
        lap("concatenate (host threads)");
        int rc;
        if ((rc = ensure(c, b->ptr, 4 * ptr_base)) || (rc = ensure(c, b->ind, 4 * std::max<int64_t>(val_base, 1))) ||
            (rc = ensure(c, b->val, 8 * std::max<int64_t>(val_base, 1))) ||
            (rc = ensure(c, b->F, 8 * std::max<int64_t>(row_base, 1))))
            return rc;
        for (size_t k = 0; k < np; k++) {
            const BatchProb &p = b->probs[k];
            HIPCHK(c, hipMemcpy((int32_t *)b->ptr.p + p.ptr_base, hptrs[k].data(), 4 * (p.ndofs + 1), hipMemcpyHostToDevice));
            if (p.nnz)
                HIPCHK(c, hipMemcpy((int32_t *)b->ind.p + p.val_base, hinds[k].data(), 4 * p.nnz, hipMemcpyHostToDevice));
        }
        for (int kind = 0; kind < BK_N; kind++) {
            KindData &K = b->kd[kind];
            K.E = koff_E[kind];
            K.Nn = koff_N[kind];
            if (!K.E) continue;
            const KindHost &H = kh[kind];
            if ((rc = ensure(c, K.conn, 4 * H.conn.size())) || (rc = ensure(c, K.xy, 8 * H.xy.size()))) return rc;
            HIPCHK(c, hipMemcpy(K.conn.p, H.conn.data(), 4 * H.conn.size(), hipMemcpyHostToDevice));
            HIPCHK(c, hipMemcpy(K.xy.p, H.xy.data(), 8 * H.xy.size(), hipMemcpyHostToDevice));
            if (H.any_coef) {
                if ((rc = ensure(c, K.coef, 8 * H.coef.size()))) return rc;
                HIPCHK(c, hipMemcpy(K.coef.p, H.coef.data(), 8 * H.coef.size(), hipMemcpyHostToDevice));
            } else {
                release(K.coef);
            }
            for (int bi = 0; bi < 2; bi++)
                if ((rc = upload_tables(c, cats[kind][bi], K.tab[bi]))) return rc;
            K.ntiles = K.tab[1].ntiles;
            K.ntasks = K.tab[1].ntasks;
        }
        if (b->total_nnz >= INT32_MAX || b->total_ptr >= INT32_MAX)
            return set_err(c, ELFEL_ENOTSUP, "batch beyond int32 value offsets");
        b->built = true;
        lap("upload");
        // launch shape: cache or autotune
        hipDeviceProp_t prop;
        HIPCHK(c, hipGetDeviceProperties(&prop, c->device));
        const int ncu = prop.multiProcessorCount;
        // the key names the batch exactly: precision, layout, and per kind the
        // tile counts at both workgroup sizes and the element count; the matrix
        // entries
        std::string key = std::string(prop.gcnArchName) + ":batch:" + (b->precision == ELFEL_FP64 ? "fp64" : "mixed") +
                          ":" + (b->layout == ELFEL_CSC ? "csc" : "csr");
        for (int k = 0; k < BK_N; k++) {
            char part[160];
            snprintf(part, sizeof part, ":%s.%lld.%lld.%lld", KS[k].name, (long long)b->kd[k].tab[0].ntiles,
                     (long long)b->kd[k].tab[1].ntiles, (long long)b->kd[k].E);
            key += part;
        }
        key += ":nnz" + std::to_string((long long)b->total_nnz);
        int block = 256, gmult = 0;
        bool cached = tune_lookup(key, &block, &gmult);
        // ELFEL_HIP_BATCH_BLOCK=128 or 256: that workgroup size, one workgroup
        // per tile, no timing (tests of each tile form)
        const int forced = env_int("ELFEL_HIP_BATCH_BLOCK", 0);
        if (forced == 128 || forced == 256) {
            block = forced;
            gmult = 0;
            cached = true;
            autotune = 0;
        }
        if (autotune || !cached) {
            const int blocks[2] = {128, 256};
            const int gms[5] = {2, 4, 8, 16, 0};
            double best = 1e30;
            for (int bl : blocks)
                for (int gm : gms) {
                    double ms = 0;
                    if ((rc = time_variant(c, b, bl, gm ? gm * ncu : 0, &ms))) return rc;
                    if (ms < best) { best = ms; block = bl; gmult = gm; }
                }
            b->tuned_ms = best;
            tune_store(key, block, gmult);
        }
        b->block = block;
        b->grid = gmult ? gmult * ncu : 0;
        // capture one assembly in a graph
        HIPCHK(c, hipStreamSynchronize(c->stream));
        HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
        rc = batch_launch(c, b, b->block, b->grid);
        hipGraph_t g = nullptr;
        hipError_t e = hipStreamEndCapture(c->stream, &g);
        if (rc) { if (g) hipGraphDestroy(g); return rc; }
        if (e != hipSuccess) return set_err(c, ELFEL_EHIP, "hipStreamEndCapture failed");
        b->graph = g;
        HIPCHK(c, hipGraphInstantiate(&b->exec, g, nullptr, nullptr, 0));
        lap("launch shape + graph");
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_batch_assemble(elfel_hip_ctx *c, int wait)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        Batch *b = (Batch *)c->batch;
        if (!b || !b->built || !b->exec) return set_err(c, ELFEL_ESTATE, "build the batch first");
        hipEvent_t ev = nullptr;
        tm_begin(c, ELFEL_PH_TOTAL, &ev);
        if (c->tm.region) c->tm.region_n++;
        HIPCHK(c, hipGraphLaunch(b->exec, c->stream));
        tm_end(c, ELFEL_PH_TOTAL, ev);
        if (wait) HIPCHK(c, hipStreamSynchronize(c->stream));
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_batch_info(elfel_hip_ctx *c, int32_t k, int64_t *ndofs, int64_t *nnz, int64_t *nelem,
                                    int *block, int *grid)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        Batch *b = (Batch *)c->batch;
        if (!b || !b->built) return set_err(c, ELFEL_ESTATE, "build the batch first");
        if (k < 0 || k >= (int32_t)b->probs.size()) return set_err(c, ELFEL_EINVAL, "problem index out of range");
        if (ndofs) *ndofs = b->probs[k].ndofs;
        if (nnz) *nnz = b->probs[k].nnz;
        if (nelem) *nelem = b->probs[k].E;
        if (block) *block = b->block;
        if (grid) *grid = b->grid;
        return ELFEL_OK;
    });
}

__global__ void k_ptr_i64(const int32_t *__restrict__ s, int64_t *__restrict__ d, int64_t n, int64_t base)
{
    int64_t i = (int64_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) d[i] = (int64_t)s[i] + base;
}

extern "C" int elfel_hip_batch_copy(elfel_hip_ctx *c, int32_t k, int base, int64_t *ptr, int64_t *ind, double *val,
                                    double *F)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        c->err.clear();
        (void)hipSetDevice(c->device);
        Batch *b = (Batch *)c->batch;
        if (!b || !b->built) return set_err(c, ELFEL_ESTATE, "build the batch first");
        if (k < 0 || k >= (int32_t)b->probs.size()) return set_err(c, ELFEL_EINVAL, "problem index out of range");
        const BatchProb &p = b->probs[k];
        int rc = ensure(c, c->stage64, 8 * std::max<int64_t>(p.nnz, p.ndofs + 1));
        if (rc) return rc;
        int64_t *st = (int64_t *)c->stage64.p;
        if (ptr) {
            hipLaunchKernelGGL(k_ptr_i64, dim3(ceil_div(p.ndofs + 1, 256)), dim3(256), 0, c->stream,
                               (const int32_t *)b->ptr.p + p.ptr_base, st, p.ndofs + 1, (int64_t)base);
            HIPCHK(c, hipMemcpyAsync(ptr, st, 8 * (p.ndofs + 1), hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        if (ind && p.nnz) {
            hipLaunchKernelGGL(k_ptr_i64, dim3(ceil_div(p.nnz, 256)), dim3(256), 0, c->stream,
                               (const int32_t *)b->ind.p + p.val_base, st, p.nnz, (int64_t)base);
            HIPCHK(c, hipMemcpyAsync(ind, st, 8 * p.nnz, hipMemcpyDeviceToHost, c->stream));
            HIPCHK(c, hipStreamSynchronize(c->stream));
        }
        if (val && p.nnz)
            HIPCHK(c, hipMemcpyAsync(val, (double *)b->val.p + p.val_base, 8 * p.nnz, hipMemcpyDeviceToHost, c->stream));
        if (F && KS[p.kind].heat)
            HIPCHK(c, hipMemcpyAsync(F, (double *)b->F.p + p.row_base, 8 * p.ndofs, hipMemcpyDeviceToHost, c->stream));
        HIPCHK(c, hipStreamSynchronize(c->stream));
        return ELFEL_OK;
    });
}

extern "C" int elfel_hip_batch_stats(elfel_hip_ctx *c, int64_t *nprob, int64_t *nelem, int64_t *nnz,
                                     int64_t *ntiles, double *tuned_ms)
{
    return host_guard(c, [&]() -> int {
        if (!c) return ELFEL_EINVAL;
        Batch *b = (Batch *)c->batch;
        if (!b || !b->built) return set_err(c, ELFEL_ESTATE, "build the batch first");
        if (nprob) *nprob = (int64_t)b->probs.size();
        if (nelem) *nelem = b->kd[0].E + b->kd[1].E + b->kd[2].E;
        if (nnz) *nnz = b->total_nnz;
        if (ntiles) *ntiles = b->kd[0].ntiles + b->kd[1].ntiles + b->kd[2].ntiles;
        if (tuned_ms) *tuned_ms = b->tuned_ms;
        return ELFEL_OK;
    });
}
