// Host side of the tile engine: context, device arenas, the plain, batched
// and served forms, and the C ABI of include/terranni.h.
//
// An Arena is one stream with its own pipeline buffers.  The context owns a
// main arena for the synchronous calls and, once tiles are submitted, a few
// slot arenas used by the serving scheduler so that batches overlap.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <map>
#include <memory>
#include <mutex>
#include <new>
#include <string>
#include <thread>
#include <unordered_map>
#include <vector>

#include "tnni_internal.h"

using namespace tnni;

// ---- scratch of the allocation-free entry point -----------------------------
// tnni_tile_run_dev_noalloc keeps everything besides its caller's inputs and
// outputs (descriptors, band histograms and scan partials, the banded and
// sorted site records, bucket offsets, distance maps) in this array: one
// copy of kNoAllocScratchBytes per device, held until the process ends.  In
// host code `scratch` is only a host-side stand-in for the symbol: the
// device address of each device's copy comes from hipGetSymbolAddress at
// context creation (noalloc_setup) and only that address is used.
constexpr size_t kNoAllocScratchBytes = (size_t)128 << 20;
static __device__ __align__(256) uint8_t scratch[kNoAllocScratchBytes];

namespace {

// Per device: the address of its copy of `scratch`, found once with the
// device current; mu also runs that device's allocation-free calls one at a
// time (they share the array).
constexpr int kNoAllocDevices = 64;
struct NoAllocScratch {
    std::mutex mu;
    bool tried = false;
    hipError_t err = hipSuccess;
    void *ptr = nullptr;
};
NoAllocScratch g_noalloc[kNoAllocDevices];

// Called by tnni_ctx_create with `device` current.
void noalloc_setup(int device)
{
    if (device < 0 || device >= kNoAllocDevices) return;
    NoAllocScratch &ns = g_noalloc[device];
    std::lock_guard<std::mutex> lk(ns.mu);
    if (ns.tried) return;
    ns.tried = true;
    void *ptr = nullptr;
    ns.err = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
    if (ns.err == hipSuccess) ns.ptr = ptr;
}

thread_local std::string g_thread_err;
// The scheduler thread reports errors through its requests only: it never
// writes the context's message, which caller threads read and write.
thread_local bool g_scheduler_thread = false;
// tnni_last_error(ctx) hands out a copy of the context's message, owned by
// the calling thread (valid until its next tnni_last_error call).
thread_local std::string g_err_copy;

double now_s()
{
    using namespace std::chrono;
    return duration<double>(steady_clock::now().time_since_epoch()).count();
}

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

struct HostBuf {
    void *p = nullptr;
    size_t cap = 0;
};

// Everything a captured pipeline graph depends on.
struct GraphKey {
    const void *ptrs[10];
    size_t lds;
    int alg, mode, main_grid, gather_grid, label_only, dense;
    size_t gather_lds;
    bool operator==(const GraphKey &o) const
    {
        return std::memcmp(ptrs, o.ptrs, sizeof ptrs) == 0 && lds == o.lds && alg == o.alg &&
               mode == o.mode && main_grid == o.main_grid && gather_grid == o.gather_grid &&
               gather_lds == o.gather_lds && label_only == o.label_only && dense == o.dense;
    }
};

struct Arena {
    hipStream_t stream = nullptr;
    hipEvent_t done = nullptr;
    DevBuf tiles, chunks, batch, hist, partials, banded, sorted, offs, hdist, edist;
    DevBuf sx, sy, sz, sh, sl, sd;  // device staging of host-pointer calls
    DevBuf st, acc;                 // space-time: point times, per-cell sums
    DevBuf cv, cd, flag;            // space-time: a slice combined from layers, sortedness flag
    DevBuf lv, ld;                  // split form: site-value and squared-distance rasters
    HostBuf h_desc;                 // pinned descriptors
    HostBuf h_in, h_out;            // pinned staging of the served form
    hipGraphExec_t gexec = nullptr;
    GraphKey gkey{};
    uint64_t graph_launches = 0;
};

// Fixed launch grids of graph-captured pipelines: the kernels read every
// size from device memory and stride over their work.
constexpr int kGraphSortGrid = 2048;
constexpr int kGraphBandGrid = 2048;
constexpr int kGraphRowGrid = 2048;
constexpr int kGraphGatherGrid = 2048;

struct Request {
    uint64_t id = 0;
    bool dev = false;  // device-resident inputs and outputs (tnni_submit_dev)
    const int32_t *X = nullptr, *Y = nullptr, *Z = nullptr;
    uint64_t n = 0;
    float *out_h = nullptr;
    int32_t *out_labels = nullptr;
    uint32_t *out_d2 = nullptr;
    TileDev t{};
    PrepInfo info;
    uint64_t footprint = 0;
    double t_submit = 0, t_done = 0;
    std::string err;  // why it failed
    tnni_status status = TNNI_PENDING;
    // host requests: pinned copies of the inputs (X|Y|Z) and of the outputs
    // (h|labels|d2), filled by tnni_submit and drained by tnni_wait
    void *pin_in = nullptr, *pin_out = nullptr;
    size_t pin_in_cap = 0, pin_out_cap = 0;
    // placement inside its batch's device staging
    size_t in_off = 0, out_h_off = 0, out_l_off = 0, out_d_off = 0;
};

// Pinned host blocks reused across requests (pinning is slow).
struct PinnedPool {
    std::mutex mu;
    std::multimap<size_t, void *> free;
    size_t held = 0;
    static constexpr size_t kMaxHeld = (size_t)8 << 30;

    void *acquire(size_t bytes, size_t &cap)
    {
        {
            std::lock_guard<std::mutex> lk(mu);
            auto it = free.lower_bound(bytes);
            if (it != free.end() && it->first <= 2 * bytes + (1 << 20)) {
                void *p = it->second;
                cap = it->first;
                held -= cap;
                free.erase(it);
                return p;
            }
        }
        size_t want = 1 << 16;
        while (want < bytes) want <<= 1;
        void *p = nullptr;
        if (hipHostMalloc(&p, want, hipHostMallocDefault) != hipSuccess) return nullptr;
        cap = want;
        return p;
    }
    void release(void *p, size_t cap)
    {
        if (!p) return;
        std::lock_guard<std::mutex> lk(mu);
        free.emplace(cap, p);
        held += cap;
        while (held > kMaxHeld && !free.empty()) {
            auto it = std::prev(free.end());
            (void)hipHostFree(it->second);
            held -= it->first;
            free.erase(it);
        }
    }
    void clear()
    {
        std::lock_guard<std::mutex> lk(mu);
        for (auto &kv : free) (void)hipHostFree(kv.second);
        free.clear();
        held = 0;
    }
};

struct Slot {
    Arena a;
    bool busy = false;
    std::vector<std::shared_ptr<Request>> reqs;
    uint64_t bytes = 0;
};

}  // namespace

// Space-time cache (what the Delta render-buffer cache saves, nni.cpp:
// 281-379, 902-921): per point time ("year"), the 2-D Voronoi of the sites
// of that time, as the site value and squared plane distance of each of the
// tile's W x H pixels.  A slice's nearest site is the least (d2 + dt2) over
// the layers of the times within its radius, dt2 being constant per layer
// and slice, so a slice is a combine of layers plus the gather, and a layer
// depends on the tile and its time only: every slice of every later call
// that reaches that time reuses it.  Values are kept, not point indices: a
// later call's points may sit at other indices (a window of a time-sorted
// stream).  Ties within a time go to the earlier point, which the window
// keeps in order; ties across times go to the earlier time, which for
// points sorted by time is the lower index (calls whose points are not
// sorted by time run without the cache).
struct tnni_time_cache {
    tnni_ctx *ctx = nullptr;  // null once the context is destroyed
    bool keyed = false;       // desc names the cached tile
    tnni_tile_desc desc{};
    int max_slots = 0;  // 0: the layers of one call
    struct Slot {
        bool full = false;
        uint32_t time_bits = 0;  // the layer's time as float bits
        uint64_t used = 0;       // last call that used it, for least-recently-used reuse
        DevBuf val, d2;
    };
    std::vector<Slot> slots;
    uint64_t clock = 0, hits = 0, misses = 0;
};

struct tnni_ctx {
    int device = 0;
    int n_cu = 0;
    std::mutex mu;  // the synchronous calls and the main arena
    mutable std::mutex err_mu;  // err: written and read by any caller thread
    std::string err;
    uint64_t budget = 0;
    std::mutex mem_mu;
    uint64_t allocated = 0;
    Arena main;

    bool prof = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t prof_launches = 0;
    double prof_ms = 0.0;

    std::vector<void *> user_allocs;
    std::vector<tnni_time_cache *> caches;  // space-time caches on this context
    tnni_time_cache *scratch_cache = nullptr;  // the layers of uncached space-time calls

    // measured bucket / cell-tile sizes per shape class (tnni_autotune)
    TuneTable tuned;
    hipEvent_t tev0 = nullptr, tev1 = nullptr;

    // served form
    int n_slots = 2;
    int max_batch_tiles = 256;
    int batch_window_us = 0;
    std::mutex smu;
    std::condition_variable scv;   // scheduler wake-up
    std::condition_variable dcv;   // request completion
    std::deque<std::shared_ptr<Request>> pending;
    std::unordered_map<uint64_t, std::shared_ptr<Request>> reqs;
    std::vector<std::unique_ptr<Slot>> slots;
    std::thread sched;
    bool sched_started = false;
    bool shutdown = false;
    uint64_t next_id = 1;
    uint64_t serve_budget = 0;
    uint64_t inflight_bytes = 0;
    // a launch found the HBM budget taken by batches in flight and cached
    // buffers: admit nothing until a batch retires
    bool budget_blocked = false;
    bool budget_full_trim = false;  // the last retry ran with nothing in flight
    tnni_serve_stats stats{};
    PinnedPool pinned;
};

namespace {

void set_ctx_error(tnni_ctx *ctx, const std::string &msg)
{
    std::lock_guard<std::mutex> lk(ctx->err_mu);
    ctx->err = msg;
}

tnni_status fail(tnni_ctx *ctx, tnni_status st, const std::string &msg)
{
    g_thread_err = msg;
    if (ctx && !g_scheduler_thread) set_ctx_error(ctx, msg);
    return st;
}

tnni_status hip_fail(tnni_ctx *ctx, hipError_t e, const char *what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(ctx, e == hipErrorOutOfMemory ? TNNI_E_NOMEM : TNNI_E_HIP, m);
}

#define HIP_TRY(ctx, expr)                                     \
    do {                                                       \
        hipError_t e_ = (expr);                                \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)

tnni_status ensure(tnni_ctx *ctx, DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (b.cap >= bytes) return TNNI_OK;
    size_t want = bytes + bytes / 4;
    std::lock_guard<std::mutex> lk(ctx->mem_mu);
    if (b.p) {
        hipError_t e = hipFree(b.p);
        ctx->allocated -= b.cap;
        b.p = nullptr;
        b.cap = 0;
        if (e != hipSuccess) return hip_fail(ctx, e, "hipFree");
    }
    if (ctx->allocated + want > ctx->budget) {
        want = bytes;
        if (ctx->allocated + want > ctx->budget)
            return fail(ctx, TNNI_E_BUDGET, "request exceeds the context's HBM budget");
    }
    HIP_TRY(ctx, hipMalloc(&b.p, want));
    b.cap = want;
    ctx->allocated += want;
    return TNNI_OK;
}

tnni_status ensure_host(tnni_ctx *ctx, HostBuf &b, size_t bytes)
{
    if (b.cap >= bytes && b.p) return TNNI_OK;
    if (b.p) (void)hipHostFree(b.p);
    b.p = nullptr;
    b.cap = 0;
    const size_t want = std::max<size_t>(bytes + bytes / 4, 4096);
    HIP_TRY(ctx, hipHostMalloc(&b.p, want, hipHostMallocDefault));
    b.cap = want;
    return TNNI_OK;
}

void release(tnni_ctx *ctx, DevBuf &b)
{
    if (b.p) {
        (void)hipFree(b.p);
        std::lock_guard<std::mutex> lk(ctx->mem_mu);
        ctx->allocated -= b.cap;
    }
    b.p = nullptr;
    b.cap = 0;
}

tnni_status arena_init(tnni_ctx *ctx, Arena &a)
{
    HIP_TRY(ctx, hipStreamCreateWithFlags(&a.stream, hipStreamNonBlocking));
    HIP_TRY(ctx, hipEventCreateWithFlags(&a.done, hipEventDisableTiming));
    return TNNI_OK;
}

void arena_free(tnni_ctx *ctx, Arena &a)
{
    if (a.stream) (void)hipStreamSynchronize(a.stream);
    DevBuf *bufs[] = {&a.tiles, &a.chunks, &a.batch, &a.hist, &a.partials, &a.banded, &a.sorted,
                      &a.offs, &a.hdist, &a.edist, &a.sx, &a.sy, &a.sz, &a.sh, &a.sl, &a.sd,
                      &a.st, &a.acc, &a.cv, &a.cd, &a.flag, &a.lv, &a.ld};
    for (DevBuf *b : bufs) release(ctx, *b);
    HostBuf *hb[] = {&a.h_desc, &a.h_in, &a.h_out};
    for (HostBuf *b : hb) {
        if (b->p) (void)hipHostFree(b->p);
        b->p = nullptr;
        b->cap = 0;
    }
    if (a.gexec) (void)hipGraphExecDestroy(a.gexec);
    a.gexec = nullptr;
    if (a.done) (void)hipEventDestroy(a.done);
    if (a.stream) (void)hipStreamDestroy(a.stream);
    a.done = nullptr;
    a.stream = nullptr;
}

// One prepared tile of a batch, before the batch layout is assigned.
struct Prepared {
    TileDev t;
    PrepInfo info;
};

enum class Launch { Direct, Graph };

// Lay the batch out (bases, chunks, items), upload the descriptors and
// enqueue the pipeline on the arena's stream.  Tiles keep their order; the
// caller orders them.  sync: wait for the stream before returning.
tnni_status run_prepared(tnni_ctx *ctx, Arena &a, std::vector<Prepared> &tiles, bool sync,
                         Launch how, hipEvent_t ev_begin = nullptr, hipEvent_t ev_end = nullptr)
{
    if (tiles.empty()) return TNNI_OK;
    const int alg = tiles[0].t.algorithm, mode = tiles[0].t.mode;
    const bool time = tiles[0].t.T != nullptr;
    if (time && how == Launch::Graph) return fail(ctx, TNNI_E_INVALID, "space-time tiles run direct");
    int64_t pts = 0, bks = 0, ebytes = 0, items = 0, rows = 0, bands = 0, hents = 0;
    int64_t gather_items = 0, split_pix = 0, split_d2_pix = 0;  // the split tiles' gather blocks and rasters
    size_t gather_lds = 0;
    size_t lds = 0;
    int wg_cap = 0;       // tuned workgroups per CU of the tile kernel
    bool wg_any = false;  // some tile has no tuned cap
    bool any_label_only = false;
    std::vector<Chunk> chunks;
    int64_t all_pts = 0;
    for (const Prepared &pr : tiles) all_pts += pr.t.n;
    const int cpts = chunk_points(all_pts, ctx->n_cu);
    for (size_t k = 0; k < tiles.size(); ++k) {
        TileDev &t = tiles[k].t;
        if (t.algorithm != alg || t.mode != mode || (t.T != nullptr) != time)
            return fail(ctx, TNNI_E_INVALID, "a batch must share algorithm, accum_mode and time mode");
        t.pt_base = pts;
        t.bk_base = bks;
        t.nbx4 = (t.nbx + 3) & ~3;
        t.e_base = ebytes;
        ebytes += (int64_t)t.nbx4 * t.nby;
        t.item_base = items;
        t.row_base = rows;
        t.band_base = bands;
        t.hist_base = hents;
        t.chunk_base = (int)chunks.size();
        // a split tile with heights wanted: its cells are gathered from its
        // value and squared-distance rasters (scratch unless the caller
        // takes the squared distances); other tiles get an empty range
        t.gitem_base = gather_items;
        any_label_only = any_label_only || t.label_only;
        if (t.label_only && t.out_h) {
            gather_items += (int64_t)t.gntx * t.gnty;
            gather_lds = std::max(gather_lds, split_pass_lds(t));
            split_pix += (int64_t)t.W * t.H;
            if (!t.out_d2) split_d2_pix += (int64_t)t.W * t.H;
        }
        // every tile has at least one (possibly empty) chunk
        for (int64_t b = 0; b < std::max<int64_t>(t.n, 1); b += cpts) {
            Chunk c;
            c.tile = (int)k;
            c.pad = 0;
            c.begin = b;
            c.end = std::min<int64_t>(t.n, b + cpts);
            chunks.push_back(c);
        }
        t.nchunks = (int)chunks.size() - t.chunk_base;
        pts += t.n;
        bks += (int64_t)t.nbx * t.nby;
        rows += t.nby;
        bands += t.nbands;
        hents += (int64_t)t.nbands * t.nchunks;
        items += (int64_t)t.ntx * t.nty;
        lds = std::max(lds, tiles[k].info.lds);
        wg_any = wg_any || tiles[k].info.wg_per_cu == 0;
        wg_cap = std::max(wg_cap, tiles[k].info.wg_per_cu);
    }
    if (bks + 1 > 0xFFFFFFFFll || pts > 0xFFFFFFFFll || hents + 1 > 0xFFFFFFFFll)
        return fail(ctx, TNNI_E_INVALID, "batch exceeds 2^32 buckets or points");
    tnni_status st;
    if ((st = ensure(ctx, a.tiles, tiles.size() * sizeof(TileDev))) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.chunks, chunks.size() * sizeof(Chunk))) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.batch, sizeof(BatchDev))) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.hist, (size_t)(hents + 1) * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.partials, kScanBlocks * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.banded, (size_t)pts * 16)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sorted, (size_t)pts * 16)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.offs, (size_t)(bks + 1) * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.hdist, (size_t)ebytes)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.edist, (size_t)ebytes)) != TNNI_OK) return st;
    if (gather_items) {
        if ((st = ensure(ctx, a.lv, (size_t)split_pix * 4)) != TNNI_OK) return st;
        if (split_d2_pix && (st = ensure(ctx, a.ld, (size_t)split_d2_pix * 4)) != TNNI_OK) return st;
        int64_t pv = 0, pd = 0;
        for (Prepared &p : tiles) {
            TileDev &t = p.t;
            if (!(t.label_only && t.out_h)) continue;
            t.out_val = (int32_t *)a.lv.p + pv;
            pv += (int64_t)t.W * t.H;
            if (!t.out_d2) {
                t.out_d2 = (uint32_t *)a.ld.p + pd;
                pd += (int64_t)t.W * t.H;
            }
        }
    }

    BatchDev bd;
    std::memset(&bd, 0, sizeof bd);
    bd.n_tiles = (int)tiles.size();
    bd.n_chunks = (int)chunks.size();
    bd.chunk_pts = (uint32_t)cpts;
    bd.n_scan = (uint32_t)(hents + 1);
    bd.n_items = items;
    bd.n_rows = rows;
    bd.n_bands = bands;
    bd.scan_chunk = (uint32_t)((bd.n_scan + kScanBlocks - 1) / kScanBlocks);
    bd.n_gather_items = gather_items;

    // descriptors through one pinned staging block; it may still feed the
    // copies of the previous batch on this stream
    const size_t sz_t = tiles.size() * sizeof(TileDev);
    const size_t sz_c = chunks.size() * sizeof(Chunk);
    HIP_TRY(ctx, hipStreamSynchronize(a.stream));
    if ((st = ensure_host(ctx, a.h_desc, sz_t + sz_c + sizeof(BatchDev))) != TNNI_OK) return st;
    char *h = (char *)a.h_desc.p;
    for (size_t k = 0; k < tiles.size(); ++k) std::memcpy(h + k * sizeof(TileDev), &tiles[k].t, sizeof(TileDev));
    std::memcpy(h + sz_t, chunks.data(), sz_c);
    std::memcpy(h + sz_t + sz_c, &bd, sizeof bd);
    if (ev_begin) HIP_TRY(ctx, hipEventRecord(ev_begin, a.stream));
    HIP_TRY(ctx, hipMemcpyAsync(a.tiles.p, h, sz_t, hipMemcpyHostToDevice, a.stream));
    HIP_TRY(ctx, hipMemcpyAsync(a.chunks.p, h + sz_t, sz_c, hipMemcpyHostToDevice, a.stream));
    HIP_TRY(ctx, hipMemcpyAsync(a.batch.p, h + sz_t + sz_c, sizeof bd, hipMemcpyHostToDevice, a.stream));

    LaunchCfg cfg;
    int occ = tile_kernel_occupancy(lds);
    if (!wg_any && wg_cap > 0) occ = std::min(occ, wg_cap);
    cfg.main_lds = lds;
    cfg.algorithm = alg;
    cfg.mode = mode;
    cfg.time = time ? 1 : 0;
    cfg.label_only = any_label_only ? 1 : 0;
    cfg.dense = 0;
    for (const Prepared &pr : tiles) cfg.dense |= pr.t.dense ? 1 : 0;
    DevBuffers db;
    db.batch = (BatchDev *)a.batch.p;
    db.tiles = (const TileDev *)a.tiles.p;
    db.chunks = (const Chunk *)a.chunks.p;
    db.hist = (uint32_t *)a.hist.p;
    db.partials = (uint32_t *)a.partials.p;
    db.banded = (int4 *)a.banded.p;
    db.sorted = (int4 *)a.sorted.p;
    db.offs = (uint32_t *)a.offs.p;
    db.hdist = (uint8_t *)a.hdist.p;
    db.edist = (uint8_t *)a.edist.p;
    const bool prof = ctx->prof && how == Launch::Direct;
    if (how == Launch::Direct) {
        cfg.sort_grid = (int)std::min<size_t>(4096, chunks.size());
        cfg.band_grid = (int)std::max<int64_t>(1, std::min<int64_t>(8192, bands));
        cfg.row_grid = (int)std::max<int64_t>(1, std::min<int64_t>(8192, rows));
        cfg.main_grid = (int)std::max<int64_t>(1, std::min<int64_t>(items, (int64_t)ctx->n_cu * occ));
        cfg.gather_grid = (int)std::min<int64_t>(gather_items, 8192);
        cfg.gather_lds = gather_lds;
        HIP_TRY(ctx, enqueue_pipeline(db, cfg, a.stream, prof ? ctx->ev0 : nullptr,
                                      prof ? ctx->ev1 : nullptr));
    } else {
        cfg.sort_grid = kGraphSortGrid;
        cfg.band_grid = kGraphBandGrid;
        cfg.row_grid = kGraphRowGrid;
        cfg.main_grid = ctx->n_cu * occ;
        cfg.gather_grid = gather_items ? kGraphGatherGrid : 0;
        cfg.gather_lds = gather_lds;
        GraphKey key;
        const void *p[10] = {a.batch.p, a.tiles.p, a.chunks.p, a.hist.p, a.partials.p,
                             a.banded.p, a.sorted.p, a.offs.p, a.hdist.p, a.edist.p};
        std::memcpy(key.ptrs, p, sizeof p);
        key.lds = lds;
        key.alg = alg;
        key.mode = mode;
        key.main_grid = cfg.main_grid;
        key.gather_grid = cfg.gather_grid;
        key.gather_lds = cfg.gather_lds;
        key.label_only = cfg.label_only;
        key.dense = cfg.dense;
        if (!a.gexec || !(key == a.gkey)) {
            if (a.gexec) HIP_TRY(ctx, hipGraphExecDestroy(a.gexec));
            a.gexec = nullptr;
            hipGraph_t g = nullptr;
            HIP_TRY(ctx, hipStreamBeginCapture(a.stream, hipStreamCaptureModeThreadLocal));
            hipError_t ce = enqueue_pipeline(db, cfg, a.stream, nullptr, nullptr);
            hipError_t ee = hipStreamEndCapture(a.stream, &g);
            if (ce != hipSuccess) {
                if (g) (void)hipGraphDestroy(g);
                return hip_fail(ctx, ce, "pipeline capture");
            }
            if (ee != hipSuccess) return hip_fail(ctx, ee, "hipStreamEndCapture");
            hipError_t ie = hipGraphInstantiate(&a.gexec, g, nullptr, nullptr, 0);
            (void)hipGraphDestroy(g);
            if (ie != hipSuccess) {
                a.gexec = nullptr;
                return hip_fail(ctx, ie, "hipGraphInstantiate");
            }
            a.gkey = key;
        }
        HIP_TRY(ctx, hipGraphLaunch(a.gexec, a.stream));
        a.graph_launches++;
    }
    if (ev_end) HIP_TRY(ctx, hipEventRecord(ev_end, a.stream));
    if (sync || prof) HIP_TRY(ctx, hipStreamSynchronize(a.stream));
    if (prof) {
        float ms = 0.f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->prof_ms += ms;
        ctx->prof_launches++;
    }
    return TNNI_OK;
}

// Heaviest tiles first: the persistent kernel drains big tiles early and
// fills the tail with small ones.
double tile_cost(const TileDev &t) { return (double)t.W * t.H + 8.0 * (double)t.n; }

// ---- served form ---------------------------------------------------------

uint64_t request_bytes(const TileDev &t, bool labels, bool d2)
{
    const uint64_t cells = (uint64_t)t.nrows * t.ncols, pix = (uint64_t)t.W * t.H;
    // points: staging, band order, bucket order; buckets: offsets and the
    // two distance maps (rows padded to 4 bytes); cells: heights
    const uint64_t nbx4 = ((uint64_t)t.nbx + 3) & ~3ull;
    uint64_t b = (uint64_t)t.n * (12 + 16 + 16) + ((uint64_t)t.nbx * t.nby + 1) * 4 +
                 nbx4 * t.nby * 2 + cells * 4;
    if (labels) b += pix * 4;
    if (d2) b += pix * 4;
    if (t.label_only) b += pix * 8;  // split form: value and squared-distance rasters
    const uint64_t chunks = (uint64_t)t.n / kChunkMin + 1;  // chunks are at least kChunkMin points
    // the band histograms: one entry per (band, chunk)
    return b + sizeof(TileDev) + chunks * sizeof(Chunk) + ((uint64_t)t.nbands * chunks + 1) * 4;
}

void finish_request(tnni_ctx *ctx, const std::shared_ptr<Request> &r, tnni_status st)
{
    r->status = st;
    r->t_done = now_s();
    ctx->stats.completed++;
}

// Drop the cached device buffers of an idle arena (its graph goes with them).
void arena_trim(tnni_ctx *ctx, Arena &a)
{
    DevBuf *bufs[] = {&a.tiles, &a.chunks, &a.batch, &a.hist, &a.partials, &a.banded, &a.sorted,
                      &a.offs, &a.hdist, &a.edist, &a.sx, &a.sy, &a.sz, &a.sh, &a.sl, &a.sd,
                      &a.st, &a.acc, &a.cv, &a.cd, &a.flag, &a.lv, &a.ld};
    for (DevBuf *b : bufs) release(ctx, *b);
    if (a.gexec) (void)hipGraphExecDestroy(a.gexec);
    a.gexec = nullptr;
}

// Launch one admitted batch on slot s (scheduler thread, smu not held; only
// this thread changes the slots' arenas and busy flags).
tnni_status launch_batch(tnni_ctx *ctx, Slot &s)
{
    Arena &a = s.a;
    uint64_t allocated;
    {
        std::lock_guard<std::mutex> lk(ctx->mem_mu);
        allocated = ctx->allocated;
    }
    if (allocated + s.bytes + s.bytes / 4 > ctx->budget)
        for (auto &q : ctx->slots)
            if (q.get() != &s && !q->busy) arena_trim(ctx, q->a);
    size_t n_in = 0, n_h = 0, n_l = 0, n_d = 0;
    for (auto &r : s.reqs) {
        if (r->dev) continue;
        r->in_off = n_in;
        n_in += r->n;
        r->out_h_off = n_h;
        n_h += (size_t)r->t.nrows * r->t.ncols;
        r->out_l_off = n_l;
        if (r->out_labels) n_l += (size_t)r->t.W * r->t.H;
        r->out_d_off = n_d;
        if (r->out_d2) n_d += (size_t)r->t.W * r->t.H;
    }
    tnni_status st;
    if ((st = ensure(ctx, a.sx, n_in * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sy, n_in * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sz, n_in * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sh, n_h * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sl, n_l * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sd, n_d * 4)) != TNNI_OK) return st;
    std::vector<Prepared> prep(s.reqs.size());
    for (size_t k = 0; k < s.reqs.size(); ++k) {
        Request &r = *s.reqs[k];
        prep[k].t = r.t;
        prep[k].info = r.info;
        if (r.dev) {
            prep[k].t.X = r.X;
            prep[k].t.Y = r.Y;
            prep[k].t.Z = r.Z;
            prep[k].t.out_h = r.out_h;
            prep[k].t.out_labels = r.out_labels;
            prep[k].t.out_d2 = r.out_d2;
            continue;
        }
        prep[k].t.X = (const int32_t *)a.sx.p + r.in_off;
        prep[k].t.Y = (const int32_t *)a.sy.p + r.in_off;
        prep[k].t.Z = (const int32_t *)a.sz.p + r.in_off;
        prep[k].t.out_h = (float *)a.sh.p + r.out_h_off;
        prep[k].t.out_labels = r.out_labels ? (int32_t *)a.sl.p + r.out_l_off : nullptr;
        prep[k].t.out_d2 = r.out_d2 ? (uint32_t *)a.sd.p + r.out_d_off : nullptr;
        if (r.n) {
            const int32_t *pin = (const int32_t *)r.pin_in;
            HIP_TRY(ctx, hipMemcpyAsync((int32_t *)a.sx.p + r.in_off, pin, r.n * 4, hipMemcpyHostToDevice, a.stream));
            HIP_TRY(ctx, hipMemcpyAsync((int32_t *)a.sy.p + r.in_off, pin + r.n, r.n * 4, hipMemcpyHostToDevice, a.stream));
            HIP_TRY(ctx, hipMemcpyAsync((int32_t *)a.sz.p + r.in_off, pin + 2 * r.n, r.n * 4, hipMemcpyHostToDevice, a.stream));
        }
    }
    if ((st = run_prepared(ctx, a, prep, false, Launch::Graph)) != TNNI_OK) return st;
    for (auto &rp : s.reqs) {
        Request &r = *rp;
        if (r.dev) continue;
        const size_t cells = (size_t)r.t.nrows * r.t.ncols, pix = (size_t)r.t.W * r.t.H;
        char *po = (char *)r.pin_out;
        HIP_TRY(ctx, hipMemcpyAsync(po, (float *)a.sh.p + r.out_h_off, cells * 4, hipMemcpyDeviceToHost, a.stream));
        po += cells * 4;
        if (r.out_labels) {
            HIP_TRY(ctx, hipMemcpyAsync(po, (int32_t *)a.sl.p + r.out_l_off, pix * 4, hipMemcpyDeviceToHost, a.stream));
            po += pix * 4;
        }
        if (r.out_d2)
            HIP_TRY(ctx, hipMemcpyAsync(po, (uint32_t *)a.sd.p + r.out_d_off, pix * 4, hipMemcpyDeviceToHost, a.stream));
    }
    HIP_TRY(ctx, hipEventRecord(a.done, a.stream));
    return TNNI_OK;
}

void scheduler(tnni_ctx *ctx)
{
    g_scheduler_thread = true;
    (void)hipSetDevice(ctx->device);
    std::unique_lock<std::mutex> lk(ctx->smu);
    for (;;) {
        // 1. retire finished batches
        bool any_busy = false;
        for (auto &sp : ctx->slots) {
            Slot &s = *sp;
            if (!s.busy) continue;
            const hipError_t q = hipEventQuery(s.a.done);
            if (q == hipErrorNotReady) {
                any_busy = true;
                continue;
            }
            for (auto &r : s.reqs) finish_request(ctx, r, q == hipSuccess ? TNNI_OK : TNNI_E_HIP);
            ctx->inflight_bytes -= s.bytes;
            s.reqs.clear();
            s.bytes = 0;
            s.busy = false;
            ctx->budget_blocked = false;
            ctx->dcv.notify_all();
        }
        if (ctx->shutdown && !any_busy) break;
        // 2. admit queued tiles into free slots: oldest first, one
        //    algorithm/mode per batch, within the HBM budget
        bool launched = false;
        for (auto &sp : ctx->slots) {
            Slot &s = *sp;
            if (s.busy || ctx->pending.empty() || ctx->shutdown || ctx->budget_blocked) continue;
            if (ctx->batch_window_us > 0 && !any_busy &&
                (int)ctx->pending.size() < ctx->max_batch_tiles &&
                now_s() - ctx->pending.front()->t_submit < ctx->batch_window_us * 1e-6)
                break;
            const int alg = ctx->pending.front()->t.algorithm, mode = ctx->pending.front()->t.mode;
            while (!ctx->pending.empty() && (int)s.reqs.size() < ctx->max_batch_tiles) {
                auto &r = ctx->pending.front();
                if (r->t.algorithm != alg || r->t.mode != mode) break;
                const bool over = ctx->inflight_bytes + s.bytes + r->footprint > ctx->serve_budget;
                // over budget: wait until running batches free memory
                if (over && (!s.reqs.empty() || any_busy)) break;
                s.bytes += r->footprint;
                s.reqs.push_back(r);
                ctx->pending.pop_front();
            }
            if (s.reqs.empty()) break;
            std::stable_sort(s.reqs.begin(), s.reqs.end(),
                             [](const std::shared_ptr<Request> &x, const std::shared_ptr<Request> &y) {
                                 return tile_cost(x->t) > tile_cost(y->t);
                             });
            ctx->inflight_bytes += s.bytes;
            s.busy = true;
            lk.unlock();
            const tnni_status st = launch_batch(ctx, s);
            lk.lock();
            bool others_busy = false;
            for (auto &q : ctx->slots)
                if (q.get() != &s && q->busy) others_busy = true;
            if (st == TNNI_E_BUDGET && (others_busy || !ctx->budget_full_trim)) {
                // The budget is held by batches in flight or by buffers cached
                // from earlier, larger batches: drop this slot's and the idle
                // slots' caches, put the requests back at the head of the
                // queue, and admit again once a batch retires (or at once if
                // none is in flight: everything is trimmed now).
                (void)hipStreamSynchronize(s.a.stream);
                arena_trim(ctx, s.a);
                for (auto &q : ctx->slots)
                    if (q.get() != &s && !q->busy) arena_trim(ctx, q->a);
                for (auto it = s.reqs.rbegin(); it != s.reqs.rend(); ++it) ctx->pending.push_front(*it);
                ctx->inflight_bytes -= s.bytes;
                s.reqs.clear();
                s.bytes = 0;
                s.busy = false;
                ctx->budget_full_trim = !others_busy;
                ctx->budget_blocked = others_busy;
                ctx->stats.budget_retries++;
                continue;
            }
            if (st == TNNI_OK) ctx->budget_full_trim = false;
            if (st != TNNI_OK) {
                // copies and the graph may already be on the slot's stream:
                // let them end before the requests' pinned buffers go back
                // to the pool (tnni_wait releases them)
                lk.unlock();
                (void)hipStreamSynchronize(s.a.stream);
                lk.lock();
                for (auto &r : s.reqs) {
                    r->err = g_thread_err;
                    finish_request(ctx, r, st);
                }
                ctx->inflight_bytes -= s.bytes;
                s.reqs.clear();
                s.bytes = 0;
                s.busy = false;
                ctx->dcv.notify_all();
                continue;
            }
            ctx->stats.batches++;
            uint64_t tiles = 0, batches = 0;
            for (auto &q : ctx->slots)
                if (q->busy) {
                    tiles += q->reqs.size();
                    batches++;
                }
            ctx->stats.max_inflight_tiles = std::max(ctx->stats.max_inflight_tiles, tiles);
            ctx->stats.max_inflight_batches = std::max(ctx->stats.max_inflight_batches, batches);
            launched = true;
            any_busy = true;
        }
        if (launched) continue;
        if (any_busy) {
            // batches in flight: poll their events, yielding the core in
            // between (a sleep would add its granularity, tens of
            // microseconds, to every request's latency)
            lk.unlock();
            std::this_thread::yield();
            lk.lock();
        } else if (!ctx->pending.empty()) {
            // a batch window open
            lk.unlock();
            std::this_thread::sleep_for(std::chrono::microseconds(20));
            lk.lock();
        } else {
            ctx->scv.wait(lk, [ctx] { return ctx->shutdown || !ctx->pending.empty(); });
        }
    }
    // shutdown: fail what never ran
    for (auto &r : ctx->pending) finish_request(ctx, r, TNNI_E_SHUTDOWN);
    ctx->pending.clear();
    ctx->dcv.notify_all();
}

tnni_status start_scheduler(tnni_ctx *ctx)
{
    if (ctx->sched_started) return TNNI_OK;
    for (int k = 0; k < ctx->n_slots; ++k) {
        auto s = std::make_unique<Slot>();
        tnni_status st = arena_init(ctx, s->a);
        ctx->slots.push_back(std::move(s));
        if (st != TNNI_OK) return st;
    }
    // admitted footprints stay below 3/4 of the budget: arenas grow with
    // 25 % headroom and the descriptors and graphs need a little more
    ctx->serve_budget = ctx->budget / 4 * 3;
    ctx->sched = std::thread(scheduler, ctx);
    ctx->sched_started = true;
    return TNNI_OK;
}

}  // namespace

// ---- C ABI -----------------------------------------------------------------

extern "C" {

int tnni_abi_version(void) { return TNNI_ABI_VERSION; }

const char *tnni_last_error(const tnni_ctx *ctx)
{
    if (!ctx) return g_thread_err.c_str();
    std::lock_guard<std::mutex> lk(ctx->err_mu);
    g_err_copy = ctx->err;
    return g_err_copy.c_str();
}

tnni_status tnni_frame(const tnni_tile_desc *desc, tnni_frame_info *out)
{
    if (!desc || !out) return fail(nullptr, TNNI_E_INVALID, "null argument");
    std::string err;
    tnni_status st = compute_frame(*desc, *out, err);
    if (st != TNNI_OK) return fail(nullptr, st, err);
    return TNNI_OK;
}

tnni_status tnni_footprint(const tnni_tile_desc *desc, uint64_t n_points, int with_labels,
                           uint64_t *bytes)
{
    if (!desc || !bytes) return fail(nullptr, TNNI_E_INVALID, "null argument");
    TileDev t;
    std::string err;
    PrepInfo info;
    tnni_status st = prepare_tile(*desc, n_points, 256, true, nullptr, t, info, err);
    if (st != TNNI_OK) return fail(nullptr, st, err);
    *bytes = request_bytes(t, with_labels != 0, with_labels != 0);
    return TNNI_OK;
}

tnni_status tnni_ctx_create(const tnni_ctx_opts *opts, tnni_ctx **out)
{
    if (!out) return fail(nullptr, TNNI_E_INVALID, "null argument");
    *out = nullptr;
    tnni_ctx_opts o;
    std::memset(&o, 0, sizeof o);
    if (opts) o = *opts;
    int ndev = 0;
    if (hipGetDeviceCount(&ndev) != hipSuccess || ndev == 0)
        return fail(nullptr, TNNI_E_NODEVICE, "no HIP device visible");
    if (o.device < 0 || o.device >= ndev) return fail(nullptr, TNNI_E_INVALID, "device ordinal out of range");
    tnni_ctx *ctx = new tnni_ctx();
    ctx->device = o.device;
    ctx->n_slots = o.n_slots > 0 ? o.n_slots : 2;
    ctx->max_batch_tiles = o.max_batch_tiles > 0 ? o.max_batch_tiles : 256;
    ctx->batch_window_us = std::max(0, o.batch_window_us);
    hipError_t e;
    if ((e = hipSetDevice(o.device)) != hipSuccess) {
        delete ctx;
        return hip_fail(nullptr, e, "hipSetDevice");
    }
    hipDeviceProp_t prop;
    if ((e = hipGetDeviceProperties(&prop, o.device)) != hipSuccess) {
        delete ctx;
        return hip_fail(nullptr, e, "hipGetDeviceProperties");
    }
    ctx->n_cu = prop.multiProcessorCount;
    size_t fr = 0, tot = 0;
    if ((e = hipMemGetInfo(&fr, &tot)) != hipSuccess) {
        delete ctx;
        return hip_fail(nullptr, e, "hipMemGetInfo");
    }
    ctx->budget = o.hbm_budget ? o.hbm_budget : (uint64_t)(0.8 * (double)fr);
    noalloc_setup(o.device);
    tnni_status st = arena_init(ctx, ctx->main);
    if (st == TNNI_OK && ((e = hipEventCreate(&ctx->ev0)) != hipSuccess ||
                          (e = hipEventCreate(&ctx->ev1)) != hipSuccess ||
                          (e = hipEventCreate(&ctx->tev0)) != hipSuccess ||
                          (e = hipEventCreate(&ctx->tev1)) != hipSuccess))
        st = hip_fail(nullptr, e, "hipEventCreate");
    if (st != TNNI_OK) {
        tnni_ctx_destroy(ctx);
        return st;
    }
    *out = ctx;
    return TNNI_OK;
}

void tnni_ctx_destroy(tnni_ctx *ctx)
{
    if (!ctx) return;
    (void)hipSetDevice(ctx->device);
    if (ctx->sched_started) {
        {
            std::lock_guard<std::mutex> lk(ctx->smu);
            ctx->shutdown = true;
        }
        ctx->scv.notify_all();
        ctx->sched.join();
    }
    for (auto &s : ctx->slots) arena_free(ctx, s->a);
    arena_free(ctx, ctx->main);
    for (auto &kv : ctx->reqs) {
        ctx->pinned.release(kv.second->pin_in, kv.second->pin_in_cap);
        ctx->pinned.release(kv.second->pin_out, kv.second->pin_out_cap);
    }
    ctx->pinned.clear();
    for (void *p : ctx->user_allocs) (void)hipFree(p);
    for (tnni_time_cache *c : ctx->caches) {
        for (auto &s : c->slots) {
            release(ctx, s.val);
            release(ctx, s.d2);
        }
        c->slots.clear();
        c->keyed = false;
        c->ctx = nullptr;
    }
    delete ctx->scratch_cache;
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->tev0) (void)hipEventDestroy(ctx->tev0);
    if (ctx->tev1) (void)hipEventDestroy(ctx->tev1);
    delete ctx;
}

tnni_status tnni_dev_alloc(tnni_ctx *ctx, uint64_t bytes, void **dptr)
{
    if (!ctx || !dptr) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    void *p = nullptr;
    HIP_TRY(ctx, hipMalloc(&p, bytes ? bytes : 1));
    ctx->user_allocs.push_back(p);
    *dptr = p;
    return TNNI_OK;
}

tnni_status tnni_dev_free(tnni_ctx *ctx, void *dptr)
{
    if (!ctx) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    auto it = std::find(ctx->user_allocs.begin(), ctx->user_allocs.end(), dptr);
    if (it == ctx->user_allocs.end()) return fail(ctx, TNNI_E_INVALID, "pointer not from tnni_dev_alloc");
    ctx->user_allocs.erase(it);
    HIP_TRY(ctx, hipStreamSynchronize(ctx->main.stream));
    HIP_TRY(ctx, hipFree(dptr));
    return TNNI_OK;
}

tnni_status tnni_h2d(tnni_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->main.stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->main.stream));
    return TNNI_OK;
}

tnni_status tnni_d2h(tnni_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->main.stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->main.stream));
    return TNNI_OK;
}

tnni_status tnni_d2d(tnni_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToDevice, ctx->main.stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->main.stream));
    return TNNI_OK;
}

tnni_status tnni_tile_run_dev(tnni_ctx *ctx, const tnni_tile_desc *desc, const int32_t *dX,
                              const int32_t *dY, const int32_t *dZ, uint64_t n, float *d_out_h,
                              int32_t *d_out_labels, uint32_t *d_out_d2)
{
    if (!ctx || !desc || !d_out_h || (n && (!dX || !dY || !dZ)))
        return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::vector<Prepared> v(1);
    std::string err;
    tnni_status st = prepare_tile(*desc, n, ctx->n_cu, false, &ctx->tuned, v[0].t, v[0].info, err);
    if (st != TNNI_OK) return fail(ctx, st, err);
    v[0].t.X = dX;
    v[0].t.Y = dY;
    v[0].t.Z = dZ;
    v[0].t.out_h = d_out_h;
    v[0].t.out_labels = d_out_labels;
    v[0].t.out_d2 = d_out_d2;
    return run_prepared(ctx, ctx->main, v, true, Launch::Direct);
}

tnni_status tnni_tile_run_dev_noalloc(tnni_ctx *ctx, const tnni_tile_desc *desc, const int32_t *dX,
                                      const int32_t *dY, const int32_t *dZ, uint64_t n, float *d_out_h,
                                      int32_t *d_out_labels, uint32_t *d_out_d2)
{
    if (!ctx || !desc || !d_out_h || (n && (!dX || !dY || !dZ)))
        return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    if (ctx->device >= kNoAllocDevices) return fail(ctx, TNNI_E_INVALID, "device ordinal above 63");
    NoAllocScratch &ns = g_noalloc[ctx->device];
    std::lock_guard<std::mutex> lk_dev(ns.mu);
    if (!ns.ptr)
        return fail(ctx, TNNI_E_HIP, std::string("no scratch address for this device: hipGetSymbolAddress: ") +
                                         hipGetErrorString(ns.err));
    Prepared p;
    std::string err;
    tnni_status st = prepare_tile(*desc, n, ctx->n_cu, false, &ctx->tuned, p.t, p.info, err);
    if (st != TNNI_OK) return fail(ctx, st, err);
    TileDev &t = p.t;
    t.X = dX;
    t.Y = dY;
    t.Z = dZ;
    t.out_h = d_out_h;
    t.out_labels = d_out_labels;
    t.out_d2 = d_out_d2;
    // the layout of a batch of this one tile (run_prepared's, with every
    // base at 0): at least one chunk, a histogram entry per (band, chunk)
    const int cpts = chunk_points(t.n, ctx->n_cu);
    const int64_t nch = std::max<int64_t>(1, (t.n + cpts - 1) / cpts);
    const int64_t bks = (int64_t)t.nbx * t.nby, hents = (int64_t)t.nbands * nch;
    t.nbx4 = (t.nbx + 3) & ~3;
    const int64_t ebytes = (int64_t)t.nbx4 * t.nby;
    t.chunk_base = 0;
    t.nchunks = (int)nch;
    if (bks + 1 > 0xFFFFFFFFll || hents + 1 > 0xFFFFFFFFll)
        return fail(ctx, TNNI_E_INVALID, "tile exceeds 2^32 buckets");
    // carve the scratch: every buffer on a 256-byte boundary
    size_t off = 0;
    auto carve = [&off](size_t bytes) {
        const size_t at = off;
        off += (std::max<size_t>(bytes, 1) + 255) & ~(size_t)255;
        return at;
    };
    const size_t o_batch = carve(sizeof(BatchDev)), o_tiles = carve(sizeof(TileDev));
    const size_t o_chunks = carve((size_t)nch * sizeof(Chunk)), o_hist = carve((size_t)(hents + 1) * 4);
    const size_t o_part = carve(kScanBlocks * 4), o_banded = carve((size_t)t.n * 16);
    const size_t o_sorted = carve((size_t)t.n * 16), o_offs = carve((size_t)(bks + 1) * 4);
    const size_t o_hdist = carve((size_t)ebytes), o_edist = carve((size_t)ebytes);
    // the split form's value raster (and squared-distance raster, unless the
    // caller takes the squared distances)
    const size_t pix = (size_t)t.W * t.H;
    const size_t o_val = t.label_only ? carve(pix * 4) : 0;
    const size_t o_sd2 = t.label_only && !d_out_d2 ? carve(pix * 4) : 0;
    if (off > kNoAllocScratchBytes)
        return fail(ctx, TNNI_E_BUDGET,
                    "tile needs " + std::to_string(off) + " bytes of scratch; the allocation-free entry point has " +
                        std::to_string(kNoAllocScratchBytes) + " (use tnni_tile_run_dev)");
    char *base = (char *)ns.ptr;
    DevBuffers db;
    db.batch = (BatchDev *)(base + o_batch);
    db.tiles = (const TileDev *)(base + o_tiles);
    db.chunks = (const Chunk *)(base + o_chunks);
    db.hist = (uint32_t *)(base + o_hist);
    db.partials = (uint32_t *)(base + o_part);
    db.banded = (int4 *)(base + o_banded);
    db.sorted = (int4 *)(base + o_sorted);
    db.offs = (uint32_t *)(base + o_offs);
    db.hdist = (uint8_t *)(base + o_hdist);
    db.edist = (uint8_t *)(base + o_edist);
    t.gitem_base = 0;
    if (t.label_only) {
        t.out_val = (int32_t *)(base + o_val);
        if (!d_out_d2) t.out_d2 = (uint32_t *)(base + o_sd2);
    }
    const int64_t gather_items = t.label_only ? (int64_t)t.gntx * t.gnty : 0;
    BatchDev bd;
    std::memset(&bd, 0, sizeof bd);
    bd.n_tiles = 1;
    bd.n_chunks = (int)nch;
    bd.chunk_pts = (uint32_t)cpts;
    bd.n_scan = (uint32_t)(hents + 1);
    bd.n_items = (int64_t)t.ntx * t.nty;
    bd.n_rows = t.nby;
    bd.n_bands = t.nbands;
    bd.scan_chunk = (uint32_t)((bd.n_scan + kScanBlocks - 1) / kScanBlocks);
    bd.n_gather_items = gather_items;
    // the array keeps the previous call's contents: the setup kernel writes
    // every descriptor and k_band_count zeroes the work counter, and the
    // passes write every histogram entry, offset and distance they read
    Arena &a = ctx->main;
    HIP_TRY(ctx, launch_single_tile_setup(t, bd, (TileDev *)db.tiles, (Chunk *)db.chunks, db.batch, a.stream));
    LaunchCfg cfg;
    int occ = tile_kernel_occupancy(p.info.lds);
    if (p.info.wg_per_cu > 0) occ = std::min(occ, p.info.wg_per_cu);
    cfg.main_lds = p.info.lds;
    cfg.algorithm = t.algorithm;
    cfg.mode = t.mode;
    cfg.time = 0;
    cfg.label_only = t.label_only;
    cfg.dense = t.dense ? 1 : 0;
    cfg.sort_grid = (int)std::min<int64_t>(4096, nch);
    cfg.band_grid = (int)std::max<int64_t>(1, std::min<int64_t>(8192, t.nbands));
    cfg.row_grid = (int)std::max<int64_t>(1, std::min<int64_t>(8192, t.nby));
    cfg.main_grid = (int)std::max<int64_t>(1, std::min<int64_t>(bd.n_items, (int64_t)ctx->n_cu * occ));
    cfg.gather_grid = (int)std::min<int64_t>(gather_items, 8192);
    cfg.gather_lds = t.label_only ? split_pass_lds(t) : 0;
    HIP_TRY(ctx, enqueue_pipeline(db, cfg, a.stream, nullptr, nullptr));
    HIP_TRY(ctx, hipStreamSynchronize(a.stream));
    return TNNI_OK;
}

tnni_status tnni_tile_run(tnni_ctx *ctx, const tnni_tile_desc *desc, const int32_t *X,
                          const int32_t *Y, const int32_t *Z, uint64_t n, float *out_h,
                          int32_t *out_labels, uint32_t *out_d2)
{
    if (!ctx || !desc || !out_h || (n && (!X || !Y || !Z)))
        return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::vector<Prepared> v(1);
    std::string err;
    tnni_status st = prepare_tile(*desc, n, ctx->n_cu, false, &ctx->tuned, v[0].t, v[0].info, err);
    if (st != TNNI_OK) return fail(ctx, st, err);
    Arena &a = ctx->main;
    TileDev &t = v[0].t;
    const size_t cells = (size_t)t.nrows * t.ncols, pix = (size_t)t.W * t.H;
    if ((st = ensure(ctx, a.sx, n * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sy, n * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sz, n * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sh, cells * 4)) != TNNI_OK) return st;
    if (out_labels && (st = ensure(ctx, a.sl, pix * 4)) != TNNI_OK) return st;
    if (out_d2 && (st = ensure(ctx, a.sd, pix * 4)) != TNNI_OK) return st;
    if (n) {
        HIP_TRY(ctx, hipMemcpyAsync(a.sx.p, X, n * 4, hipMemcpyHostToDevice, a.stream));
        HIP_TRY(ctx, hipMemcpyAsync(a.sy.p, Y, n * 4, hipMemcpyHostToDevice, a.stream));
        HIP_TRY(ctx, hipMemcpyAsync(a.sz.p, Z, n * 4, hipMemcpyHostToDevice, a.stream));
    }
    t.X = (const int32_t *)a.sx.p;
    t.Y = (const int32_t *)a.sy.p;
    t.Z = (const int32_t *)a.sz.p;
    t.out_h = (float *)a.sh.p;
    t.out_labels = out_labels ? (int32_t *)a.sl.p : nullptr;
    t.out_d2 = out_d2 ? (uint32_t *)a.sd.p : nullptr;
    if ((st = run_prepared(ctx, a, v, false, Launch::Direct)) != TNNI_OK) return st;
    HIP_TRY(ctx, hipMemcpyAsync(out_h, a.sh.p, cells * 4, hipMemcpyDeviceToHost, a.stream));
    if (out_labels)
        HIP_TRY(ctx, hipMemcpyAsync(out_labels, a.sl.p, pix * 4, hipMemcpyDeviceToHost, a.stream));
    if (out_d2) HIP_TRY(ctx, hipMemcpyAsync(out_d2, a.sd.p, pix * 4, hipMemcpyDeviceToHost, a.stream));
    HIP_TRY(ctx, hipStreamSynchronize(a.stream));
    return TNNI_OK;
}

}  // extern "C"

namespace {

// The slices of one space-time tile (performNNI's loop, nni.cpp:910-933):
// j = -(int)r .. r, slice time t = cur_time + j in float, each a full
// pipeline run over the tile's points that keeps the points within the time
// radius of t as sites; the cells' sums stay in the arena between the runs
// and the last run divides.  p.t.X/Y/Z/T and out_h are set by the caller.
// With points sorted by time (and at most kMaxLayers times within reach) the
// slices are built from per-time layers instead: the layers the cache (or,
// without one, the context's scratch cache emptied first) lacks are
// searched together as plain tiles of their runs of points, their labels
// turned into site values, then each slice combines its layers
// (k_combine_layers) and is gathered (k_gather_slice).  desc is the
// cache's key.
// A space-time tile's descriptor for its slices: the fused form (the
// slices' kernels accumulate across slices in the tile kernel's gather).
tnni_status prepare_time_tile(tnni_ctx *ctx, const tnni_tile_desc &desc, uint64_t n, Prepared &p,
                              std::string &err)
{
    tnni_tile_desc d = desc;
    d.form = TNNI_FORM_FUSED;
    return prepare_tile(d, n, ctx->n_cu, false, &ctx->tuned, p.t, p.info, err);
}

tnni_status run_time_tile(tnni_ctx *ctx, Arena &a, const Prepared &p, float cur_time,
                          float time_radius, tnni_time_cache *cache = nullptr,
                          const tnni_tile_desc *desc = nullptr)
{
    const TileDev &t = p.t;
    if (!(time_radius >= 0.0f) || time_radius > 1.0e4f)
        return fail(ctx, TNNI_E_INVALID, "time_radius must be in [0, 1e4]");
    // squared distances plus time terms stay below the "no site" key
    const uint64_t u = (uint64_t)std::ceil((double)time_radius * t.upm) + 1;
    const uint64_t dxy = t.algorithm == TNNI_ALG_PLANES ? 2ull * t.Rs * t.Rs : (uint64_t)t.Rs2;
    const uint64_t dq = t.algorithm == TNNI_ALG_PLANES ? 2ull * t.Rq * t.Rq : (uint64_t)t.Rq2;
    if (dxy + u * u >= (uint64_t)kNoD2 || dq + u * u >= (uint64_t)kNoD2)
        return fail(ctx, TNNI_E_GEOMETRY, "time radius too large for 32-bit squared distances");
    const size_t cells = (size_t)t.nrows * t.ncols;
    tnni_status st = ensure(ctx, a.acc, cells * sizeof(uint4));
    if (st != TNNI_OK) return st;
    const int j0 = -(int)time_radius;
    int j1 = j0;
    while ((float)(j1 + 1) <= time_radius) ++j1;
    const uint64_t pix = (uint64_t)t.W * t.H;
    // the slices' site filter and time distance, as the kernels take them
    // (renderInputPoints' filter nni.cpp:770, drawPlane's term nni.cpp:197)
    auto in_slice = [time_radius](float st_, float ft) {
        return (double)std::fabs(st_ - ft) <= (double)time_radius + 0.000001;
    };
    auto dt2_of = [&t](float st_, float ft) {
        const uint32_t uu = (uint32_t)std::rint((double)std::fabs(st_ - ft) * (double)t.upm);
        return uu * uu;
    };
    std::vector<float> years;  // the point times (as floats, ascending) the slices reach
    std::vector<uint32_t> runs;
    if (cache && !std::isfinite(cur_time)) return fail(ctx, TNNI_E_INVALID, "cur_time must be finite");
    // per-time layers when the call's points are sorted by time and reach
    // at most kMaxLayers times; the layers are kept between calls only with
    // a cache (without one, the context's scratch cache starts empty)
    bool layered = std::isfinite(cur_time) && t.n <= 0xFFFFFFFFll;
    if (layered) {
        for (int j = j0; j <= j1; ++j) {
            const float sj = cur_time + (float)j;
            const double lo = std::floor((double)sj - (double)time_radius) - 1.0;
            const double hi = std::ceil((double)sj + (double)time_radius) + 1.0;
            for (double y = lo; y <= hi; y += 1.0) {
                if (y < (double)INT32_MIN || y > (double)INT32_MAX) continue;
                const float ft = (float)(int32_t)y;
                if (in_slice(sj, ft) && std::find(years.begin(), years.end(), ft) == years.end())
                    years.push_back(ft);
            }
        }
        std::sort(years.begin(), years.end());
        layered = (int)years.size() <= kMaxLayers;
    }
    if (layered) {
        // whether the points are sorted by time, and each time's run of them
        LayerTimes ys;
        ys.n = (int)years.size();
        for (int k = 0; k < ys.n; ++k) ys.t[k] = years[(size_t)k];
        runs.assign(1 + 2 * kMaxLayers, (uint32_t)t.n);
        runs[0] = 0u;
        const size_t rb = runs.size() * 4;
        if ((st = ensure(ctx, a.flag, rb)) != TNNI_OK) return st;
        HIP_TRY(ctx, hipMemcpyAsync(a.flag.p, runs.data(), rb, hipMemcpyHostToDevice, a.stream));
        HIP_TRY(ctx, launch_time_runs(t.T, (uint64_t)t.n, ys, (uint32_t *)a.flag.p, a.stream));
        HIP_TRY(ctx, hipMemcpyAsync(runs.data(), a.flag.p, rb, hipMemcpyDeviceToHost, a.stream));
        HIP_TRY(ctx, hipStreamSynchronize(a.stream));
        // points out of time order: ties across times would not follow the
        // index, so this call searches every slice
        layered = runs[0] == 0u;
    }
    if (!layered) {
        // every slice searched: one pipeline run each, its gather summing
        // into the cells' running sums
        for (int j = j0; j <= j1; ++j) {
            std::vector<Prepared> v(1, p);
            TileDev &s = v[0].t;
            s.slice_t = cur_time + (float)j;
            s.time_radius = time_radius;
            const int64_t ju = (int64_t)j * t.upm;
            s.jj2 = (uint32_t)(ju * ju);
            s.acc_load = j != j0;
            s.acc_final = j == j1;
            s.acc = (uint4 *)a.acc.p;
            if ((st = run_prepared(ctx, a, v, false, Launch::Direct)) != TNNI_OK) return st;
        }
        return TNNI_OK;
    }
    if (!cache) {
        // Delta^2: nothing is kept between calls
        if (!ctx->scratch_cache) {
            ctx->scratch_cache = new (std::nothrow) tnni_time_cache;
            if (!ctx->scratch_cache) return fail(ctx, TNNI_E_NOMEM, "time cache");
            ctx->scratch_cache->ctx = ctx;
            ctx->caches.push_back(ctx->scratch_cache);
        }
        cache = ctx->scratch_cache;
        cache->keyed = false;
    }
    // another tile: start over (the reference's driver clears the buffers
    // per work package, terranni.cpp:422, nni.cpp:362-379)
    if (!cache->keyed || std::memcmp(&cache->desc, desc, sizeof *desc) != 0) {
        for (auto &c : cache->slots) c.full = false;
        cache->desc = *desc;
        cache->keyed = true;
    }
    const size_t want = std::max((size_t)cache->max_slots, years.size());
    if (cache->slots.size() < want) cache->slots.resize(want);
    // the layers: those held first (so that none of them is reused for
    // another time in this call), then the missing ones, searched together
    // into the least recently used slots
    const uint64_t now = ++cache->clock;
    std::vector<tnni_time_cache::Slot *> lay(years.size(), nullptr);
    for (size_t k = 0; k < years.size(); ++k) {
        uint32_t key;
        std::memcpy(&key, &years[k], 4);
        for (auto &c : cache->slots)
            if (c.full && c.time_bits == key) {
                lay[k] = &c;
                c.used = now;
                ++cache->hits;
                break;
            }
    }
    if ((st = ensure(ctx, a.cv, pix * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.cd, pix * 4)) != TNNI_OK) return st;
    std::vector<Prepared> v;
    std::vector<tnni_time_cache::Slot *> vslot;
    // cell blocks of the layer searches: a call that searches one or two
    // layers would leave most workgroups idle with the plain form's blocks,
    // so below 16 items per CU a layer's blocks shrink (to no less than 8
    // cells a side, and only to a size its LDS budget takes; the block size
    // never changes a result)
    size_t n_missing = 0;
    for (size_t k = 0; k < years.size(); ++k) n_missing += lay[k] ? 0 : 1;
    const double item_target = 16.0 * ctx->n_cu / (double)std::max<size_t>(n_missing, 1);
    // (label-only layer tiles: pixels a side, whole 8x8 wave blocks)
    const int tc_small = std::max(8, (int)std::floor(std::sqrt((double)t.W * t.H / item_target) / 8.0) * 8);
    for (size_t k = 0; k < years.size(); ++k) {
        if (lay[k]) continue;
        // a slot this call uses (held, or already taken for another missing
        // time; those fill only after the search) is never reused
        tnni_time_cache::Slot *victim = nullptr;
        for (auto &c : cache->slots) {
            if (c.used == now) continue;
            if (!victim || (victim->full && (!c.full || c.used < victim->used))) victim = &c;
        }
        if (!victim) return fail(ctx, TNNI_E_INVALID, "space-time cache: no free slot");
        victim->full = false;
        if ((st = ensure(ctx, victim->val, pix * 4)) != TNNI_OK) return st;
        if ((st = ensure(ctx, victim->d2, pix * 4)) != TNNI_OK) return st;
        // the 2-D Voronoi of the points of this time: the plain tile of
        // their run (its heights go to scratch that the slices overwrite)
        const uint32_t lo = runs[1 + k], hi = std::max(runs[1 + k], runs[1 + kMaxLayers + k]);
        // label-only items (no query halo, no accumulation): the layer takes
        // each pixel's site value and squared distance straight from them
        tnni_tile_desc lyr = *desc;
        lyr.form = TNNI_FORM_SPLIT;
        Prepared q;
        std::string err;
        if ((st = prepare_tile(lyr, hi - lo, ctx->n_cu, false, &ctx->tuned, q.t, q.info, err)) != TNNI_OK)
            return fail(ctx, st, err);
        if (desc->tile_cells <= 0 && (double)q.t.ntx * q.t.nty < item_target && tc_small < q.t.tc) {
            tnni_tile_desc ld = lyr;
            ld.tile_cells = tc_small;
            Prepared qs;
            if (prepare_tile(ld, hi - lo, ctx->n_cu, false, &ctx->tuned, qs.t, qs.info, err) == TNNI_OK) q = qs;
        }
        q.t.X = t.X + lo;
        q.t.Y = t.Y + lo;
        q.t.Z = t.Z + lo;
        q.t.out_h = nullptr;
        q.t.out_val = (int32_t *)victim->val.p;
        q.t.out_d2 = (uint32_t *)victim->d2.p;
        v.push_back(q);
        vslot.push_back(victim);
        std::memcpy(&victim->time_bits, &years[k], 4);
        victim->used = now;
        lay[k] = victim;
        ++cache->misses;
    }
    if (!v.empty()) {
        if ((st = run_prepared(ctx, a, v, false, Launch::Direct)) != TNNI_OK) return st;
        for (size_t i = 0; i < v.size(); ++i) vslot[i]->full = true;
    }
    // each slice: its layers combined, then gathered
    for (int j = j0; j <= j1; ++j) {
        TileDev s = p.t;
        s.slice_t = cur_time + (float)j;
        s.time_radius = time_radius;
        const int64_t ju = (int64_t)j * t.upm;
        s.jj2 = (uint32_t)(ju * ju);
        s.acc_load = j != j0;
        s.acc_final = j == j1;
        s.acc = (uint4 *)a.acc.p;
        LayerSet ls;
        ls.n = 0;
        for (size_t k = 0; k < years.size(); ++k) {
            if (!in_slice(s.slice_t, years[k])) continue;
            ls.dt2[ls.n] = dt2_of(s.slice_t, years[k]);
            ls.val[ls.n] = (const int32_t *)lay[k]->val.p;
            ls.d2[ls.n] = (const uint32_t *)lay[k]->d2.p;
            ++ls.n;
        }
        HIP_TRY(ctx, launch_combine_layers(ls, pix, (int32_t *)a.cv.p, (uint32_t *)a.cd.p, a.stream));
        HIP_TRY(ctx, launch_gather_slice(s, (const int32_t *)a.cv.p, (const uint32_t *)a.cd.p, a.stream));
    }
    return TNNI_OK;
}

}  // namespace

extern "C" {

tnni_status tnni_time_cache_create(tnni_ctx *ctx, int slots, tnni_time_cache **out)
{
    if (!ctx || !out) return fail(ctx, TNNI_E_INVALID, "null argument");
    if (slots > 4096) return fail(ctx, TNNI_E_INVALID, "at most 4096 slots");
    std::lock_guard<std::mutex> lk(ctx->mu);
    tnni_time_cache *c = new (std::nothrow) tnni_time_cache;
    if (!c) return fail(ctx, TNNI_E_NOMEM, "time cache");
    c->ctx = ctx;
    c->max_slots = std::max(slots, 0);
    ctx->caches.push_back(c);
    *out = c;
    return TNNI_OK;
}

void tnni_time_cache_destroy(tnni_time_cache *cache)
{
    if (!cache) return;
    if (tnni_ctx *ctx = cache->ctx) {
        std::lock_guard<std::mutex> lk(ctx->mu);
        (void)hipSetDevice(ctx->device);
        for (auto &s : cache->slots) {
            release(ctx, s.val);
            release(ctx, s.d2);
        }
        ctx->caches.erase(std::remove(ctx->caches.begin(), ctx->caches.end(), cache), ctx->caches.end());
    }
    delete cache;
}

tnni_status tnni_time_cache_clear(tnni_time_cache *cache)
{
    if (!cache) return fail(nullptr, TNNI_E_INVALID, "null argument");
    if (!cache->ctx) return TNNI_OK;
    std::lock_guard<std::mutex> lk(cache->ctx->mu);
    for (auto &s : cache->slots) s.full = false;
    cache->keyed = false;
    return TNNI_OK;
}

tnni_status tnni_time_cache_stats(const tnni_time_cache *cache, uint64_t *hits, uint64_t *misses)
{
    if (!cache || !hits || !misses) return fail(nullptr, TNNI_E_INVALID, "null argument");
    if (cache->ctx) {
        std::lock_guard<std::mutex> lk(cache->ctx->mu);
        *hits = cache->hits;
        *misses = cache->misses;
    } else {
        *hits = cache->hits;
        *misses = cache->misses;
    }
    return TNNI_OK;
}

tnni_status tnni_tile_run_time_dev(tnni_ctx *ctx, const tnni_tile_desc *desc, const int32_t *dX,
                                   const int32_t *dY, const int32_t *dZ, const int32_t *dT,
                                   uint64_t n, float cur_time, float time_radius, float *d_out_h)
{
    return tnni_tile_run_time_cached_dev(ctx, nullptr, desc, dX, dY, dZ, dT, n, cur_time, time_radius,
                                         d_out_h);
}

tnni_status tnni_tile_run_time_cached_dev(tnni_ctx *ctx, tnni_time_cache *cache,
                                          const tnni_tile_desc *desc, const int32_t *dX,
                                          const int32_t *dY, const int32_t *dZ, const int32_t *dT,
                                          uint64_t n, float cur_time, float time_radius, float *d_out_h)
{
    if (!ctx || !desc || !d_out_h || (n && (!dX || !dY || !dZ || !dT)))
        return fail(ctx, TNNI_E_INVALID, "null argument");
    if (cache && cache->ctx != ctx) return fail(ctx, TNNI_E_INVALID, "the cache belongs to another context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    Prepared p;
    std::string err;
    tnni_status st = prepare_time_tile(ctx, *desc, n, p, err);
    if (st != TNNI_OK) return fail(ctx, st, err);
    p.t.X = dX;
    p.t.Y = dY;
    p.t.Z = dZ;
    // with no points T is never read: any non-null pointer marks the time mode
    p.t.T = n ? dT : (const int32_t *)d_out_h;
    p.t.out_h = d_out_h;
    if ((st = run_time_tile(ctx, ctx->main, p, cur_time, time_radius, cache, desc)) != TNNI_OK) return st;
    HIP_TRY(ctx, hipStreamSynchronize(ctx->main.stream));
    return TNNI_OK;
}

tnni_status tnni_tile_run_time(tnni_ctx *ctx, const tnni_tile_desc *desc, const int32_t *X,
                               const int32_t *Y, const int32_t *Z, const int32_t *T, uint64_t n,
                               float cur_time, float time_radius, float *out_h)
{
    return tnni_tile_run_time_cached(ctx, nullptr, desc, X, Y, Z, T, n, cur_time, time_radius, out_h);
}

tnni_status tnni_tile_run_time_cached(tnni_ctx *ctx, tnni_time_cache *cache, const tnni_tile_desc *desc,
                                      const int32_t *X, const int32_t *Y, const int32_t *Z,
                                      const int32_t *T, uint64_t n, float cur_time, float time_radius,
                                      float *out_h)
{
    if (!ctx || !desc || !out_h || (n && (!X || !Y || !Z || !T)))
        return fail(ctx, TNNI_E_INVALID, "null argument");
    if (cache && cache->ctx != ctx) return fail(ctx, TNNI_E_INVALID, "the cache belongs to another context");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    Prepared p;
    std::string err;
    tnni_status st = prepare_time_tile(ctx, *desc, n, p, err);
    if (st != TNNI_OK) return fail(ctx, st, err);
    Arena &a = ctx->main;
    const size_t cells = (size_t)p.t.nrows * p.t.ncols;
    if ((st = ensure(ctx, a.sx, n * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sy, n * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sz, n * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.st, n * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, a.sh, cells * 4)) != TNNI_OK) return st;
    if (n) {
        HIP_TRY(ctx, hipMemcpyAsync(a.sx.p, X, n * 4, hipMemcpyHostToDevice, a.stream));
        HIP_TRY(ctx, hipMemcpyAsync(a.sy.p, Y, n * 4, hipMemcpyHostToDevice, a.stream));
        HIP_TRY(ctx, hipMemcpyAsync(a.sz.p, Z, n * 4, hipMemcpyHostToDevice, a.stream));
        HIP_TRY(ctx, hipMemcpyAsync(a.st.p, T, n * 4, hipMemcpyHostToDevice, a.stream));
    }
    p.t.X = (const int32_t *)a.sx.p;
    p.t.Y = (const int32_t *)a.sy.p;
    p.t.Z = (const int32_t *)a.sz.p;
    p.t.T = n ? (const int32_t *)a.st.p : (const int32_t *)a.sh.p;  // non-null: time mode
    p.t.out_h = (float *)a.sh.p;
    if ((st = run_time_tile(ctx, a, p, cur_time, time_radius, cache, desc)) != TNNI_OK) return st;
    HIP_TRY(ctx, hipMemcpyAsync(out_h, a.sh.p, cells * 4, hipMemcpyDeviceToHost, a.stream));
    HIP_TRY(ctx, hipStreamSynchronize(a.stream));
    return TNNI_OK;
}

tnni_status tnni_batch_run_dev(tnni_ctx *ctx, int n_tiles, const tnni_tile_desc *descs,
                               const int32_t *const *dX, const int32_t *const *dY,
                               const int32_t *const *dZ, const uint64_t *n,
                               float *const *d_out_h, int32_t *const *d_out_labels,
                               uint32_t *const *d_out_d2)
{
    if (!ctx || n_tiles < 0 || (n_tiles && (!descs || !dX || !dY || !dZ || !n || !d_out_h)))
        return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    std::vector<Prepared> v((size_t)n_tiles);
    std::string err;
    for (int k = 0; k < n_tiles; ++k) {
        tnni_status st = prepare_tile(descs[k], n[k], ctx->n_cu, true, &ctx->tuned, v[k].t, v[k].info, err);
        if (st != TNNI_OK) return fail(ctx, st, "tile " + std::to_string(k) + ": " + err);
        if (!d_out_h[k] || (n[k] && (!dX[k] || !dY[k] || !dZ[k])))
            return fail(ctx, TNNI_E_INVALID, "tile " + std::to_string(k) + ": null pointer");
        v[k].t.X = dX[k];
        v[k].t.Y = dY[k];
        v[k].t.Z = dZ[k];
        v[k].t.out_h = d_out_h[k];
        v[k].t.out_labels = d_out_labels ? d_out_labels[k] : nullptr;
        v[k].t.out_d2 = d_out_d2 ? d_out_d2[k] : nullptr;
    }
    std::stable_sort(v.begin(), v.end(),
                     [](const Prepared &a, const Prepared &b) { return tile_cost(a.t) > tile_cost(b.t); });
    return run_prepared(ctx, ctx->main, v, true, ctx->prof ? Launch::Direct : Launch::Graph);
}

}  // extern "C"

namespace {

tnni_status submit_common(tnni_ctx *ctx, const tnni_tile_desc *desc, bool dev, const int32_t *X,
                          const int32_t *Y, const int32_t *Z, uint64_t n, float *out_h,
                          int32_t *out_labels, uint32_t *out_d2, uint64_t *ticket)
{
    if (!ctx || !desc || !out_h || !ticket || (n && (!X || !Y || !Z)))
        return fail(ctx, TNNI_E_INVALID, "null argument");
    auto r = std::make_shared<Request>();
    std::string err;
    tnni_status st = prepare_tile(*desc, n, ctx->n_cu, true, &ctx->tuned, r->t, r->info, err);
    if (st != TNNI_OK) return fail(ctx, st, err);
    r->dev = dev;
    r->X = X;
    r->Y = Y;
    r->Z = Z;
    r->n = n;
    r->out_h = out_h;
    r->out_labels = out_labels;
    r->out_d2 = out_d2;
    r->footprint = request_bytes(r->t, out_labels != nullptr, out_d2 != nullptr);
    if (!dev) {
        // the inputs are copied now, in the caller's thread, so that the
        // scheduler only issues asynchronous copies
        const size_t pix = (size_t)r->t.W * r->t.H;
        const size_t out_b = ((size_t)r->t.nrows * r->t.ncols + (out_labels ? pix : 0) + (out_d2 ? pix : 0)) * 4;
        r->pin_in = ctx->pinned.acquire(std::max<size_t>(n * 12, 4), r->pin_in_cap);
        r->pin_out = ctx->pinned.acquire(out_b, r->pin_out_cap);
        if (!r->pin_in || !r->pin_out) {
            ctx->pinned.release(r->pin_in, r->pin_in_cap);
            ctx->pinned.release(r->pin_out, r->pin_out_cap);
            return fail(ctx, TNNI_E_NOMEM, "pinned host staging");
        }
        int32_t *p = (int32_t *)r->pin_in;
        if (n) {
            std::memcpy(p, X, n * 4);
            std::memcpy(p + n, Y, n * 4);
            std::memcpy(p + 2 * n, Z, n * 4);
        }
    }
    std::lock_guard<std::mutex> lk(ctx->smu);
    auto release = [&] {
        ctx->pinned.release(r->pin_in, r->pin_in_cap);
        ctx->pinned.release(r->pin_out, r->pin_out_cap);
    };
    if (ctx->shutdown) {
        release();
        return fail(ctx, TNNI_E_SHUTDOWN, "context is shutting down");
    }
    if ((st = start_scheduler(ctx)) != TNNI_OK) {
        release();
        return st;
    }
    if (r->footprint > ctx->serve_budget) {
        release();
        return fail(ctx, TNNI_E_BUDGET, "tile footprint exceeds the context's HBM budget");
    }
    r->id = ctx->next_id++;
    r->t_submit = now_s();
    ctx->reqs[r->id] = r;
    ctx->pending.push_back(r);
    ctx->stats.submitted++;
    ctx->scv.notify_all();
    *ticket = r->id;
    return TNNI_OK;
}

}  // namespace

extern "C" {

tnni_status tnni_submit(tnni_ctx *ctx, const tnni_tile_desc *desc, const int32_t *X,
                        const int32_t *Y, const int32_t *Z, uint64_t n, float *out_h,
                        int32_t *out_labels, uint32_t *out_d2, uint64_t *ticket)
{
    return submit_common(ctx, desc, false, X, Y, Z, n, out_h, out_labels, out_d2, ticket);
}

tnni_status tnni_submit_dev(tnni_ctx *ctx, const tnni_tile_desc *desc, const int32_t *dX,
                            const int32_t *dY, const int32_t *dZ, uint64_t n, float *d_out_h,
                            int32_t *d_out_labels, uint32_t *d_out_d2, uint64_t *ticket)
{
    return submit_common(ctx, desc, true, dX, dY, dZ, n, d_out_h, d_out_labels, d_out_d2, ticket);
}

tnni_status tnni_poll(tnni_ctx *ctx, uint64_t ticket)
{
    if (!ctx) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->smu);
    auto it = ctx->reqs.find(ticket);
    if (it == ctx->reqs.end()) return fail(ctx, TNNI_E_INVALID, "unknown ticket");
    return it->second->status;
}

tnni_status tnni_wait(tnni_ctx *ctx, uint64_t ticket, double *latency_s)
{
    if (!ctx) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::unique_lock<std::mutex> lk(ctx->smu);
    auto it = ctx->reqs.find(ticket);
    if (it == ctx->reqs.end()) return fail(ctx, TNNI_E_INVALID, "unknown ticket");
    std::shared_ptr<Request> r = it->second;
    ctx->dcv.wait(lk, [&r] { return r->status != TNNI_PENDING; });
    ctx->reqs.erase(ticket);
    lk.unlock();
    if (!r->dev) {
        // results out of the pinned staging, in the caller's thread
        if (r->status == TNNI_OK) {
            const size_t cells = (size_t)r->t.nrows * r->t.ncols, pix = (size_t)r->t.W * r->t.H;
            const char *po = (const char *)r->pin_out;
            std::memcpy(r->out_h, po, cells * 4);
            po += cells * 4;
            if (r->out_labels) {
                std::memcpy(r->out_labels, po, pix * 4);
                po += pix * 4;
            }
            if (r->out_d2) std::memcpy(r->out_d2, po, pix * 4);
        }
        ctx->pinned.release(r->pin_in, r->pin_in_cap);
        ctx->pinned.release(r->pin_out, r->pin_out_cap);
        r->pin_in = r->pin_out = nullptr;
        r->t_done = std::max(r->t_done, now_s());  // results in the caller's buffers
    }
    if (latency_s) *latency_s = r->t_done - r->t_submit;
    if (r->status != TNNI_OK) {
        const std::string m = "request failed" + (r->err.empty() ? std::string() : ": " + r->err);
        g_thread_err = m;
        set_ctx_error(ctx, m);
    }
    return r->status;
}

tnni_status tnni_serve_stats_read(tnni_ctx *ctx, tnni_serve_stats *out)
{
    if (!ctx || !out) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->smu);
    *out = ctx->stats;
    out->inflight_bytes = ctx->inflight_bytes;
    return TNNI_OK;
}

tnni_status tnni_serve_stats_reset(tnni_ctx *ctx)
{
    if (!ctx) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->smu);
    ctx->stats.max_inflight_tiles = 0;
    ctx->stats.max_inflight_batches = 0;
    return TNNI_OK;
}

tnni_status tnni_profile_enable(tnni_ctx *ctx, int enable)
{
    if (!ctx) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    ctx->prof = enable != 0;
    return TNNI_OK;
}

tnni_status tnni_profile_read(tnni_ctx *ctx, uint64_t *launches, double *total_ms, int reset)
{
    if (!ctx) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    if (launches) *launches = ctx->prof_launches;
    if (total_ms) *total_ms = ctx->prof_ms;
    if (reset) {
        ctx->prof_launches = 0;
        ctx->prof_ms = 0.0;
    }
    return TNNI_OK;
}

}  // extern "C"

// ---- autotuning ----------------------------------------------------------

namespace {

// Time one setting (bucket side bm, cell-tile side tc, at most wg
// workgroups per CU; 0 = all) on `replicas` copies of the sample tiles idx.
// Batched: one launch over all of them, GPU time between events around the
// pipeline; plain: one synchronous launch per tile, host clock.  One
// warm-up run, then the best of reps.
tnni_status time_setting(tnni_ctx *ctx, const std::vector<int> &idx, const tnni_tile_desc *descs,
                         const int32_t *const *dX, const int32_t *const *dY,
                         const int32_t *const *dZ, const uint64_t *n, bool batched, int replicas,
                         int bm, int tc, int wg, int form, int reps, float &best_ms)
{
    Arena &a = ctx->main;
    std::vector<Prepared> v;
    std::vector<size_t> out_off;
    std::string err;
    size_t cells = 0;
    for (int r = 0; r < replicas; ++r) {
        for (int k : idx) {
            tnni_tile_desc d = descs[k];
            d.bucket_px = bm;
            d.tile_cells = tc;
            d.form = form;
            Prepared p;
            tnni_status st = prepare_tile(d, n[k], ctx->n_cu, batched, nullptr, p.t, p.info, err);
            if (st != TNNI_OK) return fail(ctx, st, err);
            p.info.wg_per_cu = wg;
            p.t.X = dX[k];
            p.t.Y = dY[k];
            p.t.Z = dZ[k];
            out_off.push_back(cells);
            cells += (size_t)p.t.nrows * p.t.ncols;
            v.push_back(p);
        }
    }
    tnni_status st = ensure(ctx, a.sh, cells * 4);
    if (st != TNNI_OK) return st;
    for (size_t k = 0; k < v.size(); ++k) v[k].t.out_h = (float *)a.sh.p + out_off[k];
    if (batched)
        std::stable_sort(v.begin(), v.end(), [](const Prepared &x, const Prepared &y) {
            return tile_cost(x.t) > tile_cost(y.t);
        });
    best_ms = 3.0e38f;
    for (int rep = 0; rep <= reps; ++rep) {
        float ms = 0.f;
        if (batched) {
            std::vector<Prepared> w = v;
            if ((st = run_prepared(ctx, a, w, true, Launch::Direct, ctx->tev0, ctx->tev1)) != TNNI_OK)
                return st;
            HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->tev0, ctx->tev1));
        } else {
            const double t0 = now_s();
            for (const Prepared &p : v) {
                std::vector<Prepared> w(1, p);
                if ((st = run_prepared(ctx, a, w, true, Launch::Direct)) != TNNI_OK) return st;
            }
            ms = (float)((now_s() - t0) * 1e3);
        }
        if (rep > 0) best_ms = std::min(best_ms, ms);
    }
    return TNNI_OK;
}

}  // namespace

extern "C" {

tnni_status tnni_shape_class(const tnni_tile_desc *desc, uint64_t n_points, int batched,
                             uint64_t *cls)
{
    if (!desc || !cls) return fail(nullptr, TNNI_E_INVALID, "null argument");
    tnni_frame_info f;
    std::string err;
    tnni_status st = compute_frame(*desc, f, err);
    if (st != TNNI_OK) return fail(nullptr, st, err);
    *cls = shape_class(*desc, n_points, f, batched != 0);
    return TNNI_OK;
}

tnni_status tnni_autotune(tnni_ctx *ctx, int n_tiles, const tnni_tile_desc *descs,
                          const int32_t *const *dX, const int32_t *const *dY,
                          const int32_t *const *dZ, const uint64_t *n, int batched, int reps,
                          tnni_tune_entry *out, int *n_classes)
{
    if (!ctx || n_tiles < 0 || (n_tiles && (!descs || !dX || !dY || !dZ || !n)))
        return fail(ctx, TNNI_E_INVALID, "null argument");
    if (n_classes) *n_classes = 0;
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    reps = std::max(1, std::min(reps, 20));
    const bool bat = batched != 0;
    std::map<uint64_t, std::vector<int>> groups;
    std::vector<uint64_t> pix((size_t)n_tiles);
    for (int k = 0; k < n_tiles; ++k) {
        if (n[k] && (!dX[k] || !dY[k] || !dZ[k]))
            return fail(ctx, TNNI_E_INVALID, "tile " + std::to_string(k) + ": null pointer");
        tnni_frame_info f;
        std::string err;
        tnni_status st = compute_frame(descs[k], f, err);
        if (st != TNNI_OK) return fail(ctx, st, "tile " + std::to_string(k) + ": " + err);
        groups[shape_class(descs[k], n[k], f, bat)].push_back(k);
        pix[(size_t)k] = (uint64_t)f.W * (uint64_t)f.H;
    }
    int written = 0;
    for (const auto &g : groups) {
        const std::vector<int> &idx = g.second;
        tnni_tile_desc d0 = descs[idx[0]];
        d0.bucket_px = 0;
        d0.tile_cells = 0;
        TileDev t0;
        PrepInfo i0;
        std::string err;
        tnni_status st = prepare_tile(d0, n[idx[0]], ctx->n_cu, bat, nullptr, t0, i0, err);
        if (st != TNNI_OK) return fail(ctx, st, err);
        const int bm0 = t0.bm, tc0 = t0.tc;
        uint64_t p = 0;
        for (int k : idx) p += pix[(size_t)k];
        const int replicas =
            bat ? (int)std::max<uint64_t>(1, std::min<uint64_t>(64, ((uint64_t)1 << 25) / std::max<uint64_t>(p, 1)))
                : 1;
        // candidates: buckets around the built-in side, the most lane-efficient
        // cell tiles under the LDS limit (for a batch, the four-workgroup
        // limit: one launch runs every tile at the largest LDS among them)
        std::vector<int> bms;
        for (int b : {bm0, bm0 - 1, bm0 + 1, bm0 + 2})
            if (b >= 1 && std::find(bms.begin(), bms.end(), b) == bms.end()) bms.push_back(b);
        std::vector<int> tcs{tc0};
        for (int c : rank_tile_cells(t0, bat ? kLdsBudget4 : kLdsBudget, bat ? 0 : ctx->n_cu)) {
            if (tcs.size() >= 5) break;
            if (std::find(tcs.begin(), tcs.end(), c) == tcs.end()) tcs.push_back(c);
        }
        const int f0 = i0.form;
        float ms_default = 0.f;
        st = time_setting(ctx, idx, descs, dX, dY, dZ, n, bat, replicas, bm0, tc0, 0, f0, reps, ms_default);
        if (st != TNNI_OK) return st;
        TuneChoice best;
        best.bucket_px = bm0;
        best.tile_cells = tc0;
        best.form = f0;
        float best_ms = ms_default;
        int tried = 1;
        auto try_setting = [&](int b, int c, int form) -> tnni_status {
            float ms = 0.f;
            tnni_status s2 = time_setting(ctx, idx, descs, dX, dY, dZ, n, bat, replicas, b, c, 0, form, reps, ms);
            if (s2 == TNNI_E_GEOMETRY) return TNNI_OK;  // e.g. a bucket grid too large
            if (s2 != TNNI_OK) return s2;
            ++tried;
            if (ms < best_ms) {
                best_ms = ms;
                best.bucket_px = b;
                best.tile_cells = c;
                best.form = form;
            }
            return TNNI_OK;
        };
        for (int b : bms)
            for (int c : tcs) {
                if (b == bm0 && c == tc0) continue;
                if ((st = try_setting(b, c, f0)) != TNNI_OK) return st;
            }
        // the other form (fused or split) at the built-in and the best
        // bucket side, with its own three most lane-efficient cell tiles
        int tc_split = f0 == TNNI_FORM_SPLIT ? tc0 : 0;  // the split form's built-in label tile
        {
            const int f1 = f0 == TNNI_FORM_SPLIT ? TNNI_FORM_FUSED : TNNI_FORM_SPLIT;
            tnni_tile_desc d1 = d0;
            d1.form = f1;
            TileDev t1;
            PrepInfo i1;
            if (prepare_tile(d1, n[idx[0]], ctx->n_cu, bat, nullptr, t1, i1, err) == TNNI_OK) {
                if (f1 == TNNI_FORM_SPLIT) tc_split = t1.tc;
                std::vector<int> tcs1{t1.tc};
                for (int c : rank_tile_cells(t1, bat ? kLdsBudget4 : kLdsBudget, bat ? 0 : ctx->n_cu)) {
                    if (tcs1.size() >= 3) break;
                    if (std::find(tcs1.begin(), tcs1.end(), c) == tcs1.end()) tcs1.push_back(c);
                }
                std::vector<int> bms1{bm0};
                if (best.bucket_px != bm0) bms1.push_back(best.bucket_px);
                for (int b : bms1)
                    for (int c : tcs1)
                        if ((st = try_setting(b, c, f1)) != TNNI_OK) return st;
            }
        }
        // the scatter form (integer height modes only): the split form's
        // labelling at the best split label tile found, accumulated by
        // spreading each pixel over the cells it reaches
        if (d0.accum_mode != TNNI_ACC_FP32_2SUM) {
            const int tc_s = best.form == TNNI_FORM_SPLIT ? best.tile_cells : tc_split;
            if (tc_s > 0 && (st = try_setting(best.bucket_px, tc_s, TNNI_FORM_SCATTER)) != TNNI_OK) return st;
        }
        // fewer workgroups per CU at the best sizes (kept only if clearly faster)
        {
            tnni_tile_desc db = d0;
            db.bucket_px = best.bucket_px;
            db.tile_cells = best.tile_cells;
            db.form = best.form;
            TileDev tb;
            PrepInfo ib;
            if ((st = prepare_tile(db, n[idx[0]], ctx->n_cu, bat, nullptr, tb, ib, err)) != TNNI_OK)
                return fail(ctx, st, err);
            const int occ = tile_kernel_occupancy(ib.lds);
            for (int wg = occ - 1; wg >= std::max(1, occ - 2); --wg) {
                float ms = 0.f;
                st = time_setting(ctx, idx, descs, dX, dY, dZ, n, bat, replicas, best.bucket_px,
                                  best.tile_cells, wg, best.form, reps, ms);
                if (st != TNNI_OK) return st;
                ++tried;
                if (ms < 0.98f * best_ms) {
                    best_ms = ms;
                    best.wg_per_cu = wg;
                }
            }
        }
        ctx->tuned.put(g.first, best);
        if (out) {
            tnni_tune_entry &e = out[written];
            e.shape_class = g.first;
            e.bucket_px = best.bucket_px;
            e.tile_cells = best.tile_cells;
            e.wg_per_cu = best.wg_per_cu;
            e.form = best.form;
            e.pad = 0;
            e.candidates = tried;
            e.ms_default = ms_default;
            e.ms_tuned = best_ms;
        }
        ++written;
    }
    if (n_classes) *n_classes = written;
    return TNNI_OK;
}

tnni_status tnni_tune_read(tnni_ctx *ctx, tnni_tune_entry *out, int cap, int *n)
{
    if (!ctx || !n || (cap > 0 && !out)) return fail(ctx, TNNI_E_INVALID, "null argument");
    const auto e = ctx->tuned.entries();
    *n = (int)e.size();
    for (int i = 0; i < std::min(cap, *n); ++i) {
        std::memset(&out[i], 0, sizeof out[i]);
        out[i].shape_class = e[(size_t)i].first;
        out[i].bucket_px = e[(size_t)i].second.bucket_px;
        out[i].tile_cells = e[(size_t)i].second.tile_cells;
        out[i].wg_per_cu = e[(size_t)i].second.wg_per_cu;
        out[i].form = e[(size_t)i].second.form;
    }
    return TNNI_OK;
}

tnni_status tnni_tune_load(tnni_ctx *ctx, const tnni_tune_entry *in, int n)
{
    if (!ctx || n < 0 || (n && !in)) return fail(ctx, TNNI_E_INVALID, "null argument");
    for (int i = 0; i < n; ++i)
        if (in[i].bucket_px < 0 || in[i].tile_cells < 0 || in[i].tile_cells > 64 || in[i].wg_per_cu < 0 ||
            in[i].form < 0 || in[i].form > TNNI_FORM_SCATTER)
            return fail(ctx, TNNI_E_INVALID, "tune entry " + std::to_string(i) + " out of range");
    for (int i = 0; i < n; ++i) {
        TuneChoice c;
        c.bucket_px = in[i].bucket_px;
        c.tile_cells = in[i].tile_cells;
        c.wg_per_cu = in[i].wg_per_cu;
        c.form = in[i].form;
        ctx->tuned.put(in[i].shape_class, c);
    }
    return TNNI_OK;
}

tnni_status tnni_tune_clear(tnni_ctx *ctx)
{
    if (!ctx) return fail(ctx, TNNI_E_INVALID, "null argument");
    ctx->tuned.clear();
    return TNNI_OK;
}

}  // extern "C"
