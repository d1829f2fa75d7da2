// Host side of the tile engine: context, geometry, arenas, the plain and
// batched forms, and the C ABI of include/terranni.h.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstring>
#include <mutex>
#include <string>
#include <vector>

#include "tnni_internal.h"

using namespace tnni;

namespace {

thread_local std::string g_thread_err;

// ---- geometry ------------------------------------------------------------

int64_t isqrt64(int64_t v)
{
    if (v <= 0) return 0;
    int64_t r = (int64_t)std::sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

// The pixel frame of nni::performNNI (nni.cpp:879-886) and setupRun's
// glOrtho (nni.cpp:656-660), put on an integer grid of 1/upm metre.
tnni_status compute_frame(const tnni_tile_desc &d, tnni_frame_info &f, std::string &err)
{
    std::memset(&f, 0, sizeof f);
    if (d.algorithm != TNNI_ALG_PLANES && d.algorithm != TNNI_ALG_CONES) {
        err = "algorithm must be 0 (planes) or 1 (cones)";
        return TNNI_E_INVALID;
    }
    if (d.scale < 1 || d.nrows < 1 || d.ncols < 1 || !(d.cell_size > 0.0f) ||
        !(d.site_radius > 0.0f) || !(d.query_radius > 0.0f)) {
        err = "scale, nrows, ncols, cell_size, site_radius and query_radius must be positive";
        return TNNI_E_INVALID;
    }
    const double cs = (double)d.cell_size;
    const double qr = (double)d.query_radius;
    const double sr = (double)d.site_radius;
    const double ps = cs / d.scale;  // nni.cpp:881
    int po = d.pixel_offset >= 0 ? d.pixel_offset
                                 : (int)std::ceil((qr - std::floor(d.scale / 2.0) * ps) / ps) + 1;  // nni.cpp:883
    if (po < 0) po = 0;
    const int64_t W = (int64_t)d.ncols * d.scale + 2 * (int64_t)po;  // nni.cpp:885-886
    const int64_t H = (int64_t)d.nrows * d.scale + 2 * (int64_t)po;
    if (W > (1 << 20) || H > (1 << 20)) {
        err = "raster larger than 2^20 pixels on a side";
        return TNNI_E_GEOMETRY;
    }
    f.po = po;
    f.W = (int)W;
    f.H = (int)H;
    f.algorithm = d.algorithm;
    f.spacing_factor = (int)std::ceil((qr * 2 + 0.001) / (cs * 5));  // nni.cpp:879
    // the smallest unit of 1/(100 k) metre that puts pixel centres, queries
    // and cell edges on integers
    for (int t = 1; t <= 200; ++t) {
        int upm = 100 * t;
        if (d.units_per_m > 0) {
            if (t > 1) break;
            upm = d.units_per_m;
        }
        const double c = cs * upm;
        const int64_t ci = std::llround(c);
        if (ci <= 0 || std::fabs(c - (double)ci) > 1e-3 || ci % d.scale) continue;
        const int64_t ps_u = ci / d.scale;
        const int64_t x0n = (int64_t)(1 - 2 * po - d.scale) * ps_u;  // 2 * centre of pixel 0
        if (x0n % 2) continue;
        const int64_t Rs = (int64_t)std::floor(sr * upm / std::sqrt(2.0));  // nni.cpp:183
        const int64_t Rq = (int64_t)std::floor(qr * upm / std::sqrt(2.0));
        const double srd = sr * upm, qrd = qr * upm;
        const int64_t Rs2 = (int64_t)std::floor(srd * srd);
        const int64_t Rq2 = (int64_t)std::floor(qrd * qrd);
        const int64_t Rsb = d.algorithm == TNNI_ALG_PLANES ? Rs : isqrt64(Rs2);
        const int64_t Rqb = d.algorithm == TNNI_ALG_PLANES ? Rq : isqrt64(Rq2);
        if (d.algorithm == TNNI_ALG_PLANES ? (2 * Rs * Rs >= (int64_t)kNoD2) : (Rs2 >= (int64_t)kNoD2))
            continue;
        const int64_t ext = std::max(W, H) * ps_u + 2 * Rsb + 2 * Rqb;
        if (ext >= ((int64_t)1 << 30)) continue;
        f.upm = upm;
        f.ps_u = (int)ps_u;
        f.cs_u = (int)ci;
        f.x0 = f.y0 = (int)(x0n / 2);
        f.Rs = (int)Rs;
        f.Rq = (int)Rq;
        f.Rs2 = Rs2;
        f.Rq2 = Rq2;
        f.Rsb = (int)Rsb;
        f.Rqb = (int)Rqb;
        return TNNI_OK;
    }
    err = "cell size / scale / radii cannot be put on an exact fixed-point grid";
    return TNNI_E_GEOMETRY;
}

constexpr size_t kLdsBudget = 60 * 1024;  // two workgroups per CU, under the 64 KB per-group limit

// Fill everything of a TileDev that depends on the descriptor and the
// point count (not on the batch layout).  lds_out: bytes of LDS it needs.
tnni_status prepare_tile(const tnni_tile_desc &d, uint64_t n, int n_cu, bool batched, TileDev &t,
                         size_t &lds_out, std::string &err)
{
    tnni_frame_info f;
    tnni_status st = compute_frame(d, f, err);
    if (st != TNNI_OK) return st;
    if (d.accum_mode < TNNI_ACC_FIXED64 || d.accum_mode > TNNI_ACC_S0) {
        err = "accum_mode must be 0, 1 or 2";
        return TNNI_E_INVALID;
    }
    if (d.accum_mode == TNNI_ACC_S0 && d.multiplier < 1) {
        err = "TNNI_ACC_S0 needs multiplier >= 1";
        return TNNI_E_INVALID;
    }
    if (n > 0x7FFFFFFFull) {
        err = "more than 2^31-1 points in one tile";
        return TNNI_E_INVALID;
    }
    std::memset(&t, 0, sizeof t);
    t.n = (int64_t)n;
    t.sx = d.las_scale[0];
    t.sy = d.las_scale[1];
    t.sz = d.las_scale[2];
    t.ox = d.las_offset[0];
    t.oy = d.las_offset[1];
    t.oz = d.las_offset[2];
    t.origin_x = d.origin_x;
    t.origin_y = d.origin_y;
    t.upm = f.upm;
    t.ps_u = f.ps_u;
    t.cs_u = f.cs_u;
    t.po = f.po;
    t.W = f.W;
    t.H = f.H;
    t.x0 = f.x0;
    t.y0 = f.y0;
    t.Rs = f.Rs;
    t.Rq = f.Rq;
    t.Rsb = f.Rsb;
    t.Rqb = f.Rqb;
    t.Rs2 = (uint32_t)std::min<int64_t>(f.Rs2, 0xFFFFFFFFll);
    t.Rq2 = (uint32_t)std::min<int64_t>(f.Rq2, 0xFFFFFFFFll);
    t.algorithm = d.algorithm;
    t.mode = d.accum_mode;
    t.nrows = d.nrows;
    t.ncols = d.ncols;
    t.scale = d.scale;
    t.point_min = d.point_min;
    t.multiplier = d.multiplier;
    // sites further than the site region from every pixel centre are dropped
    t.lo_x = (double)f.x0 - f.Rsb;
    t.hi_x = (double)f.x0 + (double)(f.W - 1) * f.ps_u + f.Rsb;
    t.lo_y = (double)f.y0 - f.Rsb;
    t.hi_y = (double)f.y0 + (double)(f.H - 1) * f.ps_u + f.Rsb;
    // Bucket grid over that window: square buckets of bm x bm pixels, edges
    // on pixel edges, so pixel k sits in bucket column bK + k / bm.  bm is
    // set for about one site per bucket at the tile's mean density.
    int bm = d.bucket_px;
    if (bm <= 0) {
        const double per_px = (double)n / ((double)f.W * (double)f.H);
        bm = per_px > 0 ? (int)std::lround(std::sqrt(1.0 / per_px)) : 64;
    }
    bm = std::max(1, std::min(bm, f.Rsb / f.ps_u + 1));
    const int64_t bs = (int64_t)bm * f.ps_u;
    const int64_t half = f.ps_u / 2;
    const int64_t K = f.Rsb > half ? (f.Rsb - half + bs - 1) / bs : 0;
    const int64_t gx0 = (int64_t)f.x0 - half - K * bs;
    const int64_t last_x = (int64_t)f.x0 + (int64_t)(f.W - 1) * f.ps_u + f.Rsb;
    const int64_t last_y = (int64_t)f.y0 + (int64_t)(f.H - 1) * f.ps_u + f.Rsb;
    const int64_t nbx = (last_x - gx0) / bs + 1, nby = (last_y - gx0) / bs + 1;
    if (nbx * nby > (int64_t)1 << 30) {
        err = "bucket table too large: raise bucket_px";
        return TNNI_E_GEOMETRY;
    }
    t.gx0 = (int)gx0;
    t.gy0 = (int)gx0;
    t.bs = (int)bs;
    t.nbx = (int)nbx;
    t.nby = (int)nby;
    t.bm = bm;
    t.bK = (int)K;
    t.ecap = (int)std::min<int64_t>(250, f.Rsb / bs + 2);
    // cell tiles: the largest that fits the LDS budget; the plain form
    // shrinks them so that a lone small tile still spreads over the CUs
    auto lds_for = [&t, &d](int tc) {
        t.tc = tc;
        t.ntx = (d.ncols + tc - 1) / tc;
        t.nty = (d.nrows + tc - 1) / tc;
        int rw = 0, rh = 0;
        for (int ti = 0; ti < t.ntx; ++ti) {
            const Region g = region_of(t, ti, 0);
            rw = std::max(rw, g.rk1 - g.rk0 + 1);
        }
        for (int tj = 0; tj < t.nty; ++tj) {
            const Region g = region_of(t, 0, tj);
            rh = std::max(rh, g.rl1 - g.rl0 + 1);
        }
        return (size_t)rw * rh * 8;
    };
    size_t lds = 0;
    if (d.tile_cells > 0) {
        lds = lds_for(d.tile_cells);
        if (lds > kLdsBudget) {
            err = "tile_cells too large for the LDS budget";
            return TNNI_E_GEOMETRY;
        }
    } else {
        static const int kTc[] = {64, 48, 32, 24, 16, 12, 8, 6, 4, 3, 2, 1};
        int chosen = -1;
        for (int tc : kTc) {
            if (lds_for(tc) > kLdsBudget) continue;
            chosen = tc;
            if (!batched && tc > 16 && (int64_t)t.ntx * t.nty < 2ll * n_cu) continue;
            break;
        }
        if (chosen < 0) {
            err = "no cell-tile size fits the LDS budget for this geometry";
            return TNNI_E_GEOMETRY;
        }
        lds = lds_for(chosen);
    }
    lds_out = lds;
    return TNNI_OK;
}

// ---- device arena --------------------------------------------------------

struct DevBuf {
    void *p = nullptr;
    size_t cap = 0;
};

}  // namespace

struct tnni_ctx {
    int device = 0;
    int n_cu = 0;
    hipStream_t stream = nullptr;
    std::mutex mu;
    std::string err;
    uint64_t budget = 0;
    uint64_t allocated = 0;

    // pipeline buffers
    DevBuf tiles, chunks, batch, counts, rank, partials, sorted, hdist, edist;
    // staging for the host-pointer API
    DevBuf sx, sy, sz, sh, sl, sd;
    // pinned host mirrors of the descriptors
    void *h_desc = nullptr;
    size_t h_desc_cap = 0;

    bool prof = false;
    hipEvent_t ev0 = nullptr, ev1 = nullptr;
    uint64_t prof_launches = 0;
    double prof_ms = 0.0;

    std::vector<void *> user_allocs;
};

namespace {

tnni_status fail(tnni_ctx *ctx, tnni_status st, const std::string &msg)
{
    if (ctx) ctx->err = msg;
    g_thread_err = msg;
    return st;
}

tnni_status hip_fail(tnni_ctx *ctx, hipError_t e, const char *what)
{
    std::string m = std::string(what) + ": " + hipGetErrorString(e);
    return fail(ctx, e == hipErrorOutOfMemory ? TNNI_E_NOMEM : TNNI_E_HIP, m);
}

#define HIP_TRY(ctx, expr)                                 \
    do {                                                   \
        hipError_t e_ = (expr);                            \
        if (e_ != hipSuccess) return hip_fail(ctx, e_, #expr); \
    } while (0)

tnni_status ensure(tnni_ctx *ctx, DevBuf &b, size_t bytes)
{
    if (bytes == 0) bytes = 256;
    if (b.cap >= bytes) return TNNI_OK;
    size_t want = bytes + bytes / 4;
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

void release(DevBuf &b)
{
    if (b.p) (void)hipFree(b.p);
    b.p = nullptr;
    b.cap = 0;
}

// One prepared tile of a batch, before the batch layout is assigned.
struct Prepared {
    TileDev t;
    size_t lds;
};

// Lay the batch out (bases, chunks, items), upload the descriptors and run
// the pipeline.  Tiles keep their order; the caller orders them.
tnni_status run_prepared(tnni_ctx *ctx, std::vector<Prepared> &tiles, bool sync)
{
    if (tiles.empty()) return TNNI_OK;
    const int alg = tiles[0].t.algorithm, mode = tiles[0].t.mode;
    int64_t pts = 0, bks = 0, items = 0, rows = 0;
    size_t lds = 0;
    std::vector<Chunk> chunks;
    for (size_t k = 0; k < tiles.size(); ++k) {
        TileDev &t = tiles[k].t;
        if (t.algorithm != alg || t.mode != mode)
            return fail(ctx, TNNI_E_INVALID, "a batch must share algorithm and accum_mode");
        t.pt_base = pts;
        t.bk_base = bks;
        t.item_base = items;
        t.row_base = rows;
        rows += t.nby;
        for (int64_t b = 0; b < t.n; b += kChunkPts) {
            Chunk c;
            c.tile = (int)k;
            c.pad = 0;
            c.begin = b;
            c.end = std::min<int64_t>(t.n, b + kChunkPts);
            chunks.push_back(c);
        }
        pts += t.n;
        bks += (int64_t)t.nbx * t.nby;
        items += (int64_t)t.ntx * t.nty;
        lds = std::max(lds, tiles[k].lds);
    }
    if (bks + 1 > 0xFFFFFFFFll || pts > 0xFFFFFFFFll)
        return fail(ctx, TNNI_E_INVALID, "batch exceeds 2^32 buckets or points");
    if (chunks.empty()) {
        Chunk c{0, 0, 0, 0};
        chunks.push_back(c);
    }
    tnni_status st;
    if ((st = ensure(ctx, ctx->tiles, tiles.size() * sizeof(TileDev))) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->chunks, chunks.size() * sizeof(Chunk))) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->batch, sizeof(BatchDev))) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->counts, (size_t)(bks + 1) * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->rank, (size_t)pts * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->partials, kScanBlocks * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->sorted, (size_t)pts * 16)) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->hdist, (size_t)bks)) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->edist, (size_t)bks)) != TNNI_OK) return st;

    BatchDev bd;
    std::memset(&bd, 0, sizeof bd);
    bd.n_tiles = (int)tiles.size();
    bd.n_chunks = (int)chunks.size();
    bd.n_scan = (uint32_t)(bks + 1);
    bd.n_items = items;
    bd.n_rows = rows;
    bd.scan_chunk = (uint32_t)((bd.n_scan + kScanBlocks - 1) / kScanBlocks);

    // descriptors through one pinned staging block
    const size_t sz_t = tiles.size() * sizeof(TileDev);
    const size_t sz_c = chunks.size() * sizeof(Chunk);
    const size_t need = sz_t + sz_c + sizeof(BatchDev);
    if (ctx->h_desc_cap < need) {
        if (ctx->h_desc) (void)hipHostFree(ctx->h_desc);
        ctx->h_desc = nullptr;
        ctx->h_desc_cap = 0;
        HIP_TRY(ctx, hipHostMalloc(&ctx->h_desc, need * 2, hipHostMallocDefault));
        ctx->h_desc_cap = need * 2;
    }
    // the staging block may still feed a copy of the previous batch
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    char *h = (char *)ctx->h_desc;
    for (size_t k = 0; k < tiles.size(); ++k) std::memcpy(h + k * sizeof(TileDev), &tiles[k].t, sizeof(TileDev));
    std::memcpy(h + sz_t, chunks.data(), sz_c);
    std::memcpy(h + sz_t + sz_c, &bd, sizeof bd);
    HIP_TRY(ctx, hipMemcpyAsync(ctx->tiles.p, h, sz_t, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->chunks.p, h + sz_t, sz_c, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipMemcpyAsync(ctx->batch.p, h + sz_t + sz_c, sizeof bd, hipMemcpyHostToDevice, ctx->stream));

    LaunchCfg cfg;
    cfg.fill_grid = (int)std::min<int64_t>(1024, (bks + 1 + kBlock - 1) / kBlock);
    cfg.sort_grid = (int)std::min<size_t>(4096, chunks.size());
    cfg.row_grid = (int)std::max<int64_t>(1, std::min<int64_t>(8192, rows));
    const int occ = tile_kernel_occupancy(lds);
    cfg.main_grid = (int)std::max<int64_t>(1, std::min<int64_t>(items, (int64_t)ctx->n_cu * occ));
    cfg.main_lds = lds;
    cfg.algorithm = alg;
    cfg.mode = mode;
    DevBuffers db;
    db.batch = (BatchDev *)ctx->batch.p;
    db.tiles = (const TileDev *)ctx->tiles.p;
    db.chunks = (const Chunk *)ctx->chunks.p;
    db.counts = (uint32_t *)ctx->counts.p;
    db.rank = (uint32_t *)ctx->rank.p;
    db.partials = (uint32_t *)ctx->partials.p;
    db.sorted = (int4 *)ctx->sorted.p;
    db.hdist = (uint8_t *)ctx->hdist.p;
    db.edist = (uint8_t *)ctx->edist.p;
    HIP_TRY(ctx, enqueue_pipeline(db, cfg, ctx->stream, ctx->prof ? ctx->ev0 : nullptr,
                                  ctx->prof ? ctx->ev1 : nullptr));
    if (sync || ctx->prof) HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    if (ctx->prof) {
        float ms = 0.f;
        HIP_TRY(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
        ctx->prof_ms += ms;
        ctx->prof_launches++;
    }
    return TNNI_OK;
}

}  // namespace

// ---- C ABI -----------------------------------------------------------------

extern "C" {

int tnni_abi_version(void) { return TNNI_ABI_VERSION; }

const char *tnni_last_error(const tnni_ctx *ctx)
{
    if (ctx) return ctx->err.c_str();
    return g_thread_err.c_str();
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
    size_t lds;
    std::string err;
    tnni_status st = prepare_tile(*desc, n_points, 256, true, t, lds, err);
    if (st != TNNI_OK) return fail(nullptr, st, err);
    uint64_t b = 0;
    b += n_points * 12;                         // X, Y, Z
    b += n_points * (4 + 16);                   // rank + site records
    b += ((uint64_t)t.nbx * t.nby + 1) * 6;     // bucket offsets + distance maps
    b += (uint64_t)t.nrows * t.ncols * 4;       // heights
    if (with_labels) b += (uint64_t)t.W * t.H * 8;  // labels + d2
    b += sizeof(TileDev) + ((n_points + kChunkPts - 1) / kChunkPts + 1) * sizeof(Chunk);
    *bytes = b;
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
    if ((e = hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking)) != hipSuccess ||
        (e = hipEventCreate(&ctx->ev0)) != hipSuccess || (e = hipEventCreate(&ctx->ev1)) != hipSuccess) {
        tnni_status st = hip_fail(nullptr, e, "stream/event creation");
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
    if (ctx->stream) (void)hipStreamSynchronize(ctx->stream);
    DevBuf *bufs[] = {&ctx->tiles, &ctx->chunks, &ctx->batch, &ctx->counts, &ctx->rank,
                      &ctx->partials, &ctx->sorted, &ctx->hdist, &ctx->edist, &ctx->sx, &ctx->sy, &ctx->sz,
                      &ctx->sh, &ctx->sl, &ctx->sd};
    for (DevBuf *b : bufs) release(*b);
    for (void *p : ctx->user_allocs) (void)hipFree(p);
    if (ctx->h_desc) (void)hipHostFree(ctx->h_desc);
    if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
    if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
    if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
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
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    HIP_TRY(ctx, hipFree(dptr));
    return TNNI_OK;
}

tnni_status tnni_h2d(tnni_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
    return TNNI_OK;
}

tnni_status tnni_d2h(tnni_ctx *ctx, void *dst, const void *src, uint64_t bytes)
{
    if (!ctx || (!dst && bytes) || (!src && bytes)) return fail(ctx, TNNI_E_INVALID, "null argument");
    std::lock_guard<std::mutex> lk(ctx->mu);
    HIP_TRY(ctx, hipSetDevice(ctx->device));
    HIP_TRY(ctx, hipMemcpyAsync(dst, src, bytes, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
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
    tnni_status st = prepare_tile(*desc, n, ctx->n_cu, false, v[0].t, v[0].lds, err);
    if (st != TNNI_OK) return fail(ctx, st, err);
    v[0].t.X = dX;
    v[0].t.Y = dY;
    v[0].t.Z = dZ;
    v[0].t.out_h = d_out_h;
    v[0].t.out_labels = d_out_labels;
    v[0].t.out_d2 = d_out_d2;
    return run_prepared(ctx, v, true);
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
    tnni_status st = prepare_tile(*desc, n, ctx->n_cu, false, v[0].t, v[0].lds, err);
    if (st != TNNI_OK) return fail(ctx, st, err);
    TileDev &t = v[0].t;
    const size_t cells = (size_t)t.nrows * t.ncols, pix = (size_t)t.W * t.H;
    if ((st = ensure(ctx, ctx->sx, n * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->sy, n * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->sz, n * 4)) != TNNI_OK) return st;
    if ((st = ensure(ctx, ctx->sh, cells * 4)) != TNNI_OK) return st;
    if (out_labels && (st = ensure(ctx, ctx->sl, pix * 4)) != TNNI_OK) return st;
    if (out_d2 && (st = ensure(ctx, ctx->sd, pix * 4)) != TNNI_OK) return st;
    if (n) {
        HIP_TRY(ctx, hipMemcpyAsync(ctx->sx.p, X, n * 4, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->sy.p, Y, n * 4, hipMemcpyHostToDevice, ctx->stream));
        HIP_TRY(ctx, hipMemcpyAsync(ctx->sz.p, Z, n * 4, hipMemcpyHostToDevice, ctx->stream));
    }
    t.X = (const int32_t *)ctx->sx.p;
    t.Y = (const int32_t *)ctx->sy.p;
    t.Z = (const int32_t *)ctx->sz.p;
    t.out_h = (float *)ctx->sh.p;
    t.out_labels = out_labels ? (int32_t *)ctx->sl.p : nullptr;
    t.out_d2 = out_d2 ? (uint32_t *)ctx->sd.p : nullptr;
    if ((st = run_prepared(ctx, v, false)) != TNNI_OK) return st;
    HIP_TRY(ctx, hipMemcpyAsync(out_h, ctx->sh.p, cells * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (out_labels)
        HIP_TRY(ctx, hipMemcpyAsync(out_labels, ctx->sl.p, pix * 4, hipMemcpyDeviceToHost, ctx->stream));
    if (out_d2) HIP_TRY(ctx, hipMemcpyAsync(out_d2, ctx->sd.p, pix * 4, hipMemcpyDeviceToHost, ctx->stream));
    HIP_TRY(ctx, hipStreamSynchronize(ctx->stream));
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
        tnni_status st = prepare_tile(descs[k], n[k], ctx->n_cu, true, v[k].t, v[k].lds, err);
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
    // largest work first: the persistent kernel drains big tiles early and
    // fills the tail with small ones
    std::stable_sort(v.begin(), v.end(), [](const Prepared &a, const Prepared &b) {
        const double ca = (double)a.t.W * a.t.H + 8.0 * (double)a.t.n;
        const double cb = (double)b.t.W * b.t.H + 8.0 * (double)b.t.n;
        return ca > cb;
    });
    return run_prepared(ctx, v, true);
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
