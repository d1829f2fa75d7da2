// Tile geometry on the host: the pixel frame on the fixed-point grid, the
// bucket grid and the cell-tile decomposition of one work package.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>
#include <string>
#include <tuple>
#include <vector>

#include "tnni_internal.h"

namespace tnni {
namespace {

// ceil(2^32 / d) for d >= 2; 0 marks d == 1 (no 32-bit multiplier)
uint32_t div_magic_of(uint32_t d)
{
    return d <= 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + d - 1) / d);
}

int64_t isqrt64(int64_t v)
{
    if (v <= 0) return 0;
    int64_t r = (int64_t)std::sqrt((double)v);
    while (r * r > v) --r;
    while ((r + 1) * (r + 1) <= v) ++r;
    return r;
}

}  // namespace

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
        // the query offsets (|dx|, |dy| <= Rqb) square into 32 bits as well
        if (2 * Rqb * Rqb >= (int64_t)kNoD2) continue;
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

// Shape class of a tile: what decides its best bucket and cell-tile sizes
// (algorithm, accumulation, scale, query span and site radius in pixels,
// grid side to a power of two, point density to half a power of two, and
// whether it runs alone or in a batch).
uint64_t shape_class(const tnni_tile_desc &d, uint64_t n, const tnni_frame_info &f, bool batched)
{
    const int qa = -floordiv_i(f.Rqb + f.x0, f.ps_u), qb = floordiv_i(f.Rqb - f.x0, f.ps_u);
    const int span = std::min(255, std::max(0, qb - qa + 1));
    const int rs_px = std::min(255, f.Rsb / std::max(1, f.ps_u));
    const int side = std::max(d.nrows, d.ncols);
    int gc = 0;
    while ((1 << gc) < side && gc < 31) ++gc;
    const double per_px = (double)n / ((double)f.W * (double)f.H);
    const int dc = n == 0 ? 0 : std::max(1, std::min(127, (int)std::lround(2.0 * std::log2(per_px)) + 64));
    return (uint64_t)(d.algorithm & 1) | ((uint64_t)(d.accum_mode & 3) << 1) |
           ((uint64_t)std::min(d.scale, 255) << 3) | ((uint64_t)span << 11) | ((uint64_t)rs_px << 19) |
           ((uint64_t)gc << 27) | ((uint64_t)dc << 32) | ((uint64_t)(batched ? 1 : 0) << 39);
}

// Set the cell-tile size of t to tc and return the LDS its items need: the
// per-wave stages plus the label tile (and search list) of the widest item.
size_t tile_lds_bytes(TileDev &t, int tc)
{
    t.tc = tc;
    if (t.label_only) {  // tc x tc pixel tiles
        t.ntx = (t.W + tc - 1) / tc;
        t.nty = (t.H + tc - 1) / tc;
        const size_t rw = (size_t)std::min(tc, t.W), rh = (size_t)std::min(tc, t.H);
        return (rw * rh * 10 + 3) / 4 * 4 + tile_kernel_stage_bytes();
    }
    t.ntx = (t.ncols + tc - 1) / tc;
    t.nty = (t.nrows + tc - 1) / tc;
    // an item's pixel width depends on its column only, its height on its
    // row only.  A query halo wider than tc cells clips items other than
    // the first at the raster edge, so every item within the halo's reach
    // of either edge is measured; the full items past them (neither first
    // nor last, halo inside the raster) all share the width of item `near`.
    const int near = (std::abs(query_lo(t)) + std::abs(query_hi(t))) / std::max(1, tc * t.scale) + 2;
    auto widest = [near](int n, auto extent) {
        int m = 0;
        for (int i = 0; i < n; ++i) {
            if (i > near && i < n - 1 - near) i = n - 1 - near;
            m = std::max(m, extent(i));
        }
        return m;
    };
    const int rw = widest(t.ntx, [&](int ti) {
        const Region g = region_of(t, ti, 0);
        return g.rk1 - g.rk0 + 1;
    });
    const int rh = widest(t.nty, [&](int tj) {
        const Region g = region_of(t, 0, tj);
        return g.rl1 - g.rl0 + 1;
    });
    // per pixel: d2 and value (4 + 4 bytes), a list slot (2 bytes)
    return ((size_t)rw * rh * 10 + 3) / 4 * 4 + tile_kernel_stage_bytes();
}

int query_span(const TileDev &t) { return std::max(1, query_hi(t) - query_lo(t) + 1); }

size_t gather_window_bytes(const TileDev &t, int gb)
{
    const size_t side = (size_t)(gb - 1) * t.scale + query_span(t);
    return side * side * 8;
}

size_t split_pass_lds(const TileDev &t)
{
    // after the window: scatter, an 8-byte sum and a 4-byte count per cell
    // of the block; a gather of blocks of fewer cells than threads in an
    // integer mode, the threads' 16-byte partial sums
    if (t.scatter) return (size_t)t.gb * t.gb * 12;  // the window is read from HBM
    size_t extra = 0;
    if (t.mode != TNNI_ACC_FP32_2SUM && t.gb * t.gb < kBlock) extra = (size_t)kBlock * 16;
    return ((gather_window_bytes(t, t.gb) + 15) & ~(size_t)15) + extra;
}

// Useful share of the label lanes of an interior item of tc cells: it labels
// (tc-1)*scale + span pixels a side in 8x8 wave blocks, four waves a
// workgroup, for the tc*scale pixels a side its cells own (a label-only
// item, which has no query halo, passes span = scale).
double tile_efficiency(int tc, int scale, int span)
{
    const int P = (tc - 1) * scale + span;
    const int nb = (P + 7) / 8, blocks = nb * nb;
    const double own = (double)tc * scale;
    return own * own / (64.0 * blocks) * (double)blocks / (4.0 * ((blocks + 3) / 4));
}

// Cell-tile sizes 1..64 whose items fit `budget` bytes of LDS, best lane
// efficiency first (ties: the larger tile).  min_items_cu > 0 (plain form):
// sizes above 16 that leave fewer than 2 * min_items_cu items go last.
namespace {

std::vector<int> rank_tile_cells_uncached(TileDev &t, size_t budget, int min_items_cu);

// The ranking reads only these fields of the tile (region_of,
// tile_lds_bytes, tile_efficiency), so tiles of one shape share it: a
// batch or a stream of work packages mostly repeats a few shapes, and the
// ranking is most of a tile's host preparation.  At most kRankMemo
// entries (then it starts afresh).
using RankKey = std::tuple<int, int, int, int, int, int, int, int, int, int, size_t, int>;
constexpr size_t kRankMemo = 4096;
std::mutex g_rank_mu;
std::map<RankKey, std::vector<int>> g_rank_memo;

}  // namespace

std::vector<int> rank_tile_cells(TileDev &t, size_t budget, int min_items_cu)
{
    const RankKey key(t.label_only, t.ncols, t.nrows, t.scale, t.W, t.H, t.Rqb, t.x0, t.ps_u, t.po, budget,
                      min_items_cu);
    {
        std::lock_guard<std::mutex> lk(g_rank_mu);
        auto it = g_rank_memo.find(key);
        if (it != g_rank_memo.end()) {
            if (!it->second.empty()) tile_lds_bytes(t, it->second[0]);
            return it->second;
        }
    }
    std::vector<int> out = rank_tile_cells_uncached(t, budget, min_items_cu);
    std::lock_guard<std::mutex> lk(g_rank_mu);
    if (g_rank_memo.size() >= kRankMemo) g_rank_memo.clear();
    g_rank_memo.emplace(key, out);
    return out;
}

namespace {

std::vector<int> rank_tile_cells_uncached(TileDev &t, size_t budget, int min_items_cu)
{
    if (t.label_only) {
        // pixel tiles of 8..64 pixels a side in whole 8x8 wave blocks; the
        // largest that fits, the plain form first those leaving at least
        // twice the CUs in items
        std::vector<int> big, few;
        for (int tc = 64; tc >= 8; tc -= 8) {
            if (tile_lds_bytes(t, tc) > budget) continue;
            const bool f = min_items_cu > 0 && (int64_t)t.ntx * t.nty < 2ll * min_items_cu;
            (f ? few : big).push_back(tc);
        }
        big.insert(big.end(), few.begin(), few.end());
        if (!big.empty()) tile_lds_bytes(t, big[0]);
        return big;
    }
    const int span = query_span(t);
    struct C {
        int tc;
        double e;
        bool few;
    };
    std::vector<C> c;
    for (int tc = 1; tc <= 64; ++tc) {
        if (tile_lds_bytes(t, tc) > budget) continue;
        const bool few = min_items_cu > 0 && tc > 16 && (int64_t)t.ntx * t.nty < 2ll * min_items_cu;
        c.push_back({tc, tile_efficiency(tc, t.scale, span), few});
    }
    std::stable_sort(c.begin(), c.end(), [](const C &a, const C &b) {
        if (a.few != b.few) return !a.few;
        if (std::fabs(a.e - b.e) > 1e-9) return a.e > b.e;
        return a.tc > b.tc;
    });
    std::vector<int> out;
    for (const C &x : c) out.push_back(x.tc);
    if (!out.empty()) tile_lds_bytes(t, out[0]);
    return out;
}

}  // namespace

// Fill everything of a TileDev that depends on the descriptor and the
// point count (not on the batch layout).  tuned (nullable): bucket and
// cell-tile sizes measured for the tile's shape class, used where the
// descriptor leaves them 0.
tnni_status prepare_tile(const tnni_tile_desc &d, uint64_t n, int n_cu, bool batched,
                         const TuneTable *tuned, TileDev &t, PrepInfo &info, std::string &err)
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
    if (d.form < TNNI_FORM_AUTO || d.form > TNNI_FORM_SCATTER) {
        err = "form must be 0 (automatic), 1 (fused), 2 (split) or 3 (scatter)";
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
    info = PrepInfo();
    info.shape_class = shape_class(d, n, f, batched);
    TuneChoice tc_tuned;
    if (tuned && (d.bucket_px <= 0 || d.tile_cells <= 0 || d.form <= 0) && tuned->lookup(info.shape_class, tc_tuned))
        info.wg_per_cu = tc_tuned.wg_per_cu;
    int bm = d.bucket_px > 0 ? d.bucket_px : tc_tuned.bucket_px;
    if (bm <= 0) {
        const double per_px = (double)n / ((double)f.W * (double)f.H);
        bm = per_px > 0 ? (int)std::lround(std::sqrt(1.0 / per_px)) : 64;
    }
    bm = std::max(1, std::min(bm, f.Rsb / f.ps_u + 1));
    // the band sort keeps one bucket row (and the bands of a tile) in LDS
    {
        const int64_t span = (int64_t)std::max(f.W, f.H) * f.ps_u + 2 * (int64_t)f.Rsb + f.ps_u;
        const int64_t need = (span + (int64_t)kBandBuckets * f.ps_u - 1) / ((int64_t)kBandBuckets * f.ps_u);
        bm = std::max<int>(bm, (int)need);
    }
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
    // bands of bucket rows: at most kBandBuckets buckets (the LDS counters of
    // one band-sort workgroup) and, at the tile's mean density, about
    // kBandPts points, so that dense tiles spread over many workgroups
    const int64_t per_row = std::max<int64_t>(1, (int64_t)n / nby);
    // rows enough to keep the tile within kMaxBands bands
    const int64_t br_min = (nby + kMaxBands - 1) / kMaxBands;
    const int64_t br_pts = std::max<int64_t>(br_min, kBandPts / per_row);
    t.br = (int)std::max<int64_t>(1, std::min<int64_t>(std::min<int64_t>(nby, kBandBuckets / nbx), br_pts));
    t.nbands = (int)((nby + t.br - 1) / t.br);
    t.br_magic = div_magic_of((uint32_t)t.br);
    if (nbx > kBandBuckets || t.nbands > kMaxBands) {
        err = "bucket grid too large for the band sort: raise bucket_px";
        return TNNI_E_GEOMETRY;
    }
    t.bm_magic = bm == 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)bm - 1) / (uint64_t)bm);
    t.bs_magic = div_magic_of((uint32_t)bs);
    // A site of ring 0 or 1 around a pixel's bucket is within 2*bs of the
    // pixel on both axes (strictly less than 2*bs).
    t.ring1_covered = d.algorithm == TNNI_ALG_PLANES ? (2 * bs <= f.Rs) : (8 * bs * bs <= f.Rs2);
    // More than kDenseSites sites per bucket at the tile's mean density: a
    // pixel's own bucket often settles its search, so the searches go ring
    // by ring and stop early; otherwise ring 1 is nearly always needed and
    // the searches visit whole squares of buckets (fewer, longer ranges).
    t.dense = (double)n * (double)bm * (double)bm > kDenseSites * (double)f.W * (double)f.H ? 1 : 0;
    // the form: fused (labels with the query halo, cells accumulated in the
    // same kernel) or split (each pixel labelled once into a raster, then a
    // gather pass).  Built in: split where label-only items use their lanes
    // at least kSplitGain times better than fused ones, i.e. where the
    // query halo (wide query windows at large scale) dominates the labelling
    int form = d.form > 0 ? d.form : tc_tuned.form;
    if (form == TNNI_FORM_SCATTER && d.accum_mode == TNNI_ACC_FP32_2SUM) {
        if (d.form == TNNI_FORM_SCATTER) {
            err = "the scatter form needs an integer accumulation mode (TNNI_ACC_FIXED64 or TNNI_ACC_S0)";
            return TNNI_E_INVALID;
        }
        form = 0;  // a tuned choice for another mode: decide below
    }
    if (form != TNNI_FORM_FUSED && form != TNNI_FORM_SPLIT && form != TNNI_FORM_SCATTER) {
        // (label-only pixel tiles use every lane of their 8x8 blocks)
        TileDev a = t;
        a.label_only = 0;
        std::vector<int> order = rank_tile_cells(a, kLdsBudget4, batched ? 0 : n_cu);
        const double eff = order.empty() ? 0.0 : tile_efficiency(order[0], t.scale, query_span(t));
        form = 1.0 >= kSplitGain * eff ? TNNI_FORM_SPLIT : TNNI_FORM_FUSED;
    }
    t.label_only = form == TNNI_FORM_SPLIT || form == TNNI_FORM_SCATTER ? 1 : 0;
    t.scatter = form == TNNI_FORM_SCATTER ? 1 : 0;
    t.gitem_base = 0;
    if (t.scatter) {
        // scatter blocks: only their sums in LDS; the largest of 32 .. 4
        // cells a side that leaves two blocks per CU on a tile of its own
        // (the window overlap of neighbouring blocks falls with the side,
        // the parallelism with it), with a window under 2^24 pixels
        auto blocks = [&](int gb) { return (int64_t)((t.ncols + gb - 1) / gb) * ((t.nrows + gb - 1) / gb); };
        auto window = [&](int gb) {
            const int64_t side = (int64_t)(gb - 1) * t.scale + query_span(t);
            return side * side;
        };
        t.gb = 32;
        while (t.gb > 4 && blocks(t.gb) < 2 * (int64_t)std::max(n_cu, 1)) t.gb /= 2;
        while (t.gb > 1 && window(t.gb) >= (int64_t)1 << 24) t.gb /= 2;
        if (window(t.gb) >= (int64_t)1 << 24) {
            err = "query window too wide for the scatter form";
            return TNNI_E_GEOMETRY;
        }
        t.gntx = (t.ncols + t.gb - 1) / t.gb;
        t.gnty = (t.nrows + t.gb - 1) / t.gb;
    } else if (t.label_only) {
        // gather blocks: the largest of 16 .. 1 cells a side whose pixel
        // window fits kGatherLds
        t.gb = 16;
        while (t.gb > 1 && split_pass_lds(t) > kGatherLds) t.gb /= 2;
        if (split_pass_lds(t) > kGatherLds) {
            err = "query window too wide for the split form's gather";
            return TNNI_E_GEOMETRY;
        }
        t.gntx = (t.ncols + t.gb - 1) / t.gb;
        t.gnty = (t.nrows + t.gb - 1) / t.gb;
    }
    info.form = form;
    // cell tiles: the most lane-efficient size that fits the LDS budget
    // of four workgroups per CU; the plain form avoids sizes that leave a
    // lone small tile with fewer items than twice the CUs
    int tc_fixed = d.tile_cells > 0 ? d.tile_cells : tc_tuned.tile_cells;
    if (t.label_only && tc_fixed > 0) tc_fixed = std::max(8, tc_fixed);  // pixels: at least one wave block
    size_t lds = 0;
    if (tc_fixed > 0) {
        lds = tile_lds_bytes(t, tc_fixed);
        if (lds > kLdsBudget) {
            err = "tile_cells too large for the LDS budget";
            return TNNI_E_GEOMETRY;
        }
    } else {
        std::vector<int> order = rank_tile_cells(t, kLdsBudget4, batched ? 0 : n_cu);
        if (order.empty()) order = rank_tile_cells(t, kLdsBudget, batched ? 0 : n_cu);
        if (order.empty()) {
            err = "no cell-tile size fits the LDS budget for this geometry";
            return TNNI_E_GEOMETRY;
        }
        lds = tile_lds_bytes(t, order[0]);
    }
    info.lds = lds;
    return TNNI_OK;
}

}  // namespace tnni
