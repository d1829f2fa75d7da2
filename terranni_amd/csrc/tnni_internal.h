// Internal types shared by the gfx950 kernels and the host engine.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include <map>
#include <mutex>
#include <string>
#include <utility>
#include <vector>

#include "terranni.h"

namespace tnni {

constexpr int kBlock = 256;            // threads per workgroup (4 wave64)
constexpr int kWaves = kBlock / 64;
constexpr int kChunkPts = 16384;       // points per chunk of the sort passes (at most)
constexpr int kChunkMin = 1024;        // points per chunk of the sort passes (at least)
constexpr int kSortBlock = 1024;       // threads of a sort-pass workgroup
constexpr int kSortCap = 7680;         // records a band sorts in LDS (larger bands: two passes)
constexpr uint32_t kNoD2 = 0xFFFFFFFFu;
constexpr int kScanBlocks = 512;       // fixed grid of the device-wide scan
constexpr int kScanItems = 4;          // items per thread per scan step
constexpr int kMaxBands = 8192;        // bucket-row bands per tile (LDS histogram)
constexpr int kBandBuckets = 8192;     // buckets per band (LDS counters)
constexpr int kBandPts = 4096;         // points per band aimed at (mean density)
constexpr double kDenseSites = 2.0;    // sites per bucket above which searches go ring by ring
constexpr double kSplitGain = 1.6;     // label-lane efficiency gain above which tiles take the split form

// Everything a kernel needs to know about one tile of a batch.  Lengths are
// in the tile-local fixed-point unit (1/upm metre).
struct TileDev {
    const int32_t *X, *Y, *Z;   // LAS int32 coordinates (device)
    int64_t n;                  // points
    int64_t pt_base;            // first point of this tile in the batch
    double sx, sy, sz, ox, oy, oz;
    double origin_x, origin_y;
    double lo_x, hi_x, lo_y, hi_y;  // window of sites that can cover a pixel
    int upm, ps_u, cs_u, po, W, H, x0, y0;
    int Rs, Rq, Rsb, Rqb;
    uint32_t Rs2, Rq2;
    int algorithm, mode, nrows, ncols, scale;
    int point_min, multiplier;
    // bucket grid: square buckets of bm x bm pixels (side bs units) aligned
    // to pixel edges; pixel k lies in bucket column bK + k / bm
    int gx0, gy0, bs, nbx, nby;
    int bm, bK;
    int ecap;                   // cap of the empty-distance map
    uint32_t bm_magic;          // ceil(2^32 / bm), 0 for bm == 1
    int ring1_covered;          // every site of rings 0-1 covers the pixel
    int dense;                  // buckets hold several sites: ring-by-ring searches
    uint32_t bs_magic;          // ceil(2^32 / bs), 0 for bs == 1
    uint32_t br_magic;          // ceil(2^32 / br), 0 for br == 1
    int64_t bk_base;            // first bucket of this tile in the batch
    int64_t e_base;             // first byte of its distance maps (4-aligned)
    int nbx4;                   // bytes per row of the distance maps (nbx rounded up to 4)
    int pad3;
    int64_t row_base;           // first bucket row of this tile in the batch
    int br, nbands;             // bucket rows per band, bands
    int64_t band_base;          // first band of this tile in the batch
    int64_t hist_base;          // first band-histogram entry of this tile
    int chunk_base, nchunks;    // this tile's chunks of points
    int tc, ntx, nty;           // cell-tile side and tiles along x / y
    // label_only: items are tc x tc pixel tiles (tc in pixels, no query
    // halo) that label their pixels and write the site values (out_val) and
    // squared distances (out_d2) to device memory instead of accumulating
    // cells; the split form then accumulates the cells from those rasters
    // (k_gather_split) in blocks of gb x gb cells (gntx x gnty blocks), its
    // items [gitem_base, gitem_base + gntx * gnty) of the batch's gather
    // items, by gathering (each cell tests its query window) or, where
    // scatter is set, by scattering (each pixel adds to the cells its
    // nearest-site distance reaches)
    int label_only;
    int64_t item_base;          // first work item of this tile
    int64_t gitem_base;         // first gather item (tiles without: an empty range)
    int32_t *out_val;
    int gb, gntx, gnty;         // gather blocks: side in cells, blocks along x / y
    int scatter;                // split tile accumulated by the scatter pass
    // space-time slice (T != nullptr): the sites are the points with
    // |slice_t - T[i]| <= time_radius + 1e-6, at squared time distance dt2
    // (units^2) added to their squared distance; queries add jj2; acc holds
    // every cell's sums between the slices of one tile
    const int32_t *T;
    float slice_t, time_radius;
    uint32_t jj2;
    int acc_load, acc_final;    // start from acc / divide into out_h
    int pad2;
    uint4 *acc;
    float *out_h;
    int32_t *out_labels;
    uint32_t *out_d2;
};

// Per-wave LDS staging of the tile kernel: the sites around one 8x8 block
// of pixels, i.e. the block's buckets plus one bucket on every side (rings
// 0 and 1 of every pixel of the block).  A block spans at most 8 buckets a
// side (buckets are at least one pixel), so the window is at most kWinMax
// buckets a side; woff holds each window row's bucket offsets plus the end
// of its last bucket (kWinMax + 1 a row).
constexpr int kWinMax = 10;      // window side in buckets (8 pixels + 2)
constexpr int kStageCap = 256;   // staged sites per wave
template <int CAP>
struct WaveStageT {
    uint32_t woff[kWinMax * (kWinMax + 1)];  // bucket offsets of the window rows
    uint32_t rowpre[kWinMax];                // first staged slot of each row
    uint32_t pad[2];
    int4 sites[CAP];
};
using WaveStage = WaveStageT<kStageCap>;
static_assert(sizeof(WaveStage) % 16 == 0, "WaveStage must keep int4 alignment");

// A run of points of one tile, the unit of work of the sort passes.
struct Chunk {
    int tile;
    int pad;
    int64_t begin, end;
};

struct BatchDev {
    int n_tiles;
    int n_chunks;
    uint32_t n_scan;       // band-histogram entries of the whole batch + 1
    uint32_t work_counter; // next work item of the persistent kernel
    int64_t n_items;
    int64_t n_rows;        // bucket rows of the whole batch
    int64_t n_bands;       // bands of the whole batch
    uint32_t scan_chunk;   // elements per scan block
    uint32_t chunk_pts;    // points per chunk of the sort passes
    int64_t n_gather_items;  // cell blocks the split form's gather pass accumulates
};

// Points per chunk of the sort passes for a batch of pts points on n_cu
// CUs: the least power of two from kChunkMin up to kChunkPts that gives no
// more chunks than CUs, so that a small tile's count and scatter passes
// spread over the chip (100k points: 98 chunks of 1024 instead of 7 of
// 16384) while large batches keep kChunkPts (longer runs per band).
__host__ __device__ inline int chunk_points(int64_t pts, int n_cu)
{
    const int64_t per = pts / (n_cu > 0 ? n_cu : 1);
    int c = kChunkMin;
    while (c < kChunkPts && c < per) c <<= 1;
    return c;
}

__host__ __device__ inline int floordiv_i(int a, int b)  // b > 0
{
    int q = a / b;
    if ((a % b) != 0 && a < 0) --q;
    return q;
}

// The cells of one work item and the pixels it labels.
struct Region {
    int i0, i1, j0, j1;      // cells [i0,i1) x [j0,j1)
    int rk0, rk1, rl0, rl1;  // pixels labelled into LDS (inclusive)
    int ok0, ok1, ol0, ol1;  // pixels whose label this item writes out
};

// Pixel k lies in the query region of cell i iff k - i*scale is in
// [qa, qb]: |x0 + (k - i*scale)*ps| <= Rqb.
__host__ __device__ inline int query_lo(const TileDev &t) { return -floordiv_i(t.Rqb + t.x0, t.ps_u); }
__host__ __device__ inline int query_hi(const TileDev &t) { return floordiv_i(t.Rqb - t.x0, t.ps_u); }

__host__ __device__ inline Region region_of(const TileDev &t, int ti, int tj)
{
    Region g;
    if (t.label_only) {  // a tc x tc pixel tile: no cells, no query halo
        g.i0 = g.i1 = g.j0 = g.j1 = 0;
        g.rk0 = g.ok0 = ti * t.tc;
        g.rk1 = g.ok1 = g.rk0 + t.tc - 1 < t.W - 1 ? g.rk0 + t.tc - 1 : t.W - 1;
        g.rl0 = g.ol0 = tj * t.tc;
        g.rl1 = g.ol1 = g.rl0 + t.tc - 1 < t.H - 1 ? g.rl0 + t.tc - 1 : t.H - 1;
        return g;
    }
    g.i0 = ti * t.tc;
    g.i1 = g.i0 + t.tc < t.ncols ? g.i0 + t.tc : t.ncols;
    g.j0 = tj * t.tc;
    g.j1 = g.j0 + t.tc < t.nrows ? g.j0 + t.tc : t.nrows;
    const int qa = query_lo(t), qb = query_hi(t);
    int kg0 = g.i0 * t.scale + qa, kg1 = (g.i1 - 1) * t.scale + qb;
    int lg0 = g.j0 * t.scale + qa, lg1 = (g.j1 - 1) * t.scale + qb;
    kg0 = kg0 < 0 ? 0 : kg0;
    lg0 = lg0 < 0 ? 0 : lg0;
    kg1 = kg1 > t.W - 1 ? t.W - 1 : kg1;
    lg1 = lg1 > t.H - 1 ? t.H - 1 : lg1;
    // every pixel of the raster is owned by exactly one item: the cell
    // columns' own pixels, plus the border for the first and last items
    g.ok0 = (g.i0 == 0) ? 0 : t.po + g.i0 * t.scale;
    g.ok1 = (g.i1 == t.ncols) ? t.W - 1 : t.po + g.i1 * t.scale - 1;
    g.ol0 = (g.j0 == 0) ? 0 : t.po + g.j0 * t.scale;
    g.ol1 = (g.j1 == t.nrows) ? t.H - 1 : t.po + g.j1 * t.scale - 1;
    g.rk0 = kg0 < g.ok0 ? kg0 : g.ok0;
    g.rk1 = kg1 > g.ok1 ? kg1 : g.ok1;
    g.rl0 = lg0 < g.ol0 ? lg0 : g.ol0;
    g.rl1 = lg1 > g.ol1 ? lg1 : g.ol1;
    return g;
}

// Host-side launch helpers (tnni_kernels.hip).
struct LaunchCfg {
    int sort_grid;      // blocks of the band count/scatter passes
    int band_grid;      // blocks of the band sort
    int main_grid;      // blocks of the persistent tile kernel
    int row_grid;       // blocks of the empty-distance passes
    size_t main_lds;    // dynamic LDS bytes of the tile kernel
    int algorithm;      // TNNI_ALG_* of every tile in the batch
    int mode;           // TNNI_ACC_* of every tile in the batch
    int time;           // space-time slices (every tile has T)
    int label_only;     // some tile of the batch is a label-only tile
    int dense;          // some tile of the batch is dense (ring-by-ring searches)
    int gather_grid;    // blocks of the split form's gather pass (0: none)
    size_t gather_lds;  // its dynamic LDS: the largest pixel window of a cell block
};

struct DevBuffers {
    BatchDev *batch;
    const TileDev *tiles;
    const Chunk *chunks;
    uint32_t *hist;     // n_scan band-histogram entries, scanned in place
    uint32_t *partials; // kScanBlocks
    int4 *banded;       // one per point: sites in band order
    int4 *sorted;       // one per point: sites in bucket order
    uint32_t *offs;     // first site of every bucket, + 1 sentinel
    uint8_t *hdist;     // per bucket: distance to a non-empty bucket in its row
    uint8_t *edist;     // per bucket: Chebyshev distance to a non-empty bucket
                        // (both: rows of nbx4 bytes from the tile's e_base)
};

// Enqueue the whole pipeline on `stream`.  ev_main0/ev_main1 (nullable)
// bracket the tile kernel.
hipError_t enqueue_pipeline(const DevBuffers &b, const LaunchCfg &cfg,
                            hipStream_t stream, hipEvent_t ev_main0,
                            hipEvent_t ev_main1);

// Write the descriptors of a one-tile batch (tile, chunks, batch header)
// into device memory with a kernel (no host staging, no copies).
hipError_t launch_single_tile_setup(const TileDev &t, const BatchDev &bd, TileDev *tiles, Chunk *chunks,
                                    BatchDev *batch, hipStream_t s);

// Space-time year layers: a slice's values and squared distances combined
// from the layers of its
// times (ascending, dt2 each); the query gather of a slice (values and
// squared distances of its W x H pixels) into t.acc / t.out_h; whether the
// points' times are out of order, and the run of each layer's time.
constexpr int kMaxLayers = 64;
struct LayerSet {
    int n;
    uint32_t dt2[kMaxLayers];
    const int32_t *val[kMaxLayers];
    const uint32_t *d2[kMaxLayers];
};
hipError_t launch_combine_layers(const LayerSet &ls, uint64_t pix, int32_t *val, uint32_t *d2, hipStream_t s);
hipError_t launch_gather_slice(const TileDev &t, const int32_t *val, const uint32_t *d2, hipStream_t s);
struct LayerTimes {
    int n;
    float t[kMaxLayers];
};
// out: 1 + 2 * kMaxLayers words, filled by the caller with 0, then
// kMaxLayers times n, then kMaxLayers times n.
hipError_t launch_time_runs(const int32_t *T, uint64_t n, const LayerTimes &ys, uint32_t *out, hipStream_t s);

// Largest workgroups per CU of the tile kernel at the given LDS size.
int tile_kernel_occupancy(size_t lds_bytes);
// LDS of the per-wave site stages of the tile kernel (in front of the
// label tile).
size_t tile_kernel_stage_bytes();

// ---- host geometry (tnni_geometry.hip) ----
// LDS of one workgroup: the per-wave stages plus the label tile, under the
// 64 KB per-group limit (static LDS takes a few more bytes)
constexpr size_t kLdsBudget = 60 * 1024;
constexpr size_t kLdsBudget4 = 32 * 1024;  // five workgroups per CU (160 KB LDS)
constexpr size_t kGatherLds = 80 * 1024;   // a gather block's pixel window (two workgroups per CU)

// The pixel window (value and squared distance, 8 bytes a pixel) of an
// interior block of gb x gb cells of the split form's gather.
size_t gather_window_bytes(const TileDev &t, int gb);
// LDS of the split form's accumulation pass for one block of t.gb cells:
// the pixel window, plus the block's cell sums where it scatters.
size_t split_pass_lds(const TileDev &t);

tnni_status compute_frame(const tnni_tile_desc &d, tnni_frame_info &f, std::string &err);

// Bucket side, cell-tile side and workgroups per CU chosen for one shape
// class (0 = the built-in choice).
struct TuneChoice {
    int bucket_px = 0, tile_cells = 0, wg_per_cu = 0;
    int form = 0;  // TNNI_FORM_*, 0: the built-in rule
};

// Measured choices per shape class; shared by the forms of one context.
class TuneTable {
public:
    bool lookup(uint64_t cls, TuneChoice &c) const
    {
        std::lock_guard<std::mutex> lk(mu_);
        auto it = map_.find(cls);
        if (it == map_.end()) return false;
        c = it->second;
        return true;
    }
    void put(uint64_t cls, const TuneChoice &c)
    {
        std::lock_guard<std::mutex> lk(mu_);
        map_[cls] = c;
    }
    void clear()
    {
        std::lock_guard<std::mutex> lk(mu_);
        map_.clear();
    }
    std::vector<std::pair<uint64_t, TuneChoice>> entries() const
    {
        std::lock_guard<std::mutex> lk(mu_);
        return std::vector<std::pair<uint64_t, TuneChoice>>(map_.begin(), map_.end());
    }

private:
    mutable std::mutex mu_;
    std::map<uint64_t, TuneChoice> map_;
};

// What prepare_tile derives besides the TileDev.
struct PrepInfo {
    size_t lds = 0;            // dynamic LDS of the tile kernel for this tile
    int wg_per_cu = 0;         // tuned workgroups per CU (0 = all that fit)
    uint64_t shape_class = 0;  // the tile's shape class
    int form = 0;              // TNNI_FORM_FUSED or TNNI_FORM_SPLIT
};

uint64_t shape_class(const tnni_tile_desc &d, uint64_t n, const tnni_frame_info &f, bool batched);
size_t tile_lds_bytes(TileDev &t, int tc);
double tile_efficiency(int tc, int scale, int span);
// The query window of a cell in pixels (a side).
int query_span(const TileDev &t);
std::vector<int> rank_tile_cells(TileDev &t, size_t budget, int min_items_cu);

// Fill everything of a TileDev that depends on the descriptor and the point
// count (not on the batch layout).  batched: the tile shares the GPU with
// others, so its cell tiles need not be shrunk to fill the CUs.  tuned
// (nullable): measured sizes for the tile's shape class.
tnni_status prepare_tile(const tnni_tile_desc &d, uint64_t n, int n_cu, bool batched,
                         const TuneTable *tuned, TileDev &t, PrepInfo &info, std::string &err);

}  // namespace tnni
