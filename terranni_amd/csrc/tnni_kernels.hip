// gfx950 kernels of the discrete-Sibson tile engine.
//
// Pipeline for a batch of tiles (a single tile is a batch of one):
//   k_band_count    FP64 de-scale of every LAS point into tile-local units;
//                   per chunk of points an LDS histogram over bands of
//                   bucket rows, written band-major
//   k_scan_*        device-wide exclusive scan of those histograms
//   k_band_scatter  every chunk places its sites in band order
//   k_band_sort     per band: LDS counting sort into buckets, the bucket
//                   offsets, and each bucket's distance to a non-empty one
//                   along its row
//   k_edist_cols    the same distance across rows (Chebyshev, capped)
//   k_tiles     persistent kernel over (tile, cell-block) work items: the
//               discrete Voronoi of the block's pixels (plus the query halo)
//               into LDS, then the query gather out of LDS, then the divide.
//
// The site pass replaces the OpenGL plane rendering and depth test of
// nni.cpp:170-230, 722-790; the gather replaces the query stencil
// (nni.cpp:483-534, 794-820) and cuda_nni.cu's calcNNI / cudaCheckBits /
// divide kernels (cuda_nni.cu:407-503).  All distances are exact integers.
#include "tnni_internal.h"

namespace tnni {
namespace {

// The descriptor's arrays are device allocations.  A pointer read from
// memory is generic, and generic (flat) accesses also count against the
// LDS wait counter, so the compiler waits for all of them before any LDS
// access: these casts make them global-memory instructions.
template <class T>
__device__ __forceinline__ const __attribute__((address_space(1))) T *glob(const T *p)
{
    return (const __attribute__((address_space(1))) T *)p;
}
template <class T>
__device__ __forceinline__ __attribute__((address_space(1))) T *glob(T *p)
{
    return (__attribute__((address_space(1))) T *)p;
}

// LAS int32 -> tile-local fixed-point units, in FP64.  Mirrors the LAS
// de-scale (las_reader.inl:59) and drawPlane's shift by the work-package
// origin (nni.cpp:189-190, 546).
__device__ __forceinline__ bool point_units(const TileDev &t, int32_t X, int32_t Y,
                                            int &xu, int &yu)
{
    const double xr = rint((fma((double)X, t.sx, t.ox) - t.origin_x) * (double)t.upm);
    const double yr = rint((fma((double)Y, t.sy, t.oy) - t.origin_y) * (double)t.upm);
    const bool keep = xr >= t.lo_x && xr <= t.hi_x && yr >= t.lo_y && yr <= t.hi_y;
    xu = keep ? (int)xr : 0;
    yu = keep ? (int)yr : 0;
    return keep;
}

// Site value carried to the gather: the LAS Z integer, or for the
// reference-compatible mode floor((z + zOffset + 1) * zMult) (nni.cpp:539).
__device__ __forceinline__ int32_t site_value(const TileDev &t, int32_t Z)
{
    if (t.mode != TNNI_ACC_S0) return Z;
    const double z = fma((double)Z, t.sz, t.oz);
    const double v = floor((z + (double)(-t.point_min) + 1.0) * (double)t.multiplier);
    return (int32_t)(uint32_t)(int64_t)v;
}

// Space-time: whether a point of time Ti is a site of the slice at
// t.slice_t (renderInputPoints' filter, nni.cpp:770: |t - p.t| in float
// against timeRadius + 1e-6 in double) and its squared time distance in
// units^2 (drawPlane's timeOffset term, nni.cpp:197: one time unit is one
// metre), the distance rounded to the unit.
__device__ __forceinline__ bool slice_site(const TileDev &t, int32_t Ti, uint32_t &dt2)
{
    const float dt = fabsf(t.slice_t - (float)Ti);
    const bool in = (double)dt <= (double)t.time_radius + 0.000001;
    const uint32_t u = in ? (uint32_t)rint((double)dt * (double)t.upm) : 0u;
    dt2 = u * u;
    return in;
}

// v / d for 0 <= v < 2^30 and d < 2^30, in integers: with magic =
// ceil(2^32 / d) = (2^32 + e) / d, 0 <= e < d, v * magic / 2^32 exceeds v / d
// by v * e / (d * 2^32) < 1/4, so its integer part is the quotient or one
// more, and one multiply-compare corrects it.  magic == 0 marks d == 1.
__device__ __forceinline__ int div_magic(int v, int d, uint32_t magic)
{
    if (!magic) return v;
    int q = (int)__umulhi((uint32_t)v, magic);
    if (q * d > v) --q;
    return q;
}

__device__ __forceinline__ int bucket_col(const TileDev &t, int xu) { return div_magic(xu - t.gx0, t.bs, t.bs_magic); }
__device__ __forceinline__ int bucket_row(const TileDev &t, int yu) { return div_magic(yu - t.gy0, t.bs, t.bs_magic); }

// Bucket row band of a kept site: bands of br bucket rows.
__device__ __forceinline__ int point_band(const TileDev &t, int yu)
{
    return div_magic(bucket_row(t, yu), t.br, t.br_magic);
}

constexpr int kPtsPerStep = 4;  // points per thread with their loads in flight
// The count and scatter passes: half a thread's share of a chunk
// (kChunkPts / kSortBlock points) with its loads in flight at once (all of
// it would spill the FP64 de-scale of the count pass).
constexpr int kChunkStep = kChunkPts / kSortBlock / 2;

// Pass 1 of the site sort: per chunk of one tile, the number of its sites in
// each band of the tile, written band-major (band, chunk) so that one scan
// gives every chunk its first slot in every band.
__global__ void __launch_bounds__(kSortBlock)
k_band_count(BatchDev *b, const TileDev *tiles, const Chunk *chunks, uint32_t *hist)
{
    __shared__ uint32_t h[kMaxBands];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        b->work_counter = 0u;
        hist[b->n_scan - 1] = 0u;  // the scan's total lands here
    }
    const int nch = b->n_chunks;
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const Chunk ch = chunks[c];
        const TileDev &t = tiles[ch.tile];
        const int nb = t.nbands;
        for (int i = threadIdx.x; i < nb; i += kSortBlock) h[i] = 0u;
        __syncthreads();
        const auto gX = glob(t.X), gY = glob(t.Y), gT = glob(t.T);
        const bool st = t.T != nullptr;
        for (int64_t i0 = ch.begin + threadIdx.x; i0 < ch.end; i0 += kSortBlock * kChunkStep) {
            int32_t X[kChunkStep], Y[kChunkStep], Tm[kChunkStep];
            // every load issued before any is used: past the chunk's end a
            // lane reads its last point again, and the pass skips it
#pragma unroll
            for (int q = 0; q < kChunkStep; ++q) {
                const int64_t i = min(i0 + (int64_t)q * kSortBlock, ch.end - 1);
                X[q] = gX[i];
                Y[q] = gY[i];
                Tm[q] = st ? gT[i] : 0;
            }
#pragma unroll
            for (int q = 0; q < kChunkStep; ++q) {
                int xu, yu;
                uint32_t dt2;
                if (i0 + (int64_t)q * kSortBlock < ch.end && point_units(t, X[q], Y[q], xu, yu) &&
                    (!t.T || slice_site(t, Tm[q], dt2)))
                    atomicAdd(&h[point_band(t, yu)], 1u);
            }
        }
        __syncthreads();
        const int64_t col = c - t.chunk_base;
        for (int i = threadIdx.x; i < nb; i += kSortBlock)
            hist[t.hist_base + (int64_t)i * t.nchunks + col] = h[i];
        __syncthreads();
    }
}

// Pass 2: every chunk places its sites in band order (LDS cursors seeded
// from the scanned band-major histogram).
__global__ void __launch_bounds__(kSortBlock)
k_band_scatter(const BatchDev *b, const TileDev *tiles, const Chunk *chunks,
               const uint32_t *__restrict__ hist, int4 *__restrict__ banded)
{
    __shared__ uint32_t cur[kMaxBands];
    const int nch = b->n_chunks;
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const Chunk ch = chunks[c];
        const TileDev &t = tiles[ch.tile];
        const int nb = t.nbands;
        const int64_t col = c - t.chunk_base;
        for (int i = threadIdx.x; i < nb; i += kSortBlock)
            cur[i] = hist[t.hist_base + (int64_t)i * t.nchunks + col];
        __syncthreads();
        const auto gX = glob(t.X), gY = glob(t.Y), gW = glob(t.T ? t.T : t.Z);
        for (int64_t i0 = ch.begin + threadIdx.x; i0 < ch.end; i0 += kSortBlock * kChunkStep) {
            int32_t X[kChunkStep], Y[kChunkStep], W[kChunkStep];
            // every load issued before any is used (as in the count pass)
#pragma unroll
            for (int q = 0; q < kChunkStep; ++q) {
                const int64_t i = min(i0 + (int64_t)q * kSortBlock, ch.end - 1);
                X[q] = gX[i];
                Y[q] = gY[i];
                W[q] = gW[i];
            }
            // a site record {index, x, y, w}: the index first (the low half
            // of the search key); w the site's value, or for a space-time
            // slice its squared time distance
#pragma unroll
            for (int q = 0; q < kChunkStep; ++q) {
                const int64_t i = i0 + (int64_t)q * kSortBlock;
                int xu, yu;
                uint32_t dt2 = 0;
                if (i < ch.end && point_units(t, X[q], Y[q], xu, yu) &&
                    (!t.T || slice_site(t, W[q], dt2))) {
                    const uint32_t pos = atomicAdd(&cur[point_band(t, yu)], 1u);
                    banded[pos] = make_int4((int)i, xu, yu, t.T ? (int)dt2 : site_value(t, W[q]));
                }
            }
        }
        __syncthreads();
    }
}

// ---- device-wide exclusive scan (fixed grid, graph-capturable) ----------

template <int NT>
__device__ __forceinline__ uint32_t block_excl_scan(uint32_t v, uint32_t *s_w, uint32_t &total)
{
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    uint32_t inc = v;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const uint32_t o = __shfl_up(inc, d, 64);
        if (lane >= d) inc += o;
    }
    if (lane == 63) s_w[w] = inc;
    __syncthreads();
    uint32_t off = 0, tot = 0;
#pragma unroll
    for (int i = 0; i < NW; ++i) {
        const uint32_t x = s_w[i];
        off += (i < w) ? x : 0u;
        tot += x;
    }
    __syncthreads();
    total = tot;
    return off + inc - v;
}

__global__ void __launch_bounds__(kBlock)
k_scan_reduce(const BatchDev *b, const uint32_t *counts, uint32_t *partials)
{
    __shared__ uint32_t s_w[kWaves];
    const uint32_t n = b->n_scan, chunk = b->scan_chunk;
    const uint64_t beg = (uint64_t)blockIdx.x * chunk;
    const uint64_t end = min((uint64_t)n, beg + chunk);
    uint32_t s = 0;
    for (uint64_t i = beg + threadIdx.x; i < end; i += kBlock) s += counts[i];
    uint32_t total;
    block_excl_scan<kBlock>(s, s_w, total);
    if (threadIdx.x == 0) partials[blockIdx.x] = total;
}

__global__ void __launch_bounds__(kBlock)
k_scan_down(const BatchDev *b, uint32_t *counts, const uint32_t *partials)
{
    __shared__ uint32_t s_w[kWaves];
    const uint32_t n = b->n_scan, chunk = b->scan_chunk;
    const uint64_t beg = (uint64_t)blockIdx.x * chunk;
    const uint64_t end = min((uint64_t)n, beg + chunk);
    // the block's first output: the totals of the blocks before it (at
    // most kScanBlocks of them, summed here rather than by a launch of
    // their own)
    uint32_t before = 0;
    for (int i = threadIdx.x; i < (int)blockIdx.x; i += kBlock) before += partials[i];
    uint32_t run;
    block_excl_scan<kBlock>(before, s_w, run);
    constexpr int STEP = kBlock * kScanItems;
    for (uint64_t base = beg; base < end; base += STEP) {
        uint32_t v[kScanItems];
        uint32_t s = 0;
        const uint64_t i0 = base + (uint64_t)threadIdx.x * kScanItems;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            v[k] = (i0 + k < end) ? counts[i0 + k] : 0u;
            s += v[k];
        }
        uint32_t total;
        uint32_t ex = block_excl_scan<kBlock>(s, s_w, total) + run;
#pragma unroll
        for (int k = 0; k < kScanItems; ++k) {
            if (i0 + k < end) counts[i0 + k] = ex;
            ex += v[k];
        }
        run += total;
    }
}

// ---- empty-distance map ----------------------------------------------------
// edist[b] = min(ecap, Chebyshev distance in buckets from b to the nearest
// bucket holding a site), computed separably: first along rows (hdist), then
// along columns.

__device__ __forceinline__ int tile_of_row(const TileDev *tiles, int n_tiles, int64_t row)
{
    int lo = 0, hi = n_tiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tiles[mid].row_base <= row) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

__device__ __forceinline__ int tile_of_band(const TileDev *tiles, int n_tiles, int64_t band)
{
    int lo = 0, hi = n_tiles - 1;
    while (lo < hi) {
        const int mid = (lo + hi + 1) >> 1;
        if (tiles[mid].band_base <= band) lo = mid;
        else hi = mid - 1;
    }
    return lo;
}

// Pass 3: one workgroup per band sorts the band's sites into its buckets
// (LDS counters), writes the bucket offsets, and the distance of every
// bucket to the nearest bucket holding a site in its own row (hdist).
// A band of at most kSortCap sites is read once into registers, each
// record taking its rank in its bucket from the counter; after the scan the
// records are placed in an LDS copy of the band's output and written out in
// order, whole lines.  A larger band is read twice and its records written
// one by one to their slots.
constexpr int kSortRecs = (kSortCap + kSortBlock - 1) / kSortBlock;  // records per thread

__global__ void __launch_bounds__(kSortBlock)
k_band_sort(const BatchDev *b, const TileDev *tiles, const uint32_t *__restrict__ hist,
            const int4 *__restrict__ banded, int4 *__restrict__ sorted, uint32_t *__restrict__ offs,
            uint8_t *__restrict__ hdist)
{
    __shared__ uint32_t cnt[kBandBuckets];
    __shared__ int4 stage[kSortCap];
    __shared__ uint32_t s_w[kSortBlock / 64];
    const int64_t n_bands = b->n_bands;
    const int n_tiles = b->n_tiles;
    for (int64_t g = blockIdx.x; g < n_bands; g += gridDim.x) {
        const TileDev &t = tiles[tile_of_band(tiles, n_tiles, g)];
        const int band = (int)(g - t.band_base);
        const uint32_t s0 = hist[t.hist_base + (int64_t)band * t.nchunks];
        const uint32_t s1 = hist[t.hist_base + (int64_t)(band + 1) * t.nchunks];
        const int nbx = t.nbx, r0 = band * t.br, r1 = min(r0 + t.br, t.nby);
        const int nloc = (r1 - r0) * nbx;
        const int64_t obase = t.bk_base + (int64_t)r0 * nbx;
        const uint32_t n = s1 - s0;
        const bool in_lds = n <= (uint32_t)kSortCap;
        for (int i = threadIdx.x; i < nloc; i += kSortBlock) cnt[i] = 0u;
        __syncthreads();
        int4 rec[kSortRecs];
        uint32_t key[kSortRecs];  // bucket in the band | rank in the bucket << 16
        if (in_lds) {
#pragma unroll
            for (int q = 0; q < kSortRecs; ++q) {
                const uint32_t i = threadIdx.x + (uint32_t)q * kSortBlock;
                // (a plain guarded load: a select against other values
                // would be made a select of addresses, one dword at a time)
                rec[q] = make_int4(0, 0, 0, 0);
                if (i < n) rec[q] = banded[s0 + i];
            }
#pragma unroll
            for (int q = 0; q < kSortRecs; ++q) {
                key[q] = 0u;
                if (threadIdx.x + (uint32_t)q * kSortBlock < n) {
                    const int bx = bucket_col(t, rec[q].y);
                    const int by = bucket_row(t, rec[q].z);
                    const uint32_t bk = (uint32_t)((by - r0) * nbx + bx);
                    key[q] = bk | (atomicAdd(&cnt[bk], 1u) << 16);
                }
            }
        } else {
            // kPtsPerStep records in flight per thread
            for (uint32_t p0 = s0 + threadIdx.x; p0 < s1; p0 += kSortBlock * kPtsPerStep) {
                int4 r4[kPtsPerStep];
#pragma unroll
                for (int q = 0; q < kPtsPerStep; ++q) {
                    const uint32_t p = p0 + (uint32_t)q * kSortBlock;
                    r4[q] = make_int4(0, 0, 0, 0);
                    if (p < s1) r4[q] = banded[p];
                }
#pragma unroll
                for (int q = 0; q < kPtsPerStep; ++q) {
                    if (p0 + (uint32_t)q * kSortBlock >= s1) break;
                    const int bx = bucket_col(t, r4[q].y);
                    const int by = bucket_row(t, r4[q].z);
                    atomicAdd(&cnt[(by - r0) * nbx + bx], 1u);
                }
            }
        }
        __syncthreads();
        // row distances to a bucket holding a site, capped at ecap.  Each
        // row is cut into chunks of 64 columns and each chunk's non-empty
        // buckets become one 64-bit mask (a wave ballot); a bucket's
        // nearest non-empty bucket on either side is the nearest set bit in
        // its own chunk's mask or, walking outward, in the next non-zero
        // one, no farther than cap.  The masks go to the stage's memory,
        // which holds no records yet.
        const int cap = t.ecap;
        {
            const int nck = (nbx + 63) >> 6, nrow = r1 - r0;
            // at most one chunk per bucket of the band
            static_assert(kBandBuckets * sizeof(uint64_t) <= kSortCap * sizeof(int4), "masks in the stage");
            uint64_t *masks = reinterpret_cast<uint64_t *>(stage);
            const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
            for (int c = wave; c < nrow * nck; c += kSortBlock / 64) {
                const int rr = c / nck, x = ((c - rr * nck) << 6) + lane;
                const uint64_t m = __ballot(x < nbx && cnt[rr * nbx + x] != 0u);
                if (lane == 0) masks[c] = m;
            }
            __syncthreads();
            for (int i = threadIdx.x; i < nloc; i += kSortBlock) {
                const int rr = i / nbx, x = i - rr * nbx, k = x >> 6, bit = x & 63;
                const uint64_t *row = masks + rr * nck;
                int d = cap;
                // left: set bits at or below the bucket's own
                uint64_t m = row[k] & (bit == 63 ? ~0ull : ((2ull << bit) - 1ull));
                for (int kk = k;;) {
                    if (m) {
                        d = min(d, x - ((kk << 6) + 63 - __clzll((long long)m)));
                        break;
                    }
                    if (--kk < 0 || x - ((kk << 6) + 63) >= d) break;
                    m = row[kk];
                }
                // right: set bits at or above it
                m = row[k] & (~0ull << bit);
                for (int kk = k;;) {
                    if (m) {
                        d = min(d, (kk << 6) + __ffsll((long long)m) - 1 - x);
                        break;
                    }
                    if (++kk >= nck || (kk << 6) - x >= d) break;
                    m = row[kk];
                }
                hdist[t.e_base + (int64_t)(r0 + rr) * t.nbx4 + x] = (uint8_t)d;
            }
            __syncthreads();
        }
        const int seg = (nloc + kSortBlock - 1) / kSortBlock;
        const int sb = min(nloc, (int)threadIdx.x * seg), se = min(nloc, sb + seg);
        // exclusive scan of the counters: each thread a contiguous segment
        uint32_t sum = 0;
        for (int i = sb; i < se; ++i) sum += cnt[i];
        uint32_t total;
        uint32_t run = block_excl_scan<kSortBlock>(sum, s_w, total);
        for (int i = sb; i < se; ++i) {
            const uint32_t v = cnt[i];
            cnt[i] = run;
            run += v;
        }
        __syncthreads();
        for (int i = threadIdx.x; i < nloc; i += kSortBlock) offs[obase + i] = s0 + cnt[i];
        if (r1 == t.nby && threadIdx.x == 0) offs[t.bk_base + (int64_t)t.nby * nbx] = s1;
        if (in_lds) {
#pragma unroll
            for (int q = 0; q < kSortRecs; ++q)
                if (threadIdx.x + (uint32_t)q * kSortBlock < n)
                    stage[cnt[key[q] & 0xFFFFu] + (key[q] >> 16)] = rec[q];
            __syncthreads();
            for (uint32_t i = threadIdx.x; i < n; i += kSortBlock) sorted[s0 + i] = stage[i];
        } else {
            __syncthreads();
            for (uint32_t p0 = s0 + threadIdx.x; p0 < s1; p0 += kSortBlock * kPtsPerStep) {
                int4 r4[kPtsPerStep];
#pragma unroll
                for (int q = 0; q < kPtsPerStep; ++q) {
                    const uint32_t p = p0 + (uint32_t)q * kSortBlock;
                    r4[q] = make_int4(0, 0, 0, 0);
                    if (p < s1) r4[q] = banded[p];
                }
#pragma unroll
                for (int q = 0; q < kPtsPerStep; ++q) {
                    if (p0 + (uint32_t)q * kSortBlock >= s1) break;
                    const int bx = bucket_col(t, r4[q].y);
                    const int by = bucket_row(t, r4[q].z);
                    const uint32_t pos = atomicAdd(&cnt[(by - r0) * nbx + bx], 1u);
                    sorted[s0 + pos] = r4[q];
                }
            }
        }
        __syncthreads();
    }
}

__global__ void __launch_bounds__(kBlock)
k_edist_cols(const BatchDev *b, const TileDev *tiles, const uint8_t *__restrict__ hdist,
             uint8_t *__restrict__ edist)
{
    const int64_t nr = b->n_rows;
    const int n_tiles = b->n_tiles;
    // Workgroups go to the 8 XCDs in turn (blockIdx mod 8), each XCD with
    // its own L2, and a row reads its neighbours' rows: so XCD x takes the
    // contiguous x-th eighth of the rows, where those reads hit its L2.
    const int G = (int)gridDim.x;
    const int xcd = G >= 8 ? (int)blockIdx.x % 8 : 0;
    const int64_t lo = G >= 8 ? nr * xcd / 8 : 0, hi = G >= 8 ? nr * (xcd + 1) / 8 : nr;
    const int64_t step = G >= 8 ? (G - xcd + 7) / 8 : G;
    for (int64_t item = lo + (G >= 8 ? (int)blockIdx.x / 8 : (int)blockIdx.x); item < hi; item += step) {
        const TileDev &t = tiles[tile_of_row(tiles, n_tiles, item)];
        const int row = (int)(item - t.row_base);
        const int nbx = t.nbx, nby = t.nby;
        // rows of nbx4 bytes from a 4-aligned base: four columns per thread
        // with aligned 4-byte loads and stores (the padding columns compute
        // values nothing reads)
        const int nbx4 = t.nbx4;
        const uint8_t *h = hdist + t.e_base;
        const int cap = t.ecap;
        (void)nbx;
        for (int g = threadIdx.x; g < nbx4 / 4; g += kBlock) {
            const int x = 4 * g;
            int e[4] = {cap, cap, cap, cap};
            for (int q = 0; q < cap; ++q) {
                const int a = row - q, c = row + q;
                const uint32_t up = a >= 0 ? *(const uint32_t *)(h + (int64_t)a * nbx4 + x) : 0xFFFFFFFFu;
                const uint32_t dn = c < nby ? *(const uint32_t *)(h + (int64_t)c * nbx4 + x) : 0xFFFFFFFFu;
#pragma unroll
                for (int k = 0; k < 4; ++k) {
                    const int v = min((int)((up >> (8 * k)) & 0xFFu), (int)((dn >> (8 * k)) & 0xFFu));
                    e[k] = min(e[k], max(q, v));
                }
                if (max(max(e[0], e[1]), max(e[2], e[3])) <= q) break;
            }
            *(uint32_t *)(edist + t.e_base + (int64_t)row * nbx4 + x) =
                (uint32_t)e[0] | ((uint32_t)e[1] << 8) | ((uint32_t)e[2] << 16) | ((uint32_t)e[3] << 24);
        }
    }
}

// ---- the tile kernel -------------------------------------------------------

// One site against the pixel centre (px, py): keep the smallest
// (d2, index) among the sites whose region covers the pixel.  Branch-free,
// so the whole record is one 16-byte load.  |dx|, |dy| <= R < 2^23 for a
// covering site, so 24-bit multiplies are exact there.
// ALG == kAllCover: the caller has shown that every visited site covers the
// pixel, so the region test is skipped.
constexpr int kAllCover = 2;

// A record is {index, x, y, w}: w is the site's value, or for a space-time
// slice (TIME) its squared time distance, added to the distance (the region
// test stays in the plane, nni.cpp:117-157, 207-225; the winner's value is
// then read through its index).
template <int ALG, bool TIME>
__device__ __forceinline__ void visit_site(const int4 st, int px, int py, int R, uint32_t R2,
                                           uint64_t &best, int &bval)
{
    const int dx = px - st.y, dy = py - st.z;
    const uint32_t dxy = (uint32_t)__mul24(dx, dx) + (uint32_t)__mul24(dy, dy);
    const uint32_t dd = TIME ? dxy + (uint32_t)st.w : dxy;
    if (ALG == kAllCover) {
        const uint64_t key = ((uint64_t)dd << 32) | (uint32_t)st.x;
        const bool better = key < best;
        best = better ? key : best;
        if (!TIME) bval = better ? st.w : bval;
        return;
    }
    bool cov = max((uint32_t)(dx + R), (uint32_t)(dy + R)) <= 2u * (uint32_t)R;
    if (ALG == TNNI_ALG_CONES) cov = cov && dxy <= R2;
    // a site that does not cover the pixel gets the key of "no site"
    const uint32_t hi = cov ? dd : 0xFFFFFFFFu;
    const uint32_t lo = cov ? (uint32_t)st.x : 0xFFFFFFFFu;
    const uint64_t key = ((uint64_t)hi << 32) | lo;
    const bool better = key < best;
    best = better ? key : best;
    if (!TIME) bval = better ? st.w : bval;
}

// Visit the sites stored at [s, e) of a site array (global or LDS).
template <int ALG, bool TIME>
__device__ __forceinline__ void scan_range(uint32_t s, uint32_t e, const int4 *__restrict__ sites,
                                           int px, int py, int R, uint32_t R2,
                                           uint64_t &best, int &bval)
{
    // two records in flight per step; the loop runs on the record pointer
    const int4 *q = sites + s;
    const int4 *const qe = sites + e;
    if (q >= qe) return;
    const int4 *const qlast = qe - 1;
    for (; q < qlast; q += 2) {
        const int4 a = q[0];
        const int4 b = q[1];
        visit_site<ALG, TIME>(a, px, py, R, R2, best, bval);
        visit_site<ALG, TIME>(b, px, py, R, R2, best, bval);
    }
    if (q < qe) visit_site<ALG, TIME>(*q, px, py, R, R2, best, bval);
}

// The bucket columns and rows that can hold a site both covering the pixel
// centre (px, py) and at least as near as the best so far: a site within R
// of it in x and in y (the site region), and within floor(sqrt(bd)) (a
// site at exactly the best distance may still win on its index).
__device__ __forceinline__ void site_reach(const TileDev &t, int px, int py, int R, uint64_t bd, int &c0,
                                           int &c1, int &q0, int &q1)
{
    int D = R;
    if (bd != 0xFFFFFFFFu) {
        // floor(sqrt(bd)): the float root of a value below 2^32 is within
        // 0.01 of the true root, so truncating it is off by at most one
        int q = (int)sqrtf((float)bd);
        q = (uint64_t)q * (uint64_t)q > bd ? q - 1 : q;
        q = (uint64_t)(q + 1) * (uint64_t)(q + 1) <= bd ? q + 1 : q;
        D = min(D, q);
    }
    c0 = max(div_magic(max(px - D - t.gx0, 0), t.bs, t.bs_magic), 0);
    c1 = min(div_magic(px + D - t.gx0, t.bs, t.bs_magic), t.nbx - 1);
    q0 = max(div_magic(max(py - D - t.gy0, 0), t.bs, t.bs_magic), 0);
    q1 = min(div_magic(py + D - t.gy0, t.bs, t.bs_magic), t.nby - 1);
}

// Ring r of the bucket square around (bx, by), from the global arrays.
template <int ALG, bool TIME>
__device__ __forceinline__ void ring_global(const TileDev &t, const uint32_t *__restrict__ offs,
                                            const int4 *__restrict__ sorted, int bx, int by, int r,
                                            int px, int py, int R, uint64_t &best, int &bval)
{
    const uint32_t *base = offs + t.bk_base;
    for (int dy = -r; dy <= r; ++dy) {
        const int row = by + dy;
        if (row < 0 || row >= t.nby) continue;
        const uint32_t *rs = base + (int64_t)row * t.nbx;
        uint32_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;
        if (dy == -r || dy == r) {
            const int c0 = max(bx - r, 0), c1 = min(bx + r, t.nbx - 1);
            if (c0 <= c1) {
                a0 = rs[c0];
                a1 = rs[c1 + 1];
            }
        } else {
            if (bx - r >= 0) {
                a0 = rs[bx - r];
                a1 = rs[bx - r + 1];
            }
            if (bx + r < t.nbx) {
                b0 = rs[bx + r];
                b1 = rs[bx + r + 1];
            }
        }
        scan_range<ALG, TIME>(a0, a1, sorted, px, py, R, t.Rs2, best, bval);
        scan_range<ALG, TIME>(b0, b1, sorted, px, py, R, t.Rs2, best, bval);
    }
}

// The square of buckets within R of (bx, by), from the global arrays: one
// contiguous range per bucket row, the row offsets four rows at a time.
template <int ALG, bool TIME>
__device__ __forceinline__ void square_global(const TileDev &t, const uint32_t *__restrict__ offs,
                                              const int4 *__restrict__ sorted, int bx, int by, int R,
                                              int px, int py, int Rr, uint64_t &best, int &bval)
{
    const uint32_t *base = offs + t.bk_base;
    const int c0 = max(bx - R, 0), c1 = min(bx + R, t.nbx - 1);
    const int r0 = max(by - R, 0), r1 = min(by + R, t.nby - 1);
    for (int row = r0; row <= r1; row += 4) {
        uint32_t a[4], e[4];
#pragma unroll
        for (int q = 0; q < 4; ++q) {
            const uint32_t *rs = base + (int64_t)min(row + q, r1) * t.nbx;
            a[q] = row + q <= r1 ? rs[c0] : 0u;
            e[q] = row + q <= r1 ? rs[c1 + 1] : 0u;
        }
#pragma unroll
        for (int q = 0; q < 4; ++q)
            scan_range<ALG, TIME>(a[q], e[q], sorted, px, py, Rr, t.Rs2, best, bval);
    }
}

// Visit every G-th site of [s, e) from s + sub: the G lanes of a group
// (sub = 0 .. G-1) share one range, neighbouring lanes reading neighbouring
// records.
template <int ALG, bool TIME>
__device__ __forceinline__ void scan_range_part(uint32_t s, uint32_t e, uint32_t sub, uint32_t G,
                                                const int4 *__restrict__ sites, int px, int py, int R,
                                                uint32_t R2, uint64_t &best, int &bval)
{
    uint32_t i = s + sub;
    for (; i + G < e; i += 2 * G) {
        const int4 a = sites[i];
        const int4 b = sites[i + G];
        visit_site<ALG, TIME>(a, px, py, R, R2, best, bval);
        visit_site<ALG, TIME>(b, px, py, R, R2, best, bval);
    }
    if (i < e) visit_site<ALG, TIME>(sites[i], px, py, R, R2, best, bval);
}

// The buckets [c0, c1] x [r0, r1] (inside the grid), their sites shared by
// the G lanes of a group: lane sub visits every G-th site of each bucket
// row's range.  The row offsets come nine rows at a time (the square of a
// long search at the built-in bucket side is 9 rows or fewer: one round
// trip instead of three).
template <int ALG, bool TIME>
__device__ __forceinline__ void square_global_box(const TileDev &t, const uint32_t *__restrict__ offs,
                                                  const int4 *__restrict__ sorted, int c0, int c1, int r0,
                                                  int r1, int px, int py, int Rr, uint32_t sub, uint32_t G,
                                                  uint64_t &best, int &bval)
{
    const uint32_t *base = offs + t.bk_base;
    for (int row = r0; row <= r1; row += 9) {
        uint32_t a[9], e[9];
#pragma unroll
        for (int q = 0; q < 9; ++q) {
            const uint32_t *rs = base + (int64_t)min(row + q, r1) * t.nbx;
            a[q] = row + q <= r1 ? rs[c0] : 0u;
            e[q] = row + q <= r1 ? rs[c1 + 1] : 0u;
        }
#pragma unroll
        for (int q = 0; q < 9; ++q)
            scan_range_part<ALG, TIME>(a[q], e[q], sub, G, sorted, px, py, Rr, t.Rs2, best, bval);
    }
}

// ring_global's sites shared by the G lanes of a group (lane sub visits
// every G-th site of each range).
template <int ALG, bool TIME>
__device__ __forceinline__ void ring_global_part(const TileDev &t, const uint32_t *__restrict__ offs,
                                                 const int4 *__restrict__ sorted, int bx, int by, int r,
                                                 int px, int py, int R, uint32_t sub, uint32_t G,
                                                 uint64_t &best, int &bval, int xc0, int xc1, int xq0, int xq1)
{
    // only the ring's buckets within the columns [xc0, xc1] and rows
    // [xq0, xq1] (inside the grid) that can still hold a winner
    const uint32_t *base = offs + t.bk_base;
    for (int dy = max(-r, xq0 - by); dy <= min(r, xq1 - by); ++dy) {
        const int row = by + dy;
        const uint32_t *rs = base + (int64_t)row * t.nbx;
        uint32_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;
        if (dy == -r || dy == r) {
            const int c0 = max(bx - r, xc0), c1 = min(bx + r, xc1);
            if (c0 <= c1) {
                a0 = rs[c0];
                a1 = rs[c1 + 1];
            }
        } else {
            if (bx - r >= xc0) {
                a0 = rs[bx - r];
                a1 = rs[bx - r + 1];
            }
            if (bx + r <= xc1) {
                b0 = rs[bx + r];
                b1 = rs[bx + r + 1];
            }
        }
        scan_range_part<ALG, TIME>(a0, a1, sub, G, sorted, px, py, R, t.Rs2, best, bval);
        scan_range_part<ALG, TIME>(b0, b1, sub, G, sorted, px, py, R, t.Rs2, best, bval);
    }
}

// k / bm for a pixel index 0 <= k < 2^20: multiply by ceil(2^32 / bm)
// (exact since k * (ceil(2^32/bm) * bm - 2^32) < 2^32); bm == 1 has no
// 32-bit multiplier and is taken as is.
__device__ __forceinline__ int div_bm(const TileDev &t, int k)
{
    return t.bm_magic ? (int)__umulhi((uint32_t)k, t.bm_magic) : k;
}

__device__ __forceinline__ void wave_sync()
{
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Ring r (0 or 1) of the bucket square around (bx, by), from the stage.
template <int ALG, bool TIME>
__device__ __forceinline__ void ring_staged(const TileDev &t, const WaveStage &ws, int cx0, int cy0,
                                            int wp, int bx, int by, int r, int px, int py, int R,
                                            uint64_t &best, int &bval)
{
    for (int dy = -r; dy <= r; ++dy) {
        const int row = by + dy;
        if (row < 0 || row >= t.nby) continue;
        const int j = row - cy0;
        const uint32_t *wr = ws.woff + j * wp;
        const uint32_t rel = ws.rowpre[j] - wr[0];  // staged slot = offset + rel
        uint32_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;
        if (dy == -r || dy == r) {
            const int c0 = max(bx - r, 0), c1 = min(bx + r, t.nbx - 1);
            if (c0 <= c1) {
                a0 = wr[c0 - cx0] + rel;
                a1 = wr[c1 + 1 - cx0] + rel;
            }
        } else {
            if (bx - r >= 0) {
                a0 = wr[bx - r - cx0] + rel;
                a1 = wr[bx - r + 1 - cx0] + rel;
            }
            if (bx + r < t.nbx) {
                b0 = wr[bx + r - cx0] + rel;
                b1 = wr[bx + r + 1 - cx0] + rel;
            }
        }
        scan_range<ALG, TIME>(a0, a1, ws.sites, px, py, R, t.Rs2, best, bval);
        scan_range<ALG, TIME>(b0, b1, ws.sites, px, py, R, t.Rs2, best, bval);
    }
}

// Rings 0 and 1 of a pixel (the 3x3 square of buckets around its own,
// (bx, by)) from a wave's stage, pruned by the distances to its bucket's
// edges: every site of the row above is at least eb from the pixel in y,
// of the left column at least el in x, and so on.  The middle row first;
// then the row above only if eb^2 <= the best d2 so far, and its corner
// buckets only if el^2 + eb^2 (er^2 + eb^2) <= it; then the row below the
// same way.  A site farther than the best cannot win, nor tie on d2 (a
// tie at exactly the best distance is still visited), so the minimum is
// the one the whole square gives.
template <int ALG, bool TIME>
__device__ __forceinline__ void square3_staged(const TileDev &t, const WaveStage &ws, int cx0, int cy0, int wp,
                                               int bx, int by, int px, int py, int R, int el, int er, int eb,
                                               int et, uint64_t &best, int &bval)
{
    const int c0 = max(bx - 1, 0), c1 = min(bx + 1, t.nbx - 1);
    {
        const uint32_t *wr = ws.woff + (by - cy0) * wp;
        const uint32_t rel = ws.rowpre[by - cy0] - wr[0];
        scan_range<ALG, TIME>(wr[c0 - cx0] + rel, wr[c1 + 1 - cx0] + rel, ws.sites, px, py, R, t.Rs2, best, bval);
    }
    const uint64_t el2 = (uint64_t)el * (uint64_t)el, er2 = (uint64_t)er * (uint64_t)er;
#pragma unroll
    for (int side = 0; side < 2; ++side) {
        const int row = side == 0 ? by - 1 : by + 1;
        const int e = side == 0 ? eb : et;
        const uint64_t bd = best >> 32;
        const uint64_t e2 = (uint64_t)e * (uint64_t)e;
        if (row < 0 || row >= t.nby || e2 > bd) continue;
        const int a = e2 + el2 <= bd ? c0 : bx;
        const int b = e2 + er2 <= bd ? c1 : bx;
        const uint32_t *wr = ws.woff + (row - cy0) * wp;
        const uint32_t rel = ws.rowpre[row - cy0] - wr[0];
        scan_range<ALG, TIME>(wr[a - cx0] + rel, wr[b + 1 - cx0] + rel, ws.sites, px, py, R, t.Rs2, best, bval);
    }
}

// Ring 1 around (bx, by) within the columns [xc0, xc1] and rows [xq0,
// xq1] (inside the grid; site_reach), from the global arrays (stage ==
// nullptr) or from a wave's stage.
template <int ALG, bool TIME, bool STAGED>
__device__ __forceinline__ void ring1_clipped(const TileDev &t, const uint32_t *__restrict__ offs,
                                              const int4 *__restrict__ sorted, const WaveStage &ws,
                                              int cx0, int cy0, int wp, int bx, int by, int px, int py, int R,
                                              int xc0, int xc1, int xq0, int xq1, uint64_t &best, int &bval)
{
    // (STAGED a template argument: a pointer that may be the stage or the
    // global arrays would make every read a flat one)
    for (int row = max(by - 1, xq0); row <= min(by + 1, xq1); ++row) {
        // the row's bucket offsets and their staged slots (rel = 0: global)
        const uint32_t *wr;
        uint32_t rel = 0;
        int cb = 0;
        if (STAGED) {
            wr = ws.woff + (row - cy0) * wp;
            rel = ws.rowpre[row - cy0] - wr[0];
            cb = cx0;
        } else {
            wr = offs + t.bk_base + (int64_t)row * t.nbx;
        }
        const int4 *sites = STAGED ? ws.sites : sorted;
        uint32_t a0 = 0, a1 = 0, b0 = 0, b1 = 0;
        if (row != by) {
            const int c0 = max(bx - 1, xc0), c1 = min(bx + 1, xc1);
            if (c0 <= c1) {
                a0 = wr[c0 - cb] + rel;
                a1 = wr[c1 + 1 - cb] + rel;
            }
        } else {
            if (bx - 1 >= xc0) {
                a0 = wr[bx - 1 - cb] + rel;
                a1 = wr[bx - cb] + rel;
            }
            if (bx + 1 <= xc1) {
                b0 = wr[bx + 1 - cb] + rel;
                b1 = wr[bx + 2 - cb] + rel;
            }
        }
        scan_range<ALG, TIME>(a0, a1, sites, px, py, R, t.Rs2, best, bval);
        scan_range<ALG, TIME>(b0, b1, sites, px, py, R, t.Rs2, best, bval);
    }
}

// Nearest covering site of every pixel of one 8x8 block (one wave, one
// pixel per lane; inactive lanes still take part in the staging).  Each
// lane searches the bucket grid in rings around its pixel's bucket: rings
// closer than the bucket's empty distance r0 hold no site; rings 0 and 1
// come from the wave's LDS stage when the window fits it, as one 3x3
// square on sparse tiles and ring by ring (stopping after ring 0 where it
// settles the pixel) on dense ones; the search ends when every unvisited
// site is provably farther than the best so far (strictly, so an equal
// distance still reaches the lower index) or outside the site region, and
// otherwise goes on later (pending) from ring max(r0, 2).
template <int ALG, bool TIME, bool DENSE>
__device__ __forceinline__ void label_block(const TileDev &t, const uint32_t *__restrict__ offs,
                                            const uint8_t *__restrict__ edist,
                                            const int4 *__restrict__ sorted, WaveStage &ws,
                                            int kA, int lA, int kB, int lB, int k, int l,
                                            bool active, uint64_t &best, int &bval, bool &pending)
{
    const int lane = threadIdx.x & 63;
    best = ~0ull;
    bval = 0;
    pending = false;
    const int bs = t.bs;
    const int bx = t.bK + div_bm(t, k), by = t.bK + div_bm(t, l);
    const int r0 = active ? (int)edist[t.e_base + (int64_t)by * t.nbx4 + bx] : 1 << 20;

    // the window: the block's buckets and one more on each side
    const int cx0 = max(t.bK + div_bm(t, kA) - 1, 0), cx1 = min(t.bK + div_bm(t, kB) + 1, t.nbx - 1);
    const int cy0 = max(t.bK + div_bm(t, lA) - 1, 0), cy1 = min(t.bK + div_bm(t, lB) + 1, t.nby - 1);
    const int ww = cx1 - cx0 + 1, wh = cy1 - cy0 + 1, wp = ww + 1;
    bool staged = false;
    if (__any(r0 <= 1)) {
        wave_sync();  // the previous block's reads of the stage are done
        for (int i = lane; i < wh * wp; i += 64) {
            const int j = i / wp, c = i - j * wp;
            ws.woff[i] = offs[t.bk_base + (int64_t)(cy0 + j) * t.nbx + cx0 + c];
        }
        wave_sync();
        uint32_t cnt = 0;
        if (lane < wh) cnt = ws.woff[lane * wp + ww] - ws.woff[lane * wp];
        uint32_t inc = cnt;
#pragma unroll
        for (int d = 1; d < 16; d <<= 1) {
            const uint32_t o = __shfl_up(inc, d, 64);
            if (lane >= d) inc += o;
        }
        const uint32_t pre = inc - cnt;
        const uint32_t total = __shfl(inc, wh - 1, 64);
        if (lane < wh) ws.rowpre[lane] = pre;
        staged = total <= (uint32_t)kStageCap;
        if (staged) {
            wave_sync();
            // direct global -> LDS loads (no VGPRs): lane L of step q writes
            // slot 64q + L; all steps in flight before the one wait
            for (int q = 0; q < kStageCap / 64; ++q) {
                if (64u * q >= total) break;
                const uint32_t i = lane + 64u * q;
                if (i < total) {
                    int j = 0;
#pragma unroll
                    for (int jj = 1; jj < kWinMax; ++jj)
                        j += (jj < wh && ws.rowpre[jj] <= i) ? 1 : 0;
                    const int4 *src = sorted + (ws.woff[j * wp] + (i - ws.rowpre[j]));
                    __builtin_amdgcn_global_load_lds((const __attribute__((address_space(1))) void *)src,
                                                     (__attribute__((address_space(3))) void *)(ws.sites + 64 * q),
                                                     16, 0, 0);
                }
            }
            __builtin_amdgcn_s_waitcnt(0);
        }
        wave_sync();
    }
    if (!active) return;

    const int px = t.x0 + k * t.ps_u, py = t.y0 + l * t.ps_u;
    const int R = (ALG == TNNI_ALG_PLANES) ? t.Rs : t.Rsb;
    // distance from the pixel centre to the outside of its own bucket
    const int el = px - (t.gx0 + bx * bs) + 1, er = t.gx0 + (bx + 1) * bs - px;
    const int eb = py - (t.gy0 + by * bs) + 1, et = t.gy0 + (by + 1) * bs - py;
    const int e0 = min(min(el, er), min(eb, et));
    const bool cover = t.ring1_covered;
    if (t.dense || !staged) {
        // ring by ring: ring 1 only where it may still hold a covering site
        // nearer than the best (gap e0).  Also where the block's window
        // overflows the stage (a cluster core in a sparser tile): there the
        // 3x3 square would read hundreds of sites per pixel from global
        // memory, and ring 0 nearly always settles the pixel
        if (r0 == 0) {
            if (!staged) ring_global<ALG, TIME>(t, offs, sorted, bx, by, 0, px, py, R, best, bval);
            else if (cover) ring_staged<kAllCover, TIME>(t, ws, cx0, cy0, wp, bx, by, 0, px, py, R, best, bval);
            else ring_staged<ALG, TIME>(t, ws, cx0, cy0, wp, bx, by, 0, px, py, R, best, bval);
        }
        // ring 1, only its buckets that can still hold a winner
        if (r0 <= 1 && e0 <= t.Rsb && (uint64_t)e0 * (uint64_t)e0 <= (best >> 32)) {
            int xc0, xc1, xq0, xq1;
            site_reach(t, px, py, R, best >> 32, xc0, xc1, xq0, xq1);
            if (!staged)
                ring1_clipped<ALG, TIME, false>(t, offs, sorted, ws, cx0, cy0, wp, bx, by, px, py, R, xc0, xc1, xq0,
                                                xq1, best, bval);
            else if (cover)
                ring1_clipped<kAllCover, TIME, true>(t, offs, sorted, ws, cx0, cy0, wp, bx, by, px, py, R, xc0, xc1,
                                                     xq0, xq1, best, bval);
            else
                ring1_clipped<ALG, TIME, true>(t, offs, sorted, ws, cx0, cy0, wp, bx, by, px, py, R, xc0, xc1, xq0,
                                               xq1, best, bval);
        }
    } else if (r0 <= 1) {
        // rings 0 and 1 as one 3x3 square of buckets (three row ranges).
        // The ring loop would skip ring 1 only when e0 > Rsb (no site of it
        // covers the pixel) or e0^2 > d2 after ring 0 (none of it is nearer),
        // so visiting it anyway leaves the minimum as it is.
        if (!staged) square_global<ALG, TIME>(t, offs, sorted, bx, by, 1, px, py, R, best, bval);
        else if (cover)
            square3_staged<kAllCover, TIME>(t, ws, cx0, cy0, wp, bx, by, px, py, R, el, er, eb, et, best, bval);
        else square3_staged<ALG, TIME>(t, ws, cx0, cy0, wp, bx, by, px, py, R, el, er, eb, et, best, bval);
    }
    // every site not yet visited lies in ring max(r0, 2) or beyond, at least
    // gap away: if one may still cover the pixel and be nearer, the pixel
    // joins its item's list and is finished later with the other such
    // pixels, on full waves
    const int gap = e0 + (max(r0, 2) - 1) * bs;
    pending = gap <= t.Rsb && (uint64_t)gap * (uint64_t)gap <= (best >> 32);
}

// The rest of a pixel's search from ring max(r0, 2) on, from the global
// arrays: the pixels of an item that get there are gathered first, so that
// these long searches (sparse stretches) run side by side on full waves
// instead of one or two lanes of a wave each.  best / bval hold what rings
// 0-1 found.
template <int ALG, bool TIME, bool DENSE>
__device__ __forceinline__ void finish_pixel(const TileDev &t, const uint32_t *__restrict__ offs,
                                             const uint8_t *__restrict__ edist,
                                             const int4 *__restrict__ sorted, int k, int l,
                                             uint32_t sub, uint32_t G, uint64_t &best, int &bval)
{
    const int bs = t.bs;
    const int bx = t.bK + div_bm(t, k), by = t.bK + div_bm(t, l);
    const int r0 = (int)edist[t.e_base + (int64_t)by * t.nbx4 + bx];
    const int px = t.x0 + k * t.ps_u, py = t.y0 + l * t.ps_u;
    const int R = (ALG == TNNI_ALG_PLANES) ? t.Rs : t.Rsb;
    const int el = px - (t.gx0 + bx * bs) + 1, er = t.gx0 + (bx + 1) * bs - px;
    const int eb = py - (t.gy0 + by * bs) + 1, et = t.gy0 + (by + 1) * bs - py;
    const int e0 = min(min(el, er), min(eb, et));
    if (!DENSE && t.dense) {
        // (a batch of sparse tiles: never taken, one pixel a thread)
        for (int r = max(r0, 2);; ++r) {
            const int gap = e0 + (r - 1) * bs;
            if (gap > t.Rsb || (uint64_t)gap * (uint64_t)gap > (best >> 32)) break;
            if (r == r0) {
                square_global<ALG, TIME>(t, offs, sorted, bx, by, r + 1, px, py, R, best, bval);
                ++r;
            } else {
                ring_global<ALG, TIME>(t, offs, sorted, bx, by, r, px, py, R, best, bval);
            }
        }
        return;
    }
    if (DENSE && t.dense) {
        // ring by ring from max(r0, 2), stopping as soon as the gap passes
        // the best so far; a sparse stretch (r0 >= 2) opens with rings r0
        // and r0 + 1 as one square (every ring inside r0 is empty).  The G
        // lanes of a group visit every G-th site of each range and take the
        // group's least key after every ring, so they all stop together
        for (int r = max(r0, 2);; ++r) {
            const int gap = e0 + (r - 1) * bs;
            if (gap > t.Rsb || (uint64_t)gap * (uint64_t)gap > (best >> 32)) break;
            // the ring's part that can still hold a winner (the group's best)
            int c0, c1, q0, q1;
            site_reach(t, px, py, R, best >> 32, c0, c1, q0, q1);
            if (r == r0) {
                square_global_box<ALG, TIME>(t, offs, sorted, max(c0, bx - r - 1), min(c1, bx + r + 1),
                                             max(q0, by - r - 1), min(q1, by + r + 1), px, py, R, sub, G, best,
                                             bval);
                ++r;
            } else {
                ring_global_part<ALG, TIME>(t, offs, sorted, bx, by, r, px, py, R, sub, G, best, bval, c0, c1, q0,
                                            q1);
            }
            for (uint32_t m = G >> 1; m; m >>= 1) {
                const uint32_t ohi = (uint32_t)__shfl_xor((int)(uint32_t)(best >> 32), (int)m, 64);
                const uint32_t olo = (uint32_t)__shfl_xor((int)(uint32_t)best, (int)m, 64);
                const uint64_t o = ((uint64_t)ohi << 32) | olo;
                best = o < best ? o : best;
            }
        }
        return;
    }
    // the last ring whose sites could still beat the best so far or lie in
    // the site region: the ring loop above would stop before the first ring
    // past it
    const uint64_t bd = best >> 32;
    int rl = max(r0, 2) - 1;
    for (;;) {
        const int gap = e0 + rl * bs;  // of ring rl + 1
        if (gap > t.Rsb || (uint64_t)gap * (uint64_t)gap > bd) break;
        ++rl;
    }
    // rings max(r0, 2) .. rl as the one square of radius rl (one range per
    // bucket row, the row offsets loaded four rows at a time): the rings
    // inside it are empty (below r0) or already visited (0 and 1, a few
    // sites at this density), and a site seen twice changes no minimum
    // and within the buckets that can hold a site both covering the pixel
    // and at least as near as the best so far: |dx| and |dy| at most D =
    // min(R, floor(sqrt(best d2))) (a site at exactly the best distance may
    // still win on its index)
    if (rl >= max(r0, 2)) {
        int c0, c1, q0, q1;
        site_reach(t, px, py, R, bd, c0, c1, q0, q1);
        square_global_box<ALG, TIME>(t, offs, sorted, max(c0, bx - rl), min(c1, bx + rl), max(q0, by - rl),
                                     min(q1, by + rl), px, py, R, sub, G, best, bval);
    }
}

template <int MODE>
struct Acc {
    int64_t s64 = 0;
    float hi = 0.f, lo = 0.f;
    uint32_t numer = 0;
    uint32_t cnt = 0;

    __device__ __forceinline__ void add(int32_t v)
    {
        if (MODE == TNNI_ACC_FIXED64) {
            s64 += v;
            ++cnt;
        } else if (MODE == TNNI_ACC_FP32_2SUM) {
            // Knuth's TwoSum: hi + err == old hi + x exactly
            const float x = (float)v;
            const float s = hi + x;
            const float bb = s - hi;
            const float err = (hi - (s - bb)) + (x - bb);
            hi = s;
            lo += err;
            ++cnt;
        } else {
            // the reference skips pixels whose colour is 0 (cuda_nni.cu:423)
            if (v != 0) {
                numer += (uint32_t)v;  // int32 wrap, cuda_nni.cu:447
                ++cnt;
            }
        }
    }

    // add(v) where take, else nothing, without a branch: adding an integer
    // 0 changes no sum bit in any mode (FP32 two-sum: neither hi nor lo is
    // ever -0.0, since they start at +0.0 and an exact zero sum of
    // non-negative-zero terms rounds to +0.0), and only taken pixels count.
    __device__ __forceinline__ void add_if(bool take, int32_t v)
    {
        if (MODE == TNNI_ACC_FIXED64) {
            s64 += take ? v : 0;
            cnt += take ? 1u : 0u;
        } else if (MODE == TNNI_ACC_FP32_2SUM) {
            const float x = take ? (float)v : 0.0f;
            const float s = hi + x;
            const float bb = s - hi;
            const float err = (hi - (s - bb)) + (x - bb);
            hi = s;
            lo += err;
            cnt += take ? 1u : 0u;
        } else {
            const bool t = take && v != 0;  // cuda_nni.cu:423
            numer += t ? (uint32_t)v : 0u;
            cnt += t ? 1u : 0u;
        }
    }

    // The running sums as 16 bytes, carried between the slices of a
    // space-time tile.
    __device__ __forceinline__ void load(const uint4 &v)
    {
        if (MODE == TNNI_ACC_FIXED64) s64 = (int64_t)(((uint64_t)v.y << 32) | v.x);
        else if (MODE == TNNI_ACC_FP32_2SUM) {
            hi = __uint_as_float(v.x);
            lo = __uint_as_float(v.y);
        } else numer = v.x;
        cnt = v.z;
    }
    __device__ __forceinline__ uint4 save() const
    {
        uint4 v;
        if (MODE == TNNI_ACC_FIXED64) {
            v.x = (uint32_t)(uint64_t)s64;
            v.y = (uint32_t)((uint64_t)s64 >> 32);
        } else if (MODE == TNNI_ACC_FP32_2SUM) {
            v.x = __float_as_uint(hi);
            v.y = __float_as_uint(lo);
        } else {
            v.x = numer;
            v.y = 0u;
        }
        v.z = cnt;
        v.w = 0u;
        return v;
    }

    __device__ __forceinline__ float result(const TileDev &t) const
    {
        if (MODE == TNNI_ACC_FIXED64) {
            if (cnt == 0) return -9999.0f;
            return (float)fma((double)s64 / (double)cnt, t.sz, t.oz);
        } else if (MODE == TNNI_ACC_FP32_2SUM) {
            if (cnt == 0) return -9999.0f;
            return (float)fma(((double)hi + (double)lo) / (double)cnt, t.sz, t.oz);
        } else {
            // divide (cuda_nni.cu:497-500) and de-scale (nni.cpp:841-846)
            int32_t res;
            if (cnt == 0)
                res = (int32_t)((uint32_t)(-9999) * (uint32_t)t.multiplier);
            else
                res = (int32_t)numer / (int32_t)cnt;
            double z = 1.0 * (double)res / (double)t.multiplier;
            if (z != -9999.0) z = z - (double)(-t.point_min) - 1.0;
            return (float)z;
        }
    }
};

// LABEL_ONLY: the batch holds label-only tiles of the split form (checked
// per tile); false for batches of fused tiles only, which then carry none
// of that code.
template <int ALG, int MODE, bool TIME, bool LABEL_ONLY, bool DENSE>
__device__ __forceinline__ void process_block(const TileDev &t, int ti, int tj, uint32_t *smem,
                                              uint32_t *s_next, uint32_t *s_npend, uint32_t share,
                                              const uint32_t *__restrict__ offs,
                                              const uint8_t *__restrict__ edist,
                                              const int4 *__restrict__ sorted)
{
    const Region g = region_of(t, ti, tj);
    const int RW = g.rk1 - g.rk0 + 1, RH = g.rl1 - g.rl0 + 1;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    WaveStage &ws = reinterpret_cast<WaveStage *>(smem)[wave];
    // per labelled pixel: its squared distance and site value side by side
    // (one 8-byte LDS read per pixel in the gather)
    uint2 *s_dv = reinterpret_cast<uint2 *>(smem + kWaves * (sizeof(WaveStage) / 4));
    uint16_t *s_pend = reinterpret_cast<uint16_t *>(s_dv + RW * RH);  // pixels still searching

    // 1. discrete Voronoi of the block's pixels, 8x8 pixels per wave; the
    //    waves take blocks as they finish (an LDS counter), so one wave on
    //    dense blocks does not hold the others at the barrier below
    const int nbx8 = (RW + 7) >> 3, nby8 = (RH + 7) >> 3;
    for (int blk = wave; blk < nbx8 * nby8;) {
        const int kA = g.rk0 + (blk % nbx8) * 8, lA = g.rl0 + (blk / nbx8) * 8;
        const int kB = min(kA + 7, g.rk1), lB = min(lA + 7, g.rl1);
        const int k = kA + (lane & 7), l = lA + (lane >> 3);
        const bool active = k <= kB && l <= lB;
        uint64_t best;
        int bval;
        bool pending;
        label_block<ALG, TIME, DENSE>(t, offs, edist, sorted, ws, kA, lA, kB, lB, k, l, active, best, bval,
                               pending);
        // the pixels that still search: one list slot each, one LDS atomic
        // per wave
        const uint64_t pm = __ballot(active && pending);
        uint32_t base = 0;
        if (pm) {
            if (lane == 0) base = atomicAdd(s_npend, (uint32_t)__popcll(pm));
            base = (uint32_t)__shfl((int)base, 0, 64);
        }
        if (active && pending) {
            // what rings 0-1 found: d2 and, until the search ends, the index
            const int p = (l - g.rl0) * RW + (k - g.rk0);
            s_dv[p] = make_uint2((uint32_t)(best >> 32), (uint32_t)best);
            s_pend[base + (uint32_t)__popcll(pm & ((1ull << lane) - 1ull))] = (uint16_t)p;
        } else if (active) {
            // a space-time winner's value, through its index
            if (TIME) bval = best == ~0ull ? 0 : site_value(t, glob(t.Z)[(uint32_t)best]);
            const uint32_t d2 = (uint32_t)(best >> 32);
            s_dv[(l - g.rl0) * RW + (k - g.rk0)] = make_uint2(d2, (uint32_t)bval);
            if (k >= g.ok0 && k <= g.ok1 && l >= g.ol0 && l <= g.ol1) {
                const int64_t o = (int64_t)l * t.W + k;
                if (t.out_labels) glob(t.out_labels)[o] = (best == ~0ull) ? -1 : (int32_t)(uint32_t)best;
                if (t.out_d2) glob(t.out_d2)[o] = d2;
                if (LABEL_ONLY && t.out_val) glob(t.out_val)[o] = bval;
            }
        }
        uint32_t nxt = 0;
        if (lane == 0) nxt = atomicAdd(s_next, 1u);
        blk = __shfl((int)nxt, 0, 64);
    }
    __syncthreads();

    // 1b. the pixels whose search goes past ring 1.  Where there are few
    //     of them (a sparse stretch beside a cluster: each one's square
    //     may hold thousands of sites), G lanes of a wave (a power of two)
    //     share each pixel's square and the group takes the least key of
    //     its lanes: the least (d2, index) is the same whichever lane met
    //     it.  Groups cover up to share times the workgroup's lanes in
    //     pixels.  Ring-by-ring searches (dense tiles) are shared the
    //     same way, the group taking its least key after every ring
    const uint32_t npend = *s_npend;
    uint32_t G = 1;
    if (DENSE || !t.dense)
        while (G < 64 && npend * (2 * G) <= share * (uint32_t)kBlock) G *= 2;
    const uint32_t sub = threadIdx.x & (G - 1), ngrp = kBlock / G;
    for (uint32_t q = threadIdx.x / G; q < npend; q += ngrp) {
        const int p = s_pend[q];
        const int k = g.rk0 + p % RW, l = g.rl0 + p / RW;
        const uint2 dv = s_dv[p];
        uint64_t best = ((uint64_t)dv.x << 32) | dv.y;
        int bval = (!TIME && G == 1 && best != ~0ull) ? site_value(t, glob(t.Z)[(uint32_t)best]) : 0;
        finish_pixel<ALG, TIME, DENSE>(t, offs, edist, sorted, k, l, sub, G, best, bval);
        if (G > 1) {
            // every lane of the group is here (same q, same trip count)
            for (uint32_t m = G >> 1; m; m >>= 1) {
                const uint32_t ohi = (uint32_t)__shfl_xor((int)(uint32_t)(best >> 32), (int)m, 64);
                const uint32_t olo = (uint32_t)__shfl_xor((int)(uint32_t)best, (int)m, 64);
                const uint64_t o = ((uint64_t)ohi << 32) | olo;
                best = o < best ? o : best;
            }
            if (sub != 0) continue;
        }
        if (TIME || G > 1) bval = best == ~0ull ? 0 : site_value(t, glob(t.Z)[(uint32_t)best]);
        const uint32_t d2 = (uint32_t)(best >> 32);
        s_dv[p] = make_uint2(d2, (uint32_t)bval);
        if (k >= g.ok0 && k <= g.ok1 && l >= g.ol0 && l <= g.ol1) {
            const int64_t o = (int64_t)l * t.W + k;
            if (t.out_labels) glob(t.out_labels)[o] = (best == ~0ull) ? -1 : (int32_t)(uint32_t)best;
            if (t.out_d2) glob(t.out_d2)[o] = d2;
            if (LABEL_ONLY && t.out_val) glob(t.out_val)[o] = bval;
        }
    }
    // a label-only item has written its pixels' values and squared
    // distances: the split form's gather pass accumulates the cells
    if (LABEL_ONLY && t.label_only) return;
    __syncthreads();

    // 2. query gather: cell q steals pixel x iff x is in q's region and
    //    |x - q|^2 <= d2(x); accumulate the label's value, then divide.
    const int qa = query_lo(t), qb = query_hi(t);
    const int ncx = g.i1 - g.i0, ncy = g.j1 - g.j0;
    // A 3-pixel query window (scale 1): a cell whose window lies inside the
    // raster tests the nine pixels at the same offsets as every other, so
    // their squared query distances (plus a space-time slice's offset, or
    // "never" outside a cone's disk) are one uniform table, and a test is
    // dqt < d2 + 1 (unsigned: d2 + 1 wraps to 0 where no site covers the
    // pixel, and is at most 2^32 - 1 otherwise, so this is dq <= d2)
    const bool span3 = qb - qa == 2;
    uint32_t dqt[9];
#pragma unroll
    for (int n = 0; n < 3; ++n)
#pragma unroll
        for (int m = 0; m < 3; ++m) {
            const uint32_t ux = (uint32_t)(t.x0 + (qa + m) * t.ps_u), uy = (uint32_t)(t.y0 + (qa + n) * t.ps_u);
            const uint32_t dq = ux * ux + uy * uy;
            dqt[n * 3 + m] = (ALG == TNNI_ALG_CONES && dq > t.Rq2) ? 0xFFFFFFFFu : dq + (TIME ? t.jj2 : 0u);
        }
    // cell c = (ci, cj) of the block, stepped by kBlock cells without a
    // division per cell
    const int step_j = kBlock / ncx, step_i = kBlock - step_j * ncx;
    int ci = (int)threadIdx.x % ncx, cj = (int)threadIdx.x / ncx;
    for (int c = threadIdx.x; c < ncx * ncy; c += kBlock) {
        const int i = g.i0 + ci, j = g.j0 + cj;
        ci += step_i;
        cj += step_j;
        if (ci >= ncx) {
            ci -= ncx;
            ++cj;
        }
        const int ma = max(qa, -i * t.scale), mb = min(qb, t.W - 1 - i * t.scale);
        const int na = max(qa, -j * t.scale), nb = min(qb, t.H - 1 - j * t.scale);
        Acc<MODE> acc;
        const int64_t cell = (int64_t)j * t.ncols + i;
        if (TIME && t.acc_load) acc.load(t.acc[cell]);
        if (span3 && ma == qa && mb == qb && na == qa && nb == qb) {
            // the same pixels in the same order as the loop below
            const int p0 = (j * t.scale + qa - g.rl0) * RW - g.rk0 + i * t.scale + qa;
#pragma unroll
            for (int n = 0; n < 3; ++n)
#pragma unroll
                for (int m = 0; m < 3; ++m) {
                    const int q = p0 + n * RW + m;
                    const uint2 dv = s_dv[q];
                    acc.add_if(dqt[n * 3 + m] < dv.x + 1u, (int32_t)dv.y);
                }
            if (!TIME || t.acc_final) glob(t.out_h)[cell] = acc.result(t);
            else t.acc[cell] = acc.save();
            continue;
        }
        for (int n = na; n <= nb; ++n) {
            const uint32_t uy = (uint32_t)(t.y0 + n * t.ps_u);
            const int row = (j * t.scale + n - g.rl0) * RW - g.rk0 + i * t.scale;
            for (int m = ma; m <= mb; ++m) {
                const uint32_t ux = (uint32_t)(t.x0 + m * t.ps_u);
                // the query region is in the plane; a space-time query adds
                // its time offset to the distance only
                const uint32_t dq = ux * ux + uy * uy;
                const uint2 dv = s_dv[row + m];
                const uint32_t d2 = dv.x;
                bool steal = d2 != kNoD2 && dq + (TIME ? t.jj2 : 0u) <= d2;
                if (ALG == TNNI_ALG_CONES) steal = steal && dq <= t.Rq2;
                acc.add_if(steal, (int32_t)dv.y);
            }
        }
        if (!TIME || t.acc_final) glob(t.out_h)[cell] = acc.result(t);
        else t.acc[cell] = acc.save();
    }
}

// DENSE: the batch may hold dense tiles, whose long searches (ring by
// ring) are shared by lane groups; false builds the kernel for batches of
// sparse tiles, whose register allocation the shared ring searches would
// disturb (a dense tile there would search one pixel a thread, as exactly).
template <int ALG, int MODE, bool TIME, bool LABEL_ONLY, bool DENSE = true>
__global__ void __launch_bounds__(kBlock)
k_tiles(BatchDev *b, const TileDev *__restrict__ tiles, const uint32_t *__restrict__ offs,
        const uint8_t *__restrict__ edist, const int4 *__restrict__ sorted)
{
    extern __shared__ __align__(16) uint32_t smem[];
    __shared__ uint32_t s_item;
    __shared__ uint32_t s_next;   // next label block of the current item
    __shared__ uint32_t s_npend;  // its pixels still searching past ring 1
    const int64_t n_items = b->n_items;
    const int n_tiles = b->n_tiles;
    // Fewer than 8 items per resident workgroup: the launch's time is its
    // slowest items, so a long search is shared more widely (several
    // pixels per lane group in turn) to shorten the heaviest ones; with
    // more items the machine is full and one pass of groups does less work
    const uint32_t share = n_items < 8 * (int64_t)gridDim.x ? 4u : 1u;
    for (;;) {
        if (threadIdx.x == 0) {
            s_item = atomicAdd(&b->work_counter, 1u);
            s_next = kWaves;  // blocks 0..3 go to waves 0..3
            s_npend = 0u;
        }
        __syncthreads();
        // (readfirstlane is a signed int: read the counter back unsigned)
        const int64_t item = (int64_t)(uint32_t)__builtin_amdgcn_readfirstlane((int)s_item);
        __syncthreads();
        if (item >= n_items) break;
        int lo = 0, hi = n_tiles - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tiles[mid].item_base <= item) lo = mid;
            else hi = mid - 1;
        }
        // the descriptor read in place: the kernel never writes the tile
        // array (restrict), so its fields load where they are used rather
        // than all at once into registers for the whole item
        const TileDev &t = tiles[lo];
        const int64_t local = item - t.item_base;
        const int tj = (int)(local / t.ntx), ti = (int)(local % t.ntx);
        process_block<ALG, MODE, TIME, LABEL_ONLY, DENSE>(t, ti, tj, smem, &s_next, &s_npend, share, offs, edist,
                                                         sorted);
        __syncthreads();
    }
}

// 8-byte slots of an interior block's pixel window in the split pass's
// LDS, rounded up to an even count: what follows the window (the scatter's
// sums, the gather's partial sums) starts 16-byte aligned.
__device__ __forceinline__ size_t split_window_slots(const TileDev &t, int qa, int qb)
{
    const size_t side = (size_t)(t.gb - 1) * t.scale + (size_t)max(1, qb - qa + 1);
    return (side * side + 1) & ~(size_t)1;
}

__device__ __forceinline__ int wave_min(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = min(v, __shfl_xor(v, o, 64));
    return v;
}
__device__ __forceinline__ int wave_max(int v)
{
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) v = max(v, __shfl_xor(v, o, 64));
    return v;
}

// The scatter form's accumulation of one block of cells [i0, i1) x
// [j0, j1): every pixel of the block's window (ww x wh pixels from raster
// pixel (kx0, ky0), read once each from the squared-distance and value
// rasters) adds its site value to the block's cells whose query point it
// would give up, |x - q|^2 <= d2(x) (the steal test of the gather), the way
// cuda_nni.cu:407-488 spreads a pixel over its disk, there with two global
// atomics per stolen pair.  Here:
//   - a pixel visits only the cells of its disk's bounding box (a radius of
//     at least sqrt(d2), clipped to the query window [qa, qb]);
//   - each pixel is read once, so the window is not staged in LDS and a
//     block can be larger than a gather block (less window overlap);
//   - the block's sums live in LDS (s_sum, s_cnt), so nothing is added
//     to device memory but the finished heights;
//   - a wavefront walks the union of its lanes' boxes as offsets from
//     each lane's first possible cell, ceil((x - qb) / scale), so the
//     loop is uniform and for one offset up to `scale` neighbouring lanes
//     hit one cell.  They add to LDS directly: summing them across the
//     wavefront first (a segmented suffix sum, by shuffles or by DPP row
//     shifts) was measured slower on gfx950 than the LDS's own handling
//     of the conflicting atomics.
// Integer sums only (FIXED64 or S0), so the order of the additions does
// not change a bit of the result.
template <int ALG, int MODE>
__device__ __forceinline__ void scatter_block(const TileDev &t, int kx0, int ky0, int ww, int wh, int i0,
                                              int i1, int j0, int j1, int qa, int qb,
                                              unsigned long long *s_sum, uint32_t *s_cnt)
{
    static_assert(MODE != TNNI_ACC_FP32_2SUM, "the scatter form sums integers");
    typedef const __attribute__((address_space(1))) uint32_t gu32;
    gu32 *val = (gu32 *)t.out_val;
    gu32 *d2r = (gu32 *)t.out_d2;
    const int gb = t.gb, s = t.scale, ps = t.ps_u, x0 = t.x0, y0 = t.y0, W = t.W;
    const int lane = (int)threadIdx.x & 63;
    for (int k = threadIdx.x; k < gb * gb; k += kBlock) {
        s_sum[k] = 0ull;
        s_cnt[k] = 0u;
    }
    __syncthreads();
    constexpr int kBig = 1 << 30;
    const int npx = ww > 0 && wh > 0 ? ww * wh : 0;
    // window pixel p: row p / ww by a float reciprocal, corrected (p < 2^24)
    const float inv_ww = 1.0f / (float)max(ww, 1);
    auto fetch = [&](int p, int &wx, int &wy) -> uint2 {
        wy = (int)((float)p * inv_ww);
        wx = p - wy * ww;
        if (wx < 0) {
            --wy;
            wx += ww;
        } else if (wx >= ww) {
            ++wy;
            wx -= ww;
        }
        if (p >= npx) return make_uint2(kNoD2, 0u);
        const int64_t o = (int64_t)(ky0 + wy) * W + kx0 + wx;
        return make_uint2(d2r[o], val[o]);
    };
    int nwx, nwy;
    uint2 nxt = fetch((int)threadIdx.x, nwx, nwy);  // the next pixel, loaded a round ahead
    for (int p0 = (int)threadIdx.x - lane; p0 < npx; p0 += kBlock) {  // wave-uniform
        const int p = p0 + lane;
        const bool live = p < npx;
        const uint2 e = nxt;
        const int wx = nwx, wy = nwy;
        nxt = fetch(p + kBlock, nwx, nwy);
        const int px = kx0 + wx, py = ky0 + wy;
        bool has = live && e.x != kNoD2;
        if (MODE == TNNI_ACC_S0) has = has && (int32_t)e.y != 0;  // cuda_nni.cu:423
        // the disk's box in pixel offsets from a cell's base pixel, then in
        // cells: m = x - i * scale within [mlo, mhi]
        const int ib = -floordiv_i(qb - px, s), jb = -floordiv_i(qb - py, s);
        int ilo = kBig, ihi = -kBig, jlo = kBig, jhi = -kBig;
        if (has) {
            const int r = (int)sqrtf((float)e.x) + 2;  // >= sqrt(d2)
            const int mlo = max(qa, -floordiv_i(r + x0, ps)), mhi = min(qb, floordiv_i(r - x0, ps));
            const int nlo = max(qa, -floordiv_i(r + y0, ps)), nhi = min(qb, floordiv_i(r - y0, ps));
            ilo = max(i0, -floordiv_i(mhi - px, s));
            ihi = min(i1 - 1, floordiv_i(px - mlo, s));
            jlo = max(j0, -floordiv_i(nhi - py, s));
            jhi = min(j1 - 1, floordiv_i(py - nlo, s));
            if (ilo > ihi || jlo > jhi) {  // reaches no cell of the block
                ilo = jlo = kBig;
                ihi = jhi = -kBig;
            }
        }
        const int cx0 = wave_min(ilo - ib), cx1 = wave_max(ihi - ib);
        const int cy0 = wave_min(jlo - jb), cy1 = wave_max(jhi - jb);
        for (int cy = cy0; cy <= cy1; ++cy) {
            const int j = jb + cy;
            const uint32_t uy = (uint32_t)(y0 + (py - j * s) * ps);
            const uint32_t uy2 = uy * uy;
            const bool row_in = j >= jlo && j <= jhi;
            for (int cx = cx0; cx <= cx1; ++cx) {
                const int i = ib + cx;
                const uint32_t ux = (uint32_t)(x0 + (px - i * s) * ps);
                const uint32_t dq = ux * ux + uy2;
                bool take = row_in && i >= ilo && i <= ihi && dq <= e.x;
                if (ALG == TNNI_ALG_CONES) take = take && dq <= t.Rq2;
                if (take) {
                    const int cell = (j - j0) * gb + (i - i0);
                    if (MODE == TNNI_ACC_FIXED64) atomicAdd(&s_sum[cell], (unsigned long long)(int64_t)(int32_t)e.y);
                    else atomicAdd((uint32_t *)&s_sum[cell], e.y);
                    atomicAdd(&s_cnt[cell], 1u);
                }
            }
        }
    }
    __syncthreads();
    for (int k = threadIdx.x; k < gb * gb; k += kBlock) {
        const int i = i0 + k % gb, j = j0 + k / gb;
        if (i >= i1 || j >= j1) continue;
        Acc<MODE> acc;
        acc.s64 = (int64_t)s_sum[k];
        acc.numer = (uint32_t)s_sum[k];
        acc.cnt = s_cnt[k];
        glob(t.out_h)[(int64_t)j * t.ncols + i] = acc.result(t);
    }
}

// The split form's accumulation pass: the batch's split tiles in blocks of
// gb x gb cells, one workgroup a block.  The block's pixel window (every
// pixel some cell of the block may steal) is copied from the value and
// squared-distance rasters of the label-only items into LDS.  Then, in a
// tile with scatter set, scatter_block spreads the window's pixels over
// the block's cells; otherwise each thread gathers one cell from the
// window, in the pixel order of process_block's gather, so FP32 two-sum
// sums are bit-equal to the fused form's.  The blocks of tile k are
// [gitem_base, gitem_base + gntx * gnty) of the n_gather_items; other
// tiles' ranges are empty.
template <int ALG, int MODE>
__global__ void __launch_bounds__(kBlock)
k_gather_split(const BatchDev *b, const TileDev *tiles)
{
    extern __shared__ __align__(16) uint2 win[];  // {d2, value} per window pixel
    const int64_t n_items = b->n_gather_items;
    const int n_tiles = b->n_tiles;
    for (int64_t item = blockIdx.x; item < n_items; item += gridDim.x) {
        int lo = 0, hi = n_tiles - 1;  // the last tile whose range starts at or before item
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tiles[mid].gitem_base <= item) lo = mid;
            else hi = mid - 1;
        }
        const TileDev &t = tiles[lo];
        const int gb = t.gb, scale = t.scale, W = t.W, H = t.H;
        const int64_t local = item - t.gitem_base;
        const int bj = (int)(local / t.gntx), bi = (int)(local - (int64_t)bj * t.gntx);
        const int i0 = bi * gb, j0 = bj * gb;
        const int i1 = min(i0 + gb, t.ncols), j1 = min(j0 + gb, t.nrows);
        const int qa = query_lo(t), qb = query_hi(t);
        const int kx0 = max(0, i0 * scale + qa), kx1 = min(W - 1, (i1 - 1) * scale + qb);
        const int ky0 = max(0, j0 * scale + qa), ky1 = min(H - 1, (j1 - 1) * scale + qb);
        const int ww = kx1 - kx0 + 1, wh = ky1 - ky0 + 1;
        if constexpr (MODE != TNNI_ACC_FP32_2SUM) {
            if (t.scatter) {  // the block's sums are all its LDS
                unsigned long long *s_sum = (unsigned long long *)win;
                scatter_block<ALG, MODE>(t, kx0, ky0, ww, wh, i0, i1, j0, j1, qa, qb, s_sum,
                                         (uint32_t *)(s_sum + gb * gb));
                __syncthreads();
                continue;
            }
        }
        typedef const __attribute__((address_space(1))) int32_t gi32;
        typedef const __attribute__((address_space(1))) uint32_t gu32;
        gi32 *val = (gi32 *)t.out_val;
        gu32 *d2r = (gu32 *)t.out_d2;
        // the window, four pixels a thread in flight at once
        const int npx = ww > 0 && wh > 0 ? ww * wh : 0;
        for (int p0 = threadIdx.x; p0 < npx; p0 += 4 * kBlock) {
            uint32_t dv[4], vv[4];
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int p = p0 + q * kBlock;
                dv[q] = kNoD2;
                vv[q] = 0u;
                if (p < npx) {
                    const int r = p / ww, c = p - r * ww;
                    const int64_t o = (int64_t)(ky0 + r) * W + kx0 + c;
                    dv[q] = d2r[o];
                    vv[q] = (uint32_t)val[o];
                }
            }
#pragma unroll
            for (int q = 0; q < 4; ++q)
                if (p0 + q * kBlock < npx) win[p0 + q * kBlock] = make_uint2(dv[q], vv[q]);
        }
        __syncthreads();
        // Blocks of fewer cells than threads (wide query windows): in the
        // integer modes, whose partial sums add up exactly, the window rows
        // of a cell are shared by `parts` threads (a power of two) and the
        // partial sums added in LDS; FP32 two-sum keeps one thread a cell,
        // summing in process_block's order.
        const int cells = gb * gb;
        const int parts = MODE != TNNI_ACC_FP32_2SUM && cells < kBlock ? kBlock / cells : 1;
        const int cell = (int)threadIdx.x % cells, part = (int)threadIdx.x / cells;
        const int ci = cell % gb, cj = cell / gb;
        const int i = i0 + ci, j = j0 + cj;
        Acc<MODE> acc;
        const bool mine = part < parts && i < i1 && j < j1;
        if (mine) {
            const int ma = max(qa, -i * scale), mb = min(qb, W - 1 - i * scale);
            const int na = max(qa, -j * scale), nb = min(qb, H - 1 - j * scale);
            const int x0 = t.x0, y0 = t.y0, ps = t.ps_u;
            const uint32_t Rq2 = t.Rq2;
            for (int n = na + part; n <= nb; n += parts) {
                const uint32_t uy = (uint32_t)(y0 + n * ps);
                const uint2 *wr = win + (j * scale + n - ky0) * ww + (i * scale - kx0);
#pragma unroll 2
                for (int m = ma; m <= mb; ++m) {
                    const uint32_t ux = (uint32_t)(x0 + m * ps);
                    const uint32_t dq = ux * ux + uy * uy;
                    const uint2 e = wr[m];
                    bool steal = e.x != kNoD2 && dq <= e.x;
                    if (ALG == TNNI_ALG_CONES) steal = steal && dq <= Rq2;
                    acc.add_if(steal, (int32_t)e.y);
                }
            }
        }
        if (parts > 1) {
            // the partial sums after the window, one 16-byte slot a thread
            uint4 *s_part = (uint4 *)(win + split_window_slots(t, qa, qb));
            s_part[threadIdx.x] = acc.save();
            __syncthreads();
            if (part == 0 && mine) {
                for (int k = 1; k < parts; ++k) {
                    Acc<MODE> o;
                    o.load(s_part[k * cells + cell]);
                    acc.s64 += o.s64;
                    acc.numer += o.numer;
                    acc.cnt += o.cnt;
                }
            }
        }
        if (part == 0 && mine) glob(t.out_h)[(int64_t)j * t.ncols + i] = acc.result(t);
        __syncthreads();
    }
}

template <int ALG, int MODE, bool TIME, bool LABEL_ONLY, bool DENSE = true>
hipError_t launch_tiles(const DevBuffers &d, const LaunchCfg &cfg, hipStream_t s)
{
    hipLaunchKernelGGL((k_tiles<ALG, MODE, TIME, LABEL_ONLY, DENSE>), dim3(cfg.main_grid), dim3(kBlock), cfg.main_lds,
                       s, d.batch, d.tiles, d.offs, d.edist, d.sorted);
    if (!TIME && cfg.gather_grid > 0)
        hipLaunchKernelGGL((k_gather_split<ALG, MODE>), dim3(cfg.gather_grid), dim3(kBlock), cfg.gather_lds, s,
                           d.batch, d.tiles);
    return hipGetLastError();
}

// ---- space-time cache of per-time layers ---------------------------------

// The query gather of one cached slice (values and squared distances of
// its pixels in HBM): the cell loop of process_block's gather, in the same
// order of pixels, so the sums are the same bit for bit.
template <int ALG, int MODE>
__global__ void __launch_bounds__(kBlock)
k_gather_slice(const TileDev t, const int32_t *__restrict__ val, const uint32_t *__restrict__ d2r)
{
    const int qa = query_lo(t), qb = query_hi(t);
    const int64_t cells = (int64_t)t.nrows * t.ncols;
    for (int64_t cell = (int64_t)blockIdx.x * kBlock + threadIdx.x; cell < cells;
         cell += (int64_t)gridDim.x * kBlock) {
        const int j = (int)(cell / t.ncols), i = (int)(cell - (int64_t)j * t.ncols);
        const int ma = max(qa, -i * t.scale), mb = min(qb, t.W - 1 - i * t.scale);
        const int na = max(qa, -j * t.scale), nb = min(qb, t.H - 1 - j * t.scale);
        Acc<MODE> acc;
        if (t.acc_load) acc.load(t.acc[cell]);
        for (int n = na; n <= nb; ++n) {
            const uint32_t uy = (uint32_t)(t.y0 + n * t.ps_u);
            const int64_t row = (int64_t)(j * t.scale + n) * t.W + i * t.scale;
            for (int m = ma; m <= mb; ++m) {
                const uint32_t ux = (uint32_t)(t.x0 + m * t.ps_u);
                const uint32_t dq = ux * ux + uy * uy;
                const uint32_t d2 = d2r[row + m];
                bool steal = d2 != kNoD2 && dq + t.jj2 <= d2;
                if (ALG == TNNI_ALG_CONES) steal = steal && dq <= t.Rq2;
                acc.add_if(steal, val[row + m]);
            }
        }
        if (t.acc_final) glob(t.out_h)[cell] = acc.result(t);
        else t.acc[cell] = acc.save();
    }
}

// A space-time slice's site values and squared distances from per-year
// layers: each layer holds the 2-D Voronoi of the sites of one time (the
// nearest covering site's value and squared plane distance, kNoD2 where no
// site covers the pixel), and within the slice every site of that time is
// at the same time distance dt2.  The slice's nearest site is the least
// (d2 + dt2) over its layers; an equal sum goes to the earlier time, which
// for points sorted by time is the lower index.
__global__ void __launch_bounds__(kBlock)
k_combine_layers(const LayerSet ls, uint64_t pix, int32_t *__restrict__ val, uint32_t *__restrict__ d2)
{
    for (uint64_t p = (uint64_t)blockIdx.x * kBlock + threadIdx.x; p < pix; p += (uint64_t)gridDim.x * kBlock) {
        uint64_t best = ~0ull;
        int32_t v = 0;
        for (int k = 0; k < ls.n; ++k) {
            const uint32_t d = ls.d2[k][p];
            if (d == kNoD2) continue;
            const uint64_t key = ((uint64_t)(d + ls.dt2[k]) << 32) | (uint32_t)k;
            if (key < best) {
                best = key;
                v = ls.val[k][p];
            }
        }
        d2[p] = best == ~0ull ? kNoD2 : (uint32_t)(best >> 32);
        val[p] = v;
    }
}

// One pass over the points' times: out[0] = 1 when some T[i] > T[i + 1];
// for points sorted by time, the run of the points of time ys.t[k] (as
// float) is [out[1 + k], out[1 + kMaxLayers + k]).  The caller fills out
// with 0 and n first; every boundary (a change of time between neighbours)
// sets the runs of the times it passes.
__global__ void __launch_bounds__(kBlock)
k_time_runs(const int32_t *__restrict__ T, uint64_t n, const LayerTimes ys, uint32_t *out)
{
    for (uint64_t i = (uint64_t)blockIdx.x * kBlock + threadIdx.x; i < n; i += (uint64_t)gridDim.x * kBlock) {
        const int32_t b = T[i];
        if (i + 1 < n && b > T[i + 1]) out[0] = 1u;
        if (i > 0 && T[i - 1] == b) continue;
        const float fa = i > 0 ? (float)T[i - 1] : -__builtin_huge_valf(), fb = (float)b;
        for (int k = 0; k < ys.n; ++k) {
            const float y = ys.t[k];
            if (fa < y && y <= fb) out[1 + k] = (uint32_t)i;               // first of time >= y
            if (fa <= y && y < fb) out[1 + kMaxLayers + k] = (uint32_t)i;  // first of time > y
        }
    }
}

template <int ALG, int MODE>
hipError_t launch_gather_slice_t(const TileDev &t, const int32_t *val, const uint32_t *d2, hipStream_t s)
{
    const int64_t cells = (int64_t)t.nrows * t.ncols;
    const int grid = (int)std::max<int64_t>(1, std::min<int64_t>((cells + kBlock - 1) / kBlock, 8192));
    hipLaunchKernelGGL((k_gather_slice<ALG, MODE>), dim3(grid), dim3(kBlock), 0, s, t, val, d2);
    return hipGetLastError();
}

// The descriptors of a one-tile batch written on the device from kernel
// arguments (no host staging): the tile, the batch header, and the tile's
// chunks of bd.chunk_pts points (at least one, possibly empty).
__global__ void __launch_bounds__(kBlock)
k_single_tile_setup(const TileDev t, const BatchDev bd, TileDev *__restrict__ tiles,
                    Chunk *__restrict__ chunks, BatchDev *__restrict__ batch)
{
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        *tiles = t;
        *batch = bd;
    }
    for (int c = blockIdx.x * kBlock + threadIdx.x; c < bd.n_chunks; c += gridDim.x * kBlock) {
        Chunk ch;
        ch.tile = 0;
        ch.pad = 0;
        ch.begin = (int64_t)c * bd.chunk_pts;
        ch.end = min(t.n, ch.begin + (int64_t)bd.chunk_pts);
        chunks[c] = ch;
    }
}

}  // namespace

hipError_t launch_single_tile_setup(const TileDev &t, const BatchDev &bd, TileDev *tiles, Chunk *chunks,
                                    BatchDev *batch, hipStream_t s)
{
    const int grid = std::max(1, std::min((bd.n_chunks + kBlock - 1) / kBlock, 1024));
    hipLaunchKernelGGL(k_single_tile_setup, dim3(grid), dim3(kBlock), 0, s, t, bd, tiles, chunks, batch);
    return hipGetLastError();
}

hipError_t launch_combine_layers(const LayerSet &ls, uint64_t pix, int32_t *val, uint32_t *d2, hipStream_t s)
{
    if (ls.n < 0 || ls.n > kMaxLayers) return hipErrorInvalidValue;
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((pix + kBlock - 1) / kBlock, 8192));
    hipLaunchKernelGGL(k_combine_layers, dim3(grid), dim3(kBlock), 0, s, ls, pix, val, d2);
    return hipGetLastError();
}

hipError_t launch_time_runs(const int32_t *T, uint64_t n, const LayerTimes &ys, uint32_t *out, hipStream_t s)
{
    if (n == 0) return hipSuccess;
    if (ys.n < 0 || ys.n > kMaxLayers || n > 0xFFFFFFFFull) return hipErrorInvalidValue;
    const int grid = (int)std::max<uint64_t>(1, std::min<uint64_t>((n + kBlock - 1) / kBlock, 4096));
    hipLaunchKernelGGL(k_time_runs, dim3(grid), dim3(kBlock), 0, s, T, n, ys, out);
    return hipGetLastError();
}

hipError_t launch_gather_slice(const TileDev &t, const int32_t *val, const uint32_t *d2, hipStream_t s)
{
    switch (t.algorithm * 3 + t.mode) {
    case 0: return launch_gather_slice_t<0, 0>(t, val, d2, s);
    case 1: return launch_gather_slice_t<0, 1>(t, val, d2, s);
    case 2: return launch_gather_slice_t<0, 2>(t, val, d2, s);
    case 3: return launch_gather_slice_t<1, 0>(t, val, d2, s);
    case 4: return launch_gather_slice_t<1, 1>(t, val, d2, s);
    case 5: return launch_gather_slice_t<1, 2>(t, val, d2, s);
    default: return hipErrorInvalidValue;
    }
}

hipError_t enqueue_pipeline(const DevBuffers &d, const LaunchCfg &cfg, hipStream_t s,
                            hipEvent_t ev0, hipEvent_t ev1)
{
    hipError_t e;
    hipLaunchKernelGGL(k_band_count, dim3(cfg.sort_grid), dim3(kSortBlock), 0, s, d.batch, d.tiles,
                       d.chunks, d.hist);
    hipLaunchKernelGGL(k_scan_reduce, dim3(kScanBlocks), dim3(kBlock), 0, s, d.batch, d.hist,
                       d.partials);
    hipLaunchKernelGGL(k_scan_down, dim3(kScanBlocks), dim3(kBlock), 0, s, d.batch, d.hist,
                       d.partials);
    hipLaunchKernelGGL(k_band_scatter, dim3(cfg.sort_grid), dim3(kSortBlock), 0, s, d.batch, d.tiles,
                       d.chunks, d.hist, d.banded);
    hipLaunchKernelGGL(k_band_sort, dim3(cfg.band_grid), dim3(kSortBlock), 0, s, d.batch, d.tiles,
                       d.hist, d.banded, d.sorted, d.offs, d.hdist);
    hipLaunchKernelGGL(k_edist_cols, dim3(cfg.row_grid), dim3(kBlock), 0, s, d.batch, d.tiles,
                       d.hdist, d.edist);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev0 && (e = hipEventRecord(ev0, s)) != hipSuccess) return e;
    // space-time slices never take the split form
    if (cfg.time && cfg.label_only) return hipErrorInvalidValue;
    // fused 2-D batches without a dense tile: the kernel without the dense paths
    const int sel = (cfg.time ? 6 : cfg.label_only ? 12 : cfg.dense ? 0 : 18) + cfg.algorithm * 3 + cfg.mode;
    switch (sel) {
    case 18: e = launch_tiles<0, 0, false, false, false>(d, cfg, s); break;
    case 19: e = launch_tiles<0, 1, false, false, false>(d, cfg, s); break;
    case 20: e = launch_tiles<0, 2, false, false, false>(d, cfg, s); break;
    case 21: e = launch_tiles<1, 0, false, false, false>(d, cfg, s); break;
    case 22: e = launch_tiles<1, 1, false, false, false>(d, cfg, s); break;
    case 23: e = launch_tiles<1, 2, false, false, false>(d, cfg, s); break;
    case 0: e = launch_tiles<0, 0, false, false>(d, cfg, s); break;
    case 1: e = launch_tiles<0, 1, false, false>(d, cfg, s); break;
    case 2: e = launch_tiles<0, 2, false, false>(d, cfg, s); break;
    case 3: e = launch_tiles<1, 0, false, false>(d, cfg, s); break;
    case 4: e = launch_tiles<1, 1, false, false>(d, cfg, s); break;
    case 5: e = launch_tiles<1, 2, false, false>(d, cfg, s); break;
    case 6: e = launch_tiles<0, 0, true, false>(d, cfg, s); break;
    case 7: e = launch_tiles<0, 1, true, false>(d, cfg, s); break;
    case 8: e = launch_tiles<0, 2, true, false>(d, cfg, s); break;
    case 9: e = launch_tiles<1, 0, true, false>(d, cfg, s); break;
    case 10: e = launch_tiles<1, 1, true, false>(d, cfg, s); break;
    case 11: e = launch_tiles<1, 2, true, false>(d, cfg, s); break;
    case 12: e = launch_tiles<0, 0, false, true>(d, cfg, s); break;
    case 13: e = launch_tiles<0, 1, false, true>(d, cfg, s); break;
    case 14: e = launch_tiles<0, 2, false, true>(d, cfg, s); break;
    case 15: e = launch_tiles<1, 0, false, true>(d, cfg, s); break;
    case 16: e = launch_tiles<1, 1, false, true>(d, cfg, s); break;
    case 17: e = launch_tiles<1, 2, false, true>(d, cfg, s); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (ev1 && (e = hipEventRecord(ev1, s)) != hipSuccess) return e;
    return hipSuccess;
}

size_t tile_kernel_stage_bytes() { return (size_t)kWaves * sizeof(WaveStage); }

int tile_kernel_occupancy(size_t lds_bytes)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tiles<0, 0, false, false>, kBlock, lds_bytes) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

}  // namespace tnni
