// gfx950 kernels of the discrete-Sibson tile engine.
//
// Pipeline for a batch of tiles (a single tile is a batch of one):
//   k_zero      clear the bucket counters and the work counter
//   k_count     FP64 de-scale of every LAS point into tile-local units,
//               bucket histogram (one atomic per point, its return value is
//               the point's rank inside the bucket)
//   k_scan_*    device-wide exclusive scan of the counters -> bucket offsets
//   k_scatter   place every point's site record at offset + rank
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

__device__ __forceinline__ uint32_t bucket_of(const TileDev &t, int xu, int yu)
{
    const int bx = (xu - t.gx0) / t.bs;
    const int by = (yu - t.gy0) / t.bs;
    return (uint32_t)by * (uint32_t)t.nbx + (uint32_t)bx;
}

__global__ void __launch_bounds__(kBlock) k_zero(BatchDev *b, uint32_t *counts)
{
    const uint32_t n = b->n_scan;
    for (uint32_t i = blockIdx.x * kBlock + threadIdx.x; i < n; i += gridDim.x * kBlock)
        counts[i] = 0u;
    if (blockIdx.x == 0 && threadIdx.x == 0) b->work_counter = 0u;
}

__global__ void __launch_bounds__(kBlock)
k_count(const BatchDev *b, const TileDev *tiles, const Chunk *chunks,
        uint32_t *counts, uint32_t *rank)
{
    const int nch = b->n_chunks;
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const Chunk ch = chunks[c];
        const TileDev &t = tiles[ch.tile];
        for (int64_t i = ch.begin + threadIdx.x; i < ch.end; i += kBlock) {
            int xu, yu;
            uint32_t r = kNoD2;
            if (point_units(t, t.X[i], t.Y[i], xu, yu))
                r = atomicAdd(&counts[t.bk_base + bucket_of(t, xu, yu)], 1u);
            rank[t.pt_base + i] = r;
        }
    }
}

__global__ void __launch_bounds__(kBlock)
k_scatter(const BatchDev *b, const TileDev *tiles, const Chunk *chunks,
          const uint32_t *offsets, const uint32_t *rank, int4 *sorted)
{
    const int nch = b->n_chunks;
    for (int c = blockIdx.x; c < nch; c += gridDim.x) {
        const Chunk ch = chunks[c];
        const TileDev &t = tiles[ch.tile];
        for (int64_t i = ch.begin + threadIdx.x; i < ch.end; i += kBlock) {
            const uint32_t r = rank[t.pt_base + i];
            if (r == kNoD2) continue;
            int xu, yu;
            point_units(t, t.X[i], t.Y[i], xu, yu);
            const uint32_t pos = offsets[t.bk_base + bucket_of(t, xu, yu)] + r;
            sorted[pos] = make_int4(xu, yu, (int)i, site_value(t, t.Z[i]));
        }
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

__global__ void __launch_bounds__(kScanBlocks) k_scan_partials(uint32_t *partials)
{
    __shared__ uint32_t s_w[kScanBlocks / 64];
    const uint32_t v = partials[threadIdx.x];
    uint32_t total;
    const uint32_t ex = block_excl_scan<kScanBlocks>(v, s_w, total);
    partials[threadIdx.x] = ex;
}

__global__ void __launch_bounds__(kBlock)
k_scan_down(const BatchDev *b, uint32_t *counts, const uint32_t *partials)
{
    __shared__ uint32_t s_w[kWaves];
    const uint32_t n = b->n_scan, chunk = b->scan_chunk;
    const uint64_t beg = (uint64_t)blockIdx.x * chunk;
    const uint64_t end = min((uint64_t)n, beg + chunk);
    uint32_t run = partials[blockIdx.x];
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

__global__ void __launch_bounds__(kBlock)
k_edist_rows(const BatchDev *b, const TileDev *tiles, const uint32_t *__restrict__ offs,
             uint8_t *__restrict__ hdist)
{
    const int64_t nr = b->n_rows;
    const int n_tiles = b->n_tiles;
    for (int64_t item = blockIdx.x; item < nr; item += gridDim.x) {
        const TileDev &t = tiles[tile_of_row(tiles, n_tiles, item)];
        const int64_t rb = t.bk_base + (item - t.row_base) * t.nbx;
        const uint32_t *o = offs + rb;
        const int nbx = t.nbx, cap = t.ecap;
        for (int x = threadIdx.x; x < nbx; x += kBlock) {
            int d = cap;
            for (int q = 0; q < cap; ++q) {
                const bool l = x - q >= 0 && o[x - q + 1] > o[x - q];
                const bool r = x + q < nbx && o[x + q + 1] > o[x + q];
                if (l || r) {
                    d = q;
                    break;
                }
            }
            hdist[rb + x] = (uint8_t)d;
        }
    }
}

__global__ void __launch_bounds__(kBlock)
k_edist_cols(const BatchDev *b, const TileDev *tiles, const uint8_t *__restrict__ hdist,
             uint8_t *__restrict__ edist)
{
    const int64_t nr = b->n_rows;
    const int n_tiles = b->n_tiles;
    for (int64_t item = blockIdx.x; item < nr; item += gridDim.x) {
        const TileDev &t = tiles[tile_of_row(tiles, n_tiles, item)];
        const int row = (int)(item - t.row_base);
        const int nbx = t.nbx, nby = t.nby;
        const uint8_t *h = hdist + t.bk_base;
        for (int x = threadIdx.x; x < nbx; x += kBlock) {
            int e = t.ecap;
            for (int q = 0; q < e; ++q) {
                if (row - q >= 0) e = min(e, max(q, (int)h[(int64_t)(row - q) * nbx + x]));
                if (row + q < nby) e = min(e, max(q, (int)h[(int64_t)(row + q) * nbx + x]));
            }
            edist[t.bk_base + (int64_t)row * nbx + x] = (uint8_t)e;
        }
    }
}

// ---- the tile kernel -------------------------------------------------------

// Visit the sites stored at [s, e) for the pixel centre (px, py).
template <int ALG>
__device__ __forceinline__ void scan_range(uint32_t s, uint32_t e, const int4 *__restrict__ sorted,
                                           int px, int py, int R, uint32_t R2,
                                           uint64_t &best, int &bval)
{
    const uint32_t twoR = 2u * (uint32_t)R;
    for (uint32_t p = s; p < e; ++p) {
        const int4 st = sorted[p];
        const int dx = px - st.x, dy = py - st.y;
        const uint32_t ux = (uint32_t)dx, uy = (uint32_t)dy;
        const uint32_t dd = ux * ux + uy * uy;
        // inside the L-inf extent of the region (planes: the region itself)
        bool cov = max((uint32_t)(dx + R), (uint32_t)(dy + R)) <= twoR;
        if (ALG == TNNI_ALG_CONES) cov = cov && dd <= R2;
        const uint64_t key = cov ? (((uint64_t)dd << 32) | (uint32_t)st.z) : ~0ull;
        if (key < best) {
            best = key;
            bval = st.w;
        }
    }
}

// Nearest covering site of pixel (k, l): ring search over the bucket grid.
// Rings closer than the bucket's empty distance hold no site and are
// skipped; the search stops when every unvisited site is provably farther
// than the best so far (strictly, so an equal distance still reaches the
// lower index) or lies outside the site region.
template <int ALG>
__device__ __forceinline__ void label_pixel(const TileDev &t, const uint32_t *__restrict__ offs,
                                            const uint8_t *__restrict__ edist,
                                            const int4 *__restrict__ sorted, int k, int l,
                                            uint64_t &best, int &bval)
{
    best = ~0ull;
    bval = 0;
    const int px = t.x0 + k * t.ps_u, py = t.y0 + l * t.ps_u;
    const int bs = t.bs, nbx = t.nbx, nby = t.nby;
    const int bx = t.bK + k / t.bm;
    const int by = t.bK + l / t.bm;
    const uint32_t *base = offs + t.bk_base;
    const int R = (ALG == TNNI_ALG_PLANES) ? t.Rs : t.Rsb;
    // distances from the pixel centre to the edges of its own bucket
    const int el = px - (t.gx0 + bx * bs) + 1, er = t.gx0 + (bx + 1) * bs - px;
    const int eb = py - (t.gy0 + by * bs) + 1, et = t.gy0 + (by + 1) * bs - py;
    const int e0 = min(min(el, er), min(eb, et));
    for (int r = edist[t.bk_base + (int64_t)by * nbx + bx];; ++r) {
        if (r > 0) {
            // every site outside the visited (2r-1)^2 square is at least gap away
            const int gap = e0 + (r - 1) * bs;
            if (gap > t.Rsb) break;
            if ((uint64_t)gap * (uint64_t)gap > (best >> 32)) break;
        }
        for (int dy = -r; dy <= r; ++dy) {
            const int row = by + dy;
            if (row < 0 || row >= nby) continue;
            const uint32_t *rs = base + (int64_t)row * nbx;
            if (dy == -r || dy == r) {
                const int c0 = max(bx - r, 0), c1 = min(bx + r, nbx - 1);
                if (c0 <= c1) scan_range<ALG>(rs[c0], rs[c1 + 1], sorted, px, py, R, t.Rs2, best, bval);
            } else {
                if (bx - r >= 0) scan_range<ALG>(rs[bx - r], rs[bx - r + 1], sorted, px, py, R, t.Rs2, best, bval);
                if (bx + r < nbx) scan_range<ALG>(rs[bx + r], rs[bx + r + 1], sorted, px, py, R, t.Rs2, best, bval);
            }
        }
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

template <int ALG, int MODE>
__device__ __forceinline__ void process_block(const TileDev &t, int ti, int tj, uint32_t *smem,
                                              const uint32_t *__restrict__ offs,
                                              const uint8_t *__restrict__ edist,
                                              const int4 *__restrict__ sorted)
{
    const Region g = region_of(t, ti, tj);
    const int RW = g.rk1 - g.rk0 + 1, RH = g.rl1 - g.rl0 + 1;
    uint32_t *s_d2 = smem;
    int32_t *s_val = reinterpret_cast<int32_t *>(smem + RW * RH);

    // 1. discrete Voronoi of the block's pixels, 8x8 pixels per wave
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    const int nbx8 = (RW + 7) >> 3, nby8 = (RH + 7) >> 3;
    for (int blk = wave; blk < nbx8 * nby8; blk += kWaves) {
        const int k = g.rk0 + (blk % nbx8) * 8 + (lane & 7);
        const int l = g.rl0 + (blk / nbx8) * 8 + (lane >> 3);
        if (k > g.rk1 || l > g.rl1) continue;
        uint64_t best;
        int bval;
        label_pixel<ALG>(t, offs, edist, sorted, k, l, best, bval);
        const uint32_t d2 = (uint32_t)(best >> 32);
        s_d2[(l - g.rl0) * RW + (k - g.rk0)] = d2;
        s_val[(l - g.rl0) * RW + (k - g.rk0)] = bval;
        if (k >= g.ok0 && k <= g.ok1 && l >= g.ol0 && l <= g.ol1) {
            const int64_t o = (int64_t)l * t.W + k;
            if (t.out_labels) t.out_labels[o] = (best == ~0ull) ? -1 : (int32_t)(uint32_t)best;
            if (t.out_d2) t.out_d2[o] = d2;
        }
    }
    __syncthreads();

    // 2. query gather: cell q steals pixel x iff x is in q's region and
    //    |x - q|^2 <= d2(x); accumulate the label's value, then divide.
    const int qa = query_lo(t), qb = query_hi(t);
    const int ncx = g.i1 - g.i0, ncy = g.j1 - g.j0;
    for (int c = threadIdx.x; c < ncx * ncy; c += kBlock) {
        const int i = g.i0 + c % ncx, j = g.j0 + c / ncx;
        const int ma = max(qa, -i * t.scale), mb = min(qb, t.W - 1 - i * t.scale);
        const int na = max(qa, -j * t.scale), nb = min(qb, t.H - 1 - j * t.scale);
        Acc<MODE> acc;
        for (int n = na; n <= nb; ++n) {
            const uint32_t uy = (uint32_t)(t.y0 + n * t.ps_u);
            const int row = (j * t.scale + n - g.rl0) * RW - g.rk0 + i * t.scale;
            for (int m = ma; m <= mb; ++m) {
                const uint32_t ux = (uint32_t)(t.x0 + m * t.ps_u);
                const uint32_t dq = ux * ux + uy * uy;
                const uint32_t d2 = s_d2[row + m];
                bool steal = d2 != kNoD2 && dq <= d2;
                if (ALG == TNNI_ALG_CONES) steal = steal && dq <= t.Rq2;
                if (steal) acc.add(s_val[row + m]);
            }
        }
        t.out_h[(int64_t)j * t.ncols + i] = acc.result(t);
    }
}

template <int ALG, int MODE>
__global__ void __launch_bounds__(kBlock)
k_tiles(BatchDev *b, const TileDev *tiles, const uint32_t *__restrict__ offs,
        const uint8_t *__restrict__ edist, const int4 *__restrict__ sorted)
{
    extern __shared__ uint32_t smem[];
    __shared__ uint32_t s_item;
    const int64_t n_items = b->n_items;
    const int n_tiles = b->n_tiles;
    for (;;) {
        if (threadIdx.x == 0) s_item = atomicAdd(&b->work_counter, 1u);
        __syncthreads();
        const int64_t item = (int64_t)__builtin_amdgcn_readfirstlane(s_item);
        __syncthreads();
        if (item >= n_items) break;
        int lo = 0, hi = n_tiles - 1;
        while (lo < hi) {
            const int mid = (lo + hi + 1) >> 1;
            if (tiles[mid].item_base <= item) lo = mid;
            else hi = mid - 1;
        }
        const TileDev &t = tiles[lo];
        const int64_t local = item - t.item_base;
        const int tj = (int)(local / t.ntx), ti = (int)(local % t.ntx);
        process_block<ALG, MODE>(t, ti, tj, smem, offs, edist, sorted);
        __syncthreads();
    }
}

template <int ALG, int MODE>
hipError_t launch_tiles(const DevBuffers &d, const LaunchCfg &cfg, hipStream_t s)
{
    hipLaunchKernelGGL((k_tiles<ALG, MODE>), dim3(cfg.main_grid), dim3(kBlock), cfg.main_lds, s,
                       d.batch, d.tiles, d.counts, d.edist, d.sorted);
    return hipGetLastError();
}

}  // namespace

hipError_t enqueue_pipeline(const DevBuffers &d, const LaunchCfg &cfg, hipStream_t s,
                            hipEvent_t ev0, hipEvent_t ev1)
{
    hipError_t e;
    hipLaunchKernelGGL(k_zero, dim3(cfg.fill_grid), dim3(kBlock), 0, s, d.batch, d.counts);
    hipLaunchKernelGGL(k_count, dim3(cfg.sort_grid), dim3(kBlock), 0, s, d.batch, d.tiles,
                       d.chunks, d.counts, d.rank);
    hipLaunchKernelGGL(k_scan_reduce, dim3(kScanBlocks), dim3(kBlock), 0, s, d.batch, d.counts,
                       d.partials);
    hipLaunchKernelGGL(k_scan_partials, dim3(1), dim3(kScanBlocks), 0, s, d.partials);
    hipLaunchKernelGGL(k_scan_down, dim3(kScanBlocks), dim3(kBlock), 0, s, d.batch, d.counts,
                       d.partials);
    hipLaunchKernelGGL(k_scatter, dim3(cfg.sort_grid), dim3(kBlock), 0, s, d.batch, d.tiles,
                       d.chunks, d.counts, d.rank, d.sorted);
    hipLaunchKernelGGL(k_edist_rows, dim3(cfg.row_grid), dim3(kBlock), 0, s, d.batch, d.tiles,
                       d.counts, d.hdist);
    hipLaunchKernelGGL(k_edist_cols, dim3(cfg.row_grid), dim3(kBlock), 0, s, d.batch, d.tiles,
                       d.hdist, d.edist);
    if ((e = hipGetLastError()) != hipSuccess) return e;
    if (ev0 && (e = hipEventRecord(ev0, s)) != hipSuccess) return e;
    const int sel = cfg.algorithm * 3 + cfg.mode;
    switch (sel) {
    case 0: e = launch_tiles<0, 0>(d, cfg, s); break;
    case 1: e = launch_tiles<0, 1>(d, cfg, s); break;
    case 2: e = launch_tiles<0, 2>(d, cfg, s); break;
    case 3: e = launch_tiles<1, 0>(d, cfg, s); break;
    case 4: e = launch_tiles<1, 1>(d, cfg, s); break;
    case 5: e = launch_tiles<1, 2>(d, cfg, s); break;
    default: return hipErrorInvalidValue;
    }
    if (e != hipSuccess) return e;
    if (ev1 && (e = hipEventRecord(ev1, s)) != hipSuccess) return e;
    return hipSuccess;
}

int tile_kernel_occupancy(size_t lds_bytes)
{
    int nb = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, k_tiles<0, 0>, kBlock, lds_bytes) !=
        hipSuccess)
        return 1;
    return nb > 0 ? nb : 1;
}

}  // namespace tnni
