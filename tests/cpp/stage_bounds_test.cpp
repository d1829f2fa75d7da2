// Host emulation of the index arithmetic of the tile kernel's per-wave LDS
// stage and label tile (label_block, ring_staged/square_staged,
// process_block's lists and gather in tnni_kernels.hip), over every 8x8
// pixel block of every work item of a set of tile geometries, laid out by
// the engine's own prepare_tile / region_of.  Every index the kernel forms
// into the stage, the label tile, the bucket offsets and the distance maps
// must stay inside its buffer, and the label tile must fit the dynamic LDS
// that the launch gives the workgroup.  Also prints the stage's size and
// the LDS of one workgroup for 256- and 384-site stages against the 160 KB
// of a CU at 4 and 5 workgroups per CU.  Every geometry runs in every form:
// fused (items label their cells' pixels plus the query halo and gather
// from LDS), split and scatter (label-only items, which must own every
// pixel of the raster exactly once, then blocks of cells whose pixel
// window, and for the scatter the cells' sums after it, fit the pass's
// LDS).  Exit status 0: no violation.
//
//   stage_bounds_test [n_random_geometries]
#include <algorithm>
#include <cstdio>
#include <vector>
#include <cstdlib>
#include <random>
#include <string>

#include "tnni_internal.h"

using namespace tnni;

namespace {

long g_viol = 0, g_blocks = 0, g_items = 0, g_tiles = 0;
int g_max_ww = 0, g_max_wh = 0, g_max_rwrh = 0;

void violation(const std::string &what, const tnni_tile_desc &d)
{
    if (g_viol < 20)
        std::fprintf(stderr, "VIOLATION %s (grid %dx%d cs %g scale %d sr %g bucket %d tc %d)\n", what.c_str(),
                     d.nrows, d.ncols, d.cell_size, d.scale, d.site_radius, d.bucket_px, d.tile_cells);
    ++g_viol;
}

int div_bm(const TileDev &t, int k) { return k / t.bm; }

void check_form(const tnni_tile_desc &d, uint64_t n);

// One tile in every form (the scatter form in int64 fixed point).
void check_tile(tnni_tile_desc d, uint64_t n)
{
    for (int form : {TNNI_FORM_FUSED, TNNI_FORM_SPLIT, TNNI_FORM_SCATTER}) {
        if (form == TNNI_FORM_SCATTER) d.accum_mode = TNNI_ACC_FIXED64;
        d.form = form;
        check_form(d, n);
    }
}

// One tile in one form: every item, every block, every lane.
void check_form(const tnni_tile_desc &d, uint64_t n)
{
    TileDev t;
    PrepInfo info;
    std::string err;
    if (prepare_tile(d, n, 256, false, nullptr, t, info, err) != TNNI_OK) return;
    if (t.label_only != (d.form == TNNI_FORM_SPLIT || d.form == TNNI_FORM_SCATTER) ||
        t.scatter != (d.form == TNNI_FORM_SCATTER))
        violation("form not taken", d);
    // label-only items: how often each pixel is written
    std::vector<uint8_t> owned(t.label_only ? (size_t)t.W * t.H : 0, 0);
    ++g_tiles;
    t.nbx4 = (t.nbx + 3) & ~3;
    const int64_t n_offs = (int64_t)t.nbx * t.nby + 1;  // offsets plus the sentinel
    const int64_t n_emap = (int64_t)t.nbx4 * t.nby;
    const size_t stage_bytes = tile_kernel_stage_bytes();
    if (stage_bytes != (size_t)kWaves * sizeof(WaveStage)) violation("stage bytes", d);
    const int qa = query_lo(t), qb = query_hi(t);
    for (int tj = 0; tj < t.nty; ++tj)
        for (int ti = 0; ti < t.ntx; ++ti) {
            ++g_items;
            const Region g = region_of(t, ti, tj);
            const int RW = g.rk1 - g.rk0 + 1, RH = g.rl1 - g.rl0 + 1;
            g_max_rwrh = std::max(g_max_rwrh, RW * RH);
            // the label tile (d2, value, list slot per pixel) behind the stages
            if (stage_bytes + ((size_t)RW * RH * 10 + 3) / 4 * 4 > info.lds) violation("label tile past the LDS", d);
            if (RW * RH > 65536) violation("list slot exceeds 16 bits", d);
            const int nbx8 = (RW + 7) >> 3, nby8 = (RH + 7) >> 3;
            // finish_pixel's long-search box: the square of radius rl about
            // the pixel's bucket clipped to the buckets its site region
            // reaches. The box is a product of a column range set by the
            // pixel's column and a row range set by its row, each bound
            // monotone in rl: the item's columns and rows at the smallest and
            // largest radii cover every pixel and radius.
            {
                const int R = t.algorithm == TNNI_ALG_PLANES ? t.Rs : t.Rsb;
                int cmax = -1, qmax = -1;
                for (int rl = 2; rl <= 6; rl += 4) {
                    for (int k = g.rk0; k <= g.rk1; ++k) {
                        const int bx = t.bK + div_bm(t, k), px = t.x0 + k * t.ps_u;
                        const int c0 = std::max(std::max(bx - rl, std::max(px - R - t.gx0, 0) / t.bs), 0);
                        const int c1 = std::min(std::min(bx + rl, (px + R - t.gx0) / t.bs), t.nbx - 1);
                        if (c0 > bx || c1 < bx) violation("long-search box misses the pixel's bucket column", d);
                        if (c0 < 0 || c1 >= t.nbx) violation("long-search box off the grid's columns", d);
                        cmax = std::max(cmax, c1);
                    }
                    for (int l = g.rl0; l <= g.rl1; ++l) {
                        const int by = t.bK + div_bm(t, l), py = t.y0 + l * t.ps_u;
                        const int q0 = std::max(std::max(by - rl, std::max(py - R - t.gy0, 0) / t.bs), 0);
                        const int q1 = std::min(std::min(by + rl, (py + R - t.gy0) / t.bs), t.nby - 1);
                        if (q0 > by || q1 < by) violation("long-search box misses the pixel's bucket row", d);
                        if (q0 < 0 || q1 >= t.nby) violation("long-search box off the grid's rows", d);
                        qmax = std::max(qmax, q1);
                    }
                }
                if ((int64_t)qmax * t.nbx + cmax + 1 >= n_offs) violation("long-search box offs index", d);
            }
            for (int blk = 0; blk < nbx8 * nby8; ++blk) {
                ++g_blocks;
                const int kA = g.rk0 + (blk % nbx8) * 8, lA = g.rl0 + (blk / nbx8) * 8;
                const int kB = std::min(kA + 7, g.rk1), lB = std::min(lA + 7, g.rl1);
                const int cx0 = std::max(t.bK + div_bm(t, kA) - 1, 0);
                const int cx1 = std::min(t.bK + div_bm(t, kB) + 1, t.nbx - 1);
                const int cy0 = std::max(t.bK + div_bm(t, lA) - 1, 0);
                const int cy1 = std::min(t.bK + div_bm(t, lB) + 1, t.nby - 1);
                const int ww = cx1 - cx0 + 1, wh = cy1 - cy0 + 1, wp = ww + 1;
                g_max_ww = std::max(g_max_ww, ww);
                g_max_wh = std::max(g_max_wh, wh);
                if (ww < 1 || wh < 1 || ww > kWinMax || wh > kWinMax) violation("window side", d);
                if (wh * wp > kWinMax * (kWinMax + 1)) violation("woff overflow", d);
                // woff[i] <- offs[(cy0 + j) * nbx + cx0 + c], i = j * wp + c
                const int64_t last = (int64_t)(cy0 + wh - 1) * t.nbx + cx0 + (wp - 1);
                if ((int64_t)cy0 * t.nbx + cx0 < 0 || last >= n_offs) violation("offs index of the window", d);
                for (int lane = 0; lane < 64; ++lane) {
                    const int k = kA + (lane & 7), l = lA + (lane >> 3);
                    if (!(k <= kB && l <= lB)) continue;
                    const int bx = t.bK + div_bm(t, k), by = t.bK + div_bm(t, l);
                    if (bx < 0 || bx >= t.nbx || by < 0 || by >= t.nby) violation("pixel bucket", d);
                    if ((int64_t)by * t.nbx4 + bx >= n_emap) violation("distance-map index", d);
                    // rings 0-1 from the stage: rows by-1..by+1 inside the
                    // grid must be window rows, columns bx-1..bx+1 (and the
                    // end offset of the last) window columns
                    for (int row = std::max(by - 1, 0); row <= std::min(by + 1, t.nby - 1); ++row)
                        if (row - cy0 < 0 || row - cy0 >= wh) violation("staged row", d);
                    const int c0 = std::max(bx - 1, 0) - cx0, c1 = std::min(bx + 1, t.nbx - 1) + 1 - cx0;
                    if (c0 < 0 || c1 > ww) violation("staged column", d);
                    const int p = (l - g.rl0) * RW + (k - g.rk0);
                    if (p < 0 || p >= RW * RH) violation("label tile index", d);
                    if (t.label_only) {
                        if (k < g.ok0 || k > g.ok1 || l < g.ol0 || l > g.ol1) violation("label-only pixel not owned", d);
                        else if (k < 0 || k >= t.W || l < 0 || l >= t.H) violation("label-only pixel off the raster", d);
                        else owned[(size_t)l * t.W + k]++;
                    }
                }
            }
            if (t.label_only) continue;  // no gather in the tile kernel
            // the gather: every pixel of every cell's query window inside the
            // item's label tile
            for (int j = g.j0; j < g.j1; ++j)
                for (int i = g.i0; i < g.i1; ++i) {
                    const int ma = std::max(qa, -i * t.scale), mb = std::min(qb, t.W - 1 - i * t.scale);
                    const int na = std::max(qa, -j * t.scale), nb = std::min(qb, t.H - 1 - j * t.scale);
                    if (ma > mb || na > nb) continue;
                    if (i * t.scale + ma < g.rk0 || i * t.scale + mb > g.rk1 || j * t.scale + na < g.rl0 ||
                        j * t.scale + nb > g.rl1)
                        violation("gather window outside the label tile", d);
                }
        }
    for (uint8_t c : owned)
        if (c != 1) {
            violation("label-only items do not own every pixel once", d);
            break;
        }
    if (!t.label_only) return;
    // the split form's gather: each block's pixel window within its LDS
    // (gather_window_bytes of an interior block, under kGatherLds), every
    // cell's reads inside the window, every cell of the tile in one block
    const size_t wmax = gather_window_bytes(t, t.gb);
    if (split_pass_lds(t) > kGatherLds) violation("split pass past its LDS budget", d);
    // what follows the window (the scatter's sums, the gather's partial
    // sums) starts where k_gather_split puts it, after the widest window,
    // 16-aligned, and ends within the pass's LDS
    const size_t side = (size_t)(t.gb - 1) * t.scale + (size_t)std::max(1, qb - qa + 1);
    const size_t after = ((side * side + 1) & ~(size_t)1) * 8;
    if (side * side * 8 != wmax || after % 16 || after < wmax) violation("LDS after the window misplaced", d);
    // the scatter reads its window from HBM: only its sums in LDS, the
    // window's pixel index below 2^24 (its row is found by a float product)
    if (t.scatter && ((size_t)t.gb * t.gb * 12 > split_pass_lds(t) || side * side + kBlock >= (size_t)1 << 24))
        violation("scatter sums past the pass's LDS or window too large", d);
    const bool shared = !t.scatter && t.mode != TNNI_ACC_FP32_2SUM && t.gb * t.gb < kBlock;
    if (shared && after + (size_t)kBlock * 16 > split_pass_lds(t)) violation("partial sums past the pass's LDS", d);
    if (t.gb & (t.gb - 1)) violation("gather block side not a power of two", d);
    std::vector<uint8_t> cell_seen((size_t)t.nrows * t.ncols, 0);
    for (int bj = 0; bj < t.gnty; ++bj)
        for (int bi = 0; bi < t.gntx; ++bi) {
            const int i0 = bi * t.gb, j0 = bj * t.gb;
            const int i1 = std::min(i0 + t.gb, t.ncols), j1 = std::min(j0 + t.gb, t.nrows);
            const int kx0 = std::max(0, i0 * t.scale + qa), kx1 = std::min(t.W - 1, (i1 - 1) * t.scale + qb);
            const int ky0 = std::max(0, j0 * t.scale + qa), ky1 = std::min(t.H - 1, (j1 - 1) * t.scale + qb);
            const size_t ww = (size_t)std::max(0, kx1 - kx0 + 1), wh = (size_t)std::max(0, ky1 - ky0 + 1);
            if (ww * wh * 8 > wmax) violation("gather block window past the launch's LDS", d);
            if (!t.scatter && t.gb * t.gb > kBlock) violation("gather block larger than the workgroup", d);
            for (int j = j0; j < j1; ++j)
                for (int i = i0; i < i1; ++i) {
                    cell_seen[(size_t)j * t.ncols + i]++;
                    const int ma = std::max(qa, -i * t.scale), mb = std::min(qb, t.W - 1 - i * t.scale);
                    const int na = std::max(qa, -j * t.scale), nb = std::min(qb, t.H - 1 - j * t.scale);
                    if (ma > mb || na > nb) continue;
                    if (i * t.scale + ma < kx0 || i * t.scale + mb > kx1 || j * t.scale + na < ky0 ||
                        j * t.scale + nb > ky1)
                        violation("gather reads outside the block's window", d);
                }
        }
    for (uint8_t c : cell_seen)
        if (c != 1) {
            violation("gather blocks do not cover every cell once", d);
            break;
        }
}

tnni_tile_desc base_desc(int nrows, int ncols, float cs, int scale)
{
    tnni_tile_desc d{};
    d.las_scale[0] = d.las_scale[1] = d.las_scale[2] = 0.01;
    d.cell_size = cs;
    d.scale = scale;
    d.nrows = nrows;
    d.ncols = ncols;
    d.site_radius = 5.0f;
    d.query_radius = 5.0f * cs / 2.0f - 0.01f;  // terranni.cpp:663
    d.pixel_offset = -1;
    d.multiplier = 1;
    return d;
}

}  // namespace

int main(int argc, char **argv)
{
    const int n_random = argc > 1 ? std::atoi(argv[1]) : 300;
    // the single-tile workload (4096^2 at 0.5 m, 4M points) and the
    // reference's CPU case (512^2 at 2 m, scale 5, 100k points)
    check_tile(base_desc(4096, 4096, 0.5f, 1), 4000000);
    check_tile(base_desc(512, 512, 2.0f, 5), 100000);
    // bucket sides 1..12 and fixed cell tiles on small tiles at scales 1, 2,
    // 3, 5, with and without the border
    for (int scale : {1, 2, 3, 5})
        for (int bm = 1; bm <= 12; ++bm)
            for (int tc : {0, 1, 2, 3, 7, 30})
                for (int po : {-1, 0}) {
                    tnni_tile_desc d = base_desc(37, 53, 0.5f * scale, scale);
                    d.bucket_px = bm;
                    d.tile_cells = tc;
                    d.pixel_offset = po;
                    check_tile(d, 20000);
                }
    // random geometries
    std::mt19937 rng(12345);
    auto uni = [&rng](int a, int b) { return std::uniform_int_distribution<int>(a, b)(rng); };
    const float css[] = {0.25f, 0.5f, 1.0f, 2.0f};
    for (int r = 0; r < n_random; ++r) {
        const int scale = uni(1, 6);
        tnni_tile_desc d = base_desc(uni(1, 300), uni(1, 300), css[uni(0, 3)], scale);
        d.site_radius = (float)uni(1, 20);
        if (uni(0, 1)) d.query_radius = (float)uni(1, 8) * 0.5f;
        d.algorithm = uni(0, 1);
        d.bucket_px = uni(0, 1) ? uni(1, 16) : 0;
        d.tile_cells = uni(0, 2) == 0 ? uni(1, 64) : 0;
        if (uni(0, 3) == 0) d.pixel_offset = uni(0, 5);
        const uint64_t n = (uint64_t)uni(0, 200000);
        check_tile(d, n);
    }
    // stage and workgroup LDS at 256 and 384 sites per wave
    const size_t s256 = sizeof(WaveStageT<256>), s384 = sizeof(WaveStageT<384>);
    std::printf("WaveStage<256> %zu B (16-aligned: %s), WaveStage<384> %zu B (16-aligned: %s)\n", s256,
                s256 % 16 ? "no" : "yes", s384, s384 % 16 ? "no" : "yes");
    std::printf("stage LDS per workgroup (4 waves): 256 sites %zu B, 384 sites %zu B; LDS per workgroup at "
                "4 / 5 workgroups per CU: %d / %d B\n",
                4 * s256, 4 * s384, 160 * 1024 / 4, 160 * 1024 / 5);
    std::printf("tiles %ld, items %ld, blocks %ld; widest window %d x %d buckets (limit %d); largest label "
                "tile %d pixels; violations %ld\n",
                g_tiles, g_items, g_blocks, g_max_ww, g_max_wh, kWinMax, g_max_rwrh, g_viol);
    return g_viol ? 1 : 0;
}
