// Tile geometry on the host: the pixel frame on the fixed-point grid, the
// bucket grid and the cell-tile decomposition of one work package.
#include <algorithm>
#include <cmath>
#include <cstring>
#include <string>

#include "tnni_internal.h"

namespace tnni {
namespace {

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
    t.br = (int)std::max<int64_t>(1, std::min<int64_t>(nby, kBandBuckets / nbx));
    t.nbands = (int)((nby + t.br - 1) / t.br);
    t.inv_br = 1.0 / (double)t.br;
    if (nbx > kBandBuckets || t.nbands > kMaxBands) {
        err = "bucket grid too large for the band sort: raise bucket_px";
        return TNNI_E_GEOMETRY;
    }
    t.bm_magic = bm == 1 ? 0u : (uint32_t)((((uint64_t)1 << 32) + (uint64_t)bm - 1) / (uint64_t)bm);
    t.inv_bs = 1.0 / (double)bs;
    // A site of ring 0 or 1 around a pixel's bucket is within 2*bs of the
    // pixel on both axes (strictly less than 2*bs).
    t.ring1_covered = d.algorithm == TNNI_ALG_PLANES ? (2 * bs <= f.Rs) : (8 * bs * bs <= f.Rs2);
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
        return (size_t)rw * rh * 8 + tile_kernel_stage_bytes();
    };
    size_t lds = 0;
    if (d.tile_cells > 0) {
        lds = lds_for(d.tile_cells);
        if (lds > kLdsBudget) {
            err = "tile_cells too large for the LDS budget";
            return TNNI_E_GEOMETRY;
        }
    } else {
        // the largest cell tile that leaves room for four workgroups per CU
        // (16 waves), else the largest that fits at all
        static const int kTc[] = {64, 48, 40, 32, 24, 16, 12, 8, 6, 4, 3, 2, 1};
        int chosen = -1;
        size_t budget = kLdsBudget4;
        for (int pass = 0; pass < 2 && chosen < 0; ++pass, budget = kLdsBudget)
        for (int tc : kTc) {
            if (lds_for(tc) > budget) continue;
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

}  // namespace tnni
