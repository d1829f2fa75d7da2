// nni.h (TerraNNI's per-tile API) implemented on the C ABI of terranni.h.
//
// performNNI reads the current time slice's points from the caller's TPIE
// stream exactly as renderInputPoints does (nni.cpp:741-788: time filter,
// stop at the first later point, seek back for the next slice), puts them
// on the engine's fixed-point grid, runs the tile on the GPU and appends the
// cells to the output stream in the reference's order and filter
// (nni.cpp:836-860).
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <string>
#include <vector>

#include "nni.h"
#include "terranni.h"

namespace {

nni::Settings g_settings;
tnni_ctx *g_ctx = nullptr;
// the slices of performNNI with useDelta (nni.cpp:281-379, 902-921)
tnni_time_cache *g_cache = nullptr;
int g_mode = TNNI_ACC_FIXED64;
int g_status = TNNI_OK;
std::string g_err;
tpie::cpu_timer g_t_read, g_t_tile, g_t_write;
bool g_have_origin64 = false;
double g_origin64_x = 0.0, g_origin64_y = 0.0;
// Heights go to the engine as integers of a unit chosen from the data of
// the current work package (offset 0), as the reference picks its
// zMultiplier from the data's range and decimal precision
// (terranni.cpp:674-681):
//   - the coarsest decimal unit 10^-k (k = 0..9) of which every height read
//     so far is a whole multiple (heights that came from LAS integers with
//     a decimal scale), if the largest |height| fits in zlim units: exact;
//   - otherwise 2^e metres, e the smallest with every |height| so far within
//     zlim units: at most half a unit off, 2^-31 of the largest |height|
//     (2^-24 for FP32 two-sum).
// zlim is 2^30, or 2^23 for FP32 two-sum (exact sums below 2^24).  The unit
// only coarsens within a package, so its calls keep one unit (the
// space-time cache is keyed by the tile descriptor) unless a later call
// brings a larger or finer height.  clearRenderbuffers (called after each
// work package) and a new work package start again.
struct ZUnit {
    bool any = false;     // a height has been seen in this package
    double zmax = 0.0;    // largest |height| so far
    bool decimal = true;  // every height so far a multiple of 10^-dec_k
    int dec_k = 0;
};
ZUnit g_zu;
nni::WorkPackage g_z_wp;
bool g_z_wp_set = false;
constexpr int kMaxDecimals = 9;

// Whether z is a whole multiple of 10^-k (to 1e-6 of that unit, far below
// any rounding a decimal-scaled LAS height carries).
bool decimal_at(double z, int k)
{
    const double u = z * std::pow(10.0, k);
    return std::fabs(u - std::rint(u)) <= 1e-6;
}

void set_error(int st, const std::string &msg)
{
    g_status = st;
    g_err = msg;
    std::fprintf(stderr, "nni: %s\n", msg.c_str());
}

int32_t clamp_i32(long long v)
{
    return (int32_t)std::max<long long>(INT32_MIN, std::min<long long>(INT32_MAX, v));
}

}  // namespace

namespace nni {

int init(const Settings &s)
{
    g_settings = s;
    g_status = TNNI_OK;
    g_err.clear();
    if (const char *m = std::getenv("TERRANNI_ACCUM")) {
        if (!std::strcmp(m, "fp32")) g_mode = TNNI_ACC_FP32_2SUM;
        else if (!std::strcmp(m, "s0") || !std::strcmp(m, "reference")) g_mode = TNNI_ACC_S0;
        else g_mode = TNNI_ACC_FIXED64;
    }
    if (!g_ctx) {
        tnni_ctx_opts o;
        std::memset(&o, 0, sizeof o);
        if (const char *dev = std::getenv("TERRANNI_DEVICE")) o.device = std::atoi(dev);
        const tnni_status st = tnni_ctx_create(&o, &g_ctx);
        if (st != TNNI_OK) {
            set_error(st, std::string("engine initialisation failed: ") + tnni_last_error(nullptr));
            g_ctx = nullptr;
            return 0;
        }
    }
    // largest raster side a tile may use (the GL texture limit of the
    // reference, nni.cpp:591-597); the driver caps work packages itself
    return 1 << 16;
}

void cleanup()
{
    tnni_time_cache_destroy(g_cache);
    g_cache = nullptr;
    if (g_ctx) tnni_ctx_destroy(g_ctx);
    g_ctx = nullptr;
}

void setRenderbuffers(int, int, int) {}
void setupRun(const WorkPackage &, int, int) {}
void clearRenderbuffers()
{
    if (g_cache) (void)tnni_time_cache_clear(g_cache);
    g_zu = ZUnit();
}

void printTimes()
{
    std::fprintf(stdout, "Time Results:\n");
    std::fprintf(stdout, "\tRead site points  : \t\t%f\n", g_t_read.wall_time());
    std::fprintf(stdout, "\tTile on the GPU   : \t\t%f\n", g_t_tile.wall_time());
    std::fprintf(stdout, "\tWrite query points: \t\t%f\n", g_t_write.wall_time());
}

void setAccumulationMode(int mode) { g_mode = mode; }

void setOrigin64(double origin_x, double origin_y)
{
    g_have_origin64 = true;
    g_origin64_x = origin_x;
    g_origin64_y = origin_y;
}
int lastStatus() { return g_status; }
const char *lastError() { return g_err.c_str(); }

void performNNI(const WorkPackage &wp, tpie::ami::stream<OutPoint> &out,
                tpie::ami::stream<nni::Point4D> &vec, float timeRadius, float curTime,
                bool useDelta, bool debug)
{
    (void)debug;
    g_status = TNNI_OK;
    const bool have64 = g_have_origin64;
    g_have_origin64 = false;
    if (!g_ctx) {
        set_error(TNNI_E_NODEVICE, "performNNI before a successful nni::init");
        return;
    }
    // 1. the points (renderInputPoints, nni.cpp:741-788)
    g_t_read.start();
    const float t = curTime;
    const bool space_time = timeRadius > 0;
    std::vector<Point4D> pts;
    Point4D *p = nullptr;
    if (!space_time) {
        // one slice: the reference's filter and seeks
        TPIE_OS_OFFSET seekto = vec.tell(), finalSeekTo = vec.tell();
        bool found = false;
        while (vec.read_item(&p) == tpie::ami::NO_ERROR) {
            if (!found && p->t >= t + 1 - timeRadius) {
                seekto = vec.tell() - 1;
                finalSeekTo = seekto;
                found = true;
            }
            if ((double)std::fabs(t - (float)p->t) <= (double)timeRadius + 0.000001) pts.push_back(*p);
            else if (p->t > t + timeRadius) break;
        }
        vec.seek(useDelta ? seekto : finalSeekTo);
    } else {
        // slices curTime + j, j = -(int)r..r, each with the points within r
        // of it: every point within 2r + 1 of curTime is read (the engine
        // filters each slice exactly).  The stream is left at the first
        // point of that window, where a later call with a later curTime
        // finds all of its points.  (The reference leaves it after the
        // slices it drew, relying on its render-buffer cache for the rest,
        // nni.cpp:321-379,782-788.  Here every slice's sites are read again:
        // a slice's sites do not depend on curTime, and with useDelta the
        // engine's space-time cache rebuilds the slices from the per-time
        // layers it still holds without searching them.)
        const double w = 2.0 * (double)timeRadius + 1.0;
        TPIE_OS_OFFSET start = -1;
        for (;;) {
            const TPIE_OS_OFFSET at = vec.tell();
            if (vec.read_item(&p) != tpie::ami::NO_ERROR) {
                if (start < 0) start = at;
                break;
            }
            const double dt = (double)p->t - (double)t;
            if (dt > w) {
                if (start < 0) start = at;
                break;
            }
            if (dt >= -w) {
                if (start < 0) start = at;
                pts.push_back(*p);
            }
        }
        vec.seek(start);
    }

    // 2. the tile on the engine's grid: coordinates in its own unit
    //    relative to the tile origin, heights in the package's unit
    tnni_tile_desc d;
    std::memset(&d, 0, sizeof d);
    d.origin_x = have64 ? g_origin64_x : (double)wp.origin_x;
    d.origin_y = have64 ? g_origin64_y : (double)wp.origin_y;
    d.cell_size = g_settings.cell_size;
    d.scale = g_settings.scale;
    d.nrows = wp.nrows;
    d.ncols = wp.ncols;
    d.site_radius = g_settings.site_radius;
    d.query_radius = (float)g_settings.query_radius;
    d.algorithm = g_settings.algorithm;
    d.accum_mode = g_mode;
    d.pixel_offset = -1;
    d.point_min = g_settings.pointMin;
    d.multiplier = g_settings.multiplier;
    tnni_frame_info f;
    tnni_status st = tnni_frame(&d, &f);
    if (st != TNNI_OK) {
        g_t_read.stop();
        set_error(st, tnni_last_error(nullptr));
        return;
    }
    const double upm = (double)f.upm;
    const size_t n = pts.size();
    if (g_z_wp_set && std::memcmp(&g_z_wp, &wp, sizeof wp) != 0) g_zu = ZUnit();
    g_z_wp = wp;
    g_z_wp_set = true;
    for (size_t i = 0; i < n; ++i) {
        const double z = pts[i].p[2];
        if (!std::isfinite(z)) {
            g_t_read.stop();
            set_error(TNNI_E_INVALID, "a point height is not finite");
            return;
        }
        g_zu.any = true;
        g_zu.zmax = std::max(g_zu.zmax, std::fabs(z));
        while (g_zu.decimal && !decimal_at(z, g_zu.dec_k))
            if (++g_zu.dec_k > kMaxDecimals) g_zu.decimal = false;
    }
    const double zlim = g_mode == TNNI_ACC_FP32_2SUM ? 8388608.0 : 1073741824.0;
    const bool use_dec = g_zu.decimal && g_zu.zmax * std::pow(10.0, g_zu.dec_k) <= zlim;
    int zexp = -64;
    while (!use_dec && zexp < 64 && g_zu.zmax > std::ldexp(zlim, zexp)) ++zexp;
    const double zmul = use_dec ? std::pow(10.0, g_zu.dec_k) : std::ldexp(1.0, -zexp);
    d.las_scale[0] = d.las_scale[1] = 1.0 / upm;
    d.las_scale[2] = use_dec ? 1.0 / zmul : std::ldexp(1.0, zexp);
    d.las_offset[0] = d.origin_x;
    d.las_offset[1] = d.origin_y;
    d.las_offset[2] = 0.0;
    std::vector<int32_t> X(n), Y(n), Z(n), T(space_time ? n : 0);
    for (size_t i = 0; i < n; ++i) {
        X[i] = clamp_i32(std::llround((pts[i].p[0] - d.origin_x) * upm));
        Y[i] = clamp_i32(std::llround((pts[i].p[1] - d.origin_y) * upm));
        // |Z| <= zlim; rint rounds to nearest, ties to even
        Z[i] = (int32_t)std::rint(pts[i].p[2] * zmul);
        if (space_time) T[i] = pts[i].t;
    }
    g_t_read.stop();

    g_t_tile.start();
    std::vector<float> h((size_t)wp.nrows * wp.ncols);
    // useDelta keeps the slices for the next curTime; Delta^2 searches
    // every slice again (nni.cpp:915-917).  The results are the same.
    if (space_time && useDelta && !g_cache && (st = tnni_time_cache_create(g_ctx, 0, &g_cache)) != TNNI_OK) {
        g_t_tile.stop();
        set_error(st, tnni_last_error(g_ctx));
        return;
    }
    if (space_time)
        st = tnni_tile_run_time_cached(g_ctx, useDelta ? g_cache : nullptr, &d, X.data(), Y.data(),
                                       Z.data(), T.data(), n, curTime, timeRadius, h.data());
    else
        st = tnni_tile_run(g_ctx, &d, X.data(), Y.data(), Z.data(), n, h.data(), nullptr, nullptr);
    g_t_tile.stop();
    if (st != TNNI_OK) {
        set_error(st, tnni_last_error(g_ctx));
        return;
    }

    // 3. the cells, column by column, inside the total grid (nni.cpp:836-860)
    g_t_write.start();
    for (int i = 0; i < wp.ncols; ++i) {
        for (int j = 0; j < wp.nrows; ++j) {
            const OutPoint o((int)((float)i + wp.grid_origin_x), (int)((float)j + wp.grid_origin_y),
                             h[(size_t)j * wp.ncols + i]);
            if (o.i < g_settings.ncols && o.j < g_settings.nrows) out.write_item(o);
        }
    }
    g_t_write.stop();
}

}  // namespace nni
