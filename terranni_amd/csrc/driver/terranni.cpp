// terranni: TerraNNI's streaming DEM driver (terranni.cpp of the reference)
// over the MI355X tile engine.
//
// The steps of the reference's main (terranni.cpp:491-731):
//   1. find the .las files under --tilepath and their bounding box
//      (terranni.cpp:86-113, 585-597), here in FP64;
//   2. derive the settings: border, query radius, zMultiplier, work-package
//      size (terranni.cpp:656-714);
//   3. bin the points into one stream per work package, every point within
//      `border` of a package copied into it, and sort each stream by time
//      (terranni.cpp:198-368), unless --wp-computed;
//   4. interpolate every package and time slice (binRun, terranni.cpp:394-426)
//      and write the grids (outputASC, terranni.cpp:430-488).
//
// Two engines run step 4:
//   --engine plain (default): the reference's loop, nni::performNNI per
//     package and time slice into per-slice OutPoint streams, sorted and
//     written as ESRI ASCII afterwards.  Packages are streams of
//     nni::Point4D, r<row>c<col>.tpie, as in the reference.
//   --engine batched: packages hold the LAS integers (X, Y, Z, time; 16 bytes
//     a point) in one common LAS scale/offset; every package of a row of
//     packages is submitted to the engine's served form at once, the GPU
//     de-scales in FP64 in each package's frame, and the rasters go straight
//     to the output rows (no OutPoint streams, no sort).  The next row of
//     packages is on the GPU while the previous one is written.
// Outputs: ESRI ASCII grid (<slice>.tpie.asc, the reference's name) and/or
// an ESRI GridFloat raster (<slice>.flt + .hdr, float32, rows top down).
#include <algorithm>
#include <chrono>
#include <cinttypes>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <filesystem>
#include <fstream>
#include <iomanip>
#include <iostream>
#include <limits>
#include <map>
#include <memory>
#include <sstream>
#include <string>
#include <vector>

#include <fcntl.h>
#include <omp.h>
#include <climits>
#include <sys/mman.h>
#include <sys/uio.h>
#include <thread>
#include <unistd.h>

#include "las_reader.h"
#include "nni.h"
#include "terranni.h"
#include "tpie/sort.h"
#include "tpie/stream.h"
#include "tpie/tpie.h"

namespace fs = std::filesystem;

namespace {

[[noreturn]] void die(const std::string &msg)
{
    std::fprintf(stderr, "terranni: %s\n", msg.c_str());
    std::exit(1);
}

double wall()
{
    return std::chrono::duration<double>(std::chrono::steady_clock::now().time_since_epoch()).count();
}

// ---- options (terranni.cpp:522-553, plus this driver's own) --------------

struct Options {
    std::string tilepath, output = "./", tmp_path = ".", wp_path = ".";
    int scale = 5;
    float cell_size = 2.0f, site_radius = 5.0f;
    bool has_qr = false;
    float qr = 0.f;
    bool has_wpw = false, has_wph = false;
    int wp_width = 0, wp_height = 0;
    bool wp_computed = false;
    bool has_ox = false, has_oy = false;
    double ox = 0, oy = 0;
    bool has_gc = false, has_gr = false;
    int grid_cols = 0, grid_rows = 0;
    int time_radius = 0;
    float time_start = 0.f, time_length = 1.f, time_step = 1.f;
    bool two_d = false, delta_squared = false, draw_cones = false, help = false;
    std::string engine = "plain", format = "asc", accum = "fixed64";
    int max_size = 5000;
    bool bin_only = false;
    long bin_block = 1L << 26;
};

const char *kUsage =
    "Allowed options (the reference's, terranni.cpp:522-553):\n"
    "  --tilepath P        path to the .las tiles (required)\n"
    "  --output P          output folder (default ./)\n"
    "  --scaling N         pixels per cell side (default 5)\n"
    "  --cell-size C       cell size in metres (default 2.0)\n"
    "  --site-radius R     site region radius in metres (default 5.0)\n"
    "  --query-radius R    query region radius (default 5*cell/2 - 0.01)\n"
    "  --wp-width W        work package width in metres\n"
    "  --wp-height H       work package height in metres\n"
    "  --tmp-path P        temporary files (default .)\n"
    "  --wp-path P         work package streams (default .)\n"
    "  --wp-computed       reuse the work package streams of an earlier run\n"
    "  --origin-x X        grid origin, x\n"
    "  --origin-y Y        grid origin, y\n"
    "  --grid-cols N       grid columns\n"
    "  --grid-rows N       grid rows\n"
    "  --time-radius N     time radius (default 0)\n"
    "  --time-start T      first time slice (default 0)\n"
    "  --time-length L     time slices to interpolate (default 1)\n"
    "  --time-step S       time step (default 1)\n"
    "  --2d                ignore the points' times: one 2-D slice\n"
    "  --delta-squared     the Delta^2 label cache instead of Delta\n"
    "  --draw-cones        cones instead of planes\n"
    "  --help              this message\n"
    "This driver's own:\n"
    "  --engine plain|batched   nni.h per package (default) or served batches\n"
    "  --format asc|flt|both    ESRI ASCII (default), GridFloat, or both\n"
    "  --accum fixed64|fp32|s0  height arithmetic (default fixed64)\n"
    "  --max-size N        largest raster side in pixels (default 5000, as the\n"
    "                      reference forces at terranni.cpp:697-698)\n"
    "  --bin-only          stop after binning\n"
    "  --bin-block N       points binned at a time: files of at most N points\n"
    "                      go side by side, N points in all, larger files in\n"
    "                      blocks of N (default 67108864)\n";

Options parse(int argc, char **argv)
{
    Options o;
    static const char *flags[] = {"wp-computed", "2d", "delta-squared", "draw-cones", "help", "bin-only"};
    for (int i = 1; i < argc; ++i) {
        std::string a = argv[i];
        if (a.rfind("--", 0) != 0) die("unexpected argument " + a + "\n" + kUsage);
        a = a.substr(2);
        std::string name = a, val;
        bool has_val = false;
        const size_t eq = a.find('=');
        if (eq != std::string::npos) {
            name = a.substr(0, eq);
            val = a.substr(eq + 1);
            has_val = true;
        }
        const bool is_flag = std::find_if(std::begin(flags), std::end(flags),
                                          [&](const char *f) { return name == f; }) != std::end(flags);
        if (is_flag) {
            if (name == "wp-computed") o.wp_computed = true;
            else if (name == "2d") o.two_d = true;
            else if (name == "delta-squared") o.delta_squared = true;
            else if (name == "draw-cones") o.draw_cones = true;
            else if (name == "help") o.help = true;
            else if (name == "bin-only") o.bin_only = true;
            continue;
        }
        if (!has_val) {
            if (i + 1 >= argc) die("--" + name + " needs a value");
            val = argv[++i];
        }
        auto num = [&](double &d) {
            char *end = nullptr;
            d = std::strtod(val.c_str(), &end);
            if (end == val.c_str() || *end) die("--" + name + ": not a number: " + val);
        };
        double d = 0;
        if (name == "tilepath") o.tilepath = val;
        else if (name == "output") o.output = val;
        else if (name == "tmp-path") o.tmp_path = val;
        else if (name == "wp-path") o.wp_path = val;
        else if (name == "engine") o.engine = val;
        else if (name == "format") o.format = val;
        else if (name == "accum") o.accum = val;
        else {
            num(d);
            if (name == "scaling") o.scale = (int)d;
            else if (name == "cell-size") o.cell_size = (float)d;
            else if (name == "site-radius") o.site_radius = (float)d;
            else if (name == "query-radius") { o.has_qr = true; o.qr = (float)d; }
            else if (name == "wp-width") { o.has_wpw = true; o.wp_width = (int)d; }
            else if (name == "wp-height") { o.has_wph = true; o.wp_height = (int)d; }
            else if (name == "origin-x") { o.has_ox = true; o.ox = d; }
            else if (name == "origin-y") { o.has_oy = true; o.oy = d; }
            else if (name == "grid-cols") { o.has_gc = true; o.grid_cols = (int)d; }
            else if (name == "grid-rows") { o.has_gr = true; o.grid_rows = (int)d; }
            else if (name == "time-radius") o.time_radius = (int)d;
            else if (name == "time-start") o.time_start = (float)d;
            else if (name == "time-length") o.time_length = (float)d;
            else if (name == "time-step") o.time_step = (float)d;
            else if (name == "max-size") o.max_size = (int)d;
            else if (name == "bin-block") o.bin_block = (long)d;
            else die("unknown option --" + name + "\n" + kUsage);
        }
    }
    if (o.engine != "plain" && o.engine != "batched") die("--engine must be plain or batched");
    if (o.format != "asc" && o.format != "flt" && o.format != "both") die("--format must be asc, flt or both");
    if (o.accum != "fixed64" && o.accum != "fp32" && o.accum != "s0") die("--accum must be fixed64, fp32 or s0");
    if (o.bin_block < 1 || o.bin_block > (1L << 30)) die("--bin-block must be in 1 .. 2^30");
    return o;
}

// ---- the grid and its work packages ---------------------------------------

struct GridBox {
    double minx, miny, maxx, maxy, minz, maxz;
};

// Work packages: wp_nc x wp_nr cells each, cols x rows of them from the grid
// origin.  Sizes are whole cells, so package origins sit on cell edges.
struct Layout {
    double minx, miny;   // grid origin
    double cs;           // cell size
    int nrcols, nrrows;  // cells of the grid
    int wp_nc, wp_nr;    // cells of one package
    int cols, rows;      // packages
    double border;       // metres of neighbouring points each package takes
    double wpw() const { return wp_nc * cs; }
    double wph() const { return wp_nr * cs; }
    double origin_x(int col) const { return minx + col * wpw(); }
    double origin_y(int row) const { return miny + row * wph(); }
};

// The same packages from a = (v - b) / w - 1 and c = (v + b) / w:
// max(0, ceil(a)) and min(n - 1, floor(c)) by truncation, which is exact
// where it is used (0 < a < n, 0 <= c < n).
bool wp_span_of(double a, double c, int n, int &lo, int &hi)
{
    if (!(c >= 0.0) || !(a < (double)n)) return false;
    lo = 0;
    if (a > 0.0) {
        lo = (int)a;
        lo += (double)lo < a;
    }
    hi = c < (double)n ? std::min(n - 1, (int)c) : n - 1;
    return lo <= hi;
}

// Packages c on one axis whose extent [c*w - b, (c+1)*w + b] holds v
// (metres from the grid origin): wp_span_of the quotients by w, each taken
// as a product with inv_w = 1 / w where that cannot change an integer
// part.  The product differs from the quotient by a few units in the last
// place, so a product farther than 1e-9 (relative) from every integer lies
// between the same two integers as the quotient, and every decision
// wp_span_of makes depends only on that; nearer, the quotient is taken.
bool wp_span(double v, double w, double inv_w, double b, int n, int &lo, int &hi)
{
    auto near_int = [](double t) {
        if (!(std::fabs(t) < 1e15)) return true;
        const double f = std::fabs(t - (double)(long long)t);
        const double m = 1e-9 * (1.0 + std::fabs(t));
        return f < m || f > 1.0 - m;
    };
    double a = (v - b) * inv_w - 1.0, c = (v + b) * inv_w;
    if (near_int(a)) a = (v - b) / w - 1.0;
    if (near_int(c)) c = (v + b) / w;
    return wp_span_of(a, c, n, lo, hi);
}

std::vector<std::string> find_tiles(const std::string &dir)
{
    std::vector<std::string> out;
    if (!fs::exists(dir)) return out;
    for (const auto &e : fs::recursive_directory_iterator(dir))
        if (e.is_regular_file() && e.path().extension() == ".las") out.push_back(e.path().string());
    std::sort(out.begin(), out.end());  // a fixed order: point indices decide ties
    return out;
}

// The file name of a time slice (terranni.cpp:409,443).
std::string slice_name(const std::string &outdir, float t)
{
    std::stringstream ss;
    ss << outdir << "/" << std::setw(5) << std::left << std::showpoint << t;
    return ss.str();
}

std::vector<float> time_slices(const Options &o)
{
    std::vector<float> v;
    for (float t = o.time_start; t <= o.time_start + o.time_length; t += o.time_step) v.push_back(t);
    return v;
}

// ---- binning ---------------------------------------------------------------

// A point of the batched engine's packages: LAS integers in the common
// scale/offset, and the time.
struct RawPoint {
    int32_t X, Y, Z, t;
};

struct Timecmp {  // terranni.cpp:119-126
    int compare(const nni::Point4D &a, const nni::Point4D &b) { return a.t - b.t; }
    int compare(const RawPoint &a, const RawPoint &b) { return a.t - b.t; }
};

std::string wp_file(const std::string &dir, int row, int col, const char *ext)
{
    return dir + "/r" + std::to_string(row) + "c" + std::to_string(col) + ext;
}

// The common LAS frame of the batched engine's packages: the first file's
// scale and offset; points of files with another frame are re-quantised to
// it (rounded to its nearest integer).
struct Frame {
    double scale[3], offset[3];
};

// Bin every point of `files` into the streams of the packages whose extent
// holds it, in one pass over the input, then sort each stream by time
// (stable, so equal times keep the file order; a stream whose times already
// arrived in order is left as it is, which is what the stable sort would
// give).  The points are binned into slots, each holding its records
// package by package, and the slots are appended to the streams in slot
// order, so every stream holds its points in file order, as a serial pass
// would:
//   - files of at most kSuper points go in groups of at most one per thread
//     and kSuper points in all, each file read and binned by one thread
//     into its own slot;
//   - a larger file is read in blocks of kSuper points, the threads binning
//     contiguous parts of a block into their slots.
// make(header) gives the file's record maker, maker(X, Y, Z, intensity, x,
// y) a record.  A stream is a file of raw records, so the slots are
// appended package by package, the packages in parallel, each with one
// positioned write per slot at the stream's end.  Any number of packages
// takes one pass (the reference bins in recursive levels of at most
// streamsPerPass streams, terranni.cpp:142-152, 324-363).
template <class Rec, class Make>
void bin_points(const std::vector<std::string> &files, const Layout &L, const std::string &dir,
                const char *ext, size_t kSuper, Make make)
{
    const int n_wp = L.cols * L.rows;
    if (L.border / std::min(L.wpw(), L.wph()) > 30000.0) die("the border spans too many work packages");
    std::vector<std::string> path((size_t)n_wp);
    for (int k = 0; k < n_wp; ++k) {
        // every package starts empty
        path[(size_t)k] = wp_file(dir, k / L.cols, k % L.cols, ext);
        const int fd = ::open(path[(size_t)k].c_str(), O_WRONLY | O_CREAT | O_TRUNC, 0644);
        if (fd < 0) die("cannot create " + path[(size_t)k]);
        ::close(fd);
    }
    std::vector<uint64_t> len((size_t)n_wp, 0);  // records in each stream
    std::vector<int32_t> last_t((size_t)n_wp, INT32_MIN);
    std::vector<char> in_order((size_t)n_wp, 1), seen((size_t)n_wp, 0);
    const int n_par = std::max(1, omp_get_max_threads());
    // A slot's records lie package by package in one buffer, kept from one
    // use to the next, so binning allocates only when a slot outgrows it.
    // a point's packages: rows r0..r0+dr, columns c0..c0+dc, k0 = r0 * cols + c0
    struct Span {
        uint32_t k0;  // UINT32_MAX: none
        uint16_t dc, dr;
    };
    struct Slot {
        std::vector<int32_t> X, Y, Z;  // a whole file's points (group form)
        std::vector<uint16_t> I;
        std::vector<Span> span;
        std::unique_ptr<Rec[]> recs;
        size_t cap = 0;
        std::vector<uint64_t> end;  // per package: one past its last record in recs
        std::vector<uint32_t> cnt;  // per package: its records in this slot
        std::vector<int> touched;   // the packages with records, in first-touch order
        std::string err;
    };
    std::vector<Slot> slots((size_t)n_par);
    for (Slot &sl : slots) {
        sl.end.assign((size_t)n_wp, 0);
        sl.cnt.assign((size_t)n_wp, 0);
    }

    // Bin points [b0, b1) into an empty slot: find each point's packages
    // (sp[i - b0]) and count the records per package, then place them.
    auto bin_range = [&](int slot, const nni_io::las_header &h, const int32_t *X, const int32_t *Y,
                         const int32_t *Z, const uint16_t *I, size_t b0, size_t b1, Span *sp) {
        Slot &sl = slots[(size_t)slot];
        const double wpw = L.wpw(), wph = L.wph(), iw = 1.0 / wpw, ih = 1.0 / wph;
        for (size_t i = b0; i < b1; ++i) {
            const double x = (double)X[i] * h.scale[0] + h.offset[0];
            const double y = (double)Y[i] * h.scale[1] + h.offset[1];
            int c0, c1, r0, r1;
            Span &p = sp[i - b0];
            if (!wp_span(x - L.minx, wpw, iw, L.border, L.cols, c0, c1) ||
                !wp_span(y - L.miny, wph, ih, L.border, L.rows, r0, r1)) {
                p.k0 = UINT32_MAX;
                continue;
            }
            p.k0 = (uint32_t)(r0 * L.cols + c0);
            p.dc = (uint16_t)(c1 - c0);
            p.dr = (uint16_t)(r1 - r0);
            for (int r = r0; r <= r1; ++r)
                for (int c = c0; c <= c1; ++c)
                    if (sl.cnt[(size_t)(r * L.cols + c)]++ == 0) sl.touched.push_back(r * L.cols + c);
        }
        uint64_t total = 0;
        for (int k : sl.touched) {
            sl.end[(size_t)k] = total;  // the fill cursor first
            total += sl.cnt[(size_t)k];
        }
        if (total > sl.cap) {
            sl.recs.reset();
            sl.recs.reset(new Rec[total]);
            sl.cap = total;
        }
        Rec *recs = sl.recs.get();
        const auto conv = make(h);
        for (size_t i = b0; i < b1; ++i) {
            const Span p = sp[i - b0];
            if (p.k0 == UINT32_MAX) continue;
            const double x = (double)X[i] * h.scale[0] + h.offset[0];
            const double y = (double)Y[i] * h.scale[1] + h.offset[1];
            const Rec rec = conv(X[i], Y[i], Z[i], I[i], x, y);
            for (size_t r = 0, k = p.k0; r <= p.dr; ++r, k += (size_t)L.cols)
                for (size_t c = 0; c <= p.dc; ++c) recs[sl.end[k + c]++] = rec;
        }
    };
    // append slots [0, n_slots) to the streams, in slot order, and empty them
    auto flush = [&](int n_slots) {
        for (int s = 0; s < n_slots; ++s)
            if (!slots[(size_t)s].err.empty()) die(slots[(size_t)s].err);
        std::vector<int> wps;
        for (int s = 0; s < n_slots; ++s)
            for (int wp : slots[(size_t)s].touched)
                if (!seen[(size_t)wp]) {
                    seen[(size_t)wp] = 1;
                    wps.push_back(wp);
                }
        bool ok = true;
#pragma omp parallel for schedule(dynamic, 1) reduction(&& : ok)
        for (size_t i = 0; i < wps.size(); ++i) {
            const size_t wp = (size_t)wps[i];
            seen[wp] = 0;
            const int fd = ::open(path[wp].c_str(), O_WRONLY);
            if (fd < 0) {
                ok = false;
                continue;
            }
            for (int s = 0; s < n_slots; ++s) {
                Slot &sl = slots[(size_t)s];
                const size_t n = sl.cnt[wp];
                if (n == 0) continue;
                const Rec *v = sl.recs.get() + (sl.end[wp] - n);
                const size_t bytes = n * sizeof(Rec);
                size_t done = 0;
                while (ok && done < bytes) {
                    const ssize_t w = ::pwrite(fd, (const char *)v + done, bytes - done,
                                               (off_t)(len[wp] * sizeof(Rec) + done));
                    if (w <= 0) ok = false;
                    else done += (size_t)w;
                }
                len[wp] += n;
                for (size_t j = 0; j < n; ++j) {
                    if ((int32_t)v[j].t < last_t[wp]) in_order[wp] = 0;
                    last_t[wp] = (int32_t)v[j].t;
                }
                sl.cnt[wp] = 0;
            }
            if (::close(fd) != 0) ok = false;
        }
        if (!ok) die("writing a work package failed");
        for (int s = 0; s < n_slots; ++s) slots[(size_t)s].touched.clear();
    };

    const size_t fstep = std::max<size_t>(1, files.size() / 20);
    size_t fi = 0;
    std::unique_ptr<nni_io::las_file> next;
    while (fi < files.size()) {
        // the next group of small files, or one large file alone
        std::vector<std::unique_ptr<nni_io::las_file>> grp;
        uint64_t total = 0;
        while (fi + grp.size() < files.size() && (int)grp.size() < n_par) {
            if (!next) {
                next.reset(new nni_io::las_file());
                next->open(files[fi + grp.size()]);
            }
            const uint64_t n = next->header().n_points;
            if (n > kSuper ? !grp.empty() : total + n > kSuper && !grp.empty()) break;
            total += n;
            grp.push_back(std::move(next));
            if (n > kSuper) break;
        }
        for (size_t k = fi; k < fi + grp.size(); ++k)
            if (k % fstep == 0) {
                std::fprintf(stderr, "binning: file %zu of %zu\n", k + 1, files.size());
                std::fflush(stderr);
            }
        if (grp.size() == 1 && grp[0]->header().n_points > kSuper) {
            nni_io::las_file &lf = *grp[0];
            const nni_io::las_header &h = lf.header();
            std::vector<int32_t> X(kSuper), Y(kSuper), Z(kSuper);
            std::vector<uint16_t> I(kSuper);
            std::vector<Span> sp(kSuper);
            size_t got;
            while ((got = lf.read_raw(kSuper, X.data(), Y.data(), Z.data(), I.data())) > 0) {
                int used = 1;
#pragma omp parallel num_threads(n_par)
                {
                    const int k = omp_get_thread_num(), nt = omp_get_num_threads();
                    if (k == 0) used = nt;
                    const size_t b0 = got * (size_t)k / (size_t)nt, b1 = got * (size_t)(k + 1) / (size_t)nt;
                    try {
                        bin_range(k, h, X.data(), Y.data(), Z.data(), I.data(), b0, b1, sp.data() + b0);
                    } catch (const std::exception &e) {
                        slots[(size_t)k].err = e.what();
                    }
                }
                flush(used);
            }
        } else {
            const int g = (int)grp.size();
#pragma omp parallel for schedule(dynamic, 1) num_threads(std::min(n_par, g))
            for (int s = 0; s < g; ++s) {
                Slot &sl = slots[(size_t)s];
                try {
                    nni_io::las_file &lf = *grp[(size_t)s];
                    const size_t n = (size_t)lf.header().n_points;
                    if (sl.X.size() < n) {
                        sl.X.resize(n);
                        sl.Y.resize(n);
                        sl.Z.resize(n);
                        sl.I.resize(n);
                        sl.span.resize(n);
                    }
                    size_t got = 0, k;
                    while (got < n && (k = lf.read_raw(n - got, sl.X.data() + got, sl.Y.data() + got,
                                                       sl.Z.data() + got, sl.I.data() + got)) > 0)
                        got += k;
                    bin_range(s, lf.header(), sl.X.data(), sl.Y.data(), sl.Z.data(), sl.I.data(), 0, got,
                              sl.span.data());
                } catch (const std::exception &e) {
                    sl.err = e.what();
                }
            }
            flush(g);
        }
        fi += grp.size();
    }
    Timecmp cmp;
    for (int k = 0; k < n_wp; ++k) {
        if (in_order[(size_t)k]) continue;
        tpie::ami::stream<Rec> st(path[(size_t)k]);
        if (!st.is_valid() || tpie::ami::sort(&st, &cmp) != tpie::ami::NO_ERROR) die("sorting a work package failed");
    }
}

// ---- output ----------------------------------------------------------------

// ESRI ASCII grid and / or GridFloat raster, rows written top down.
class RasterOut {
public:
    RasterOut(const std::string &base, const std::string &format, int nrows, int ncols, double xll,
              double yll, double cs)
        : ncols_(ncols)
    {
        if (format == "asc" || format == "both") {
            asc_ = std::fopen((base + ".tpie.asc").c_str(), "wb");
            if (!asc_) die("cannot write " + base + ".tpie.asc");
            // the reference's header (terranni.cpp:458-464), coordinates at
            // full precision
            std::fprintf(asc_, "NROWS %d\nNCOLS %d\nXLLCORNER %.15g\nYLLCORNER %.15g\nCELLSIZE %.15g\nNODATA_VALUE -9999\n",
                         nrows, ncols, xll, yll, cs);
        }
        if (format == "flt" || format == "both") {
            // the whole GridFloat file mapped for writing: rows are copied
            // in side by side (positioned writes to one file take its inode
            // lock in turn; page faults on a shared mapping do not)
            const std::string path = base + ".flt";
            flt_fd_ = ::open(path.c_str(), O_RDWR | O_CREAT | O_TRUNC, 0644);
            if (flt_fd_ < 0) die("cannot write " + path);
            flt_bytes_ = (size_t)nrows * (size_t)ncols * sizeof(float);
            if (flt_bytes_ > 0) {
                if (::ftruncate(flt_fd_, (off_t)flt_bytes_) != 0) die("cannot size " + path);
                void *m = ::mmap(nullptr, flt_bytes_, PROT_WRITE, MAP_SHARED, flt_fd_, 0);
                if (m == MAP_FAILED) die("cannot map " + path);
                flt_map_ = (char *)m;
            }
            std::FILE *h = std::fopen((base + ".hdr").c_str(), "wb");
            if (!h) die("cannot write " + base + ".hdr");
            std::fprintf(h, "ncols %d\nnrows %d\nxllcorner %.15g\nyllcorner %.15g\ncellsize %.15g\nNODATA_value -9999\nbyteorder LSBFIRST\n",
                         ncols, nrows, xll, yll, cs);
            std::fclose(h);
        }
    }
    ~RasterOut()
    {
        if (asc_) std::fclose(asc_);
        if (flt_map_) ::munmap(flt_map_, flt_bytes_);
        if (flt_fd_ >= 0) ::close(flt_fd_);
    }

    bool text() const { return asc_ != nullptr; }

    // The next n rows (top down) of a GridFloat-only raster without
    // assembling them: seg(r, iov) gives row r as pieces in column order,
    // copied into the mapped file, the rows side by side.
    template <class Seg>
    void write_rows_direct(int n, Seg seg)
    {
        const size_t row_bytes = (size_t)ncols_ * sizeof(float);
        if (flt_off_ + (size_t)n * row_bytes > flt_bytes_) die("raster rows past the grid");
#pragma omp parallel for schedule(dynamic, 16)
        for (int r = 0; r < n; ++r) {
            std::vector<iovec> iov;
            seg(r, iov);
            char *dst = flt_map_ + flt_off_ + (size_t)r * row_bytes;
            for (const iovec &v : iov) {
                std::memcpy(dst, v.iov_base, v.iov_len);
                dst += v.iov_len;
            }
        }
        flt_off_ += (size_t)n * row_bytes;
    }

    // Rows of ncols heights, the first row the top one of the batch.
    void write_rows(const float *rows, int n)
    {
        if (flt_fd_ >= 0) {
            const size_t bytes = (size_t)n * (size_t)ncols_ * sizeof(float);
            if (flt_off_ + bytes > flt_bytes_) die("raster rows past the grid");
            const int np = std::max(1, std::min(omp_get_max_threads(), (int)(bytes >> 22)));
#pragma omp parallel for num_threads(np)
            for (int k = 0; k < np; ++k) {
                const size_t b0 = bytes * (size_t)k / (size_t)np, b1 = bytes * (size_t)(k + 1) / (size_t)np;
                std::memcpy(flt_map_ + flt_off_ + b0, (const char *)rows + b0, b1 - b0);
            }
            flt_off_ += bytes;
        }
        if (!asc_) return;
        // each value as the reference's `ostream << float << " "`: %g of the
        // value, six significant digits (terranni.cpp:471)
        std::vector<std::string> text((size_t)n);
#pragma omp parallel for schedule(dynamic, 4)
        for (int r = 0; r < n; ++r) {
            std::string &s = text[(size_t)r];
            s.reserve((size_t)ncols_ * 10 + 1);
            char b[48];
            const float *row = rows + (size_t)r * ncols_;
            for (int i = 0; i < ncols_; ++i) {
                const int k = std::snprintf(b, sizeof b, "%g ", (double)row[i]);
                s.append(b, (size_t)k);
            }
            s.push_back('\n');
        }
        for (const std::string &s : text)
            if (std::fwrite(s.data(), 1, s.size(), asc_) != s.size()) die("write failed");
    }

private:
    int ncols_;
    std::FILE *asc_ = nullptr;
    int flt_fd_ = -1;
    char *flt_map_ = nullptr;  // the GridFloat file, mapped
    size_t flt_bytes_ = 0;
    size_t flt_off_ = 0;  // bytes of it written
};

// ---- plain engine: the reference's binRun + outputASC ----------------------

struct XYcmp {  // terranni.cpp:128-138
    int compare(const nni::OutPoint &a, const nni::OutPoint &b)
    {
        if (a.j != b.j) return b.j - a.j;
        return a.i - b.i;
    }
};

void run_plain(const Options &o, const Layout &L, const nni::Settings &s, const std::vector<float> &slices,
               double &t_nni, double &t_out)
{
    const double t0 = wall();
    bool first = true;
    for (int row = 0; row < L.rows; ++row) {
        for (int col = 0; col < L.cols; ++col) {
            tpie::ami::stream<nni::Point4D> wps(wp_file(o.wp_path, row, col, ".tpie"), tpie::ami::READ_STREAM);
            if (!wps.is_valid()) die("missing work package " + wp_file(o.wp_path, row, col, ".tpie"));
            nni::WorkPackage wp;
            wp.origin_x = (float)L.origin_x(col);
            wp.origin_y = (float)L.origin_y(row);
            wp.ncols = L.wp_nc;
            wp.nrows = L.wp_nr;
            wp.grid_origin_x = (float)(col * L.wp_nc);
            wp.grid_origin_y = (float)(row * L.wp_nr);
            for (float t : slices) {
                tpie::ami::stream<nni::OutPoint> out(slice_name(o.output, t) + ".tpie", tpie::ami::APPEND_STREAM);
                out.persist();
                if (first) out.truncate(0);
                nni::setOrigin64(L.origin_x(col), L.origin_y(row));
                nni::performNNI(wp, out, wps, (float)o.time_radius, t, !o.delta_squared);
                if (nni::lastStatus() != 0) die(std::string("performNNI: ") + nni::lastError());
            }
            first = false;
            nni::clearRenderbuffers();
        }
    }
    (void)s;
    t_nni = wall() - t0;
    const double t1 = wall();
    XYcmp cmp;
    std::vector<float> rowbuf((size_t)L.nrcols);
    for (float t : slices) {
        const std::string base = slice_name(o.output, t);
        tpie::ami::stream<nni::OutPoint> out(base + ".tpie", tpie::ami::APPEND_STREAM);
        out.seek(0);
        if (tpie::ami::sort(&out, &cmp) != tpie::ami::NO_ERROR) die("sorting the output failed");
        out.seek(0);
        RasterOut w(base, o.format, L.nrrows, L.nrcols, L.minx, L.miny, L.cs);
        nni::OutPoint *p;
        for (int j = L.nrrows - 1; j >= 0; --j) {
            for (int i = 0; i < L.nrcols; ++i) {
                if (out.read_item(&p) != tpie::ami::NO_ERROR) die("output stream ends early");
                if (p->i != i || p->j != j)
                    std::fprintf(stderr, "Error in output order: %d,%d -- %d,%d\n", i, j, p->i, p->j);
                rowbuf[(size_t)i] = p->h;
            }
            w.write_rows(rowbuf.data(), 1);
        }
    }
    t_out = wall() - t1;
}

// ---- batched engine ----------------------------------------------------------

struct Meta {
    Frame frame;
    double minx, miny, cs, border;
    int wp_nc, wp_nr, cols, rows, two_d;
};

void write_meta(const std::string &dir, const Meta &m)
{
    std::FILE *f = std::fopen((dir + "/wp.meta").c_str(), "wb");
    if (!f) die("cannot write " + dir + "/wp.meta");
    std::fprintf(f, "format=raw1\n");
    for (int k = 0; k < 3; ++k)
        std::fprintf(f, "scale%d=%.17g\noffset%d=%.17g\n", k, m.frame.scale[k], k, m.frame.offset[k]);
    std::fprintf(f, "minx=%.17g\nminy=%.17g\ncell_size=%.17g\nborder=%.17g\nwp_ncols=%d\nwp_nrows=%d\ncols=%d\nrows=%d\ntwo_d=%d\n",
                 m.minx, m.miny, m.cs, m.border, m.wp_nc, m.wp_nr, m.cols, m.rows, m.two_d);
    std::fclose(f);
}

Meta read_meta(const std::string &dir)
{
    std::ifstream in(dir + "/wp.meta");
    if (!in) die("--wp-computed: no " + dir + "/wp.meta (bin with --engine batched first)");
    std::map<std::string, std::string> kv;
    std::string line;
    while (std::getline(in, line)) {
        const size_t eq = line.find('=');
        if (eq != std::string::npos) kv[line.substr(0, eq)] = line.substr(eq + 1);
    }
    if (kv["format"] != "raw1") die(dir + "/wp.meta: unknown format");
    Meta m;
    auto d = [&](const char *k) { return std::strtod(kv[k].c_str(), nullptr); };
    for (int k = 0; k < 3; ++k) {
        m.frame.scale[k] = d(("scale" + std::to_string(k)).c_str());
        m.frame.offset[k] = d(("offset" + std::to_string(k)).c_str());
    }
    m.minx = d("minx");
    m.miny = d("miny");
    m.cs = d("cell_size");
    m.border = d("border");
    m.wp_nc = (int)d("wp_ncols");
    m.wp_nr = (int)d("wp_nrows");
    m.cols = (int)d("cols");
    m.rows = (int)d("rows");
    m.two_d = (int)d("two_d");
    return m;
}

// One package of a row in flight: its points of the slice and its raster.
struct Flight {
    std::vector<int32_t> X, Y, Z;
    std::vector<float> h;
    uint64_t ticket = 0;
};

// The grid rows of one row of packages, top down, from the packages' cell
// rasters (h[col], wp_nr x wp_nc, bottom row first), appended to w.
void write_band(RasterOut &w, const Layout &L, int row, const std::vector<const float *> &h,
                std::vector<float> &band)
{
    const int j_hi = std::min(L.nrrows, (row + 1) * L.wp_nr) - 1, j_lo = row * L.wp_nr;
    const int nb = j_hi - j_lo + 1;
    if (nb <= 0) return;
    if (!w.text()) {
        // GridFloat only: straight from the packages' rasters
        w.write_rows_direct(nb, [&](int jj, std::vector<iovec> &iov) {
            const int lj = j_hi - jj - j_lo;
            for (int col = 0; col < L.cols; ++col) {
                const int i0 = col * L.wp_nc, n = std::min(L.wp_nc, L.nrcols - i0);
                if (n > 0) iov.push_back({(void *)(h[(size_t)col] + (size_t)lj * L.wp_nc), (size_t)n * 4});
            }
        });
        return;
    }
    band.assign((size_t)nb * L.nrcols, -9999.0f);
#pragma omp parallel for schedule(static)
    for (int jj = 0; jj < nb; ++jj) {
        const int j = j_hi - jj;       // global row
        const int lj = j - j_lo;       // row inside the packages
        float *dst = band.data() + (size_t)jj * L.nrcols;
        for (int col = 0; col < L.cols; ++col) {
            const int i0 = col * L.wp_nc, n = std::min(L.wp_nc, L.nrcols - i0);
            if (n > 0) std::memcpy(dst + i0, h[(size_t)col] + (size_t)lj * L.wp_nc, (size_t)n * 4);
        }
    }
    w.write_rows(band.data(), nb);
}

std::vector<RawPoint> read_package(const Options &o, int row, int col)
{
    tpie::ami::stream<RawPoint> st(wp_file(o.wp_path, row, col, ".raw"), tpie::ami::READ_STREAM);
    if (!st.is_valid()) die("missing work package " + wp_file(o.wp_path, row, col, ".raw"));
    std::vector<RawPoint> all((size_t)st.stream_len());
    TPIE_OS_SIZE_T got = all.size();
    if (!all.empty() && st.read_array(all.data(), &got) != tpie::ami::NO_ERROR)
        die("cannot read " + wp_file(o.wp_path, row, col, ".raw"));
    return all;
}

// Space-time (timeRadius > 0) on the batched engine's packages: rows of
// packages outer and slice times inner, so that each package's space-time
// cache (the Delta algorithm, nni.cpp:902-921) still holds the per-time
// Voronoi layers that one call shares with the next; --delta-squared
// searches every slice.  Each
// call takes the points within 2r + 1 of its time (every site of its
// slices), one run of the time-sorted package.  One output raster per slice
// time is open throughout; each gets the band of a package row in turn.
void run_batched_time(tnni_ctx *ctx, const Options &o, const Layout &L, const tnni_tile_desc &base,
                      const std::vector<float> &slices, double &nni_s, double &out_s)
{
    const float tr = (float)o.time_radius;
    const size_t cells = (size_t)L.wp_nr * L.wp_nc;
    std::vector<std::unique_ptr<RasterOut>> outs;
    for (float t : slices)
        outs.emplace_back(new RasterOut(slice_name(o.output, t), o.format, L.nrrows, L.nrcols, L.minx, L.miny, L.cs));
    std::vector<tnni_time_cache *> caches((size_t)L.cols, nullptr);
    if (!o.delta_squared)
        for (auto &c : caches)
            if (tnni_time_cache_create(ctx, 0, &c) != TNNI_OK) die(std::string("cache: ") + tnni_last_error(ctx));
    std::vector<std::vector<RawPoint>> pkg((size_t)L.cols);
    std::vector<std::vector<float>> h((size_t)L.cols, std::vector<float>(cells));
    std::vector<const float *> hp((size_t)L.cols);
    for (int col = 0; col < L.cols; ++col) hp[(size_t)col] = h[(size_t)col].data();
    std::vector<int32_t> X, Y, Z, T;
    std::vector<float> band;
    const double w = 2.0 * (double)tr + 1.0;
    for (int row = L.rows - 1; row >= 0; --row) {
        const double ts = wall();
        for (int col = 0; col < L.cols; ++col) pkg[(size_t)col] = read_package(o, row, col);
        for (size_t si = 0; si < slices.size(); ++si) {
            const float t = slices[si];
            for (int col = 0; col < L.cols; ++col) {
                const std::vector<RawPoint> &all = pkg[(size_t)col];
                size_t b = 0;
                while (b < all.size() && (double)all[b].t - (double)t < -w) ++b;
                size_t e = b;
                while (e < all.size() && (double)all[e].t - (double)t <= w) ++e;
                const size_t k = e - b;
                X.resize(k);
                Y.resize(k);
                Z.resize(k);
                T.resize(k);
                for (size_t q = 0; q < k; ++q) {
                    const RawPoint &p = all[b + q];
                    X[q] = p.X;
                    Y[q] = p.Y;
                    Z[q] = p.Z;
                    T[q] = p.t;
                }
                tnni_tile_desc d = base;
                d.origin_x = L.origin_x(col);
                d.origin_y = L.origin_y(row);
                if (tnni_tile_run_time_cached(ctx, caches[(size_t)col], &d, X.data(), Y.data(), Z.data(), T.data(),
                                              k, t, tr, h[(size_t)col].data()) != TNNI_OK)
                    die(std::string("tile: ") + tnni_last_error(ctx));
            }
            const double tw = wall();
            write_band(*outs[si], L, row, hp, band);
            out_s += wall() - tw;
        }
        nni_s += wall() - ts;
    }
    for (tnni_time_cache *c : caches) tnni_time_cache_destroy(c);
}

void run_batched(tnni_ctx *ctx, const Options &o, const Layout &L, const nni::Settings &s,
                 const Meta &m, const std::vector<float> &slices, double &t_nni, double &t_out)
{
    tnni_tile_desc base;
    std::memset(&base, 0, sizeof base);
    for (int k = 0; k < 3; ++k) {
        base.las_scale[k] = m.frame.scale[k];
        base.las_offset[k] = m.frame.offset[k];
    }
    base.cell_size = s.cell_size;
    base.scale = s.scale;
    base.nrows = L.wp_nr;
    base.ncols = L.wp_nc;
    base.site_radius = s.site_radius;
    base.query_radius = (float)s.query_radius;
    base.algorithm = s.algorithm;
    base.accum_mode = o.accum == "fp32" ? TNNI_ACC_FP32_2SUM : o.accum == "s0" ? TNNI_ACC_S0 : TNNI_ACC_FIXED64;
    base.pixel_offset = -1;
    base.point_min = s.pointMin;
    base.multiplier = s.multiplier;

    double nni_s = 0, out_s = 0;
    const size_t cells = (size_t)L.wp_nr * L.wp_nc;
    if (o.time_radius > 0) {
        run_batched_time(ctx, o, L, base, slices, nni_s, out_s);
        t_nni = nni_s - out_s;
        t_out = out_s;
        return;
    }
    for (float t : slices) {
        RasterOut w(slice_name(o.output, t), o.format, L.nrrows, L.nrcols, L.minx, L.miny, L.cs);
        const float tr = 0.0f;
        // submit every package of one row of packages, the packages read and
        // submitted side by side (the served form takes submits from any
        // thread)
        auto submit_row = [&](int row) {
            std::vector<std::unique_ptr<Flight>> fl((size_t)L.cols);
#pragma omp parallel for schedule(dynamic, 1)
            for (int col = 0; col < L.cols; ++col) {
                std::unique_ptr<Flight> f(new Flight());
                // the whole package in one read; it is sorted by time, so a
                // slice's points are one run of it
                const std::vector<RawPoint> all = read_package(o, row, col);
                // the slice's points (nni.cpp:770 filter)
                size_t b = 0;
                while (b < all.size() && !((double)std::fabs(t - (float)all[b].t) <= (double)tr + 0.000001) &&
                       !((float)all[b].t > t + tr))
                    ++b;
                size_t e = b;
                while (e < all.size() && (double)std::fabs(t - (float)all[e].t) <= (double)tr + 0.000001) ++e;
                const size_t k = e - b;
                f->X.resize(k);
                f->Y.resize(k);
                f->Z.resize(k);
                for (size_t q = 0; q < k; ++q) {
                    const RawPoint &p = all[b + q];
                    f->X[q] = p.X;
                    f->Y[q] = p.Y;
                    f->Z[q] = p.Z;
                }
                f->h.resize(cells);
                tnni_tile_desc d = base;
                d.origin_x = L.origin_x(col);
                d.origin_y = L.origin_y(row);
                if (tnni_submit(ctx, &d, f->X.data(), f->Y.data(), f->Z.data(), f->X.size(), f->h.data(), nullptr,
                                nullptr, &f->ticket) != TNNI_OK)
                    die(std::string("tile: ") + tnni_last_error(ctx));
                fl[(size_t)col] = std::move(f);
            }
            return fl;
        };
        // A writer thread puts each finished row of packages into the
        // raster while the next rows are read, interpolated and waited for.
        std::vector<float> band;
        std::thread writer;
        std::vector<std::unique_ptr<Flight>> writing;  // the row the writer has
        double blocked = 0;  // the main thread waiting for the writer
        auto join_writer = [&] {
            if (!writer.joinable()) return;
            const double a = wall();
            writer.join();
            blocked += wall() - a;
        };
        const double ts = wall();
        auto pending = submit_row(L.rows - 1);
        const int rstep = std::max(1, L.rows / 10);
        for (int row = L.rows - 1; row >= 0; --row) {
            if ((L.rows - 1 - row) % rstep == 0) {
                std::fprintf(stderr, "interpolating: package row %d of %d\n", L.rows - row, L.rows);
                std::fflush(stderr);
            }
            std::vector<std::unique_ptr<Flight>> next;
            if (row > 0) next = submit_row(row - 1);
            // the waits copy the results out of the engine's staging, side by side
#pragma omp parallel for schedule(dynamic, 1)
            for (int col = 0; col < L.cols; ++col)
                if (tnni_wait(ctx, pending[(size_t)col]->ticket, nullptr) != TNNI_OK)
                    die(std::string("tile: ") + tnni_last_error(ctx));
            join_writer();  // bands go into the raster in order
            writing = std::move(pending);
            writer = std::thread([&w, &L, &writing, &band, &out_s, row] {
                const double tw = wall();
                std::vector<const float *> hp;
                for (auto &f : writing) hp.push_back(f->h.data());
                write_band(w, L, row, hp, band);
                out_s += wall() - tw;
            });
            pending = std::move(next);
        }
        join_writer();
        nni_s += wall() - ts - blocked;
    }
    // output overlaps the interpolation: t_out is the writer's time, t_nni
    // the rest of the loop's (reading, interpolating, waiting for results)
    t_nni = nni_s;
    t_out = out_s;
}

}  // namespace

int run(int argc, char **argv)
{
    std::cout << "TerraNNI driver on the MI355X tile engine\n";
    Options o = parse(argc, argv);
    if (o.help || o.tilepath.empty()) {
        if (!o.help) std::cout << "No tile path specified.\n";
        std::cout << kUsage;
        return EXIT_SUCCESS;
    }
    tpie::tpie_init();
    tpie::MM_manager.set_memory_limit((size_t)2000 * 1024 * 1024);

    o.time_length -= 1;  // one slice means no step (terranni.cpp:560)
    if (o.two_d) {
        o.time_radius = 0;
        o.time_start = 0;
        o.time_length = 0;
        o.time_step = 1;
    }
    if (o.time_step == 0) {
        std::cerr << "Cannot set timeStep = 0.  Resetting it to 1.\n";
        o.time_step = 1;
    }
    tpie::tempname::set_default_path(o.tmp_path);

    // 1. the tiles and their bounding box
    const double t_find0 = wall();
    const std::vector<std::string> files = find_tiles(o.tilepath);
    if (files.empty()) {
        std::cout << o.tilepath << " contains no .las files, aborting.";
        return EXIT_SUCCESS;
    }
    GridBox g;
    g.minx = g.miny = g.minz = std::numeric_limits<double>::max();
    g.maxx = g.maxy = g.maxz = std::numeric_limits<double>::lowest();
    Frame frame{};
    for (size_t k = 0; k < files.size(); ++k) {
        nni_io::las_file lf;
        lf.open(files[k]);
        const nni_io::las_header &h = lf.header();
        g.minx = std::min(g.minx, h.GetMinX());
        g.miny = std::min(g.miny, h.GetMinY());
        g.minz = std::min(g.minz, h.GetMinZ());
        g.maxx = std::max(g.maxx, h.GetMaxX());
        g.maxy = std::max(g.maxy, h.GetMaxY());
        g.maxz = std::max(g.maxz, h.GetMaxZ());
        if (k == 0)
            for (int a = 0; a < 3; ++a) {
                frame.scale[a] = h.scale[a];
                frame.offset[a] = h.offset[a];
            }
    }
    std::cout << std::setprecision(15) << "LAS BBOX: x:[" << g.minx << ", " << g.maxx << "]. y:[" << g.miny
              << ", " << g.maxy << "]\n";
    if (o.has_ox) g.minx = o.ox;
    if (o.has_oy) g.miny = o.oy;

    // 2. settings (terranni.cpp:585-690)
    nni::Settings s;
    std::memset(&s, 0, sizeof s);
    s.scale = o.scale;
    s.cell_size = o.cell_size;
    s.site_radius = o.site_radius;
    if (!(s.cell_size > 0) || s.scale < 1) die("--cell-size and --scaling must be positive");
    int nrcols = (int)((g.maxx - g.minx) / s.cell_size);
    int nrrows = (int)((g.maxy - g.miny) / s.cell_size);
    if (o.has_gc) {
        g.maxx = o.grid_cols * (double)s.cell_size + g.minx;
        nrcols = o.grid_cols;
    }
    if (o.has_gr) {
        g.maxy = o.grid_rows * (double)s.cell_size + g.miny;
        nrrows = o.grid_rows;
    }
    if (nrcols < 1 || nrrows < 1) die("the grid has no cells");
    s.pointMin = (int)std::floor(g.minz);
    s.ncols = nrcols;
    s.nrows = nrrows;

    std::error_code ec;
    if (!fs::is_directory(o.output, ec)) die("Output path is not a folder.");
    {
        std::ofstream info(o.output + "/tpie.info");
        if (!info) die("Output path is not a folder.");
        info << std::setprecision(15) << "origin-x=" << g.minx << "\n"
             << "origin-y=" << g.miny + (double)s.cell_size * nrrows << "\n"
             << "nrrows=" << s.nrows << "\n"
             << "nrcols=" << s.ncols << "\n"
             << "grid-length=" << s.cell_size << "\n";
    }
    const int blockWidth = 5;
    const float border = blockWidth * s.cell_size + s.site_radius;
    double qr = blockWidth * s.cell_size / 2.0 - 0.01;
    if (o.has_qr) qr = o.qr;
    s.query_radius = qr;
    s.algorithm = o.draw_cones ? 1 : 0;
    const int pw = (int)std::ceil(2.0 * s.site_radius * s.scale / s.cell_size);
    const double max_sum = (double)pw * pw * (g.maxz - s.pointMin);
    s.multiplier = 1;
    if (max_sum < std::numeric_limits<int>::max()) {
        s.multiplier = (int)std::floor(1.0 * std::numeric_limits<int>::max() / max_sum);
        s.multiplier = (int)std::pow(10, std::floor(std::log10(s.multiplier)));
    }

    // work package size (terranni.cpp:697-708), in whole cells
    const float max_meters =
        std::floor((o.max_size - (s.scale / s.cell_size) * 2 * border - 100) / s.scale) * s.cell_size;
    int wp_width = o.wp_width, wp_height = o.wp_height;
    if (!o.has_wpw || wp_width > max_meters) wp_width = (int)max_meters;
    if (!o.has_wph || wp_height > max_meters) wp_height = (int)max_meters;
    Layout L;
    L.minx = g.minx;
    L.miny = g.miny;
    L.cs = s.cell_size;
    L.nrcols = nrcols;
    L.nrrows = nrrows;
    L.wp_nc = std::max(1, (int)(wp_width / s.cell_size));
    L.wp_nr = std::max(1, (int)(wp_height / s.cell_size));
    L.cols = (nrcols + L.wp_nc - 1) / L.wp_nc;
    L.rows = (nrrows + L.wp_nr - 1) / L.wp_nr;
    L.border = border;
    std::fprintf(stderr, "Grid %d x %d cells; work packages %d x %d cells, %d x %d of them; border %g m\n",
                 nrcols, nrrows, L.wp_nc, L.wp_nr, L.cols, L.rows, (double)border);
    const double t_find = wall() - t_find0;

    // 3. binning
    double t_bin = 0;
    const bool batched = o.engine == "batched";
    Meta m;
    m.frame = frame;
    m.minx = L.minx;
    m.miny = L.miny;
    m.cs = L.cs;
    m.border = L.border;
    m.wp_nc = L.wp_nc;
    m.wp_nr = L.wp_nr;
    m.cols = L.cols;
    m.rows = L.rows;
    m.two_d = o.two_d ? 1 : 0;
    if (!o.wp_computed) {
        const double tb = wall();
        const bool twoD = o.two_d;
        if (batched) {
            bin_points<RawPoint>(files, L, o.wp_path, ".raw", (size_t)o.bin_block,
                                 [&frame, twoD](const nni_io::las_header &h) {
                                     const bool same = std::memcmp(h.scale, frame.scale, sizeof frame.scale) == 0 &&
                                                       std::memcmp(h.offset, frame.offset, sizeof frame.offset) == 0;
                                     return [&frame, &h, twoD, same](int32_t X, int32_t Y, int32_t Z, uint16_t I,
                                                                     double x, double y) {
                                         RawPoint r;
                                         r.t = twoD ? 0 : (int)I;
                                         if (same) {
                                             r.X = X;
                                             r.Y = Y;
                                             r.Z = Z;
                                         } else {
                                             const double z = (double)Z * h.scale[2] + h.offset[2];
                                             auto q = [](double v, double sc, double of) {
                                                 const double u = std::nearbyint((v - of) / sc);
                                                 if (!(std::fabs(u) < 2147483647.0)) die("a point does not fit the common LAS frame");
                                                 return (int32_t)u;
                                             };
                                             r.X = q(x, frame.scale[0], frame.offset[0]);
                                             r.Y = q(y, frame.scale[1], frame.offset[1]);
                                             r.Z = q(z, frame.scale[2], frame.offset[2]);
                                         }
                                         return r;
                                     };
                                 });
            write_meta(o.wp_path, m);
        } else {
            bin_points<nni::Point4D>(files, L, o.wp_path, ".tpie", (size_t)o.bin_block,
                                     [twoD](const nni_io::las_header &h) {
                                         return [&h, twoD](int32_t, int32_t, int32_t Z, uint16_t I, double x, double y) {
                                             return nni::Point4D(x, y, (double)Z * h.scale[2] + h.offset[2],
                                                                 twoD ? 0 : (int)I);
                                         };
                                     });
        }
        t_bin = wall() - tb;
    } else if (batched) {
        const Meta w = read_meta(o.wp_path);
        if (w.wp_nc != m.wp_nc || w.wp_nr != m.wp_nr || w.cols != m.cols || w.rows != m.rows ||
            w.minx != m.minx || w.miny != m.miny || w.cs != m.cs || w.two_d != m.two_d)
            die("--wp-computed: the work packages in " + o.wp_path + " were binned for another grid");
        m.frame = w.frame;
    }
    if (o.bin_only) {
        std::cerr << "Total Find Tile time: " << t_find << "\nTotal Tile time: " << t_bin << "\n";
        return EXIT_SUCCESS;
    }

    // 4. interpolation and output
    const std::vector<float> slices = time_slices(o);
    double t_nni = 0, t_out = 0;
    if (batched) {
        // the engine context is made before the timed part, like nni::init
        // in the reference (terranni.cpp:694)
        tnni_ctx_opts co;
        std::memset(&co, 0, sizeof co);
        if (const char *dev = std::getenv("TERRANNI_DEVICE")) co.device = std::atoi(dev);
        co.n_slots = 3;
        tnni_ctx *ctx = nullptr;
        if (tnni_ctx_create(&co, &ctx) != TNNI_OK) die(std::string("engine: ") + tnni_last_error(nullptr));
        run_batched(ctx, o, L, s, m, slices, t_nni, t_out);
        tnni_ctx_destroy(ctx);
    } else {
        nni::init(s);
        if (nni::lastStatus() != 0) die(std::string("nni::init: ") + nni::lastError());
        if (o.accum == "fp32") nni::setAccumulationMode(TNNI_ACC_FP32_2SUM);
        else if (o.accum == "s0") nni::setAccumulationMode(TNNI_ACC_S0);
        else nni::setAccumulationMode(TNNI_ACC_FIXED64);
        run_plain(o, L, s, slices, t_nni, t_out);
        nni::cleanup();
    }
    std::cerr << "Total Find Tile time: " << t_find << "\n"
              << "Total Tile time: " << t_bin << "\n"
              << "Total NNI time: " << t_nni << "\n"
              << "Total Print Output time: " << t_out << "\n";
    return EXIT_SUCCESS;
}

int main(int argc, char **argv)
{
    try {
        return run(argc, argv);
    } catch (const std::exception &e) {
        die(e.what());
    }
}
