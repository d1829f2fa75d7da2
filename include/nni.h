/*
 * nni.h -- the per-tile natural-neighbour API of TerraNNI (nni.h:32-99 of
 * the reference), served by the MI355X engine of terranni.h.
 *
 * Same namespace, types, layouts and functions, so the reference's callers
 * (terranni.cpp: nni::init at :694, performNNI at :419, clearRenderbuffers
 * at :422, cleanup at :721) compile and run unchanged.  The OpenGL render
 * buffers are gone: setRenderbuffers / setupRun are kept as no-ops for
 * source compatibility.  performNNI with timeRadius > 0 and useDelta keeps
 * per point time the 2-D Voronoi layer in the engine's space-time cache (in
 * place of the Delta render buffers, nni.cpp:281-379) and builds each slice
 * from it; clearRenderbuffers empties it; without useDelta (Delta^2) every
 * slice is searched.  Both give the same heights.
 */
#ifndef LIBNNI_H
#define LIBNNI_H

#include <cstdint>
#include <vector>

#include "tpie/cpu_timer.h"
#include "tpie/stream.h"
#include "tpie_config.h"

namespace nni {

// An input point: x, y, z in metres (nni.h:34-42).
struct Point3D {
    double p[3];
    Point3D() {}
    Point3D(double x, double y, double z, double t = 0)
    {
        (void)t;
        p[0] = x;
        p[1] = y;
        p[2] = z;
    }
};

// An input point with its time (LAS intensity, las_reader.inl:59), nni.h:44-53.
struct Point4D : public Point3D {
    int t;
    Point4D(double x, double y, double z, int time = 0)
    {
        p[0] = x;
        p[1] = y;
        p[2] = z;
        t = time;
    }
    Point4D() {}
};

// One interpolated grid cell: global column i, row j, height h (nni.h:55-64).
struct OutPoint {
    int i, j;
    float h;
    OutPoint(int x, int y, float z)
    {
        i = x;
        j = y;
        h = z;
    }
    OutPoint() {}
};

// One tile of the grid (nni.h:66-70).
struct WorkPackage {
    float origin_x, origin_y;           // tile origin in metres
    float grid_origin_x, grid_origin_y; // tile origin in grid cells
    int nrows, ncols;
};

// Global settings (nni.h:72-81).
struct Settings {
    int scale;           // pixels per cell side
    float cell_size;     // metres
    int nrows, ncols;    // cells of the whole grid
    float site_radius;   // metres
    double query_radius; // metres
    int algorithm;       // 0 planes, 1 cones
    int pointMin;        // floor of the lowest height (integer path only)
    int multiplier;      // fixed-point height scale (integer path only)
};

void cleanup();

int init(const Settings &s);

void setRenderbuffers(int w_width, int w_height, int tC = 1);

void printTimes();

void setupRun(const WorkPackage &wp, int pixel_offset = -1, int timeCount = 1);

void performNNI(const WorkPackage &wp, tpie::ami::stream<OutPoint> &out,
                tpie::ami::stream<nni::Point4D> &vec, float timeRadius, float curTime,
                bool useDelta, bool debug = false);

void clearRenderbuffers();

// ---- additions of this engine (not in the reference) ------------------
// Height arithmetic: 0 exact int64 sums of the heights (default), 1 FP32
// two-sum, 2 the reference's zMultiplier integer arithmetic (needs
// Settings::pointMin / multiplier).  Takes effect at the next performNNI.
void setAccumulationMode(int mode);
// FP64 origin (metres) of the next performNNI's work package.  The
// WorkPackage's float origin rounds projected coordinates (UTM eastings of
// 10^5-10^6 m) to centimetres or worse; a driver that calls this first gets
// the tile placed exactly.  performNNI consumes it (later calls fall back to
// the float origin until it is set again).
void setOrigin64(double origin_x, double origin_y);
// Status of the last call (0 = TNNI_OK) and its message.
int lastStatus();
const char *lastError();

}  // namespace nni

#endif
