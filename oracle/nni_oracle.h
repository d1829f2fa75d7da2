/*
 * CPU oracle for the discrete-Sibson (natural-neighbour) tile path.
 *
 * TEST INFRASTRUCTURE ONLY.  Nothing in the product library links, loads or
 * calls this code.  It is used by tests/, by __graft_entry__.smoke() and by
 * bench.py's cpu_baseline leg, and there only to check results (or as the timed
 * CPU baseline), never as the thing measured or shipped.
 *
 * It restates, on the CPU and in exact integer arithmetic, what TerraNNI's
 * per-tile path computes with OpenGL + CUDA:
 *
 *   - the pixel frame of a work package          nni.cpp:879-886, 656-660
 *   - the site pass (discrete Voronoi by depth)  nni.cpp:170-230, 538-549,
 *                                                nni.cpp:722-790 (GL_LEQUAL)
 *   - the query stencil + steal test             nni.cpp:483-534, 794-820
 *   - the accumulation and the divide            cuda_nni.cu:407-503
 *   - the de-scale of the result                 nni.cpp:824-866
 *
 * under the exact semantics used by the engine:
 *   labels   = among the sites whose region of influence covers a pixel
 *              centre, the one with the smallest integer squared distance,
 *              ties to the lowest input index;
 *   steal    = query covers the pixel and |x - q|^2 <= |x - s(x)|^2;
 *   heights  = sum of the stolen pixels' site heights / their count.
 *
 * Coordinates are integers in a tile-local fixed-point unit (1/upm metre).
 */
#ifndef TERRANNI_ORACLE_H
#define TERRANNI_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* Inputs of one tile: nni::Settings + nni::WorkPackage (nni.h:66-81) plus
 * the LAS scale/offset that turn int32 coordinates into metres. */
typedef struct {
    double origin_x, origin_y;   /* WorkPackage::origin_x/y, metres        */
    double las_scale[3];         /* LAS X/Y/Z scale                         */
    double las_offset[3];        /* LAS X/Y/Z offset                        */
    float cell_size;             /* Settings::cell_size (float)             */
    int scale;                   /* Settings::scale                         */
    int nrows, ncols;            /* WorkPackage::nrows/ncols                */
    float site_radius;           /* Settings::site_radius (float)           */
    float query_radius;          /* query radius, held as float nni.cpp:69  */
    int algorithm;               /* 0 planes (square ROI), 1 cones (disk)   */
    int units_per_m;             /* fixed-point units per metre, 0 = auto   */
    int pixel_offset;            /* -1 = formula of nni.cpp:883            */
    int point_min;               /* Settings::pointMin (S0-compat heights)  */
    int multiplier;              /* Settings::multiplier (S0-compat)        */
} orc_desc;

typedef struct {
    int upm;                 /* units per metre                            */
    int ps_u, cs_u;          /* pixel size, cell size in units             */
    int po;                  /* pixel offset (border pixels on each side)  */
    int W, H;                /* raster width / height in pixels            */
    int x0, y0;              /* centre of pixel (0,0), tile-local units    */
    int Rs, Rq;              /* half-side of the square regions (planes)   */
    int64_t Rs2, Rq2;        /* squared radii of the disk regions (cones)  */
    int Rsb, Rqb;            /* L-inf bound of the region actually used    */
    int spacing_factor;      /* nni.cpp:879 (rounds of the GL query pass)  */
    int algorithm;           /* copied from the descriptor                 */
} orc_frame;

/* Returns 0 on success, nonzero if the geometry cannot be represented
 * exactly in any supported fixed-point unit. */
int orc_frame_compute(const orc_desc *d, orc_frame *f);

/* LAS int32 X/Y -> tile-local units.  keep[i] = 0 for points that cannot
 * cover any pixel of the raster (further than the site region from it). */
void orc_points_to_units(const orc_desc *d, const orc_frame *f,
                         const int32_t *X, const int32_t *Y, int64_t n,
                         int32_t *xu, int32_t *yu, uint8_t *keep);

/* Site pass, site-major like the GL rasteriser: every site writes the
 * pixels of its region, smaller d2 wins, equal d2 keeps the earlier site.
 * d2[p] = 0xFFFFFFFF and label[p] = -1 where no site covers pixel p.
 * Rasters are row-major, H rows of W pixels.  nthreads <= 0: OpenMP default. */
void orc_labels_raster(const orc_frame *f, const int32_t *xu, const int32_t *yu,
                       const uint8_t *keep, int64_t n,
                       uint32_t *d2, int32_t *label, int nthreads);

/* Same result by brute force, pixel-major, O(P*N): tiny inputs only. */
void orc_labels_brute(const orc_frame *f, const int32_t *xu, const int32_t *yu,
                      const uint8_t *keep, int64_t n,
                      uint32_t *d2, int32_t *label);

/* Query pass + accumulation + divide, query-major.
 * Outputs (any may be NULL):
 *   h_fixed  float  (sum Z)/count * sz + oz, NODATA -9999  (engine formula)
 *   h64      double mean of the FP64 site heights fma(Z,sz,oz), NODATA -9999
 *   h_s0     float  the reference's integer path (zMultiplier quantisation,
 *                   int32 wrap, truncating division, de-scale), nni.cpp:539,
 *                   cuda_nni.cu:447-448,491-503, nni.cpp:841-846
 *   count    uint32 stolen pixels per cell (the denominator)
 *   pairs[0] = K, candidate (pixel, cell) pairs with a labelled pixel inside
 *              the query region; pairs[1] = S, stolen pairs. */
void orc_heights(const orc_desc *d, const orc_frame *f,
                 const uint32_t *d2, const int32_t *label, const int32_t *Z,
                 float *h_fixed, double *h64, float *h_s0, uint32_t *count,
                 uint64_t *pairs, int nthreads);

/* S0 site value floor((z + zOffset + 1) * zMult), nni.cpp:539. */
uint32_t orc_s0_value(const orc_desc *d, int32_t Z);

/* ---- space-time (timeRadius > 0), nni.cpp:902-936 ----------------------
 * performNNI runs the slices t = curTime + j, j = -(int)timeRadius ..
 * timeRadius.  Slice t's sites are the points with |t - T| <= timeRadius +
 * 1e-6 (nni.cpp:770), each at squared time distance dt2 = rint(|t - T| *
 * upm)^2 (units^2: one time unit is one metre, drawPlane's timeOffset term,
 * nni.cpp:197) added to its squared distance; the region tests stay in the
 * plane.  The queries of slice j add (j * upm)^2 (performNNIQueries with
 * timeOffset j, nni.cpp:931,810).  Every cell sums over all slices, then
 * divides once (cudaSetup zeroes once, nni.cpp:906; completeBufferAnalysis
 * after the loop, :936). */

/* The slice's sites (keep_slice) and their dt2 (0 where not a site). */
void orc_slice_sites(const orc_frame *f, const int32_t *T, const uint8_t *keep, int64_t n,
                     float t, float time_radius, uint8_t *keep_slice, uint32_t *dt2);

/* The site passes with per-site squared time distances (dt2 NULL: none). */
void orc_labels_raster_t(const orc_frame *f, const int32_t *xu, const int32_t *yu,
                         const uint8_t *keep, const uint32_t *dt2, int64_t n,
                         uint32_t *d2, int32_t *label, int nthreads);
void orc_labels_brute_t(const orc_frame *f, const int32_t *xu, const int32_t *yu,
                        const uint8_t *keep, const uint32_t *dt2, int64_t n,
                        uint32_t *d2, int32_t *label);

/* A cell's running sums over the slices. */
typedef struct {
    int64_t sumZ;     /* sum of the stolen sites' LAS Z integers          */
    double sum64;     /* sum of their FP64 heights                         */
    uint32_t numer0;  /* S0: sum of the zMultiplier values (u32 wrap)      */
    uint32_t denom0;  /* S0: their count (colour 0 skipped)                */
    uint32_t cnt;     /* stolen pixels                                     */
    uint32_t pad;
} orc_acc;

/* One slice's query pass into acc (cells row-major), the queries at time
 * offset jj2 = (j * upm)^2; pairs (nullable) accumulates K and S. */
void orc_heights_accumulate(const orc_desc *d, const orc_frame *f,
                            const uint32_t *d2, const int32_t *label, const int32_t *Z,
                            int64_t jj2, orc_acc *acc, uint64_t *pairs, int nthreads);

/* The divide of every cell (outputs as orc_heights, any may be NULL). */
void orc_heights_finish(const orc_desc *d, const orc_acc *acc, float *h_fixed, double *h64,
                        float *h_s0, uint32_t *count);

#ifdef __cplusplus
}
#endif
#endif
