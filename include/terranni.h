/*
 * terranni.h -- C ABI of the MI355X (gfx950) discrete-Sibson tile engine.
 *
 * This is the thin device boundary that replaces TerraNNI's internal
 * cuda_nni.h interface (cuda_nni.h:29-42: init, setup_run, cudaSetup,
 * bufferAnalysis, completeAnalysis, cleanupCUDA, cudaSwitchVoronoiRB,
 * cudaRegisterAllBuffers) together with the OpenGL passes of nni.cpp that
 * produce the discrete Voronoi diagram and the query stencil.  The public C++
 * API of the reference (nni.h:32-99, namespace nni) is provided on top of
 * this ABI by include/nni.h.
 *
 * Plain C: plain pointers and sizes, no C++ or framework types.  Every call
 * returns a tnni_status; nothing calls exit() or abort().  Host buffers stay
 * owned by the caller; device arenas are owned by the context.
 *
 * Exact semantics (shared with the CPU oracle under oracle/):
 *   - point (X,Y) is de-scaled in FP64, x = fma(X, sx, ox) - origin_x, then
 *     rounded to the tile-local fixed-point unit 1/upm metre;
 *   - pixel (k,l) of the W x H raster has its centre at x0 + k*ps, y0 + l*ps
 *     (the glOrtho frame of nni.cpp:656-660); query (i,j) sits at
 *     (i*cs, j*cs) (nni.cpp:488);
 *   - a site covers a pixel if it lies in its region (planes: square of
 *     half-side r/sqrt(2), nni.cpp:183; cones: disk of radius r);
 *   - label(pixel) = the covering site with the smallest integer squared
 *     distance, ties to the lowest input index;
 *   - query q steals pixel x if q's region covers x and
 *     |x - q|^2 <= |x - label(x)|^2 (the depth test of nni.cpp:794-820);
 *   - height(q) = mean height of the stolen pixels' sites, -9999 if none
 *     (cuda_nni.cu:491-503).
 */
#ifndef TERRANNI_H
#define TERRANNI_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

#define TNNI_ABI_VERSION 4

typedef enum tnni_status {
    TNNI_OK = 0,
    TNNI_PENDING = 1,        /* tnni_poll: request still running            */
    TNNI_E_INVALID = -1,     /* bad argument                                 */
    TNNI_E_GEOMETRY = -2,    /* tile geometry not representable exactly      */
    TNNI_E_NODEVICE = -3,    /* no HIP device / library built without one    */
    TNNI_E_HIP = -4,         /* a HIP runtime call failed                    */
    TNNI_E_NOMEM = -5,       /* device allocation failed                     */
    TNNI_E_BUDGET = -6,      /* request larger than the context's HBM budget */
    TNNI_E_SHUTDOWN = -7     /* context is being destroyed                   */
} tnni_status;

enum { TNNI_ALG_PLANES = 0, TNNI_ALG_CONES = 1 }; /* nni::Settings::algorithm */

enum {
    TNNI_ACC_FIXED64 = 0,   /* int64 sum of the LAS Z integers (default)     */
    TNNI_ACC_FP32_2SUM = 1, /* FP32 two-sum of the Z integers                */
    TNNI_ACC_S0 = 2         /* the reference's zMultiplier int32 arithmetic  */
};

typedef struct tnni_ctx tnni_ctx;

typedef struct tnni_ctx_opts {
    int device;               /* HIP device ordinal                         */
    int n_slots;              /* served form: batches in flight (0 = 2)     */
    uint64_t hbm_budget;      /* bytes the context may allocate (0 = 80% of
                                 free device memory)                        */
    int max_batch_tiles;      /* served form: tiles per batch (0 = 256)     */
    int batch_window_us;      /* served form: wait for more arrivals (0)    */
} tnni_ctx_opts;

/* One work package (nni::WorkPackage nni.h:66-70 + nni::Settings nni.h:72-81)
 * plus the LAS scale/offset of its points. */
typedef struct tnni_tile_desc {
    double origin_x, origin_y;  /* tile origin in metres (FP64)            */
    double las_scale[3];        /* X, Y, Z scale                           */
    double las_offset[3];       /* X, Y, Z offset                          */
    float cell_size;            /* Settings::cell_size                     */
    int scale;                  /* Settings::scale (pixels per cell side)  */
    int nrows, ncols;           /* cells in the tile                       */
    float site_radius;          /* Settings::site_radius                   */
    float query_radius;         /* Settings::query_radius (as float)       */
    int algorithm;              /* TNNI_ALG_*                              */
    int accum_mode;             /* TNNI_ACC_*                              */
    int units_per_m;            /* fixed-point unit, 0 = auto (100 first)  */
    int pixel_offset;           /* -1 = the formula of nni.cpp:883         */
    int point_min;              /* Settings::pointMin   (TNNI_ACC_S0 only) */
    int multiplier;             /* Settings::multiplier (TNNI_ACC_S0 only) */
    int bucket_px;              /* bucket side in pixels, 0 = tuned        */
    int tile_cells;             /* cell-tile side per workgroup (in the
                                   split form: the label tile side in
                                   pixels, at least 8), 0 = auto           */
    int form;                   /* TNNI_FORM_*: how labels reach the
                                   accumulation, 0 = tuned / automatic     */
} tnni_tile_desc;

/* tnni_tile_desc::form.  Every form gives the same results bit for bit.
 *  FUSED: each work item labels its cells' pixels plus the query halo into
 *    LDS and accumulates its cells from there in the same kernel (the halo
 *    is labelled by every item that needs it);
 *  SPLIT: each pixel is labelled once (items without a halo) into a raster
 *    of site values and squared distances in device memory, then a second
 *    kernel gathers every cell from that raster (each cell tests every
 *    pixel of its query window).  Pays where the query window is wide in
 *    pixels (large scale), so that the halo dominates;
 *  SCATTER: the split form's labelling, then the accumulation the other
 *    way round (cuda_nni.cu:407-488): each pixel adds its site value to
 *    the cells within its nearest-site distance, visiting only the cells of
 *    its disk's bounding box, into per-block LDS sums with lanes that hit
 *    the same cell summed across the wavefront first.  Integer modes only
 *    (TNNI_ACC_FIXED64, TNNI_ACC_S0), whose sums do not depend on order;
 *    TNNI_E_INVALID with TNNI_ACC_FP32_2SUM.  Never chosen automatically
 *    unless the autotuner measured it faster. */
enum { TNNI_FORM_AUTO = 0, TNNI_FORM_FUSED = 1, TNNI_FORM_SPLIT = 2, TNNI_FORM_SCATTER = 3 };

/* Derived geometry of a tile (all lengths in fixed-point units). */
typedef struct tnni_frame_info {
    int upm;                    /* units per metre                         */
    int ps_u, cs_u;             /* pixel and cell size                     */
    int po;                     /* border pixels on each side              */
    int W, H;                   /* raster size in pixels                   */
    int x0, y0;                 /* centre of pixel (0,0)                   */
    int Rs, Rq;                 /* square half-sides (planes)              */
    int64_t Rs2, Rq2;           /* squared disk radii (cones)              */
    int Rsb, Rqb;               /* L-inf extent of the regions in use      */
    int spacing_factor;         /* GL query rounds of the reference        */
    int algorithm;
} tnni_frame_info;

/* ---- context ---------------------------------------------------------- */
/* Replaces cuda_nni::init (cuda_nni.cu:641-669): picks the device, creates
 * the stream and arenas.  No GL context. */
tnni_status tnni_ctx_create(const tnni_ctx_opts *opts, tnni_ctx **out);
/* Replaces nni::cleanup / cuda_nni::cleanupCUDA (nni.cpp:563-566,
 * cuda_nni.cu:624-632).  Waits for outstanding requests. */
void tnni_ctx_destroy(tnni_ctx *ctx);
/* Last error message of this thread (or of ctx when ctx != NULL: the last
 * error any caller thread met on ctx).  The string belongs to the calling
 * thread and stays valid until its next tnni_last_error call. */
const char *tnni_last_error(const tnni_ctx *ctx);
int tnni_abi_version(void);

/* ---- geometry and admission ------------------------------------------ */
/* The pixel frame of performNNI (nni.cpp:879-886) and setupRun's glOrtho
 * (nni.cpp:656-660), in fixed-point units.  Host only. */
tnni_status tnni_frame(const tnni_tile_desc *desc, tnni_frame_info *out);
/* Device bytes one tile of n points needs (used for admission). */
tnni_status tnni_footprint(const tnni_tile_desc *desc, uint64_t n_points,
                           int with_labels, uint64_t *bytes);

/* ---- plain form: one tile per call, host buffers ---------------------- */
/* The whole of nni::performNNI's device work for one 2-D time slice
 * (nni.cpp:870-944): site pass, query pass, accumulate, divide.
 * X, Y, Z: n LAS int32 coordinates.  out_h: nrows*ncols float heights,
 * row-major (row j = y index).  out_labels (nullable): W*H int32 labels,
 * -1 = no covering site.  out_d2 (nullable): W*H uint32 squared distance
 * of the label, 0xFFFFFFFF where none.  Synchronous. */
tnni_status tnni_tile_run(tnni_ctx *ctx, const tnni_tile_desc *desc,
                          const int32_t *X, const int32_t *Y, const int32_t *Z,
                          uint64_t n, float *out_h, int32_t *out_labels,
                          uint32_t *out_d2);

/* ---- device-resident buffers ------------------------------------------ */
tnni_status tnni_dev_alloc(tnni_ctx *ctx, uint64_t bytes, void **dptr);
tnni_status tnni_dev_free(tnni_ctx *ctx, void *dptr);
tnni_status tnni_h2d(tnni_ctx *ctx, void *dst, const void *src, uint64_t bytes);
tnni_status tnni_d2h(tnni_ctx *ctx, void *dst, const void *src, uint64_t bytes);
tnni_status tnni_d2d(tnni_ctx *ctx, void *dst, const void *src, uint64_t bytes);

/* Plain form on device pointers: one launch sequence for one tile on the
 * context stream, then a stream synchronisation. */
tnni_status tnni_tile_run_dev(tnni_ctx *ctx, const tnni_tile_desc *desc,
                              const int32_t *dX, const int32_t *dY,
                              const int32_t *dZ, uint64_t n, float *d_out_h,
                              int32_t *d_out_labels, uint32_t *d_out_d2);

/* The same as tnni_tile_run_dev (same arguments, same results bit for bit)
 * for callers that cannot allocate device memory: code under HIP graph
 * capture, tight serving loops, harnesses that forbid hipMalloc.  It has no
 * counterpart in the reference.  It calls no hipMalloc, hipFree,
 * hipHostMalloc or other device- or pinned-memory allocation: besides the
 * caller's inputs and outputs, its scratch (descriptors, histograms, the
 * sorted site records, bucket offsets, distance maps) lives in a fixed
 * array of 128 MiB per device, set up when a context is first created on
 * that device and held until the process ends.  A tile that needs more
 * scratch (roughly 3M points and up, or a very fine bucket grid) fails with
 * TNNI_E_BUDGET and a message giving the bytes it needs.  All calls on one
 * device share its array and would race if they overlapped, so the library
 * runs them one at a time per device within the process; each call writes
 * everything it later reads, so the previous call's contents never matter.
 * Like tnni_tile_run_dev it runs on the context stream and synchronises. */
tnni_status tnni_tile_run_dev_noalloc(tnni_ctx *ctx, const tnni_tile_desc *desc,
                                      const int32_t *dX, const int32_t *dY,
                                      const int32_t *dZ, uint64_t n, float *d_out_h,
                                      int32_t *d_out_labels, uint32_t *d_out_d2);

/* ---- batched form ----------------------------------------------------- */
/* n_tiles uneven tiles in one pipeline: one bucket sort over all points and
 * one persistent launch over a device work queue of (tile, cell-block)
 * items, replayed from a captured hipGraph when the batch shape repeats.
 * Arrays are indexed by tile; pointers are device pointers; d_out_labels and
 * d_out_d2 may be NULL (or hold NULL entries). */
tnni_status tnni_batch_run_dev(tnni_ctx *ctx, int n_tiles,
                               const tnni_tile_desc *descs,
                               const int32_t *const *dX,
                               const int32_t *const *dY,
                               const int32_t *const *dZ, const uint64_t *n,
                               float *const *d_out_h,
                               int32_t *const *d_out_labels,
                               uint32_t *const *d_out_d2);

/* ---- served form: continuous batching ---------------------------------- */
/* Queue one tile (host buffers, which must stay valid until tnni_wait
 * returns).  A scheduler thread admits queued tiles as batch slots and HBM
 * free up (admission by tnni_footprint against the context budget), runs
 * each admitted batch as one graph-replayed pipeline on its own stream, so
 * that batches overlap on the device, and copies results back.  The result
 * of a tile does not depend on which batch it ran in.  Fails at once with
 * TNNI_E_BUDGET if the tile alone exceeds the budget. */
tnni_status tnni_submit(tnni_ctx *ctx, const tnni_tile_desc *desc,
                        const int32_t *X, const int32_t *Y, const int32_t *Z,
                        uint64_t n, float *out_h, int32_t *out_labels,
                        uint32_t *out_d2, uint64_t *ticket);
/* The same with device-resident inputs and outputs (device pointers that
 * stay valid until tnni_wait returns): no host copies; latency is submit ->
 * results in device memory. */
tnni_status tnni_submit_dev(tnni_ctx *ctx, const tnni_tile_desc *desc,
                            const int32_t *dX, const int32_t *dY,
                            const int32_t *dZ, uint64_t n, float *d_out_h,
                            int32_t *d_out_labels, uint32_t *d_out_d2,
                            uint64_t *ticket);
/* TNNI_OK when the result is in host memory, TNNI_PENDING while queued or
 * running, or the request's error. */
tnni_status tnni_poll(tnni_ctx *ctx, uint64_t ticket);
/* Block until done, copy a host request's results into its buffers; returns
 * the request's status and releases the ticket.  latency_s (nullable):
 * submit -> results on the host side of the engine (device memory for
 * tnni_submit_dev). */
tnni_status tnni_wait(tnni_ctx *ctx, uint64_t ticket, double *latency_s);

typedef struct tnni_serve_stats {
    uint64_t submitted, completed, batches;
    uint64_t max_inflight_tiles;   /* most tiles on the device at once   */
    uint64_t max_inflight_batches; /* most batches on the device at once */
    uint64_t inflight_bytes;       /* footprint admitted right now       */
    uint64_t budget_retries;       /* launches put back to wait for HBM  */
} tnni_serve_stats;
tnni_status tnni_serve_stats_read(tnni_ctx *ctx, tnni_serve_stats *out);
/* Restart the max_inflight_* high-water marks (the counts are kept). */
tnni_status tnni_serve_stats_reset(tnni_ctx *ctx);

/* ---- space-time (timeRadius > 0) ----------------------------------------
 * One work package at time cur_time with time radius r, as performNNI with
 * timeRadius > 0 (nni.cpp:870-944, replaces its slice loop :910-933).
 * Nothing is kept between calls (the Delta^2 algorithm; the Delta
 * algorithm's render-buffer cache is tnni_time_cache below).  Points sorted
 * by time (at most 64 point times within reach) are searched once per time
 * and each slice combined from those per-time layers (their buffers, 8
 * bytes per pixel per time reached, stay in the context for the next call);
 * other calls search every slice.  Both give the same result.  T[i] is point i's time (the LAS
 * intensity, las_reader.inl:59).  Slices t = cur_time + j (float), j =
 * -(int)r .. r: slice t's sites are the points with |t - T| <= r + 1e-6
 * (nni.cpp:770), each at squared distance |x - s|^2 + (|t - T| in metres,
 * rounded to the unit)^2, regions in the plane; the queries of slice j add
 * j^2 (nni.cpp:810); every cell sums over all slices and divides once
 * (nni.cpp:906,936).  Label outputs are per slice and not returned.  r = 0
 * is the 2-D tile of the points with T == cur_time. */
tnni_status tnni_tile_run_time(tnni_ctx *ctx, const tnni_tile_desc *desc, const int32_t *X,
                               const int32_t *Y, const int32_t *Z, const int32_t *T, uint64_t n,
                               float cur_time, float time_radius, float *out_h);
tnni_status tnni_tile_run_time_dev(tnni_ctx *ctx, const tnni_tile_desc *desc, const int32_t *dX,
                                   const int32_t *dY, const int32_t *dZ, const int32_t *dT,
                                   uint64_t n, float cur_time, float time_radius, float *d_out_h);

/* Space-time cache: performNNI with useDelta (the Delta algorithm,
 * nni.cpp:902-921) keeps the Voronoi diagram of each slice time it has drawn
 * and draws only the slices it does not hold (switchRenderbuffers,
 * nni.cpp:321-359, one buffer per slice of a call, the oldest reused,
 * getMinIndex nni.cpp:293-300).
 * Here a cache holds per-time layers instead: for each point time T, the
 * 2-D Voronoi of the points of time T (site value and squared distance of
 * every pixel, 8 bytes per pixel per layer).  A slice's nearest site is the
 * least (d2 + dt2) over the layers within its radius, so a slice is a
 * combine of layers plus the gather, and a layer serves every slice and
 * every later cur_time that reaches its time; only the layers a call lacks
 * are searched (one per call when cur_time steps by 1, or by 0.5).  The
 * result is bit for bit that of the uncached call.  The points must be
 * sorted by time (a call whose points are not, or that reaches more than 64
 * point times, runs without the cache, searching every slice); the caller
 * keeps the points of each time the
 * same, and in the same relative order, between the calls that use one
 * cache (a time-sorted stream read in windows around cur_time does); a call
 * with another tile descriptor empties the cache first.  A cache belongs to
 * one context and is used under its lock; destroy caches before their
 * context (a cache outliving it holds nothing and only accepts destroy). */
typedef struct tnni_time_cache tnni_time_cache;
/* slots <= 0: as many layers as a call reaches (9 at r = 2: the times
 * within 2r of cur_time); more slots keep layers longer (a call reaching
 * more layers than the slots still gets them).  At most 4096. */
tnni_status tnni_time_cache_create(tnni_ctx *ctx, int slots, tnni_time_cache **out);
void tnni_time_cache_destroy(tnni_time_cache *cache);
/* Empty the cache, keeping its device memory (nni::clearRenderbuffers,
 * nni.cpp:362-379). */
tnni_status tnni_time_cache_clear(tnni_time_cache *cache);
/* Layers taken from the cache and layers searched, since creation. */
tnni_status tnni_time_cache_stats(const tnni_time_cache *cache, uint64_t *hits, uint64_t *misses);
/* tnni_tile_run_time(_dev) through a cache (NULL: uncached). */
tnni_status tnni_tile_run_time_cached(tnni_ctx *ctx, tnni_time_cache *cache, const tnni_tile_desc *desc,
                                      const int32_t *X, const int32_t *Y, const int32_t *Z,
                                      const int32_t *T, uint64_t n, float cur_time, float time_radius,
                                      float *out_h);
tnni_status tnni_tile_run_time_cached_dev(tnni_ctx *ctx, tnni_time_cache *cache,
                                          const tnni_tile_desc *desc, const int32_t *dX,
                                          const int32_t *dY, const int32_t *dZ, const int32_t *dT,
                                          uint64_t n, float cur_time, float time_radius, float *d_out_h);

/* ---- autotuning --------------------------------------------------------
 * New (the reference has fixed GL raster sizes, nni.cpp:601-668): the bucket
 * side, the cell-tile side and the workgroups per CU of the tile kernel are
 * measured per shape class on sample tiles and kept in the context.  A shape
 * class groups tiles by algorithm, accumulation mode, scale, query span and
 * site radius in pixels, grid side (to a power of two), point density (to
 * half a power of two) and form (plain or batched/served).  Descriptors
 * whose bucket_px and tile_cells are 0 then use the tuned sizes of their
 * class; classes never tuned use the built-in choice. */
typedef struct tnni_tune_entry {
    uint64_t shape_class;
    int32_t bucket_px;   /* bucket side in pixels                           */
    int32_t tile_cells;  /* cell-tile side                                  */
    int32_t wg_per_cu;   /* workgroups per CU of the tile kernel (0 = all)  */
    int32_t candidates;  /* settings timed (tnni_autotune only)             */
    float ms_default;    /* sample time, built-in choice (tnni_autotune)    */
    float ms_tuned;      /* sample time, chosen setting (tnni_autotune)     */
    int32_t form;        /* TNNI_FORM_FUSED, _SPLIT or _SCATTER (0: built-in) */
    int32_t pad;
} tnni_tune_entry;

/* Shape class of a tile (batched != 0: the batched and served forms). */
tnni_status tnni_shape_class(const tnni_tile_desc *desc, uint64_t n_points, int batched,
                             uint64_t *shape_class);

/* Time candidate settings on device-resident sample tiles, grouped by shape
 * class, and keep the fastest per class.  batched != 0 tunes the batched and
 * served forms (each class's samples are replicated to about 2^25 pixels
 * and timed as one launch between HIP events); 0 tunes the plain form (one
 * synchronous launch per tile, host clock).  Each setting runs once to warm
 * up and then reps times (best kept).  Outputs go to scratch rasters the
 * context owns.  out (nullable) receives one entry per class, at most
 * n_tiles; *n_classes their number. */
tnni_status tnni_autotune(tnni_ctx *ctx, int n_tiles, const tnni_tile_desc *descs,
                          const int32_t *const *dX, const int32_t *const *dY,
                          const int32_t *const *dZ, const uint64_t *n, int batched, int reps,
                          tnni_tune_entry *out, int *n_classes);

/* The tuned table: read (up to cap entries; *n = entries held), load (for
 * example a table saved by an earlier run) and clear. */
tnni_status tnni_tune_read(tnni_ctx *ctx, tnni_tune_entry *out, int cap, int *n);
tnni_status tnni_tune_load(tnni_ctx *ctx, const tnni_tune_entry *in, int n);
tnni_status tnni_tune_clear(tnni_ctx *ctx);

/* ---- profiling -------------------------------------------------------- */
/* When enabled, HIP events bracket every launch of the main tile kernel on
 * the stream it runs on. */
tnni_status tnni_profile_enable(tnni_ctx *ctx, int enable);
tnni_status tnni_profile_read(tnni_ctx *ctx, uint64_t *launches,
                              double *total_ms, int reset);

#ifdef __cplusplus
}
#endif
#endif
