/*
 * CPU oracle for the discrete-Sibson tile path -- TEST INFRASTRUCTURE ONLY.
 * See nni_oracle.h for what is restated from which reference lines.
 * Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg use it.
 */
#include "nni_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define NOD2 0xFFFFFFFFu

static int64_t floordiv64(int64_t a, int64_t b) /* b > 0 */
{
    int64_t q = a / b;
    if ((a % b) != 0 && (a < 0)) q--;
    return q;
}
static int64_t ceildiv64(int64_t a, int64_t b) { return -floordiv64(-a, b); }
static int64_t imax64(int64_t a, int64_t b) { return a > b ? a : b; }
static int64_t imin64(int64_t a, int64_t b) { return a < b ? a : b; }

/* Largest r >= 0 with r*r <= v. */
static int64_t isqrt64(int64_t v)
{
    if (v <= 0) return 0;
    int64_t r = (int64_t)sqrt((double)v);
    while (r * r > v) r--;
    while ((r + 1) * (r + 1) <= v) r++;
    return r;
}

int orc_frame_compute(const orc_desc *d, orc_frame *f)
{
    memset(f, 0, sizeof *f);
    if (d->algorithm < 0 || d->algorithm > 1 || d->scale < 1 || d->nrows < 1 || d->ncols < 1 || !(d->cell_size > 0.0f) ||
        !(d->site_radius > 0.0f) || !(d->query_radius > 0.0f))
        return 1;
    double cs = (double)d->cell_size;     /* query_spacing, nni.cpp:572 */
    double qr = (double)d->query_radius;  /* float query_radius, nni.cpp:69,575 */
    double sr = (double)d->site_radius;   /* float site_radius, nni.cpp:68,574 */
    double ps = cs / d->scale;            /* pixel_size, nni.cpp:881 */
    int po = d->pixel_offset >= 0
                 ? d->pixel_offset
                 : (int)ceil((qr - floor(d->scale / 2.0) * ps) / ps) + 1; /* nni.cpp:883 */
    if (po < 0) po = 0;
    f->po = po;
    f->algorithm = d->algorithm;
    f->W = d->ncols * d->scale + 2 * po; /* nni.cpp:885-886 */
    f->H = d->nrows * d->scale + 2 * po;
    f->spacing_factor = (int)ceil((qr * 2 + 0.001) / (cs * 5)); /* nni.cpp:879 */

    /* the smallest unit of 1/(100 k) metre that puts pixel centres,
     * queries and cell edges on integers */
    for (int t = 1; t <= 200; ++t) {
        int upm = 100 * t;
        if (d->units_per_m > 0) {
            if (t > 1) break;
            upm = d->units_per_m;
        }
        double c = cs * upm;
        int64_t ci = llround(c);
        if (ci <= 0 || fabs(c - (double)ci) > 1e-3) continue;
        if (ci % d->scale) continue;
        int64_t ps_u = ci / d->scale;
        /* Pixel k's centre is -cs/2 - po*ps + (k + 1/2)*ps (glOrtho of
         * nni.cpp:656-660); it must land on an integer unit. */
        int64_t x0n = (int64_t)(1 - 2 * po - d->scale) * ps_u;
        if (x0n % 2) continue;
        int64_t Rs = (int64_t)floor(sr * upm / sqrt(2.0)); /* nni.cpp:183 */
        int64_t Rq = (int64_t)floor(qr * upm / sqrt(2.0));
        double srd = sr * upm, qrd = qr * upm;
        int64_t Rs2 = (int64_t)floor(srd * srd);
        int64_t Rq2 = (int64_t)floor(qrd * qrd);
        int64_t Rsb = d->algorithm == 0 ? Rs : isqrt64(Rs2);
        int64_t Rqb = d->algorithm == 0 ? Rq : isqrt64(Rq2);
        /* squared distances inside a region must fit in 32 bits */
        if (d->algorithm == 0 ? (2 * Rs * Rs >= (int64_t)NOD2) : (Rs2 >= (int64_t)NOD2)) continue;
        if (2 * Rqb * Rqb >= (int64_t)NOD2) continue; /* query offsets square into 32 bits */
        int64_t ext = (int64_t)(f->W > f->H ? f->W : f->H) * ps_u + 2 * Rsb + 2 * Rqb;
        if (ext >= ((int64_t)1 << 30)) continue;
        f->upm = upm;
        f->ps_u = (int)ps_u;
        f->cs_u = (int)ci;
        f->x0 = f->y0 = (int)(x0n / 2);
        f->Rs = (int)Rs;
        f->Rq = (int)Rq;
        f->Rs2 = Rs2;
        f->Rq2 = Rq2;
        f->Rsb = (int)Rsb;
        f->Rqb = (int)Rqb;
        return 0;
    }
    return 2;
}

void orc_points_to_units(const orc_desc *d, const orc_frame *f,
                         const int32_t *X, const int32_t *Y, int64_t n,
                         int32_t *xu, int32_t *yu, uint8_t *keep)
{
    const double lo_x = (double)f->x0 - f->Rsb;
    const double hi_x = (double)f->x0 + (double)(f->W - 1) * f->ps_u + f->Rsb;
    const double lo_y = (double)f->y0 - f->Rsb;
    const double hi_y = (double)f->y0 + (double)(f->H - 1) * f->ps_u + f->Rsb;
    for (int64_t i = 0; i < n; ++i) {
        /* LAS de-scale (las_reader.inl:59) then the tile-local shift of
         * drawPlane's x_offset (nni.cpp:189-190, 546), all in FP64. */
        double xm = fma((double)X[i], d->las_scale[0], d->las_offset[0]) - d->origin_x;
        double ym = fma((double)Y[i], d->las_scale[1], d->las_offset[1]) - d->origin_y;
        double xr = rint(xm * f->upm);
        double yr = rint(ym * f->upm);
        int k = xr >= lo_x && xr <= hi_x && yr >= lo_y && yr <= hi_y;
        keep[i] = (uint8_t)k;
        xu[i] = k ? (int32_t)xr : 0;
        yu[i] = k ? (int32_t)yr : 0;
    }
}

/* Does the region of the site at (sx,sy) cover the pixel centre (px,py)?
 * Planes: square of half-side r/sqrt(2) (nni.cpp:183,207-225).
 * Cones: disk of radius r (nni.cpp:117-157). */
static int covers(const orc_frame *f, int alg, int64_t dx, int64_t dy, int site)
{
    if (alg == 0) {
        int64_t R = site ? f->Rs : f->Rq;
        return dx <= R && dx >= -R && dy <= R && dy >= -R;
    }
    return dx * dx + dy * dy <= (site ? f->Rs2 : f->Rq2);
}

/* Site-major pass.  Writes only rows [l_lo, l_hi). */
static void raster_band(const orc_frame *f, int alg, const int32_t *xu, const int32_t *yu,
                        const uint8_t *keep, const uint32_t *dt2, const int64_t *ids, int64_t nid,
                        int l_lo, int l_hi, uint32_t *d2, int32_t *label)
{
    const int64_t ps = f->ps_u;
    for (int64_t t = 0; t < nid; ++t) {
        int64_t i = ids ? ids[t] : t;
        if (!keep[i]) continue;
        int64_t sx = xu[i], sy = yu[i];
        int64_t k0 = imax64(0, ceildiv64(sx - f->Rsb - f->x0, ps));
        int64_t k1 = imin64(f->W - 1, floordiv64(sx + f->Rsb - f->x0, ps));
        int64_t l0 = imax64(l_lo, ceildiv64(sy - f->Rsb - f->y0, ps));
        int64_t l1 = imin64(l_hi - 1, floordiv64(sy + f->Rsb - f->y0, ps));
        for (int64_t l = l0; l <= l1; ++l) {
            int64_t dy = f->y0 + l * ps - sy;
            for (int64_t k = k0; k <= k1; ++k) {
                int64_t dx = f->x0 + k * ps - sx;
                if (!covers(f, alg, dx, dy, 1)) continue;
                uint32_t dd = (uint32_t)(dx * dx + dy * dy + (dt2 ? (int64_t)dt2[i] : 0));
                size_t p = (size_t)l * f->W + k;
                /* depth test: nearer wins; equal depth keeps the lower index */
                if (dd < d2[p] || (dd == d2[p] && (int32_t)i < label[p])) {
                    d2[p] = dd;
                    label[p] = (int32_t)i;
                }
            }
        }
    }
}

void orc_labels_raster(const orc_frame *f, const int32_t *xu, const int32_t *yu,
                       const uint8_t *keep, int64_t n,
                       uint32_t *d2, int32_t *label, int nthreads)
{
    orc_labels_raster_t(f, xu, yu, keep, NULL, n, d2, label, nthreads);
}

void orc_labels_raster_t(const orc_frame *f, const int32_t *xu, const int32_t *yu,
                         const uint8_t *keep, const uint32_t *dt2, int64_t n,
                         uint32_t *d2, int32_t *label, int nthreads)
{
    const int alg = f->algorithm;
    const size_t P = (size_t)f->W * f->H;
    for (size_t p = 0; p < P; ++p) { d2[p] = NOD2; label[p] = -1; }
    if (n <= 0) return;

    /* Bands of pixel rows, processed independently.  Each site is listed
     * under the band of its first row; band b visits the lists of the bands
     * its sites can reach from above. */
    const int band_h = 32;
    const int nb = (f->H + band_h - 1) / band_h;
    const int64_t span_rows = (2 * (int64_t)f->Rsb) / f->ps_u + 1;
    const int reach = (int)((span_rows + band_h - 1) / band_h);
    int64_t *cnt = (int64_t *)calloc((size_t)nb + 1, sizeof(int64_t));
    int64_t *ids = (int64_t *)malloc((size_t)(n > 0 ? n : 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        if (!keep[i]) continue;
        int64_t l0 = imax64(0, ceildiv64((int64_t)yu[i] - f->Rsb - f->y0, f->ps_u));
        if (l0 >= f->H) continue;
        cnt[l0 / band_h + 1]++;
    }
    for (int b = 0; b < nb; ++b) cnt[b + 1] += cnt[b];
    int64_t *cur = (int64_t *)malloc((size_t)(nb + 1) * sizeof(int64_t));
    memcpy(cur, cnt, (size_t)(nb + 1) * sizeof(int64_t));
    for (int64_t i = 0; i < n; ++i) {
        if (!keep[i]) continue;
        int64_t l0 = imax64(0, ceildiv64((int64_t)yu[i] - f->Rsb - f->y0, f->ps_u));
        if (l0 >= f->H) continue;
        ids[cur[l0 / band_h]++] = i;
    }
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 1)
#endif
    for (int b = 0; b < nb; ++b) {
        int l_lo = b * band_h;
        int l_hi = l_lo + band_h < f->H ? l_lo + band_h : f->H;
        int bfirst = b - reach < 0 ? 0 : b - reach;
        raster_band(f, alg, xu, yu, keep, dt2, ids + cnt[bfirst], cnt[b + 1] - cnt[bfirst],
                    l_lo, l_hi, d2, label);
    }
    free(cur);
    free(ids);
    free(cnt);
}

void orc_labels_brute(const orc_frame *f, const int32_t *xu, const int32_t *yu,
                      const uint8_t *keep, int64_t n,
                      uint32_t *d2, int32_t *label)
{
    orc_labels_brute_t(f, xu, yu, keep, NULL, n, d2, label);
}

void orc_labels_brute_t(const orc_frame *f, const int32_t *xu, const int32_t *yu,
                        const uint8_t *keep, const uint32_t *dt2, int64_t n,
                        uint32_t *d2, int32_t *label)
{
    const int alg = f->algorithm;
    for (int64_t l = 0; l < f->H; ++l) {
        for (int64_t k = 0; k < f->W; ++k) {
            int64_t px = f->x0 + k * f->ps_u, py = f->y0 + l * f->ps_u;
            uint32_t best = NOD2;
            int32_t bi = -1;
            for (int64_t i = 0; i < n; ++i) {
                if (!keep[i]) continue;
                int64_t dx = px - xu[i], dy = py - yu[i];
                if (!covers(f, alg, dx, dy, 1)) continue;
                uint32_t dd = (uint32_t)(dx * dx + dy * dy + (dt2 ? (int64_t)dt2[i] : 0));
                if (dd < best) { best = dd; bi = (int32_t)i; }
            }
            d2[l * f->W + k] = best;
            label[l * f->W + k] = bi;
        }
    }
}

uint32_t orc_s0_value(const orc_desc *d, int32_t Z)
{
    /* uint index = floor((pnt.p[2] + zOffset + 1) * zMultiplier), nni.cpp:539,
     * zOffset = -pointMin (nni.cpp:577) */
    double z = fma((double)Z, d->las_scale[2], d->las_offset[2]);
    int zOffset = -d->point_min;
    double v = floor((z + zOffset + 1) * (double)d->multiplier);
    return (uint32_t)(int64_t)v;
}

void orc_slice_sites(const orc_frame *f, const int32_t *T, const uint8_t *keep, int64_t n,
                     float t, float time_radius, uint8_t *keep_slice, uint32_t *dt2)
{
    for (int64_t i = 0; i < n; ++i) {
        /* abs(t - p.t) <= timeRadius + 0.000001 (nni.cpp:770): float
         * difference, compared in double */
        float dt = fabsf(t - (float)T[i]);
        int in = keep[i] && (double)dt <= (double)time_radius + 0.000001;
        keep_slice[i] = (uint8_t)in;
        /* addPoint(wp, p, abs(t - p.t)) -> timeOffset^2 (nni.cpp:197,779) */
        uint32_t u = in ? (uint32_t)rint((double)dt * (double)f->upm) : 0u;
        dt2[i] = u * u;
    }
}

/* One cell's divide (orc_heights' outputs, any may be NULL). */
static void finish_cell(const orc_desc *d, const orc_acc *a, size_t o, float *h_fixed,
                        double *h64, float *h_s0, uint32_t *count)
{
    const double sz = d->las_scale[2], oz = d->las_offset[2];
    if (count) count[o] = a->cnt;
    if (h_fixed)
        h_fixed[o] = a->cnt ? (float)fma((double)a->sumZ / (double)a->cnt, sz, oz) : -9999.0f;
    if (h64) h64[o] = a->cnt ? a->sum64 / (double)a->cnt : -9999.0;
    if (h_s0) {
        /* divide, cuda_nni.cu:497-500, then de-scale nni.cpp:841-846 */
        int32_t res;
        if (a->denom0 == 0)
            res = (int32_t)((uint32_t)(-9999) * (uint32_t)d->multiplier);
        else
            res = (int32_t)a->numer0 / (int32_t)a->denom0;
        double z = 1.0 * res / d->multiplier;
        if (z != -9999) z = z - (-d->point_min) - 1;
        h_s0[o] = (float)z;
    }
}

/* The query pass over every cell, query-major.  acc (nullable): the cells'
 * running sums, read and written back; without it each cell starts from
 * zero and is divided at once into the outputs. */
static void heights_cells(const orc_desc *d, const orc_frame *f, const uint32_t *d2,
                          const int32_t *label, const int32_t *Z, int64_t jj2, orc_acc *acc,
                          int want_s0, float *h_fixed, double *h64, float *h_s0, uint32_t *count,
                          uint64_t *pairs, int nthreads)
{
    const int alg = f->algorithm;
    const int64_t ps = f->ps_u;
    const double sz = d->las_scale[2], oz = d->las_offset[2];
    uint64_t K = 0, S = 0;
    (void)nthreads;
#ifdef _OPENMP
    if (nthreads > 0) omp_set_num_threads(nthreads);
#pragma omp parallel for schedule(dynamic, 4) reduction(+ : K, S)
#endif
    for (int j = 0; j < d->nrows; ++j) {
        for (int i = 0; i < d->ncols; ++i) {
            /* query (i,j) sits at (i*cs, j*cs), nni.cpp:488 */
            int64_t qx = (int64_t)i * f->cs_u, qy = (int64_t)j * f->cs_u;
            int64_t k0 = imax64(0, ceildiv64(qx - f->Rqb - f->x0, ps));
            int64_t k1 = imin64(f->W - 1, floordiv64(qx + f->Rqb - f->x0, ps));
            int64_t l0 = imax64(0, ceildiv64(qy - f->Rqb - f->y0, ps));
            int64_t l1 = imin64(f->H - 1, floordiv64(qy + f->Rqb - f->y0, ps));
            const size_t o = (size_t)j * d->ncols + i;
            orc_acc a;
            if (acc) a = acc[o];
            else memset(&a, 0, sizeof a);
            for (int64_t l = l0; l <= l1; ++l) {
                int64_t dy = f->y0 + l * ps - qy;
                for (int64_t k = k0; k <= k1; ++k) {
                    size_t p = (size_t)l * f->W + k;
                    int32_t s = label[p];
                    if (s < 0) continue;
                    int64_t dx = f->x0 + k * ps - qx;
                    if (!covers(f, alg, dx, dy, 0)) continue;
                    K++;
                    /* query depth <= site depth (GL_LEQUAL, nni.cpp:804); a
                     * space-time query at time offset j adds j^2 (nni.cpp:810,
                     * 197) */
                    if ((uint64_t)(dx * dx + dy * dy + jj2) > (uint64_t)d2[p]) continue;
                    S++;
                    a.cnt++;
                    a.sumZ += Z[s];
                    a.sum64 += fma((double)Z[s], sz, oz);
                    if (want_s0) {
                        uint32_t v = orc_s0_value(d, Z[s]);
                        if (v > 0) { a.numer0 += v; a.denom0++; } /* cuda_nni.cu:423,447-448 */
                    }
                }
            }
            if (acc) acc[o] = a;
            else finish_cell(d, &a, o, h_fixed, h64, h_s0, count);
        }
    }
    if (pairs) { pairs[0] += K; pairs[1] += S; }
}

void orc_heights_accumulate(const orc_desc *d, const orc_frame *f,
                            const uint32_t *d2, const int32_t *label, const int32_t *Z,
                            int64_t jj2, orc_acc *acc, uint64_t *pairs, int nthreads)
{
    heights_cells(d, f, d2, label, Z, jj2, acc, 1, NULL, NULL, NULL, NULL, pairs, nthreads);
}

void orc_heights_finish(const orc_desc *d, const orc_acc *acc, float *h_fixed, double *h64,
                        float *h_s0, uint32_t *count)
{
    const size_t C = (size_t)d->nrows * d->ncols;
    for (size_t o = 0; o < C; ++o) finish_cell(d, &acc[o], o, h_fixed, h64, h_s0, count);
}

void orc_heights(const orc_desc *d, const orc_frame *f,
                 const uint32_t *d2, const int32_t *label, const int32_t *Z,
                 float *h_fixed, double *h64, float *h_s0, uint32_t *count,
                 uint64_t *pairs, int nthreads)
{
    if (pairs) { pairs[0] = 0; pairs[1] = 0; }
    heights_cells(d, f, d2, label, Z, 0, NULL, h_s0 != NULL, h_fixed, h64, h_s0, count, pairs,
                  nthreads);
}
