/* ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * A serial C restatement of rang-core's WSClean imaging weights and Stokes I
 * formation (core/src/weights.rs), the checker of the GPU weights: the
 * padded size (:36-41), the grid shape (:98-110), gridding in row and
 * channel order with float weights added into doubles (:131-184), the
 * Briggs / uniform / natural finish (:186-223), the rank filter (:225-265),
 * the lookup as floats (:267-304), Stokes I from linear correlations
 * (:507-550) and the header sums (:552-572). */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

size_t or_padded_size(size_t trimmed, double padding) {
    size_t padded = (size_t)ceil((double)trimmed * padding);
    padded += (4 - (padded % 4)) % 4;
    return padded;
}

int or_grid_shape(size_t width, size_t height, double super_weight, size_t *out) {
    if (!(super_weight > 0.0)) return 1;
    size_t w = (size_t)round((double)width / super_weight);
    size_t h = (size_t)round((double)height / super_weight);
    if (w % 2) w += 1;
    if (h % 2) h += 1;
    out[0] = w;
    out[1] = h;
    return 0;
}

typedef struct {
    size_t width, height;
    double sx, sy;
    double *grid;
    double total;
} wgrid;

static void uv_to_xy(const wgrid *g, double u, double v, long long *x, long long *y) {
    if (v < 0.0) {
        u = -u;
        v = -v;
    }
    *x = (long long)floor(u * (double)g->width * g->sx + (double)(g->width / 2));
    *y = (long long)floor(v * (double)g->height * g->sy);
}

static int within(const wgrid *g, long long x, long long y) {
    return x >= 0 && x < (long long)g->width && y < (long long)(g->height / 2);
}

static double window_mean(const wgrid *g, size_t x, size_t y, size_t window) {
    const size_t d = window / 2, half = g->height / 2;
    const size_t x1 = x > d ? x - d : 0, y1 = y > d ? y - d : 0;
    const size_t x2 = x + d >= g->width ? g->width : x + d;
    const size_t y2 = y + d >= half ? half : y + d;
    size_t count = 0;
    double sum = 0.0;
    for (size_t yi = y1; yi < y2; ++yi)
        for (size_t xi = x1; xi < x2; ++xi) {
            const double w = g->grid[yi * g->width + xi];
            if (w != 0.0) {
                count += 1;
                sum += w;
            }
        }
    return sum / (double)count;
}

/* cls 0 natural, 1 uniform, 2 Briggs(robust). out: nrow * nchan floats;
 * grid (may be NULL): the finished half-plane; sums (may be NULL): visibility
 * weight sum, total weight, max gridded weight, count. */
int or_imaging_weights(size_t nrow, size_t nchan, const double *uvw, const double *wavelength,
                       const float *weights, const uint8_t *selected, int cls, double robust,
                       size_t width, size_t height, double sx, double sy, double super_weight,
                       double rank_level, size_t rank_size, float *out, double *grid_out,
                       double *sums) {
    if (width == 0 || height == 0 || !(super_weight > 0.0) || sx <= 0.0 || sy <= 0.0) return 1;
    if (cls < 0 || cls > 2) return 1;
    size_t shape[2];
    or_grid_shape(width, height, super_weight, shape);
    wgrid g = {shape[0], shape[1], sx, sy, NULL, 0.0};
    const size_t ncells = g.width * g.height / 2;
    g.grid = calloc(ncells ? ncells : 1, sizeof(double));
    for (size_t row = 0; row < nrow; ++row) {
        if (!selected[row]) continue;
        double um = uvw[3 * row], vm = uvw[3 * row + 1];
        if (vm < 0.0) {
            um = -um;
            vm = -vm;
        }
        for (size_t ch = 0; ch < nchan; ++ch) {
            long long x, y;
            uv_to_xy(&g, um / wavelength[ch], vm / wavelength[ch], &x, &y);
            if (within(&g, x, y)) {
                const double w = (double)weights[row * nchan + ch];
                g.grid[(size_t)x + (size_t)y * g.width] += w;
                g.total += w;
            }
        }
    }
    if (cls == 2) {
        double avg = 0.0;
        for (size_t i = 0; i < ncells; ++i) avg += g.grid[i] * g.grid[i];
        avg /= g.total;
        const double num = 5.0 * exp(-log(10.0) * robust);
        const double s_sq = num * num / avg;
        for (size_t i = 0; i < ncells; ++i)
            if (g.grid[i] != 0.0) g.grid[i] = 1.0 / (1.0 + g.grid[i] * s_sq);
    } else if (cls == 1) {
        for (size_t i = 0; i < ncells; ++i)
            if (g.grid[i] != 0.0) g.grid[i] = 1.0 / g.grid[i];
    } else {
        for (size_t i = 0; i < ncells; ++i)
            if (g.grid[i] != 0.0) g.grid[i] = 1.0;
    }
    if (rank_level >= 1.0) {
        double *ng = malloc(sizeof(double) * (ncells ? ncells : 1));
        memcpy(ng, g.grid, sizeof(double) * ncells);
        for (size_t y = 0; y < g.height / 2; ++y)
            for (size_t x = 0; x < g.width; ++x) {
                const double w = g.grid[y * g.width + x];
                if (w != 0.0) {
                    const double mean = window_mean(&g, x, y, rank_size);
                    if (w > mean * rank_level) ng[y * g.width + x] = mean * rank_level;
                }
            }
        free(g.grid);
        g.grid = ng;
    }
    for (size_t row = 0; row < nrow; ++row)
        for (size_t ch = 0; ch < nchan; ++ch) {
            float v = 0.0f;
            if (selected[row]) {
                long long x, y;
                uv_to_xy(&g, uvw[3 * row] / wavelength[ch], uvw[3 * row + 1] / wavelength[ch], &x, &y);
                v = within(&g, x, y) ? (float)g.grid[(size_t)x + (size_t)y * g.width] : 0.0f;
            }
            out[row * nchan + ch] = v;
        }
    if (grid_out) memcpy(grid_out, g.grid, sizeof(double) * ncells);
    if (sums) {
        double vsum = 0.0, total = 0.0, mx = 0.0;
        uint64_t count = 0;
        for (size_t i = 0; i < nrow * nchan; ++i) {
            const float cum = weights[i] * out[i];
            if (cum != 0.0f) {
                vsum += (double)weights[i];
                count += 1;
            }
            mx = fmax(mx, (double)cum);
            total += (double)cum;
        }
        sums[0] = vsum;
        sums[1] = total;
        sums[2] = mx;
        sums[3] = (double)count;
    }
    free(g.grid);
    return 0;
}

static int finite2(float re, float im) { return isfinite(re) && isfinite(im); }

/* data: ncell * 4 interleaved complex floats (XX, XY, YX, YY); weights and
 * flags ncell * 4; vis: ncell interleaved complex floats; wout: ncell. */
int or_stokes_i(size_t ncell, const float *data, const float *weights, const uint8_t *flags, float *vis,
                float *wout) {
    for (size_t c = 0; c < ncell; ++c) {
        const float *d = data + 8 * c;
        const float xr = d[0], xi = d[1], yr = d[6], yi = d[7];
        const float re = (yr + xr) * 0.5f, im = (yi + xi) * 0.5f;
        vis[2 * c] = finite2(re, im) ? re : 0.0f;
        vis[2 * c + 1] = finite2(re, im) ? im : 0.0f;
        const float *w = weights + 4 * c;
        const uint8_t *f = flags + 4 * c;
        float weight = (f[0] == 0 && finite2(xr, xi)) ? w[0] * 4.0f : 0.0f;
        if (f[3] == 0 && finite2(yr, yi)) {
            const float yy = w[3] * 4.0f;
            weight = yy < weight ? yy : weight;
        } else {
            weight = 0.0f;
        }
        wout[c] = weight;
    }
    return 0;
}
