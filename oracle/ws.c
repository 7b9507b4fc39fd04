/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of rang-core's w-stacking gridder (core/src/grid.rs, a
 * port of WSClean 3.6's inversion), the checker of the HIP w-stacking path.
 * Only tests/ and __graft_entry__.smoke() load it; the product library never
 * links or calls it.
 *
 * Double precision throughout, serial, in the reference's orders: samples
 * placed in row-channel order and gridded in that order per layer; each
 * layer's grid transformed by an unnormalised backward 2-D DFT (rows, then
 * columns) and projected; the projected terms added in layer order. The FFT
 * here is this file's own mixed-radix transform (any length; factors above
 * 7 by a direct sum), not rustfft, so results agree to rounding.
 */
#include <complex.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define WS_C_LIGHT 299792458.0
#define WS_KB_ALPHA 8.6
#define WS_N_STATS 8

/* Same layouts as the reference's #[repr(C)] GridInputs / GridOutputs
 * (core/src/ffi_grid.rs:7-47). */
typedef struct {
    size_t nrow, nchan, width, height, kernel_size, oversampling, n_w_layers;
    double pixel_size_x, pixel_size_y, padding;
    const double *uvw, *frequencies;
    const float *data, *weights, *imaging_weights;
} ws_inputs;

typedef struct {
    double *dirty, *psf, *stats;
} ws_outputs;

/* bessel0 (grid.rs:148-163): I0 by its series, stopped at relative
 * `precision`. */
static double ws_bessel0(double x, double precision) {
    double d = 0.0, ds = 1.0, sum = 1.0;
    for (;;) {
        d += 2.0;
        ds *= x * x / (d * d);
        sum += ds;
        if (ds <= sum * precision) break;
    }
    return sum;
}

/* kaiser_bessel_kernel (grid.rs:165-188): the windowed sinc of
 * kernel_size * oversampling samples, symmetric about n/2. */
int ws_kernel(size_t ks, size_t os, double *out) {
    const size_t n = ks * os, mid = n / 2;
    if (n == 0 || !out) return 1;
    const double fr = 1.0 / (double)os;
    const double norm = (double)os / ws_bessel0(WS_KB_ALPHA, 1e-8);
    for (size_t i = 0; i <= mid && mid + i < n; ++i) {
        const double x = (double)i;
        const double s = i == 0 ? fr : sin(M_PI * fr * x) / (M_PI * x);
        const double t = (double)i / (double)mid;
        out[mid + i] = s * ws_bessel0(WS_KB_ALPHA * sqrt(1.0 - t * t), 1e-8) * norm;
    }
    for (size_t i = 0; i < mid; ++i) out[i] = out[n - 1 - i];
    return 0;
}

/* ---------------------------------------------------------------- FFT */

/* Unnormalised inverse DFT (exp(+2 pi i j k / n)) of `x` (length n, stride
 * 1) by recursive decimation in time over the smallest factor; factors of 2
 * and 3 by butterflies, others by the direct sum. `tmp` holds n values. */
static void ws_fft_rec(double complex *x, size_t n, size_t stride, double complex *out,
                       double complex *tmp) {
    if (n == 1) {
        out[0] = x[0];
        return;
    }
    size_t p = 2;
    while (n % p) ++p;
    const size_t m = n / p;
    /* sub-transforms of the p decimated sequences into tmp */
    for (size_t r = 0; r < p; ++r) ws_fft_rec(x + r * stride, m, stride * p, tmp + r * m, out);
    for (size_t k = 0; k < m; ++k) {
        for (size_t q = 0; q < p; ++q) {
            double complex acc = 0;
            for (size_t r = 0; r < p; ++r) {
                /* exp(2 pi i r (k + q m) / n), the exponent reduced mod n */
                const size_t e = (r * (k + q * m)) % n;
                const double ang = 2.0 * M_PI * (double)e / (double)n;
                acc += tmp[r * m + k] * (cos(ang) + I * sin(ang));
            }
            out[k + q * m] = acc;
        }
    }
}

static void ws_fft(double complex *x, size_t n, double complex *work) {
    /* work: 2n values */
    double complex *in = work, *tmp = work + n;
    memcpy(in, x, n * sizeof(double complex));
    ws_fft_rec(in, n, 1, x, tmp);
}

/* Fft2::backward (grid.rs:252-261): every row, then every column. */
static void ws_fft2(double complex *g, size_t w, size_t h) {
    const size_t nmax = w > h ? w : h;
    double complex *work = malloc(sizeof(double complex) * 3 * nmax);
    double complex *col = work + 2 * nmax;
    for (size_t y = 0; y < h; ++y) ws_fft(g + y * w, w, work);
    for (size_t x = 0; x < w; ++x) {
        for (size_t y = 0; y < h; ++y) col[y] = g[x + y * w];
        ws_fft(col, h, work);
        for (size_t y = 0; y < h; ++y) g[x + y * w] = col[y];
    }
    free(work);
}

/* Exported for the FFT's own test. */
void ws_fft2_backward(double *re_im, size_t w, size_t h) { ws_fft2((double complex *)re_im, w, h); }

/* ---------------------------------------------------------------- gridder */

typedef struct {
    size_t w, h, ks, os, nl;
    double psx, psy, min_w, max_w;
} ws_geo;

/* padded_size (grid.rs:80-89). */
static size_t ws_pad(size_t n, double padding) {
    size_t p = (size_t)ceil((double)n * padding);
    return p + (4 - p % 4) % 4;
}

/* w_to_layer (grid.rs:334-347): -1 where WSClean's size_t would wrap. */
static long ws_layer(const ws_geo *g, double w) {
    if (g->nl == 1) return 0;
    const double v = round((fabs(w) - g->min_w) * (double)(g->nl - 1) / (g->max_w - g->min_w));
    return v < 0.0 ? -1 : (long)v;
}

static double ws_layer_w(const ws_geo *g, size_t layer) {
    if (g->nl == 1) return 0.0;
    return (double)layer * (g->max_w - g->min_w) / (double)(g->nl - 1) + g->min_w;
}

/* sqrt(1 - l^2 - m^2) - 1 of grid index (x, y) (grid.rs:364-391). */
static double ws_sqrt_lm(const ws_geo *g, size_t x, size_t y) {
    size_t ys = (g->h - y) + g->h / 2;
    if (ys >= g->h) ys -= g->h;
    const double m = ((double)ys - (double)(g->h / 2)) * g->psy;
    size_t xs = x + g->w / 2;
    if (xs >= g->w) xs -= g->w;
    const double l = ((double)(g->w / 2) - (double)xs) * g->psx;
    return l * l + m * m < 1.0 ? sqrt(1.0 - l * l - m * m) - 1.0 : 0.0;
}

/* dct3_prefix (grid.rs:263-279). */
static void ws_dct3(const double *in, size_t len, size_t n, size_t count, double *out) {
    for (size_t k = 0; k < count; ++k) {
        const double kk = (double)k + 0.5;
        double s = in[0];
        for (size_t j = 1; j < len; ++j) s += 2.0 * in[j] * cos(M_PI * (double)j * kk / (double)n);
        out[k] = s;
    }
}

typedef struct {
    int32_t x, y;
    uint16_t xk, yk;
    float re, im, weight;
} ws_sample;

/* The inversion (grid.rs:652-885): limits, placement, per-layer gridding,
 * transforms and projection, normalisation, kernel correction, trim and
 * PSF normalisation. Returns 0, or 1 with invalid arguments. */
int ws_invert(const ws_inputs *in, ws_outputs *out) {
    if (!in || !out || !in->uvw || !in->frequencies || !in->data || !in->weights ||
        !in->imaging_weights || !out->dirty || !out->psf || !out->stats)
        return 1;
    if (in->width == 0 || in->height == 0 || in->kernel_size == 0 || in->oversampling == 0 ||
        in->n_w_layers == 0 || in->kernel_size % 2 == 0 || isnan(in->padding) || in->padding < 1.0 ||
        in->oversampling > 65535)
        return 1;
    const size_t nrow = in->nrow, nchan = in->nchan;
    ws_geo g = {ws_pad(in->width, in->padding), ws_pad(in->height, in->padding), in->kernel_size,
                in->oversampling, in->n_w_layers, in->pixel_size_x, in->pixel_size_y, 0, 0};
    /* w_limits (grid.rs:652-693), both bounds from u as in WSClean 3.6 */
    const double hw = 0.5 * (double)g.w, hh = 0.5 * (double)g.h;
    double min_w = 1e100, max_w = 0.0;
    for (size_t r = 0; r < nrow; ++r)
        for (size_t c = 0; c < nchan; ++c) {
            const size_t i = r * nchan + c;
            if (in->weights[i] == 0.0f) continue;
            const double lam = WS_C_LIGHT / in->frequencies[c];
            const double u = in->uvw[3 * r] / lam, w = in->uvw[3 * r + 2] / lam;
            const double x = u * g.psx * (double)g.w, y = u * g.psy * (double)g.h;
            if (in->imaging_weights[i] != 0.0f && floor(x) > -hw && ceil(x) < hw && floor(y) > -hh &&
                ceil(y) < hh) {
                max_w = fmax(max_w, fabs(w));
                min_w = fmin(min_w, fabs(w));
            }
        }
    if (min_w == 1e100) min_w = max_w = 0.0;
    g.min_w = min_w;
    g.max_w = min_w == max_w ? max_w + 1.0 : max_w;
    /* placement (grid.rs:757-810, 393-448) */
    ws_sample *smp = malloc(sizeof(ws_sample) * (nrow * nchan + 1));
    uint32_t *lay = malloc(sizeof(uint32_t) * (nrow * nchan + 1));
    size_t *cnt = calloc(g.nl + 1, sizeof(size_t));
    size_t ns = 0;
    double image_weight = 0, vsum = 0, maxw = 0;
    uint64_t gridded = 0;
    const long W = (long)g.w, H = (long)g.h, OS = (long)g.os;
    for (size_t r = 0; r < nrow; ++r)
        for (size_t c = 0; c < nchan; ++c) {
            const size_t i = r * nchan + c;
            const double lam = WS_C_LIGHT / in->frequencies[c];
            double u = in->uvw[3 * r] / lam, v = in->uvw[3 * r + 1] / lam;
            const double w = in->uvw[3 * r + 2] / lam;
            const long l0 = ws_layer(&g, w);
            const int inrange = l0 >= 0 && (size_t)l0 < g.nl;
            const float weight = inrange ? in->weights[i] : 0.0f;
            const float cw = weight * in->imaging_weights[i];
            if (cw != 0.0f) {
                vsum += (double)weight;
                gridded += 1;
            }
            maxw = fmax(maxw, (double)cw);
            image_weight += (double)cw;
            const float re = in->data[2 * i] * cw;
            float im = in->data[2 * i + 1] * cw;
            if (cw == 0.0f && re == 0.0f && im == 0.0f) continue;
            if (!inrange) continue;
            if (w < 0.0) {
                u = -u;
                v = -v;
                im = -im;
            }
            const double xe = u * g.psx * (double)g.w, ye = v * g.psy * (double)g.h;
            long x = (long)round(xe), y = (long)round(ye);
            long xk = (long)round((xe - (double)x) * (double)g.os);
            long yk = (long)round((ye - (double)y) * (double)g.os);
            xk = (xk + (OS * 3) / 2) % OS;
            yk = (yk + (OS * 3) / 2) % OS;
            if (!(x > -W / 2 && y > -H / 2 && x <= W / 2 && y <= H / 2)) continue;
            if (x < 0) x += W;
            if (y < 0) y += H;
            smp[ns] = (ws_sample){(int32_t)x, (int32_t)y, (uint16_t)xk, (uint16_t)yk, re, im, cw};
            lay[ns] = (uint32_t)l0;
            cnt[l0 + 1] += 1;
            ++ns;
        }
    int status = 0;
    if (image_weight == 0.0) {
        status = 1;
        goto done;
    }
    for (size_t l = 0; l < g.nl; ++l) cnt[l + 1] += cnt[l];
    {
        /* stable counting sort by layer */
        ws_sample *sorted = malloc(sizeof(ws_sample) * (ns + 1));
        size_t *cur = malloc(sizeof(size_t) * (g.nl + 1));
        memcpy(cur, cnt, sizeof(size_t) * (g.nl + 1));
        for (size_t s = 0; s < ns; ++s) sorted[cur[lay[s]]++] = smp[s];
        free(cur);
        /* kernels (grid.rs:190-202), reversed sub-cell order */
        const size_t ks = g.ks, os = g.os, mid = ks / 2;
        double *k1 = malloc(sizeof(double) * ks * os), *kern = malloc(sizeof(double) * ks * os);
        ws_kernel(ks, os, k1);
        for (size_t i = 0; i < os; ++i)
            for (size_t x = 0; x < ks; ++x) kern[(os - i - 1) * ks + x] = k1[x * os + i];
        const size_t npix = g.w * g.h;
        double complex *gd = malloc(sizeof(double complex) * npix);
        double complex *gp = malloc(sizeof(double complex) * npix);
        double *sd = calloc(npix, sizeof(double)), *sp = calloc(npix, sizeof(double));
        for (size_t l = 0; l < g.nl; ++l) {
            if (cnt[l + 1] == cnt[l]) continue;
            memset(gd, 0, sizeof(double complex) * npix);
            memset(gp, 0, sizeof(double complex) * npix);
            /* grid_layer (grid.rs:450-495) */
            for (size_t s = cnt[l]; s < cnt[l + 1]; ++s) {
                const ws_sample *q = &sorted[s];
                const double *xk = kern + (size_t)q->xk * ks, *yk = kern + (size_t)q->yk * ks;
                for (size_t j = 0; j < ks; ++j) {
                    const long cy = (((long)q->y + (long)j + H - (long)mid) % H) * W;
                    for (size_t i = 0; i < ks; ++i) {
                        const long cx = ((long)q->x + (long)i + W - (long)mid) % W;
                        const double k = yk[j] * xk[i];
                        gd[cx + cy] += (double)q->re * k + I * ((double)q->im * k);
                        gp[cx + cy] += (double)q->weight * k;
                    }
                }
            }
            ws_fft2(gd, g.w, g.h);
            ws_fft2(gp, g.w, g.h);
            /* project (grid.rs:497-529), the term added in layer order */
            const double tpw = -2.0 * M_PI * ws_layer_w(&g, l);
            for (size_t y = 0; y < g.h; ++y) {
                size_t ys = (g.h - y) + g.h / 2;
                if (ys >= g.h) ys -= g.h;
                for (size_t x = 0; x < g.w; ++x) {
                    size_t xs = x + g.w / 2;
                    if (xs >= g.w) xs -= g.w;
                    const double rad = tpw * ws_sqrt_lm(&g, x, y);
                    const double sn = sin(rad), cs = cos(rad);
                    const size_t t = xs + ys * g.w, idx = x + y * g.w;
                    sd[t] += creal(gd[idx]) * cs - cimag(gd[idx]) * sn;
                    sp[t] += creal(gp[idx]) * cs - cimag(gp[idx]) * sn;
                }
            }
        }
        /* normalise and correct (grid.rs:597-650) */
        const double f = 1.0 / image_weight;
        for (size_t i = 0; i < npix; ++i) {
            sd[i] *= f;
            sp[i] *= f;
        }
        const size_t half = ks * os / 2, rlen = ks * os / 2 + 1;
        double *ox = malloc(sizeof(double) * (g.w / 2 + 1)), *oy = malloc(sizeof(double) * (g.h / 2 + 1));
        ws_dct3(k1 + half, rlen, g.w * os / 2, g.w / 2 + 1, ox);
        ws_dct3(k1 + half, rlen, g.h * os / 2, g.h / 2 + 1, oy);
        const double nrm = 1.0 / ((double)os * (double)os);
        for (size_t y = 0; y < g.h; ++y) {
            const double yv = y >= g.h / 2 ? oy[y - g.h / 2] : oy[g.h / 2 - y];
            for (size_t x = 0; x < g.w; ++x) {
                const double xv = x >= g.w / 2 ? ox[x - g.w / 2] : ox[g.w / 2 - x];
                const double div = xv * yv * nrm;
                sd[x + y * g.w] /= div;
                sp[x + y * g.w] /= div;
            }
        }
        /* trim and PSF normalisation (grid.rs:837-873) */
        const size_t ow = in->width, oh = in->height, sx = (g.w - ow) / 2, sy = (g.h - oh) / 2;
        for (size_t y = 0; y < oh; ++y)
            for (size_t x = 0; x < ow; ++x) {
                out->dirty[x + y * ow] = sd[(sx + x) + (sy + y) * g.w];
                out->psf[x + y * ow] = sp[(sx + x) + (sy + y) * g.w];
            }
        const size_t centre = ow / 2 + (oh / 2) * ow;
        const double pn = out->psf[centre] != 0.0 ? 1.0 / out->psf[centre] : 0.0;
        for (size_t i = 0; i < ow * oh; ++i) {
            out->dirty[i] *= pn;
            out->psf[i] *= pn;
        }
        const double st[WS_N_STATS] = {image_weight, (double)gridded, vsum, maxw,
                                       min_w, max_w, pn, (double)g.nl};
        memcpy(out->stats, st, sizeof(st));
        free(ox);
        free(oy);
        free(sd);
        free(sp);
        free(gd);
        free(gp);
        free(k1);
        free(kern);
        free(sorted);
    }
done:
    free(smp);
    free(lay);
    free(cnt);
    return status;
}
