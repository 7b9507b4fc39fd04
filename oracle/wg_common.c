/*
 * ORACLE — TEST INFRASTRUCTURE ONLY (see wg_common.h).
 *
 * Precision-independent part of the CPU restatement of rang-core's exact
 * w-gridder: the kernel table, the ES kernel and its polynomial
 * approximation, the Gauss-Legendre correction function, the cost-model
 * kernel choice and the call geometry. Each function names the reference
 * lines it follows (core/src/wgridder.rs).
 */
#include <omp.h>
#include "wg_common.h"

#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

/* The kernel parameter table is data shared with the product library
 * (rang-core_amd/csrc/kernel_db.inc); tests pin it to the reference's
 * SHA256 of the table text. */
static const or_krow KTAB[] = {
#include "kernel_db.inc"
};

static _Thread_local char or_err[256];

void or_set_error(const char *m) {
    snprintf(or_err, sizeof or_err, "%s", m);
}

const char *or_last_error(void) {
    return or_err;
}

/* wgridder.rs:133-150 */
size_t or_kernel_count(void) {
    return sizeof KTAB / sizeof KTAB[0];
}

int or_kernel_row(size_t index, or_krow *out) {
    if (index >= or_kernel_count()) return 1;
    *out = KTAB[index];
    return 0;
}

/* wgridder.rs:154-181 (getAvailableKernels): per support the row with the
 * smallest oversampling reaching eps, supports ascending. */
int or_available_kernels(double eps, int ndim, int single, double ofmin, double ofmax,
                         int *out, int *nout) {
    int wlim = single ? 8 : 16;
    double ofc[20];
    int idx[20];
    for (int i = 0; i < 20; ++i) {
        ofc[i] = ofmax;
        idx[i] = -1;
    }
    for (size_t i = 0; i < or_kernel_count(); ++i) {
        const or_krow *r = &KTAB[i];
        if (r->ndim == ndim && r->single == single && r->w <= wlim && r->epsilon <= eps &&
            r->ofactor <= ofc[r->w] && r->ofactor >= ofmin) {
            ofc[r->w] = r->ofactor;
            idx[r->w] = (int)i;
        }
    }
    int n = 0;
    for (int i = 0; i < 20; ++i)
        if (idx[i] >= 0) out[n++] = idx[i];
    *nout = n;
    if (n == 0) {
        or_set_error("no appropriate kernel for this epsilon, sigma range and precision");
        return 1;
    }
    return 0;
}

/* wgridder.rs:185-222 (good_size_cmplx) */
/* The OpenMP threads of later calls (the baseline's 1- and all-thread
 * timings in one process). */
void or_set_threads(int n) {
    if (n > 0) omp_set_num_threads(n);
}

size_t or_good_size_complex(size_t n) {
    if (n <= 12) return n;
    size_t best = 2 * n;
    for (size_t f11 = 1; f11 < best; f11 *= 11)
        for (size_t f117 = f11; f117 < best; f117 *= 7)
            for (size_t f1175 = f117; f1175 < best; f1175 *= 5) {
                size_t x = f1175;
                while (x < n) x *= 2;
                for (;;) {
                    if (x < n) {
                        x *= 3;
                    } else if (x > n) {
                        if (x < best) best = x;
                        if (x & 1) break;
                        x >>= 1;
                    } else {
                        return n;
                    }
                }
            }
    return best;
}

/* wgridder.rs:225-232 (esknew) */
double or_es_kernel(double v, double beta, double e0) {
    double tmp = (1.0 - v) * (1.0 + v);
    if (tmp >= 0.0) return exp(beta * (pow(tmp, e0) - 1.0));
    return 0.0;
}

/* wgridder.rs:237-272: non-negative abscissas ascending from the centre,
 * weights doubled for positive abscissas. x and w hold ceil(n/2). */
void or_gauss_legendre_symmetric(int n, double *x, double *w) {
    int m = (n + 1) / 2;
    for (int i = 0; i < m; ++i) {
        double z = cos(M_PI * (i + 0.75) / (n + 0.5));
        double pp = 0.0;
        for (int iter = 0; iter < 100; ++iter) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                double p3 = p2;
                p2 = p1;
                p1 = ((double)(2 * j - 1) * z * p2 - (double)(j - 1) * p3) / (double)j;
            }
            pp = (double)n * (z * p1 - p2) / (z * z - 1.0);
            double dz = p1 / pp;
            z -= dz;
            if (fabs(dz) <= 1e-16 * fmax(fabs(z), 1.0) && iter > 0) break;
        }
        int k = m - 1 - i;
        x[k] = z;
        w[k] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
    if (n % 2 == 1) {
        x[0] = 0.0;
        w[0] *= 0.5;
    }
    for (int i = 0; i < m; ++i) w[i] *= 2.0;
}

/* wgridder.rs:359-412 (FunctionApproximator::getCoeffs on [-1, 1]):
 * Chebyshev interpolation of each of w pieces at degree d, converted to
 * monomials, highest power first. */
static void approximate(int w, int d, double beta, double e0, double *coeff) {
    const double lo = -1.0, hi = 1.0;
    int d1 = d + 1;
    double *chebroot = malloc(sizeof(double) * d1);
    double *cmat = calloc((size_t)d1 * d1, sizeof(double));
    double *y = malloc(sizeof(double) * d1);
    double *lcf = malloc(sizeof(double) * d1);
    double *lcf2 = malloc(sizeof(double) * d1);
    for (int i = 0; i <= d; ++i) chebroot[i] = cos((double)(2 * i + 1) * M_PI / (double)(2 * d + 2));
    cmat[0] = 1.0;
    if (d >= 1) cmat[d1 + 1] = 1.0;
    for (int j = 2; j <= d; ++j) {
        cmat[j * d1] = -cmat[(j - 2) * d1];
        for (int k = 1; k <= j; ++k)
            cmat[j * d1 + k] = 2.0 * cmat[(j - 1) * d1 + k - 1] - cmat[(j - 2) * d1 + k];
    }
    for (int i = 0; i < w; ++i) {
        double l = -1.0 + 2.0 * i / (double)w;
        double r = -1.0 + 2.0 * (i + 1) / (double)w;
        double sum = 0.0;
        for (int k = 0; k <= d; ++k) {
            double c = chebroot[k];
            double loc = ((c * (r - l) * 0.5 + (r + l) * 0.5) + 1.0) * 0.5 * (hi - lo) + lo;
            y[k] = or_es_kernel(loc, beta, e0);
            sum += y[k];
        }
        double avg = sum / (double)d1;
        for (int k = 0; k <= d; ++k) y[k] -= avg;
        for (int j = 0; j <= d; ++j) {
            double v = 0.0;
            for (int k = 0; k <= d; ++k)
                v += 2.0 / (double)d1 * y[k] * cos((double)(j * (2 * k + 1)) * M_PI / (double)(2 * d + 2));
            lcf[j] = v;
        }
        lcf[0] *= 0.5;
        for (int k = 0; k <= d; ++k) lcf2[k] = 0.0;
        for (int j = 0; j <= d; ++j)
            for (int k = 0; k <= d; ++k) lcf2[k] += cmat[j * d1 + k] * lcf[j];
        lcf2[0] += avg;
        for (int j = 0; j <= d; ++j) coeff[i * d1 + j] = lcf2[d - j];
    }
    free(chebroot);
    free(cmat);
    free(y);
    free(lcf);
    free(lcf2);
}

/* wgridder.rs:293-317 (EsKernel::new) */
int or_kernel_new(int index, or_kernel *k) {
    or_krow row;
    if (index < 0 || or_kernel_row((size_t)index, &row)) {
        or_set_error("kernel index out of range");
        return 1;
    }
    int supp = row.w;
    k->index = index;
    k->support = supp;
    k->beta = row.beta * (double)supp;
    k->e0 = row.e0;
    k->degree = supp + 3;
    k->coeff = malloc(sizeof(double) * (size_t)supp * (k->degree + 1));
    approximate(supp, k->degree, k->beta, k->e0, k->coeff);
    int p = (int)(1.5 * (double)supp) + 2;
    int m = (2 * p + 1) / 2;
    k->ncorr = m;
    k->corr_x = malloc(sizeof(double) * m);
    k->corr_w = malloc(sizeof(double) * m);
    or_gauss_legendre_symmetric(2 * p, k->corr_x, k->corr_w);
    for (int i = 0; i < m; ++i) {
        k->corr_w[i] *= or_es_kernel(k->corr_x[i], k->beta, k->e0) * (double)supp * 0.5;
        k->corr_x[i] *= M_PI * (double)supp;
    }
    return 0;
}

void or_kernel_free(or_kernel *k) {
    free(k->coeff);
    free(k->corr_x);
    free(k->corr_w);
    k->coeff = k->corr_x = k->corr_w = NULL;
}

double or_kernel_phi(const or_kernel *k, double v) {
    return or_es_kernel(v, k->beta, k->e0);
}

/* wgridder.rs:326-340 (PolynomialKernel::eval) */
double or_kernel_eval(const or_kernel *k, double v) {
    if (fabs(v) >= 1.0) return 0.0;
    int w = k->support;
    double xrel = (double)w * 0.5 * (v + 1.0);
    int nth = (int)xrel;
    if (nth > w - 1) nth = w - 1;
    double locx = ((xrel - nth) - 0.5) * 2.0;
    const double *c = k->coeff + (size_t)nth * (k->degree + 1);
    double res = c[0];
    for (int i = 1; i <= k->degree; ++i) res = res * locx + c[i];
    return res;
}

/* wgridder.rs:344-350 (corfunc) */
double or_corfunc(const or_kernel *k, double v) {
    double tmp = 0.0;
    for (int i = 0; i < k->ncorr; ++i) tmp += k->corr_w[i] * cos(k->corr_x[i] * v);
    return 1.0 / tmp;
}

/* wgridder.rs:1102-1123 */
int or_check_common(const or_options *o, size_t nrow, size_t nchan) {
    if (o->nx == 0 || o->ny == 0) {
        or_set_error("image dimensions must be positive");
        return 1;
    }
    if (isnan(o->epsilon) || o->epsilon <= 0.0) {
        or_set_error("epsilon must be positive");
        return 1;
    }
    if (isnan(o->pixsize_x) || isnan(o->pixsize_y) || o->pixsize_x <= 0.0 || o->pixsize_y <= 0.0) {
        or_set_error("pixel sizes must be positive");
        return 1;
    }
    if (nchan == 0) {
        or_set_error("uvw must have three values per row");
        return 1;
    }
    if (nrow >= ((size_t)1 << 32) || nchan >= ((size_t)1 << 32)) {
        or_set_error("too many rows or channels");
        return 1;
    }
    return 0;
}

/* wgridder.rs:1080-1100 */
void or_recentring(const or_options *o, double *r_l, double *r_m) {
    double lshift = o->flip_u ? -o->center_x : o->center_x;
    double mshift = o->flip_v ? -o->center_y : o->center_y;
    *r_l = 0.0;
    *r_m = 0.0;
    if (o->do_wgridding && (lshift != 0.0 || mshift != 0.0)) {
        double xext = fabs(lshift) + 0.5 * (double)o->nx * o->pixsize_x;
        double yext = fabs(mshift) + 0.5 * (double)o->ny * o->pixsize_y;
        if (xext * xext + yext * yext < 1.0) {
            double n0 = sqrt(1.0 - lshift * lshift - mshift * mshift);
            *r_l = lshift / n0;
            *r_m = mshift / n0;
        }
    }
}

/* wgridder.rs:553-579 (Baselines::new) */
int or_baselines_new(const double *uvw, const double *freq, size_t nrow, size_t nchan, int fu,
                     int fv, int fw, double r_l, double r_m, or_baselines *bl) {
    if (nchan == 0 || freq[0] <= 0.0) {
        or_set_error("channel frequencies must be positive");
        return 1;
    }
    for (size_t i = 1; i < nchan; ++i)
        if (freq[i] < freq[i - 1]) {
            or_set_error("channel frequencies must be sorted in ascending order");
            return 1;
        }
    double su = fu ? -1.0 : 1.0, sv = fv ? -1.0 : 1.0, sw = fw ? -1.0 : 1.0;
    bl->nrow = nrow;
    bl->nchan = nchan;
    bl->coord = malloc(sizeof(double) * 3 * (nrow ? nrow : 1));
    bl->f_over_c = malloc(sizeof(double) * nchan);
    for (size_t r = 0; r < nrow; ++r) {
        double tu = su * uvw[3 * r], tv = sv * uvw[3 * r + 1], tw = sw * uvw[3 * r + 2];
        bl->coord[3 * r] = tu + r_l * tw;
        bl->coord[3 * r + 1] = tv + r_m * tw;
        bl->coord[3 * r + 2] = tw;
    }
    for (size_t c = 0; c < nchan; ++c) bl->f_over_c[c] = freq[c] / OR_C_LIGHT;
    return 0;
}

void or_baselines_free(or_baselines *bl) {
    free(bl->coord);
    free(bl->f_over_c);
}

/* wgridder.rs:583-586 (FixW) */
void or_fix_w(const double *c, double *out, double *sign) {
    double res = 1.0 - 2.0 * (c[2] < 0.0 ? 1.0 : 0.0);
    out[0] = c[0] * res;
    out[1] = c[1] * res;
    out[2] = c[2] * res;
    *sign = res;
}

/* wgridder.rs:589-599 */
double or_nm1_of(double x, double y, double r_l, double r_m) {
    double xsq = x * x, ysq = y * y;
    double tmp = 1.0 - xsq - ysq;
    double nm1 = tmp >= 0.0 ? (-xsq - ysq) / (sqrt(tmp) + 1.0) : -sqrt(-tmp) - 1.0;
    return nm1 + r_l * x + r_m * y;
}

/* wgridder.rs:876-926 (getNuNv with one thread and vlen 8) */
static int choose_kernel(const or_options *o, int single, size_t nvis, double wmin_d, double wmax_d,
                         double nm1min, double nm1max, double nshift, int *kidx, size_t *nu,
                         size_t *nv) {
    int ndim = o->do_wgridding ? 3 : 2;
    int idx[20], nidx;
    if (or_available_kernels(o->epsilon, ndim, single, o->sigma_min, o->sigma_max, idx, &nidx))
        return 1;
    const double NREF_FFT = 2048.0, COSTREF_FFT = 0.0693;
    const int vlen = OR_COST_VLEN;
    double mincost = 1e300;
    *kidx = -1;
    for (int a = 0; a < nidx; ++a) {
        const or_krow *k = &KTAB[idx[a]];
        int supp = k->w;
        int nvec = (supp + vlen - 1) / vlen;
        double ofactor = k->ofactor;
        size_t tnu = 2 * or_good_size_complex((size_t)((double)o->nx * ofactor * 0.5) + 1);
        size_t tnv = 2 * or_good_size_complex((size_t)((double)o->ny * ofactor * 0.5) + 1);
        if (tnu < 16) tnu = 16;
        if (tnv < 16) tnv = 16;
        double logterm = log((double)(tnu * tnv)) / log(NREF_FFT * NREF_FFT);
        double fftcost = (double)tnu / NREF_FFT * (double)tnv / NREF_FFT * logterm * COSTREF_FFT;
        double gridcost = 2.2e-10 * (double)nvis *
                          ((double)(supp * nvec * vlen) + (double)((2 * nvec + 1) * (supp + 3) * vlen));
        if (o->do_wgridding) {
            double dw = 0.5 / ofactor / fmax(fabs(nm1max + nshift), fabs(nm1min + nshift));
            size_t nplanes = (size_t)((wmax_d - wmin_d) / dw + (double)supp);
            fftcost *= (double)nplanes;
            gridcost *= (double)supp;
        }
        double cost = fftcost + gridcost;
        if (cost < mincost) {
            mincost = cost;
            *kidx = idx[a];
            *nu = tnu;
            *nv = tnv;
        }
    }
    return 0;
}

/* wgridder.rs:604-754 (Setup::new) */
int or_setup_new(const or_options *o, int single, int wmax_taps, size_t nvis, double wmin_d,
                 double wmax_d, double r_l, double r_m, or_setup *s) {
    memset(s, 0, sizeof *s);
    size_t nx = o->nx, ny = o->ny;
    double px = o->pixsize_x, py = o->pixsize_y;
    double lshift = o->flip_u ? -o->center_x : o->center_x;
    double mshift = o->flip_v ? -o->center_y : o->center_y;
    int lmshift = lshift != 0.0 || mshift != 0.0;
    double xmin = lshift - 0.5 * (double)nx * px;
    double xmax = xmin + ((double)nx - 1.0) * px;
    double ymin = mshift - 0.5 * (double)ny * py;
    double ymax = ymin + ((double)ny - 1.0) * py;
    double xext[3] = {xmin, xmax, 0}, yext[3] = {ymin, ymax, 0};
    int nxe = 2, nye = 2;
    if (r_l != 0.0 || r_m != 0.0) {
        xext[nxe++] = lshift;
        yext[nye++] = mshift;
    } else {
        if (xmin * xmax < 0.0) xext[nxe++] = 0.0;
        if (ymin * ymax < 0.0) yext[nye++] = 0.0;
    }
    double nm1min = 1e300, nm1max = -1e300;
    for (int a = 0; a < nxe; ++a)
        for (int b = 0; b < nye; ++b) {
            double xc = xext[a], yc = yext[b];
            double tmp = xc * xc + yc * yc;
            double nval = tmp <= 1.0 ? sqrt(1.0 - tmp) - 1.0 : -sqrt(tmp - 1.0) - 1.0;
            nval += r_l * xc + r_m * yc;
            nm1min = fmin(nm1min, nval);
            nm1max = fmax(nm1max, nval);
        }
    int wgrid = o->do_wgridding;
    double nshift = (!o->allow_nshift || !wgrid) ? 0.0 : -0.5 * (nm1max + nm1min);
    int kidx = -1;
    size_t nu = 0, nv = 0;
    if (o->kernel_index >= 0) {
        or_krow row;
        if (or_kernel_row((size_t)o->kernel_index, &row)) {
            or_set_error("kernel index out of range");
            return 1;
        }
        if (row.w > wmax_taps) {
            or_set_error("kernel support too large for this precision");
            return 1;
        }
        kidx = (int)o->kernel_index;
        nu = o->nu;
        nv = o->nv;
    } else if (choose_kernel(o, single, nvis, wmin_d, wmax_d, nm1min, nm1max, nshift, &kidx, &nu, &nv)) {
        return 1;
    }
    if ((nu >> 4) >= ((size_t)1 << 16) || (nv >> 4) >= ((size_t)1 << 16)) {
        or_set_error("grid too large");
        return 1;
    }
    if (or_kernel_new(kidx, &s->kernel)) return 1;
    int supp = s->kernel.support;
    size_t nsafe = (size_t)(supp + 1) / 2;
    if (nu < 2 * nsafe || nv < 2 * nsafe) {
        or_kernel_free(&s->kernel);
        or_set_error("grid too small");
        return 1;
    }
    if (nx % 2 || ny % 2 || nu % 2 || nv % 2) {
        or_kernel_free(&s->kernel);
        or_set_error("image and grid sizes must be even");
        return 1;
    }
    double ofactor = fmin((double)nu / (double)nx, (double)nv / (double)ny);
    s->nx = nx;
    s->ny = ny;
    s->pixsize_x = px;
    s->pixsize_y = py;
    s->lshift = lshift;
    s->mshift = mshift;
    s->lmshift = lmshift;
    s->shifting = lmshift || nshift != 0.0;
    s->r_l = r_l;
    s->r_m = r_m;
    s->nshift = nshift;
    s->do_wgridding = wgrid;
    s->divide_by_n = o->divide_by_n;
    s->nu = nu;
    s->nv = nv;
    s->supp = supp;
    s->ushift = (double)supp * (-0.5) + 1.0 + (double)nu;
    s->vshift = (double)supp * (-0.5) + 1.0 + (double)nv;
    s->maxiu0 = (long)(nu + nsafe) - supp;
    s->maxiv0 = (long)(nv + nsafe) - supp;
    if (wgrid) {
        double dw = 0.5 / ofactor / fmax(fabs(nm1max + nshift), fabs(nm1min + nshift));
        size_t nplanes = (size_t)((wmax_d - wmin_d) / dw + (double)supp);
        if (nplanes >= ((size_t)1 << 16)) {
            or_kernel_free(&s->kernel);
            or_set_error("too many w planes");
            return 1;
        }
        s->dw = dw;
        s->xdw = 1.0 / dw;
        s->nplanes = nplanes;
        s->wmin = (wmin_d + wmax_d) * 0.5 - 0.5 * ((double)nplanes - 1.0) * dw;
        s->wshift = dw - (0.5 * (double)supp * dw) - s->wmin;
    }
    double info[OR_N_INFO] = {(double)kidx, (double)supp, (double)nu, (double)nv,
                              (double)s->nplanes, (double)nvis, wmin_d, wmax_d,
                              nm1min, nm1max, nshift, s->dw, s->wmin, ofactor};
    memcpy(s->info, info, sizeof info);
    return 0;
}

void or_setup_free(or_setup *s) {
    or_kernel_free(&s->kernel);
}

/* wgridder.rs:757-767 (getpix) */
void or_getpix(const or_setup *s, double uin, double vin, double *u, double *v, long *iu0,
               long *iv0) {
    double uu = uin * s->pixsize_x;
    uu = (uu - floor(uu)) * (double)s->nu;
    long i0 = (long)(uu + s->ushift) - (long)s->nu;
    if (i0 > s->maxiu0) i0 = s->maxiu0;
    *u = uu - (double)i0;
    *iu0 = i0;
    double vv = vin * s->pixsize_y;
    vv = (vv - floor(vv)) * (double)s->nv;
    long j0 = (long)(vv + s->vshift) - (long)s->nv;
    if (j0 > s->maxiv0) j0 = s->maxiv0;
    *v = vv - (double)j0;
    *iv0 = j0;
}

/* wgridder.rs:770-776 (get_uvwidx's plane) */
size_t or_minplane(const or_setup *s, double w) {
    if (!s->do_wgridding) return 0;
    double p = (w + s->wshift) * s->xdw;
    return p > 0.0 ? (size_t)p : 0;
}

/* Test hooks: the polynomial kernel and the ES kernel of row `index` at n
 * points, and the correction function at n points. */
int or_kernel_eval_many(int index, size_t n, const double *v, double *poly, double *phi,
                        double *corr) {
    or_kernel k;
    if (or_kernel_new(index, &k)) return 1;
    for (size_t i = 0; i < n; ++i) {
        if (poly) poly[i] = or_kernel_eval(&k, v[i]);
        if (phi) phi[i] = or_kernel_phi(&k, v[i]);
        if (corr) corr[i] = or_corfunc(&k, v[i]);
    }
    or_kernel_free(&k);
    return 0;
}

/* The plan a call would choose (wgridder.rs:3527-3551 `plan`) from the
 * number of active visibilities and their |w| range. */
int or_plan(const or_options *o, int single, size_t nvis, double wmin_d, double wmax_d,
            double *info) {
    double r_l, r_m;
    or_recentring(o, &r_l, &r_m);
    or_setup s;
    if (or_setup_new(o, single, single ? 8 : 16, nvis, wmin_d, wmax_d, r_l, r_m, &s)) return 1;
    for (int i = 0; i < OR_N_INFO; ++i) info[i] = s.info[i];
    or_setup_free(&s);
    return 0;
}
