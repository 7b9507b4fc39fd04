/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * CPU restatement of rang-core's exact w-gridder (core/src/wgridder.rs, a
 * port of ducc0 0.42's wgridder) used as the checker of the HIP product
 * path and as the timed CPU baseline. Only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load it. The product library never
 * links, calls or falls back to anything in this directory.
 *
 * This header holds the precision-independent part (kernel table, ES kernel
 * machinery, plan); wg_impl.h holds the gridding itself, instantiated for
 * float and double.
 */
#ifndef RANG_ORACLE_WG_COMMON_H
#define RANG_ORACLE_WG_COMMON_H

#include <stddef.h>
#include <stdint.h>

#define OR_C_LIGHT 299792458.0
#define OR_N_INFO 14
#define OR_COST_VLEN 8

/* Same field order and layout as the reference's #[repr(C)] WgridderOptions
 * (core/src/ffi_wgridder.rs:22-43), so one struct serves both libraries. */
typedef struct {
    size_t nx, ny;
    double pixsize_x, pixsize_y, epsilon, center_x, center_y, sigma_min, sigma_max;
    int32_t do_wgridding, divide_by_n, flip_u, flip_v, flip_w, allow_nshift;
    int64_t kernel_index; /* < 0: cost model */
    size_t nu, nv;        /* with kernel_index >= 0 */
} or_options;

typedef struct {
    int w;
    double ofactor, epsilon, beta, e0;
    int ndim, single;
} or_krow;

/* ES kernel: piecewise polynomial taps and the Gauss-Legendre correction. */
typedef struct {
    int index, support, degree;
    double beta, e0; /* beta already times W */
    double *coeff;   /* W * (degree + 1), highest power first */
    int ncorr;
    double *corr_x, *corr_w;
} or_kernel;

/* The fixed geometry of one call (wgridder.rs `Setup`). */
typedef struct {
    size_t nx, ny;
    double pixsize_x, pixsize_y, lshift, mshift;
    int lmshift, shifting;
    double r_l, r_m, nshift;
    int do_wgridding, divide_by_n;
    size_t nu, nv;
    int supp;
    double ushift, vshift;
    long maxiu0, maxiv0;
    or_kernel kernel;
    size_t nplanes;
    double wmin, dw, xdw, wshift;
    double info[OR_N_INFO];
} or_setup;

/* Coordinates after flips and recentring, and f/c per channel. */
typedef struct {
    size_t nrow, nchan;
    double *coord;    /* nrow * 3 */
    double *f_over_c; /* nchan */
} or_baselines;

size_t or_kernel_count(void);
int or_kernel_row(size_t index, or_krow *out);
size_t or_good_size_complex(size_t n);
double or_es_kernel(double v, double beta, double e0);
void or_gauss_legendre_symmetric(int n, double *x, double *w);
int or_kernel_new(int index, or_kernel *k);
void or_kernel_free(or_kernel *k);
double or_kernel_eval(const or_kernel *k, double v);
double or_kernel_phi(const or_kernel *k, double v);
double or_corfunc(const or_kernel *k, double v);
int or_available_kernels(double eps, int ndim, int single, double ofmin, double ofmax,
                         int *out, int *nout);
const char *or_last_error(void);
void or_set_error(const char *m);

int or_check_common(const or_options *o, size_t nrow, size_t nchan);
void or_recentring(const or_options *o, double *r_l, double *r_m);
int or_baselines_new(const double *uvw, const double *freq, size_t nrow, size_t nchan,
                     int fu, int fv, int fw, double r_l, double r_m, or_baselines *bl);
void or_baselines_free(or_baselines *bl);
int or_setup_new(const or_options *o, int single, int wmax_taps, size_t nvis, double wmin_d,
                 double wmax_d, double r_l, double r_m, or_setup *s);
void or_setup_free(or_setup *s);
void or_fix_w(const double *c, double *out, double *sign);
double or_nm1_of(double x, double y, double r_l, double r_m);
void or_getpix(const or_setup *s, double uin, double vin, double *u, double *v, long *iu0,
               long *iv0);
size_t or_minplane(const or_setup *s, double w);

#endif
