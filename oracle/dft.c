/*
 * ORACLE — TEST INFRASTRUCTURE ONLY.
 *
 * FP64 direct Fourier transforms, the closed form the gridder approximates
 * (core/src/wgridder.rs:30-36; the reference's own DFT checks at
 * core/src/wgridder.rs:4595-4617 and tests/test_core_wgridder.py:44-53):
 *   dirty(l, m) = sum_k wgt_k Re[V_k exp(+2 pi i (u_k l + v_k m + w_k (n-1)))]
 *   V_k = wgt_k sum_p I_p exp(-2 pi i (u_k l_p + v_k m_p + w_k (n_p-1)))
 * with n - 1 = sqrt(1 - l^2 - m^2) - 1 and u, v, w in wavelengths. With
 * divn the image is divided by n (ducc0's divide_by_n). Callers pass the
 * pixels (l, m) and visibilities they want, so subsets cost what they hold.
 */
#include <math.h>
#include <stddef.h>

static inline double nm1_of(double l, double m) {
    return sqrt(1.0 - l * l - m * m) - 1.0;
}

void or_dft_dirty(size_t nvis, const double *u, const double *v, const double *w,
                  const double *vis, const double *wgt, size_t npix, const double *l,
                  const double *m, int divn, double *out) {
#pragma omp parallel for schedule(dynamic, 4)
    for (size_t p = 0; p < npix; ++p) {
        double n1 = nm1_of(l[p], m[p]);
        double acc = 0.0;
        for (size_t k = 0; k < nvis; ++k) {
            double ph = 2.0 * M_PI * (u[k] * l[p] + v[k] * m[p] + w[k] * n1);
            double wk = wgt ? wgt[k] : 1.0;
            acc += wk * (vis[2 * k] * cos(ph) - vis[2 * k + 1] * sin(ph));
        }
        out[p] = divn ? acc / (n1 + 1.0) : acc;
    }
}

void or_dft_vis(size_t nvis, const double *u, const double *v, const double *w, size_t npix,
                const double *l, const double *m, const double *img, int divn, const double *wgt,
                double *out) {
#pragma omp parallel for schedule(dynamic, 64)
    for (size_t k = 0; k < nvis; ++k) {
        double re = 0.0, im = 0.0;
        for (size_t p = 0; p < npix; ++p) {
            double n1 = nm1_of(l[p], m[p]);
            double val = divn ? img[p] / (n1 + 1.0) : img[p];
            double ph = -2.0 * M_PI * (u[k] * l[p] + v[k] * m[p] + w[k] * n1);
            re += val * cos(ph);
            im += val * sin(ph);
        }
        double wk = wgt ? wgt[k] : 1.0;
        out[2 * k] = wk * re;
        out[2 * k + 1] = wk * im;
    }
}
