// Host-side FP64 planning of the exact w-gridder (see wg_plan.h).
#include "wg_plan.h"

#include <algorithm>
#include <cmath>

namespace rang {

namespace {

// ducc0 0.42.0's KernelDB rows (ndim 2 and 3) as rang-core keeps them
// (core/src/wgridder.rs:3597-4539); data file shared with the oracle.
const KernelRow kRows[] = {
#include "kernel_db.inc"
};

constexpr double kPi = 3.14159265358979323846;

}  // namespace

std::size_t kernel_count() { return sizeof(kRows) / sizeof(kRows[0]); }

const KernelRow* kernel_row(std::size_t index) {
    return index < kernel_count() ? &kRows[index] : nullptr;
}

// Smallest 2^a 3^b 5^c 7^d 11^e >= n, n itself for n <= 12
// (wgridder.rs:185-222, ducc0's good_size_cmplx).
std::size_t good_size_complex(std::size_t n) {
    if (n <= 12) return n;
    std::size_t best = 2 * n;
    for (std::size_t a = 1; a < best; a *= 11)
        for (std::size_t b = a; b < best; b *= 7)
            for (std::size_t c = b; c < best; c *= 5) {
                std::size_t x = c;
                while (x < n) x *= 2;
                while (true) {
                    if (x < n) {
                        x *= 3;
                        continue;
                    }
                    if (x == n) return n;
                    best = std::min(best, x);
                    if (x & 1) break;
                    x >>= 1;
                }
            }
    return best;
}

// exp(beta ((1 - v^2)^e0 - 1)) on [-1, 1] (wgridder.rs:225-232).
double es_kernel(double v, double beta, double e0) {
    double t = (1.0 - v) * (1.0 + v);
    return t >= 0.0 ? std::exp(beta * (std::pow(t, e0) - 1.0)) : 0.0;
}

// Non-negative Gauss-Legendre abscissas ascending from the centre, weights
// doubled for the positive ones (wgridder.rs:237-272).
void gauss_legendre_symmetric(int n, std::vector<double>& x, std::vector<double>& w) {
    const int m = (n + 1) / 2;
    x.assign(m, 0.0);
    w.assign(m, 0.0);
    for (int i = 0; i < m; ++i) {
        double z = std::cos(kPi * (i + 0.75) / (n + 0.5));
        double pp = 0.0;
        for (int it = 0; it < 100; ++it) {
            double p1 = 1.0, p2 = 0.0;
            for (int j = 1; j <= n; ++j) {
                const double p3 = p2;
                p2 = p1;
                p1 = (double(2 * j - 1) * z * p2 - double(j - 1) * p3) / double(j);
            }
            pp = double(n) * (z * p1 - p2) / (z * z - 1.0);
            const double dz = p1 / pp;
            z -= dz;
            if (it > 0 && std::fabs(dz) <= 1e-16 * std::max(std::fabs(z), 1.0)) break;
        }
        x[m - 1 - i] = z;
        w[m - 1 - i] = 2.0 / ((1.0 - z * z) * pp * pp);
    }
    if (n % 2 == 1) {
        x[0] = 0.0;
        w[0] *= 0.5;
    }
    for (double& v : w) v *= 2.0;
}

namespace {

// Chebyshev interpolation of each of `w` pieces of [-1, 1] at degree `d`,
// turned into monomial coefficients, highest power first
// (wgridder.rs:359-412, ducc0's FunctionApproximator).
std::vector<double> piecewise_poly(int w, int d, double beta, double e0) {
    const int d1 = d + 1;
    std::vector<double> out(std::size_t(w) * d1), root(d1), cm(std::size_t(d1) * d1, 0.0);
    for (int i = 0; i <= d; ++i) root[i] = std::cos(double(2 * i + 1) * kPi / double(2 * d + 2));
    // Chebyshev polynomials in monomials: cm[j][k] = coefficient of x^k in T_j.
    cm[0] = 1.0;
    if (d >= 1) cm[d1 + 1] = 1.0;
    for (int j = 2; j <= d; ++j) {
        cm[std::size_t(j) * d1] = -cm[std::size_t(j - 2) * d1];
        for (int k = 1; k <= j; ++k)
            cm[std::size_t(j) * d1 + k] =
                2.0 * cm[std::size_t(j - 1) * d1 + k - 1] - cm[std::size_t(j - 2) * d1 + k];
    }
    std::vector<double> y(d1), cheb(d1), mono(d1);
    for (int piece = 0; piece < w; ++piece) {
        const double lo = -1.0 + 2.0 * piece / double(w);
        const double hi = -1.0 + 2.0 * (piece + 1) / double(w);
        double sum = 0.0;
        for (int k = 0; k <= d; ++k) {
            const double at = ((root[k] * (hi - lo) * 0.5 + (hi + lo) * 0.5) + 1.0) * 0.5 * 2.0 - 1.0;
            y[k] = es_kernel(at, beta, e0);
            sum += y[k];
        }
        const double mean = sum / double(d1);
        for (double& v : y) v -= mean;
        for (int j = 0; j <= d; ++j) {
            double acc = 0.0;
            for (int k = 0; k <= d; ++k)
                acc += 2.0 / double(d1) * y[k] *
                       std::cos(double(j * (2 * k + 1)) * kPi / double(2 * d + 2));
            cheb[j] = acc;
        }
        cheb[0] *= 0.5;
        std::fill(mono.begin(), mono.end(), 0.0);
        for (int j = 0; j <= d; ++j)
            for (int k = 0; k <= d; ++k) mono[k] += cm[std::size_t(j) * d1 + k] * cheb[j];
        mono[0] += mean;
        for (int j = 0; j <= d; ++j) out[std::size_t(piece) * d1 + j] = mono[d - j];
    }
    return out;
}

}  // namespace

bool EsKernel::init(int idx, std::string& err) {
    const KernelRow* row = idx >= 0 ? kernel_row(std::size_t(idx)) : nullptr;
    if (!row) {
        err = "kernel index out of range";
        return false;
    }
    index = idx;
    support = row->w;
    beta = row->beta * support;
    e0 = row->e0;
    degree = support + 3;
    coeff = piecewise_poly(support, degree, beta, e0);
    // Gauss-Legendre correction with p = 1.5 W + 2 and 2p nodes
    // (wgridder.rs:300-306, ducc0's GLFullCorrection).
    const int p = int(1.5 * support) + 2;
    gauss_legendre_symmetric(2 * p, corr_x, corr_w);
    for (std::size_t i = 0; i < corr_x.size(); ++i) {
        corr_w[i] *= es_kernel(corr_x[i], beta, e0) * support * 0.5;
        corr_x[i] *= kPi * support;
    }
    return true;
}

// wgridder.rs:326-340
double EsKernel::eval(double v) const {
    if (std::fabs(v) >= 1.0) return 0.0;
    const double xrel = support * 0.5 * (v + 1.0);
    const int nth = std::min(int(xrel), support - 1);
    const double loc = ((xrel - nth) - 0.5) * 2.0;
    const double* c = &coeff[std::size_t(nth) * (degree + 1)];
    double r = c[0];
    for (int i = 1; i <= degree; ++i) r = r * loc + c[i];
    return r;
}

// 1 / sum_k w_k cos(x_k v) (wgridder.rs:344-350)
double EsKernel::corfunc(double v) const {
    double t = 0.0;
    for (std::size_t i = 0; i < corr_x.size(); ++i) t += corr_w[i] * std::cos(corr_x[i] * v);
    return 1.0 / t;
}

// Row k of the even part then of the odd part, tap i at [k * W + i]:
// p_i(x) = Pe_i(x^2) + x Po_i(x^2), degree padded to odd (wgridder.rs:1485-1507).
std::vector<double> EsKernel::even_odd(int& ne) const {
    const int d1 = degree + 1, dodd = degree | 1;
    ne = (dodd + 1) / 2;
    std::vector<double> out(std::size_t(2) * ne * support, 0.0);
    auto a = [&](int tap, int power) {
        return power > degree ? 0.0 : coeff[std::size_t(tap) * d1 + (degree - power)];
    };
    for (int tap = 0; tap < support; ++tap)
        for (int k = 0; k < ne; ++k) {
            out[std::size_t(k) * support + tap] = a(tap, dodd - 1 - 2 * k);
            out[std::size_t(ne + k) * support + tap] = a(tap, dodd - 2 * k);
        }
    return out;
}

// wgridder.rs:1080-1100
void recentring(const Options& o, double& r_l, double& r_m) {
    r_l = r_m = 0.0;
    const double ls = o.flip_u ? -o.center_x : o.center_x;
    const double ms = o.flip_v ? -o.center_y : o.center_y;
    if (!o.do_wgridding || (ls == 0.0 && ms == 0.0)) return;
    const double xe = std::fabs(ls) + 0.5 * double(o.nx) * o.pixsize_x;
    const double ye = std::fabs(ms) + 0.5 * double(o.ny) * o.pixsize_y;
    if (xe * xe + ye * ye < 1.0) {
        const double n0 = std::sqrt(1.0 - ls * ls - ms * ms);
        r_l = ls / n0;
        r_m = ms / n0;
    }
}

// wgridder.rs:1102-1123
bool check_common(const Options& o, std::size_t nrow, std::size_t nchan, std::string& err) {
    if (o.nx == 0 || o.ny == 0) {
        err = "image dimensions must be positive";
    } else if (std::isnan(o.epsilon) || o.epsilon <= 0.0) {
        err = "epsilon must be positive";
    } else if (std::isnan(o.pixsize_x) || std::isnan(o.pixsize_y) || o.pixsize_x <= 0.0 ||
               o.pixsize_y <= 0.0) {
        err = "pixel sizes must be positive";
    } else if (nchan == 0) {
        err = "uvw must have three values per row";
    } else if (nrow >= (std::size_t(1) << 32) || nchan >= (std::size_t(1) << 32)) {
        err = "too many rows or channels";
    } else {
        return true;
    }
    return false;
}

// wgridder.rs:560-565
bool check_freq(const double* freq, std::size_t nchan, std::string& err) {
    if (nchan == 0 || !(freq[0] > 0.0)) {
        err = "channel frequencies must be positive";
        return false;
    }
    for (std::size_t i = 1; i < nchan; ++i)
        if (freq[i] < freq[i - 1]) {
            err = "channel frequencies must be sorted in ascending order";
            return false;
        }
    return true;
}

namespace {

// Per support the row with the smallest oversampling reaching eps
// (wgridder.rs:154-181, ducc0's getAvailableKernels).
bool candidates(double eps, int ndim, bool single, double ofmin, double ofmax,
                std::vector<int>& out, std::string& err) {
    const int wlim = single ? 8 : 16;
    double best_of[20];
    int best[20];
    std::fill(best_of, best_of + 20, ofmax);
    std::fill(best, best + 20, -1);
    for (std::size_t i = 0; i < kernel_count(); ++i) {
        const KernelRow& r = kRows[i];
        if (r.ndim != ndim || (r.single != 0) != single || r.w > wlim || r.epsilon > eps) continue;
        if (r.ofactor <= best_of[r.w] && r.ofactor >= ofmin) {
            best_of[r.w] = r.ofactor;
            best[r.w] = int(i);
        }
    }
    out.clear();
    for (int v : best)
        if (v >= 0) out.push_back(v);
    if (out.empty()) {
        err = "no appropriate kernel for this epsilon, sigma range and precision";
        return false;
    }
    return true;
}

double dw_of(double ofactor, double nm1min, double nm1max, double nshift) {
    return 0.5 / ofactor / std::max(std::fabs(nm1max + nshift), std::fabs(nm1min + nshift));
}

// ducc0's getNuNv cost model with one thread and a SIMD width of eight
// (wgridder.rs:876-926).
bool cost_model(const Options& o, bool single, std::size_t nvis, double wmin_d, double wmax_d,
                double nm1min, double nm1max, double nshift, int& kidx, std::size_t& nu,
                std::size_t& nv, std::string& err, std::vector<KernelChoice>* all = nullptr) {
    std::vector<int> cand;
    if (!candidates(o.epsilon, o.do_wgridding ? 3 : 2, single, o.sigma_min, o.sigma_max, cand, err))
        return false;
    constexpr double nref = 2048.0, costref = 0.0693;
    constexpr int vlen = 8;
    double mincost = 1e300;
    if (all) all->clear();
    for (int i : cand) {
        const KernelRow& k = kRows[i];
        const int supp = k.w, nvec = (supp + vlen - 1) / vlen;
        const std::size_t tnu =
            std::max<std::size_t>(2 * good_size_complex(std::size_t(double(o.nx) * k.ofactor * 0.5) + 1), 16);
        const std::size_t tnv =
            std::max<std::size_t>(2 * good_size_complex(std::size_t(double(o.ny) * k.ofactor * 0.5) + 1), 16);
        const double logterm = std::log(double(tnu * tnv)) / std::log(nref * nref);
        double fft = double(tnu) / nref * double(tnv) / nref * logterm * costref;
        double grid = 2.2e-10 * double(nvis) *
                      (double(supp * nvec * vlen) + double((2 * nvec + 1) * (supp + 3) * vlen));
        if (o.do_wgridding) {
            const double dw = dw_of(k.ofactor, nm1min, nm1max, nshift);
            fft *= double(std::size_t((wmax_d - wmin_d) / dw + supp));
            grid *= supp;
        }
        if (all) all->push_back(KernelChoice{i, supp, tnu, tnv, fft + grid});
        if (fft + grid < mincost) {
            mincost = fft + grid;
            kidx = i;
            nu = tnu;
            nv = tnv;
        }
    }
    return true;
}

// The extremes of n - 1 over the image and the w shift (wgridder.rs:616-650).
void nm1_range(const Options& o, double& nm1min, double& nm1max, double& nshift) {
    double r_l, r_m;
    recentring(o, r_l, r_m);
    const double px = o.pixsize_x, py = o.pixsize_y;
    const double ls = o.flip_u ? -o.center_x : o.center_x;
    const double ms = o.flip_v ? -o.center_y : o.center_y;
    const double xmin = ls - 0.5 * double(o.nx) * px, xmax = xmin + (double(o.nx) - 1.0) * px;
    const double ymin = ms - 0.5 * double(o.ny) * py, ymax = ymin + (double(o.ny) - 1.0) * py;
    std::vector<double> xs{xmin, xmax}, ys{ymin, ymax};
    if (r_l != 0.0 || r_m != 0.0) {
        xs.push_back(ls);
        ys.push_back(ms);
    } else {
        if (xmin * xmax < 0.0) xs.push_back(0.0);
        if (ymin * ymax < 0.0) ys.push_back(0.0);
    }
    nm1min = 1e300;
    nm1max = -1e300;
    for (double x : xs)
        for (double y : ys) {
            const double r2 = x * x + y * y;
            double n = r2 <= 1.0 ? std::sqrt(1.0 - r2) - 1.0 : -std::sqrt(r2 - 1.0) - 1.0;
            n += r_l * x + r_m * y;
            nm1min = std::min(nm1min, n);
            nm1max = std::max(nm1max, n);
        }
    nshift = (!o.allow_nshift || !o.do_wgridding) ? 0.0 : -0.5 * (nm1max + nm1min);
}

}  // namespace

bool kernel_choices(const Options& o, bool single, std::size_t nvis, double wmin_d, double wmax_d,
                    std::vector<KernelChoice>& out, std::string& err) {
    double nm1min, nm1max, nshift;
    nm1_range(o, nm1min, nm1max, nshift);
    int kidx = -1;
    std::size_t nu = 0, nv = 0;
    if (!cost_model(o, single, nvis, wmin_d, wmax_d, nm1min, nm1max, nshift, kidx, nu, nv, err, &out))
        return false;
    std::stable_sort(out.begin(), out.end(),
                     [](const KernelChoice& a, const KernelChoice& b) { return a.cost < b.cost; });
    return true;
}

// Setup::new (wgridder.rs:604-754): n-1 extremes, nshift, kernel and grid,
// and the w-plane range.
bool make_geometry(const Options& o, bool single, std::size_t nvis, double wmin_d,
                   double wmax_d, Geometry& g, std::string& err) {
    g = Geometry();
    double r_l, r_m;
    recentring(o, r_l, r_m);
    const double px = o.pixsize_x, py = o.pixsize_y;
    const double ls = o.flip_u ? -o.center_x : o.center_x;
    const double ms = o.flip_v ? -o.center_y : o.center_y;
    double nm1min, nm1max, nshift;
    nm1_range(o, nm1min, nm1max, nshift);
    int kidx = -1;
    std::size_t nu = 0, nv = 0;
    if (o.kernel_index >= 0) {
        const KernelRow* row = kernel_row(std::size_t(o.kernel_index));
        if (!row) {
            err = "kernel index out of range";
            return false;
        }
        if (row->w > (single ? 8 : 16)) {
            err = "kernel support too large for this precision";
            return false;
        }
        kidx = int(o.kernel_index);
        nu = o.nu;
        nv = o.nv;
    } else if (!cost_model(o, single, nvis, wmin_d, wmax_d, nm1min, nm1max, nshift, kidx, nu, nv,
                           err)) {
        return false;
    }
    if ((nu >> 4) >= (std::size_t(1) << 16) || (nv >> 4) >= (std::size_t(1) << 16)) {
        err = "grid too large";
        return false;
    }
    if (!g.kernel.init(kidx, err)) return false;
    const int supp = g.kernel.support;
    const std::size_t nsafe = std::size_t(supp + 1) / 2;
    if (nu < 2 * nsafe || nv < 2 * nsafe) {
        err = "grid too small";
        return false;
    }
    if (o.nx % 2 || o.ny % 2 || nu % 2 || nv % 2) {
        err = "image and grid sizes must be even";
        return false;
    }
    const double ofactor = std::min(double(nu) / double(o.nx), double(nv) / double(o.ny));
    g.nx = o.nx;
    g.ny = o.ny;
    g.nu = nu;
    g.nv = nv;
    g.pixsize_x = px;
    g.pixsize_y = py;
    g.lshift = ls;
    g.mshift = ms;
    g.lmshift = ls != 0.0 || ms != 0.0;
    g.shifting = g.lmshift || nshift != 0.0;
    g.r_l = r_l;
    g.r_m = r_m;
    g.nshift = nshift;
    g.do_wgridding = o.do_wgridding;
    g.divide_by_n = o.divide_by_n;
    g.supp = supp;
    g.ushift = supp * (-0.5) + 1.0 + double(nu);
    g.vshift = supp * (-0.5) + 1.0 + double(nv);
    g.maxiu0 = (long long)(nu + nsafe) - supp;
    g.maxiv0 = (long long)(nv + nsafe) - supp;
    if (o.do_wgridding) {
        g.dw = dw_of(ofactor, nm1min, nm1max, nshift);
        g.xdw = 1.0 / g.dw;
        g.nplanes = std::size_t((wmax_d - wmin_d) / g.dw + supp);
        if (g.nplanes >= (std::size_t(1) << 16)) {
            err = "too many w planes";
            return false;
        }
        g.wmin = (wmin_d + wmax_d) * 0.5 - 0.5 * (double(g.nplanes) - 1.0) * g.dw;
        g.wshift = g.dw - (0.5 * supp * g.dw) - g.wmin;
    }
    g.fu = o.flip_u;
    g.fv = o.flip_v;
    g.fw = !o.flip_w;
    const double info[kNInfo] = {double(kidx), double(supp), double(nu), double(nv),
                                 double(g.nplanes), double(nvis), wmin_d, wmax_d,
                                 nm1min, nm1max, nshift, g.dw, g.wmin, ofactor};
    std::copy(info, info + kNInfo, g.info);
    return true;
}

}  // namespace rang
