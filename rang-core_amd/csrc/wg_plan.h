// Host-side planning of the exact w-gridder, in FP64: the ES kernel table,
// the kernel's piecewise-polynomial approximation and correction function,
// ducc0's cost-model choice of kernel and grid, and the w-plane geometry of
// one call. Everything here runs once per call on the host; the device
// pipeline (wg_device.hip) consumes the resulting Geometry.
//
// Behaviour follows rang-core core/src/wgridder.rs (cited per item).
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace rang {

constexpr double kCLight = 299792458.0;  // wgridder.rs:58
constexpr int kNInfo = 14;               // wgridder.rs:490

struct KernelRow {
    int w;
    double ofactor, epsilon, beta, e0;
    int ndim, single;
};

std::size_t kernel_count();
const KernelRow* kernel_row(std::size_t index);
std::size_t good_size_complex(std::size_t n);
double es_kernel(double v, double beta, double e0);
void gauss_legendre_symmetric(int n, std::vector<double>& x, std::vector<double>& w);

// The reference's WgridderOptions fields (core/src/ffi_wgridder.rs:22-43).
struct Options {
    std::size_t nx = 0, ny = 0;
    double pixsize_x = 0, pixsize_y = 0, epsilon = 1e-6;
    double center_x = 0, center_y = 0, sigma_min = 1.1, sigma_max = 2.6;
    bool do_wgridding = true, divide_by_n = false, flip_u = false, flip_v = false,
         flip_w = false, allow_nshift = true;
    long long kernel_index = -1;
    std::size_t nu = 0, nv = 0;
};

// One ES kernel (wgridder.rs:277-356).
struct EsKernel {
    int index = -1, support = 0, degree = 0;
    double beta = 0, e0 = 0;
    std::vector<double> coeff;  // support * (degree + 1), highest power first
    std::vector<double> corr_x, corr_w;
    bool init(int index, std::string& err);
    double eval(double v) const;
    double phi(double v) const { return es_kernel(v, beta, e0); }
    double corfunc(double v) const;
    // Even/odd split (wgridder.rs:1485-1507): 2 * ne * support values.
    std::vector<double> even_odd(int& ne) const;
};

// The call's geometry (wgridder.rs:516-544 Setup), plus the tile shape the
// device gridding uses.
struct Geometry {
    std::size_t nx = 0, ny = 0, nu = 0, nv = 0;
    double pixsize_x = 0, pixsize_y = 0, lshift = 0, mshift = 0;
    bool lmshift = false, shifting = false, do_wgridding = true, divide_by_n = false;
    double r_l = 0, r_m = 0, nshift = 0;
    int supp = 0;
    double ushift = 0, vshift = 0;
    long long maxiu0 = 0, maxiv0 = 0;
    EsKernel kernel;
    std::size_t nplanes = 0;
    double wmin = 0, dw = 0, xdw = 0, wshift = 0;
    double info[kNInfo] = {};
    // flips applied to the uvw rows: u, v and w (w's is the negation of
    // flip_w, wgridder.rs:2461)
    bool fu = false, fv = false, fw = false;
};

void recentring(const Options& o, double& r_l, double& r_m);
bool check_common(const Options& o, std::size_t nrow, std::size_t nchan, std::string& err);
bool check_freq(const double* freq, std::size_t nchan, std::string& err);
bool make_geometry(const Options& o, bool single, std::size_t nvis, double wmin_d,
                   double wmax_d, Geometry& g, std::string& err);

// The kernel rows the cost model weighs (one per support, the smallest
// oversampling reaching epsilon), with their grids and modelled cost,
// cheapest first.
struct KernelChoice {
    int index, support;
    std::size_t nu, nv;
    double cost;
};
bool kernel_choices(const Options& o, bool single, std::size_t nvis, double wmin_d, double wmax_d,
                    std::vector<KernelChoice>& out, std::string& err);

}  // namespace rang
