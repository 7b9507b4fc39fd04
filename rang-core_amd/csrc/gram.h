// The normal operator of the measurement without visibilities (rang-core
// core/src/gridder_gram.rs): a sum of Toeplitz operators, one per w-node,
// each applied by FFTs on the doubled grid, and its one-node circulant
// inverse used as a preconditioner.
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>
#include <functional>
#include <string>
#include <vector>

#include "wg_device.h"
#include "wg_plan.h"

namespace rang {

// The compact w-expansion (gridder_gram.rs:36-183): exp(i theta xi s) ~
// sum_j psi_j(s) exp(i theta xi t_j) for |xi| <= 1, s in [-1/2, 1/2], with
// `stencil` nodes t_j = j - (stencil - 1)/2, the coefficients tabulated at
// kWexpBins + 1 values of s.
constexpr int kWexpBins = 8192;
struct WExpansion {
    int stencil = 0;
    double theta = 0;
    std::vector<double> table;  // (kWexpBins + 1) * stencil
    double sup_error = 0, lebesgue = 0;
    double coefficient(double s, int j) const;
};
bool wexpansion_build(int stencil, double theta, double ridge, WExpansion& out, std::string& err);

// Counters of a built operator (gridder_gram.rs:189-214).
constexpr int kGramStats = 13;
// planes, stencil, theta, wmin, wmax, x_range, sup_error, lebesgue,
// active_pairs, cube_bytes, hermitian_check, spectral_sparsity, and the
// largest |K^| of the nodes (unscaled), the scale of an inverse filter's rho

struct GramBase {
    virtual ~GramBase() = default;
    virtual bool single() const = 0;
    virtual int side() const = 0;
    virtual void stats(double* out) const = 0;
};

// GramOperator::new (w-aware, gridder_gram.rs:226-407) when `ex` is given,
// GramOperator::no_w (:414-478) otherwise. Device uvw (nrow x 3), weights
// (kind f32 / f64), freq on the host. Returns null with status and message.
template <class T>
GramBase* gram_new(const Options& o, const DevInputs& in, const double* freq_host, const WExpansion* ex,
                   double epsilon, hipStream_t st, int& status, std::string& err);
// GramOperator::inverse_filter (:488-509).
GramBase* gram_inverse_filter(const GramBase* g, double rho, hipStream_t st, int& status, std::string& err);
// out = G x (n x n, device), gridder_gram.rs:519-647.
template <class T>
int gram_apply_dev(GramBase* g, const T* x, T* out, hipStream_t st, std::string& err);

// Preconditioned conjugate gradients on the prepared plan's exact normal
// operator, z = precond(r) (symmetric positive definite).
template <class T>
int prepared_pcg(PreparedBase* p, const std::function<void(const T*, T*, hipStream_t)>& precond,
                 const T* rhs, T* x, int iterations, double tolerance, int* used, double* rel,
                 hipStream_t st, std::string& err);

}  // namespace rang
