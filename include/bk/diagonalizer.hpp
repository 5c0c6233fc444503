// bk/diagonalizer.hpp -- C++17 drop-in for the reference's
// src/subspace/diagonalizer.cuh, implemented over libbk.so.
//
//   DiagMethod / DiagConfig / Eigenpair / Eigensolution ... diagonalizer.cuh:122-144
//   FullSpaceMatVec: the MatVecFunc of :400-406, callable as
//       void(const double* d_x, double* d_y)  on device pointers (FP64 real),
//     computing y = (E/N) x - g sum_j X_j x (fullspace_matvec_kernel_d, :60-77)
//   Diagonalizer::solve(n, g, cfg)
//     Lanczos ....... :377-416 (GPU Lanczos of :420-554: v0 from mt19937_64(42),
//                     uniform(-0.5, 0.5), no reorthogonalisation, breakdown at
//                     |beta| < tolerance, Ritz vectors)
//     ExactDense .... :250-294 (n <= 14), rocSOLVER on the device instead of Eigen
//   Diagonalizer::solve(basis, g, n, cfg): H restricted to the basis states
//     Lanczos ....... :299-372 (the same Lanczos over the subspace H.v, :79-114)
//     ExactDense .... :188-245 (rocSOLVER on the device, at most 32768 states)
//   Eigenvectors of a subspace solve are in the order of the given basis.
#pragma once

#include <stdexcept>
#include <string>
#include <vector>

#include "../bk_capi.h"
#include "sampler.hpp"

enum class DiagMethod { Lanczos, ExactDense };

struct DiagConfig {
    DiagMethod method = DiagMethod::Lanczos;
    int num_eigenpairs = 1;
    int max_krylov_dim = 100;
    double tolerance = 1e-12;
    bool normalize_by_n = true;
};

struct Eigenpair {
    double energy = 0.0;
    std::vector<double> state_vector;
};

struct Eigensolution {
    std::vector<Eigenpair> pairs;
    double ground_energy() const { return pairs.empty() ? 0.0 : pairs[0].energy; }
    const std::vector<double>& ground_state() const { return pairs[0].state_vector; }
};

namespace bk {

// Device-pointer H.v with the reference's MatVecFunc call shape.
class FullSpaceMatVec {
public:
    FullSpaceMatVec(bk_ctx* ctx, int n, double g, bool normalize_by_n, void* hip_stream = nullptr)
        : ctx_(ctx), n_(n), g_(g), norm_(normalize_by_n), stream_(hip_stream)
    {
    }
    void operator()(const double* d_x, double* d_y) const
    {
        check(bk_hv(ctx_, n_, g_, norm_ ? 1 : 0, 0, d_x, d_y, stream_), "bk_hv");
    }

private:
    bk_ctx* ctx_;
    int n_;
    double g_;
    bool norm_;
    void* stream_;
};

}  // namespace bk

class Diagonalizer {
public:
    explicit Diagonalizer(int device = 0) : ctx_(device) {}

    Eigensolution solve(int n, double g, DiagConfig config = {}) const
    {
        if (config.method == DiagMethod::ExactDense) {
            if (n > 14) throw std::runtime_error("Full space dense ED exceeds feasible memory for N > 14. Use Lanczos.");
        } else if (n > 30) {
            throw std::runtime_error("Full space dimension 2^N exceeds memory limit for N > 30.");
        }
        const size_t dim = size_t(1) << n;
        std::vector<double> ev(static_cast<size_t>(config.num_eigenpairs));
        std::vector<double> vec(static_cast<size_t>(config.num_eigenpairs) * dim);
        const int k = config.method == DiagMethod::ExactDense
                          ? bk_fullspace_dense(ctx_.get(), n, g, config.normalize_by_n ? 1 : 0, config.num_eigenpairs,
                                               ev.data(), vec.data())
                          : bk_lanczos(ctx_.get(), n, g, config.normalize_by_n ? 1 : 0, config.max_krylov_dim,
                                       config.tolerance, config.num_eigenpairs, ev.data(), vec.data());
        bk::check(k, config.method == DiagMethod::ExactDense ? "bk_fullspace_dense" : "bk_lanczos");
        return unpack(k, ev, vec, dim);
    }

    Eigensolution solve(const std::vector<SpinState>& basis, double g, int n, DiagConfig config = {}) const
    {
        const size_t dim = basis.size();
        if (dim == 0) return {};
        std::vector<uint64_t> bits(dim);
        for (size_t i = 0; i < dim; ++i) bits[i] = basis[i].bits;
        std::vector<double> ev(static_cast<size_t>(config.num_eigenpairs));
        std::vector<double> vec(static_cast<size_t>(config.num_eigenpairs) * dim);
        const int norm = config.normalize_by_n ? 1 : 0;
        const int k = config.method == DiagMethod::ExactDense
                          ? bk_subspace_dense(ctx_.get(), n, g, norm, bits.data(), static_cast<int64_t>(dim),
                                              config.num_eigenpairs, ev.data(), vec.data())
                          : bk_subspace_lanczos(ctx_.get(), n, g, norm, bits.data(), static_cast<int64_t>(dim),
                                                config.max_krylov_dim, config.tolerance, config.num_eigenpairs,
                                                ev.data(), vec.data());
        bk::check(k, config.method == DiagMethod::ExactDense ? "bk_subspace_dense" : "bk_subspace_lanczos");
        return unpack(k, ev, vec, dim);
    }

    bk_ctx* context() const { return ctx_.get(); }

private:
    static Eigensolution unpack(int k, const std::vector<double>& ev, const std::vector<double>& vec, size_t dim)
    {
        Eigensolution sol;
        sol.pairs.resize(static_cast<size_t>(k));
        for (int i = 0; i < k; ++i) {
            sol.pairs[i].energy = ev[static_cast<size_t>(i)];
            sol.pairs[i].state_vector.assign(vec.begin() + static_cast<std::ptrdiff_t>(i * dim),
                                             vec.begin() + static_cast<std::ptrdiff_t>((i + 1) * dim));
        }
        return sol;
    }

    bk::Context ctx_;
};
