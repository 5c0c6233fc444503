// bk_subspace.cpp -- the subspace ("sampled-Krylov") solvers and the dense ED of
// diagonalizer.cuh on the GPU.  H restricted to a set of basis states (E(s)/n or
// E(s) on the diagonal, -g between states one spin flip apart) is solved by
// Lanczos over a neighbour-table matvec (bk_subspace_lanczos, the role of
// solve_subspace_lanczos) or densely by rocSOLVER's divide-and-conquer
// eigensolver (bk_subspace_dense, bk_fullspace_dense: the roles of the Eigen
// solvers solve_subspace_dense / solve_fullspace_dense).
#include <hip/hip_runtime.h>
#include <rocblas/rocblas.h>
#include <rocsolver/rocsolver.h>

#include <algorithm>
#include <climits>
#include <utility>
#include <vector>

#include "../../include/bk_capi.h"
#include "bk_ctx.hpp"
#include "bk_host.hpp"
#include "bk_kernels.hpp"

using bkc::ScopedBuf;
using bkc::fail;

namespace {

// the basis in ascending order, with each state's position in the caller's
// list and its diagonal element
struct SortedBasis {
    std::vector<uint64_t> bits;
    std::vector<int64_t> orig;
    std::vector<double> diag;
};

int sort_basis(int n, const uint64_t* basis, int64_t dim, int normalize_by_n, SortedBasis* s)
{
    if (n < 1 || n > 64 || dim < 1 || dim > INT32_MAX || !basis)
        return fail(BK_E_INVALID, "subspace: n=%d, %lld states", n, static_cast<long long>(dim));
    const uint64_t mask = n == 64 ? ~0ull : (1ull << n) - 1ull;
    std::vector<std::pair<uint64_t, int64_t>> v(static_cast<size_t>(dim));
    for (int64_t i = 0; i < dim; ++i) {
        if (basis[i] & ~mask) return fail(BK_E_INVALID, "subspace: state 0x%llx has bits above n=%d",
                                          static_cast<unsigned long long>(basis[i]), n);
        v[static_cast<size_t>(i)] = {basis[i], i};
    }
    std::sort(v.begin(), v.end());
    for (size_t i = 1; i < v.size(); ++i)
        if (v[i].first == v[i - 1].first)
            return fail(BK_E_INVALID, "subspace: state 0x%llx appears twice", static_cast<unsigned long long>(v[i].first));
    const double scale = normalize_by_n ? 1.0 / static_cast<double>(n) : 1.0;
    s->bits.resize(v.size());
    s->orig.resize(v.size());
    s->diag.resize(v.size());
    for (size_t i = 0; i < v.size(); ++i) {
        s->bits[i] = v[i].first;
        s->orig[i] = v[i].second;
        s->diag[i] = static_cast<double>(bk::host_energy(n, v[i].first)) * scale;
    }
    return BK_OK;
}

// device copies of the sorted basis, its diagonal and its neighbour table
struct DeviceBasis {
    ScopedBuf bits, diag, nbr;
};

int upload(bk_ctx* c, int n, const SortedBasis& s, DeviceBasis* d)
{
    const size_t dim = s.bits.size();
    HIPCALL(d->bits.ensure(sizeof(uint64_t) * dim));
    HIPCALL(d->diag.ensure(sizeof(double) * dim));
    HIPCALL(d->nbr.ensure(sizeof(int32_t) * dim * static_cast<size_t>(n)));
    HIPCALL(hipMemcpyAsync(d->bits.p, s.bits.data(), sizeof(uint64_t) * dim, hipMemcpyHostToDevice, c->stream));
    HIPCALL(hipMemcpyAsync(d->diag.p, s.diag.data(), sizeof(double) * dim, hipMemcpyHostToDevice, c->stream));
    HIPCALL(bk::launch_sub_neighbors(d->bits.as<uint64_t>(), static_cast<int64_t>(dim), n, d->nbr.as<int32_t>(),
                                     c->stream));
    return BK_OK;
}

constexpr int64_t kMaxDense = 32768;  // dense H of 8.6 GB

// all eigenpairs of the dense H on the device; the lowest k_pairs returned,
// eigenvectors in the caller's basis order
int dense_solve(bk_ctx* c, int n, double g, const SortedBasis& s, int k_pairs, double* evals, double* evecs)
{
    const int64_t dim = static_cast<int64_t>(s.bits.size());
    if (dim > kMaxDense) return fail(BK_E_INVALID, "dense ED of %lld states (at most %lld)",
                                     static_cast<long long>(dim), static_cast<long long>(kMaxDense));
    DeviceBasis d;
    int rc = upload(c, n, s, &d);
    if (rc) return rc;
    ScopedBuf H, D, E, info;
    HIPCALL(H.ensure(sizeof(double) * static_cast<size_t>(dim) * static_cast<size_t>(dim)));
    HIPCALL(D.ensure(sizeof(double) * static_cast<size_t>(dim)));
    HIPCALL(E.ensure(sizeof(double) * static_cast<size_t>(dim)));
    HIPCALL(info.ensure(sizeof(rocblas_int)));
    HIPCALL(hipMemsetAsync(H.p, 0, sizeof(double) * static_cast<size_t>(dim) * static_cast<size_t>(dim), c->stream));
    HIPCALL(bk::launch_sub_dense(d.diag.as<double>(), d.nbr.as<int32_t>(), dim, n, g, H.as<double>(), c->stream));
    rocblas_handle h = nullptr;
    if (rocblas_create_handle(&h) != rocblas_status_success) return fail(BK_E_HIP, "rocblas_create_handle failed");
    rocblas_status st = rocblas_set_stream(h, c->stream);
    if (st == rocblas_status_success)
        st = rocsolver_dsyevd(h, rocblas_evect_original, rocblas_fill_lower, static_cast<rocblas_int>(dim),
                              H.as<double>(), static_cast<rocblas_int>(dim), D.as<double>(), E.as<double>(),
                              info.as<rocblas_int>());
    const hipError_t sync = hipStreamSynchronize(c->stream);
    rocblas_destroy_handle(h);
    if (st != rocblas_status_success) return fail(BK_E_HIP, "rocsolver_dsyevd: status %d", static_cast<int>(st));
    HIPCALL(sync);
    rocblas_int inf = 0;
    HIPCALL(hipMemcpy(&inf, info.p, sizeof(inf), hipMemcpyDeviceToHost));
    if (inf != 0) return fail(BK_E_INVALID, "rocsolver_dsyevd did not converge (info %d)", static_cast<int>(inf));
    const int kp = static_cast<int>(std::min<int64_t>(k_pairs, dim));
    HIPCALL(hipMemcpy(evals, D.p, sizeof(double) * static_cast<size_t>(kp), hipMemcpyDeviceToHost));
    if (evecs) {
        std::vector<double> col(static_cast<size_t>(dim));
        for (int k = 0; k < kp; ++k) {
            HIPCALL(hipMemcpy(col.data(), H.as<double>() + static_cast<size_t>(k) * static_cast<size_t>(dim),
                              sizeof(double) * static_cast<size_t>(dim), hipMemcpyDeviceToHost));
            double* out = evecs + static_cast<size_t>(k) * static_cast<size_t>(dim);
            for (int64_t i = 0; i < dim; ++i) out[s.orig[static_cast<size_t>(i)]] = col[static_cast<size_t>(i)];
        }
    }
    return kp;
}

}  // namespace

extern "C" {

static int bk_subspace_lanczos_impl(bk_ctx* c, int n, double g, int normalize_by_n, const uint64_t* basis, int64_t dim,
                                    int max_krylov_dim, double tol, int k_pairs, double* evals, double* evecs)
{
    int rc = bkc::check_ctx(c);
    if (rc) return rc;
    if (max_krylov_dim < 1 || k_pairs < 1 || !evals || !bkc::is_fin(g) || !bkc::is_fin(tol))
        return fail(BK_E_INVALID, "bk_subspace_lanczos: bad argument");
    SortedBasis s;
    if ((rc = sort_basis(n, basis, dim, normalize_by_n, &s))) return rc;
    DeviceBasis d;
    if ((rc = upload(c, n, s, &d))) return rc;
    std::vector<double> vec(evecs ? static_cast<size_t>(k_pairs) * static_cast<size_t>(dim) : 0);
    const int kp = bkc::lanczos_core(
        c, dim,
        [&](const double* x, double* y) {
            return bk::launch_sub_hv(d.diag.as<double>(), d.nbr.as<int32_t>(), dim, n, g, x, y, c->stream);
        },
        max_krylov_dim, tol, k_pairs, evals, evecs ? vec.data() : nullptr);
    if (kp < 0 || !evecs) return kp;
    for (int k = 0; k < kp; ++k)
        for (int64_t i = 0; i < dim; ++i)
            evecs[static_cast<size_t>(k) * static_cast<size_t>(dim) + static_cast<size_t>(s.orig[static_cast<size_t>(i)])] =
                vec[static_cast<size_t>(k) * static_cast<size_t>(dim) + static_cast<size_t>(i)];
    return kp;
}

int bk_subspace_lanczos(bk_ctx* c, int n, double g, int normalize_by_n, const uint64_t* basis, int64_t dim,
                        int max_krylov_dim, double tol, int k_pairs, double* evals, double* evecs)
{
    return bkc::guard_abi(
        [&] { return bk_subspace_lanczos_impl(c, n, g, normalize_by_n, basis, dim, max_krylov_dim, tol, k_pairs, evals,
                                              evecs); });
}

static int bk_subspace_dense_impl(bk_ctx* c, int n, double g, int normalize_by_n, const uint64_t* basis, int64_t dim,
                                  int k_pairs, double* evals, double* evecs)
{
    int rc = bkc::check_ctx(c);
    if (rc) return rc;
    if (k_pairs < 1 || !evals || !bkc::is_fin(g)) return fail(BK_E_INVALID, "bk_subspace_dense: bad argument");
    SortedBasis s;
    if ((rc = sort_basis(n, basis, dim, normalize_by_n, &s))) return rc;
    return dense_solve(c, n, g, s, k_pairs, evals, evecs);
}

int bk_subspace_dense(bk_ctx* c, int n, double g, int normalize_by_n, const uint64_t* basis, int64_t dim, int k_pairs,
                      double* evals, double* evecs)
{
    return bkc::guard_abi(
        [&] { return bk_subspace_dense_impl(c, n, g, normalize_by_n, basis, dim, k_pairs, evals, evecs); });
}

static int bk_fullspace_dense_impl(bk_ctx* c, int n, double g, int normalize_by_n, int k_pairs, double* evals,
                                   double* evecs)
{
    int rc = bkc::check_ctx(c);
    if (rc) return rc;
    if (n < 1 || n > 14) return fail(BK_E_INVALID, "Full space dense ED exceeds feasible memory for N > 14 (n=%d)", n);
    if (k_pairs < 1 || !evals || !bkc::is_fin(g)) return fail(BK_E_INVALID, "bk_fullspace_dense: bad argument");
    std::vector<uint64_t> all(size_t(1) << n);
    for (size_t i = 0; i < all.size(); ++i) all[i] = i;
    SortedBasis s;
    if ((rc = sort_basis(n, all.data(), static_cast<int64_t>(all.size()), normalize_by_n, &s))) return rc;
    return dense_solve(c, n, g, s, k_pairs, evals, evecs);
}

int bk_fullspace_dense(bk_ctx* c, int n, double g, int normalize_by_n, int k_pairs, double* evals, double* evecs)
{
    return bkc::guard_abi([&] { return bk_fullspace_dense_impl(c, n, g, normalize_by_n, k_pairs, evals, evecs); });
}

}  // extern "C"
