// bk_host.cpp -- host-side math of the engine; see bk_host.hpp.
#include "bk_host.hpp"

#include <algorithm>
#include <cmath>
#include <random>

namespace bk {

static inline uint64_t low_mask(int n) { return n >= 64 ? ~0ull : ((1ull << n) - 1ull); }

int64_t host_energy(int n, uint64_t s)
{
    const uint64_t b = s & low_mask(n);
    int64_t e = 0;
    for (int k = 1; k < n; ++k) {
        const int64_t c = static_cast<int64_t>(n - k) -
                          2 * static_cast<int64_t>(__builtin_popcountll((b ^ (b >> k)) & low_mask(n - k)));
        e += c * c;
    }
    return e;
}

uint64_t host_canonical(int n, uint64_t s)
{
    const uint64_t mask = low_mask(n);
    const auto reverse = [n](uint64_t x) {
        uint64_t r = 0;
        for (int i = 0; i < n; ++i) r |= ((x >> i) & 1ull) << (n - 1 - i);
        return r;
    };
    // the order-8 group: {identity, alternating flip} x {identity, reversal} x {identity, negation}
    const uint64_t a = s & mask, b = a ^ (0x5555555555555555ull & mask), c = reverse(a), d = reverse(b);
    uint64_t best = a;
    for (uint64_t x : {b, c, d, ~a & mask, ~b & mask, ~c & mask, ~d & mask}) best = x < best ? x : best;
    return best;
}

int64_t host_known_optimal(int n)
{
    static const int64_t tab[33] = {-1, -1, -1, 1,  2,  2,  7,  3,  8,  12, 13, 5,  10, 6,  19, 15, 24,
                                    32, 25, 29, 26, 26, 39, 47, 36, 36, 45, 37, 50, 62, 59, 67, 64};
    return (n < 0 || n > 32) ? -1 : tab[n];
}

int host_d_min(int n, uint64_t s0, int64_t target, int max_d)
{
    if (target < 0) target = host_energy(n, s0);
    const int none = (max_d > 0) ? max_d + 1 : n;
    const int64_t opt = host_known_optimal(n);
    if (opt >= 0 && target <= opt) return none;
    const uint64_t base = s0 & low_mask(n);
    const int lim = (max_d > 0 && max_d <= n) ? max_d : n;
    for (int d = 1; d <= lim; ++d) {
        if (d == n) {
            if (host_energy(n, base ^ low_mask(n)) < target) return n;
            break;
        }
        // next word with the same popcount, in increasing order
        for (uint64_t w = (1ull << d) - 1ull; w < (1ull << n);) {
            if (host_energy(n, base ^ w) < target) return d;
            const uint64_t lsb = w & (0ull - w);
            const uint64_t up = w + lsb;
            if (up == 0) break;
            w = up | (((up ^ w) / lsb) >> 2);
        }
    }
    return none;
}

void host_steepest_descent(int n, uint64_t start, uint64_t* fin, int64_t* fe)
{
    uint64_t cur = start & low_mask(n);
    int64_t e = host_energy(n, cur);
    for (;;) {
        int best = -1;
        int64_t be = e;
        for (int p = 0; p < n; ++p) {
            const int64_t en = host_energy(n, cur ^ (1ull << p));
            if (en < be) {
                be = en;
                best = p;
            }
        }
        if (best < 0) break;
        cur ^= 1ull << best;
        e = be;
    }
    *fin = cur;
    *fe = e;
}

void host_make_trap(int n, uint64_t seed, uint64_t* start, uint64_t* s0, int64_t* e0)
{
    std::mt19937_64 rng(seed);
    *start = rng() & low_mask(n);
    host_steepest_descent(n, *start, s0, e0);
}

// Implicit symmetric tridiagonal QR with Wilkinson shifts, deflating from the
// bottom (Golub & Van Loan, Alg. 8.3.2/8.3.3); rotations accumulated into z.
bool tridiag_eigen(std::vector<double>& d, std::vector<double> e, std::vector<double>* z)
{
    const int m = static_cast<int>(d.size());
    if (m == 0) return true;
    e.resize(m, 0.0);
    if (z) {
        z->assign(static_cast<size_t>(m) * m, 0.0);
        for (int i = 0; i < m; ++i) (*z)[static_cast<size_t>(i) * m + i] = 1.0;
    }
    const double eps = std::numeric_limits<double>::epsilon();
    int hi = m - 1;
    int guard = 0;
    while (hi > 0) {
        if (++guard > 100 * m + 1000) return false;
        // deflate negligible off-diagonals
        for (int i = 0; i < hi; ++i)
            if (std::fabs(e[i]) <= eps * (std::fabs(d[i]) + std::fabs(d[i + 1]))) e[i] = 0.0;
        if (e[hi - 1] == 0.0) {
            --hi;
            continue;
        }
        int lo = hi - 1;
        while (lo > 0 && e[lo - 1] != 0.0) --lo;
        // Wilkinson shift from the trailing 2x2 block
        const double a = d[hi - 1], b = e[hi - 1], c = d[hi];
        const double delta = 0.5 * (a - c);
        const double den = std::fabs(delta) + std::hypot(delta, b);
        const double mu = c - (delta >= 0 ? 1.0 : -1.0) * b * b / (den == 0.0 ? 1.0 : den);
        double x = d[lo] - mu, y = e[lo];
        for (int k = lo; k < hi; ++k) {
            // Givens rotation zeroing y against x
            const double r = std::hypot(x, y);
            const double cs = (r == 0.0) ? 1.0 : x / r, sn = (r == 0.0) ? 0.0 : -y / r;
            if (k > lo) e[k - 1] = r;
            const double dk = d[k], dk1 = d[k + 1], ek = e[k];
            d[k] = cs * cs * dk - 2.0 * cs * sn * ek + sn * sn * dk1;
            d[k + 1] = sn * sn * dk + 2.0 * cs * sn * ek + cs * cs * dk1;
            e[k] = cs * sn * (dk - dk1) + (cs * cs - sn * sn) * ek;
            if (k + 1 < hi) {
                x = e[k];
                y = -sn * e[k + 1];
                e[k + 1] *= cs;
            }
            if (z) {
                double* u = z->data() + static_cast<size_t>(k) * m;
                double* v = z->data() + static_cast<size_t>(k + 1) * m;
                for (int i = 0; i < m; ++i) {
                    const double ui = u[i], vi = v[i];
                    u[i] = cs * ui - sn * vi;
                    v[i] = sn * ui + cs * vi;
                }
            }
        }
    }
    // sort ascending, carrying the vectors
    std::vector<int> idx(m);
    for (int i = 0; i < m; ++i) idx[i] = i;
    std::stable_sort(idx.begin(), idx.end(), [&](int p, int q) { return d[p] < d[q]; });
    std::vector<double> ds(m);
    for (int i = 0; i < m; ++i) ds[i] = d[idx[i]];
    d.swap(ds);
    if (z) {
        std::vector<double> zs(static_cast<size_t>(m) * m);
        for (int i = 0; i < m; ++i)
            std::copy(z->begin() + static_cast<size_t>(idx[i]) * m, z->begin() + static_cast<size_t>(idx[i] + 1) * m,
                      zs.begin() + static_cast<size_t>(i) * m);
        z->swap(zs);
    }
    return true;
}

int build_schedule(const ScheduleIn& in, std::vector<unsigned char>& flags)
{
    const int total = static_cast<int>(std::round(in.t_max / in.dt));
    flags.assign(static_cast<size_t>(total) + 1, 0);
    const double interval = (in.dt_sample > 0.0) ? in.dt_sample
                                                 : (in.t_max / std::max(1, in.num_sample_points));
    double next_sample = in.t_start;
    const double traj_interval = (in.dt_trajectory > 0.0) ? in.dt_trajectory : in.dt;
    double next_traj = 0.0;
    const double sample_end = (in.num_sample_points > 1 && in.dt_sample > 0.0)
                                  ? in.t_start + (in.num_sample_points - 1) * in.dt_sample
                                  : in.t_max;
    int taken = 0;
    for (int step = 0; step <= total; ++step) {
        const double t = step * in.dt;
        unsigned char f = 0;
        if (in.record_trajectory) {
            if (t >= next_traj - 1e-9 && next_traj <= in.t_max + 1e-9) {
                f |= 1;
                next_traj += traj_interval;
            }
        } else if (in.track_p_less && t >= next_sample - 1e-9 && next_sample <= in.t_max + 1e-9) {
            f |= 1;
        }
        if (in.shots_per_point > 0 && t >= in.t_start - 1e-9 && t >= next_sample - 1e-9 &&
            next_sample <= sample_end + 1e-9 && (in.num_sample_points <= 0 || taken < in.num_sample_points)) {
            ++taken;
            f |= 2;
            next_sample += interval;
        } else if (in.shots_per_point == 0 && !in.record_trajectory && in.track_p_less &&
                   t >= next_sample - 1e-9 && next_sample <= in.t_max + 1e-9) {
            next_sample += interval;
        }
        flags[static_cast<size_t>(step)] = f;
    }
    return total;
}

Derived derived_metrics(const std::vector<double>& t, const std::vector<double>& p,
                        const std::vector<double>& dh, double avg_start, double avg_end)
{
    Derived r{0.0, 0.0, 0.0, 0.0, 0.0};
    const size_t len = p.size();
    if (len > 0) {
        size_t first_drop = len - 1;
        for (size_t i = 1; i < len; ++i)
            if (p[i] < p[i - 1]) {
                first_drop = i - 1;
                break;
            }
        r.rise_avg = (t[first_drop] > 0.0) ? p[first_drop] / t[first_drop] : 0.0;
        double ms = (t[0] > 1e-9) ? p[0] / t[0] : 0.0;
        for (size_t i = 0; i < first_drop; ++i) {
            const double step = t[i + 1] - t[i];
            if (step > 1e-9) ms = std::max(ms, (p[i + 1] - p[i]) / step);
        }
        r.rise_max = ms;
        for (size_t i = 0; i < len; ++i)
            if (p[i] > r.peak_p) {
                r.peak_p = p[i];
                r.peak_t = t[i];
            }
    }
    double sum = 0.0;
    size_t cnt = 0;
    for (size_t i = 0; i < t.size() && i < dh.size(); ++i)
        if (t[i] >= avg_start - 1e-9 && t[i] <= avg_end + 1e-9) {
            sum += dh[i];
            ++cnt;
        }
    r.window_dh = cnt > 0 ? sum / static_cast<double>(cnt) : (dh.empty() ? 0.0 : dh.back());
    return r;
}

}  // namespace bk
