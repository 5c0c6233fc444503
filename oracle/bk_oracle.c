/*
 * bk_oracle.c -- CPU restatement (FP64, C11 + OpenMP) of the reference's path.
 *
 * TEST INFRASTRUCTURE ONLY: the checker for the HIP product path and the timed
 * CPU baseline.  Nothing in bernasconi-krylov_amd/ links or calls this file.
 * See bk_oracle.h for the list of checks that pin each function.
 */
#include "bk_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

static inline uint64_t mask_n(int n) { return n >= 64 ? ~0ULL : ((1ULL << n) - 1ULL); }

/* ------------------------------------------------------------------------ */
/* L0: energies.  labs_energy.hpp:121-132 (calcEnergy), 139-149 (calcCk).    */
/* C_k = sum_i s_i s_{i+k} counted as (N-k) - 2 * #(i : bit_i != bit_{i+k}). */
/* ------------------------------------------------------------------------ */
int64_t bko_energy(int n, uint64_t s)
{
    const uint64_t b = s & mask_n(n);
    int64_t e = 0;
    for (int k = 1; k < n; ++k) {
        const uint64_t pairs = (b ^ (b >> k)) & ((1ULL << (n - k)) - 1ULL);
        const int64_t c = (int64_t)(n - k) - 2 * (int64_t)__builtin_popcountll(pairs);
        e += c * c;
    }
    return e;
}

void bko_ck(int n, uint64_t s, int64_t *ck)
{
    const uint64_t b = s & mask_n(n);
    ck[0] = 0;
    for (int k = 1; k < n; ++k) {
        const uint64_t pairs = (b ^ (b >> k)) & ((1ULL << (n - k)) - 1ULL);
        ck[k] = (int64_t)(n - k) - 2 * (int64_t)__builtin_popcountll(pairs);
    }
}

/* labs_energy.hpp:40-48: reverse the lowest n bits. */
static uint64_t rev_n(uint64_t x, int n)
{
    uint64_t r = 0;
    for (int i = 0; i < n; ++i) r |= ((x >> i) & 1ULL) << (n - 1 - i);
    return r;
}

/* labs_energy.hpp:51-78: minimum over {id, alternate-negate, reverse,
 * reverse(alternate)} x {identity, global negation}. */
uint64_t bko_canonical(int n, uint64_t s)
{
    const uint64_t m = mask_n(n);
    const uint64_t alt = 0x5555555555555555ULL & m;
    uint64_t c[8];
    c[0] = s & m;
    c[1] = c[0] ^ alt;
    c[2] = rev_n(c[0], n);
    c[3] = rev_n(c[1], n);
    for (int i = 0; i < 4; ++i) c[4 + i] = (~c[i]) & m;
    uint64_t best = c[0];
    for (int i = 1; i < 8; ++i)
        if (c[i] < best) best = c[i];
    return best;
}

/* labs_energy.hpp:156-168: known optimal LABS energies E*(N), N = 3..32. */
int64_t bko_known_optimal_energy(int n)
{
    static const int64_t tab[33] = {-1, -1, -1, 1,  2,  2,  7,  3,  8,  12, 13,
                                    5,  10, 6,  19, 15, 24, 32, 25, 29, 26, 26,
                                    39, 47, 36, 36, 45, 37, 50, 62, 59, 67, 64};
    if (n < 0 || n > 32) return -1;
    return tab[n];
}

/* labs_energy.hpp:175-207: smallest Hamming radius d at which some state has
 * E < target.  Spheres are walked with the next-combination-of-d-bits rule;
 * radius n is the single complement state.  Returns (max_d>0 ? max_d+1 : n)
 * when nothing is found, and immediately when target <= E*(N). */
int bko_d_min(int n, uint64_t s0, int64_t target, int max_d)
{
    if (target < 0) target = bko_energy(n, s0);
    const int64_t opt = bko_known_optimal_energy(n);
    const int not_found = (max_d > 0) ? (max_d + 1) : n;
    if (opt >= 0 && target <= opt) return not_found;
    const uint64_t base = s0 & mask_n(n);
    const int limit_d = (max_d > 0 && max_d <= n) ? max_d : n;
    for (int d = 1; d <= limit_d; ++d) {
        if (d == n) {
            if (bko_energy(n, base ^ mask_n(n)) < target) return n;
            break;
        }
        /* enumerate all n-bit words with exactly d ones in increasing order */
        uint64_t w = (1ULL << d) - 1ULL;
        const uint64_t end = 1ULL << n;
        while (w < end) {
            if (bko_energy(n, base ^ w) < target) return d;
            const uint64_t low = w & (~w + 1ULL);  /* lowest set bit */
            const uint64_t ripple = w + low;
            if (ripple == 0) break;
            w = ripple | (((w ^ ripple) >> 2) / low);
        }
    }
    return not_found;
}

void bko_all_energies(int n, int32_t *out)
{
    const int64_t dim = 1LL << n;
#pragma omp parallel for schedule(static)
    for (int64_t s = 0; s < dim; ++s) out[s] = (int32_t)bko_energy(n, (uint64_t)s);
}

int64_t bko_count_below(int n, int64_t e0)
{
    const int64_t dim = 1LL << n;
    int64_t cnt = 0;
#pragma omp parallel for schedule(static) reduction(+ : cnt)
    for (int64_t s = 0; s < dim; ++s) cnt += (bko_energy(n, (uint64_t)s) < e0);
    return cnt;
}

/* ------------------------------------------------------------------------ */
/* mt19937_64 (the generator std::mt19937_64 is defined to be: C++11        */
/* [rand.predef]) and libstdc++'s uniform_real_distribution over it.        */
/* ------------------------------------------------------------------------ */
#define MT_N 312
#define MT_M 156
void bko_mt64_seed(bko_mt64 *g, uint64_t seed)
{
    g->mt[0] = seed;
    for (int i = 1; i < MT_N; ++i)
        g->mt[i] = 6364136223846793005ULL * (g->mt[i - 1] ^ (g->mt[i - 1] >> 62)) + (uint64_t)i;
    g->idx = MT_N;
}

uint64_t bko_mt64_next(bko_mt64 *g)
{
    if (g->idx >= MT_N) {
        const uint64_t upper = 0xFFFFFFFF80000000ULL, lower = 0x7FFFFFFFULL;
        for (int i = 0; i < MT_N; ++i) {
            const uint64_t y = (g->mt[i] & upper) | (g->mt[(i + 1) % MT_N] & lower);
            uint64_t v = g->mt[(i + MT_M) % MT_N] ^ (y >> 1);
            if (y & 1ULL) v ^= 0xB5026F5AA96619E9ULL;
            g->mt[i] = v;
        }
        g->idx = 0;
    }
    uint64_t x = g->mt[g->idx++];
    x ^= (x >> 29) & 0x5555555555555555ULL;
    x ^= (x << 17) & 0x71D67FFFEDA60000ULL;
    x ^= (x << 37) & 0xFFF7EEE000000000ULL;
    x ^= x >> 43;
    return x;
}

/* libstdc++ generate_canonical<double,53> takes one 64-bit draw x and returns
 * (double)x / 2^64, clamped below 1; uniform_real_distribution returns
 * u*(b-a)+a. */
double bko_mt64_uniform(bko_mt64 *g, double a, double b)
{
    double u = (double)bko_mt64_next(g) / 18446744073709551616.0;
    if (u >= 1.0) u = nextafter(1.0, 0.0);
    return u * (b - a) + a;
}

/* classical.cuh:602-604 with a freshly seeded generator (one draw). */
uint64_t bko_random_state(int n, uint64_t seed)
{
    bko_mt64 g;
    bko_mt64_seed(&g, seed);
    return bko_mt64_next(&g) & mask_n(n);
}

/* classical.cuh:625-713: best-improvement single flips, strict '<' so the
 * lowest flipped index wins ties; one evaluation for the start plus n per
 * pass (the final pass that finds no improvement included). */
void bko_steepest_descent(int n, uint64_t start, uint64_t *final_state, int64_t *final_energy,
                          uint64_t *evals, uint64_t *steps)
{
    uint64_t cur = start & mask_n(n);
    int64_t e = bko_energy(n, cur);
    uint64_t ev = 1, st = 0;
    for (;;) {
        int best = -1;
        int64_t best_e = e;
        for (int p = 0; p < n; ++p) {
            const int64_t en = bko_energy(n, cur ^ (1ULL << p));
            if (en < best_e) {
                best_e = en;
                best = p;
            }
        }
        ev += (uint64_t)n;
        if (best < 0) break;
        cur ^= 1ULL << best;
        e = best_e;
        ++st;
    }
    *final_state = cur;
    *final_energy = e;
    if (evals) *evals = ev;
    if (steps) *steps = st;
}

/* ------------------------------------------------------------------------ */
/* FP64 statevector Trotter loop.                                           */
/*  phase: sampler.cuh:73-91, theta = -(dt/n) E(s)                           */
/*  mixer: sampler.cuh:451-457 + 612-617, per qubit [[c, i s],[i s, c]],     */
/*         c = cos(gdt), s = sin(gdt) (= Rx(-2 g dt))                        */
/*  observables: sampler.cuh:149-195                                         */
/* ------------------------------------------------------------------------ */
static void apply_phase(int n, double *psi, const int32_t *en, double dtp)
{
    const int64_t dim = 1LL << n;
    const double w = -(dtp / (double)n);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < dim; ++i) {
        const double th = w * (double)en[i];
        const double pr = cos(th), pi = sin(th);
        const double ar = psi[2 * i], ai = psi[2 * i + 1];
        psi[2 * i] = ar * pr - ai * pi;
        psi[2 * i + 1] = ar * pi + ai * pr;
    }
}

static void apply_mixer(int n, double *psi, double g, double dt)
{
    const double rx = -2.0 * g * dt;
    const double c = cos(rx / 2.0), s = -sin(rx / 2.0); /* matrix = [[c, i s],[i s, c]] */
    const int64_t half = 1LL << (n - 1);
    for (int q = 0; q < n; ++q) {
        const int64_t st = 1LL << q;
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < half; ++i) {
            const int64_t lo = ((i >> q) << (q + 1)) | (i & (st - 1));
            const int64_t hi = lo | st;
            const double ar = psi[2 * lo], ai = psi[2 * lo + 1];
            const double br = psi[2 * hi], bi = psi[2 * hi + 1];
            psi[2 * lo] = c * ar - s * bi;
            psi[2 * lo + 1] = c * ai + s * br;
            psi[2 * hi] = c * br - s * ai;
            psi[2 * hi + 1] = c * bi + s * ar;
        }
    }
}

static void observables_e(int n, const double *psi, const int32_t *en, uint64_t s0, int64_t e0,
                          double *pl, double *dh, double *nm)
{
    const int64_t dim = 1LL << n;
    double a = 0.0, b = 0.0, c = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : a, b, c)
    for (int64_t i = 0; i < dim; ++i) {
        const double p = psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        const int64_t e = en ? (int64_t)en[i] : bko_energy(n, (uint64_t)i);
        if (e < e0) a += p;
        b += p * (double)__builtin_popcountll((uint64_t)i ^ s0);
        c += p;
    }
    if (pl) *pl = a;
    if (dh) *dh = b;
    if (nm) *nm = c;
}

void bko_observables(int n, const double *psi, uint64_t s0, int64_t e0, double *p_less,
                     double *dh, double *norm)
{
    observables_e(n, psi, NULL, s0, e0, p_less, dh, norm);
}

int bko_trotter(int n, uint64_t s0, int64_t e0, double g, const double *g_steps, double dt,
                int steps, int order, int init_from_s0, double *psi, double *p_less, double *dh,
                double *norm, int threads)
{
    if (n < 1 || n > 34 || steps < 0 || (order != 1 && order != 2) || !psi) return -1;
#ifdef _OPENMP
    int saved = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const int64_t dim = 1LL << n;
    s0 &= mask_n(n);
    int32_t *en = (int32_t *)malloc(sizeof(int32_t) * (size_t)dim);
    if (!en) return -2;
    bko_all_energies(n, en);
    if (init_from_s0) {
#pragma omp parallel for schedule(static)
        for (int64_t i = 0; i < 2 * dim; ++i) psi[i] = 0.0;
        psi[2 * s0] = 1.0;
    }
    for (int k = 0; k <= steps; ++k) {
        if (p_less || dh || norm) {
            double a, b, c;
            observables_e(n, psi, en, s0, e0, &a, &b, &c);
            if (p_less) p_less[k] = a;
            if (dh) dh[k] = b;
            if (norm) norm[k] = c;
        }
        if (k == steps) break;
        const double gk = g_steps ? g_steps[k] : g;
        if (order == 1) {
            apply_phase(n, psi, en, dt);
            apply_mixer(n, psi, gk, dt);
        } else {
            apply_phase(n, psi, en, 0.5 * dt);
            apply_mixer(n, psi, gk, dt);
            apply_phase(n, psi, en, 0.5 * dt);
        }
    }
    free(en);
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return 0;
}

/* y = sum_q X_q x (every spin flipped once) */
static void apply_sx(int n, const double *x, double *y)
{
    const int64_t dim = 1LL << n;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < dim; ++i) {
        double yr = 0.0, yi = 0.0;
        for (int q = 0; q < n; ++q) {
            const int64_t j = i ^ (1LL << q);
            yr += x[2 * j];
            yi += x[2 * j + 1];
        }
        y[2 * i] = yr;
        y[2 * i + 1] = yi;
    }
}

/* U^+ of one step (order 1: U = X P(dt); order 2: U = P(dt/2) X P(dt/2)) */
static void step_adjoint(int n, double *psi, const int32_t *en, double g, double dt, int order)
{
    if (order == 1) {
        apply_mixer(n, psi, -g, dt);
        apply_phase(n, psi, en, -dt);
    } else {
        apply_phase(n, psi, en, -0.5 * dt);
        apply_mixer(n, psi, -g, dt);
        apply_phase(n, psi, en, -0.5 * dt);
    }
}

int bko_adjoint_grad(int n, uint64_t s0, int64_t e0, const double *g_steps, double dt,
                     int steps, int order, double wp, double wd, double *grad, double *obs2,
                     int threads)
{
    if (n < 1 || n > 30 || steps < 1 || (order != 1 && order != 2) || !g_steps || !grad) return -1;
#ifdef _OPENMP
    int saved = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const int64_t dim = 1LL << n;
    s0 &= mask_n(n);
    int32_t *en = (int32_t *)malloc(sizeof(int32_t) * (size_t)dim);
    double *psi = (double *)malloc(sizeof(double) * 2 * (size_t)dim);
    double *lam = (double *)malloc(sizeof(double) * 2 * (size_t)dim);
    double *tmp = (double *)malloc(sizeof(double) * 2 * (size_t)dim);
    double *dpsi = (double *)malloc(sizeof(double) * 2 * (size_t)dim);
    if (!en || !psi || !lam || !tmp || !dpsi) {
        free(en);
        free(psi);
        free(lam);
        free(tmp);
        free(dpsi);
        return -2;
    }
    bko_all_energies(n, en);
    /* forward */
    for (int64_t i = 0; i < 2 * dim; ++i) psi[i] = 0.0;
    psi[2 * s0] = 1.0;
    for (int k = 0; k < steps; ++k) {
        if (order == 1) {
            apply_phase(n, psi, en, dt);
            apply_mixer(n, psi, g_steps[k], dt);
        } else {
            apply_phase(n, psi, en, 0.5 * dt);
            apply_mixer(n, psi, g_steps[k], dt);
            apply_phase(n, psi, en, 0.5 * dt);
        }
    }
    if (obs2) observables_e(n, psi, en, s0, e0, &obs2[0], &obs2[1], NULL);
    /* lambda_T = W psi_T */
    for (int64_t i = 0; i < dim; ++i) {
        const double w = (en[i] < e0 ? wp : 0.0) + wd * (double)__builtin_popcountll((uint64_t)i ^ s0);
        lam[2 * i] = w * psi[2 * i];
        lam[2 * i + 1] = w * psi[2 * i + 1];
    }
    for (int k = steps - 1; k >= 0; --k) {
        const double g = g_steps[k];
        step_adjoint(n, psi, en, g, dt, order); /* psi <- psi_{k} (state before step k) */
        /* dpsi = dU/dg psi: the step's factors with i dt sum_q X_q after the mixer */
        for (int64_t i = 0; i < 2 * dim; ++i) tmp[i] = psi[i];
        if (order == 1) {
            apply_phase(n, tmp, en, dt);
            apply_mixer(n, tmp, g, dt);
            apply_sx(n, tmp, dpsi);
        } else {
            apply_phase(n, tmp, en, 0.5 * dt);
            apply_mixer(n, tmp, g, dt);
            apply_sx(n, tmp, dpsi);
            apply_phase(n, dpsi, en, 0.5 * dt);
        }
        /* grad = 2 Re <lam| i dt dpsi> = -2 dt Im <lam|dpsi> */
        double im = 0.0;
        for (int64_t i = 0; i < dim; ++i) im += lam[2 * i] * dpsi[2 * i + 1] - lam[2 * i + 1] * dpsi[2 * i];
        grad[k] = -2.0 * dt * im;
        step_adjoint(n, lam, en, g, dt, order);
    }
    free(en);
    free(psi);
    free(lam);
    free(tmp);
    free(dpsi);
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return 0;
}

int bko_trotter_bench(int n, uint64_t s0, double g, double dt, int steps, int order, int threads,
                      double *seconds)
{
    if (n < 1 || n > 34 || steps < 1 || (order != 1 && order != 2) || !seconds) return -1;
#ifdef _OPENMP
    int saved = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const int64_t dim = 1LL << n;
    int32_t *en = (int32_t *)malloc(sizeof(int32_t) * (size_t)dim);
    double *psi = (double *)malloc(sizeof(double) * 2 * (size_t)dim);
    if (!en || !psi) {
        free(en);
        free(psi);
        return -2;
    }
    bko_all_energies(n, en);
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < 2 * dim; ++i) psi[i] = 0.0;
    psi[2 * (s0 & mask_n(n))] = 1.0;
#ifdef _OPENMP
    const double t0 = omp_get_wtime();
#endif
    for (int k = 0; k < steps; ++k) {
        if (order == 1) {
            apply_phase(n, psi, en, dt);
            apply_mixer(n, psi, g, dt);
        } else {
            apply_phase(n, psi, en, 0.5 * dt);
            apply_mixer(n, psi, g, dt);
            apply_phase(n, psi, en, 0.5 * dt);
        }
    }
#ifdef _OPENMP
    *seconds = omp_get_wtime() - t0;
    omp_set_num_threads(saved);
#else
    *seconds = -1.0;
#endif
    free(en);
    free(psi);
    return 0;
}

/* ------------------------------------------------------------------------ */
/* Matrix-free H.v: diagonalizer.cuh:60-77 with diag = E/N (:391-394).      */
/* ------------------------------------------------------------------------ */
void bko_hv_real(int n, double g, int normalize_by_n, const double *x, double *y)
{
    const int64_t dim = 1LL << n;
    const double scale = normalize_by_n ? 1.0 / (double)n : 1.0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < dim; ++i) {
        double yi = ((double)bko_energy(n, (uint64_t)i) * scale) * x[i];
        for (int j = 0; j < n; ++j) yi -= g * x[i ^ (1LL << j)];
        y[i] = yi;
    }
}

void bko_hv_complex(int n, double g, int normalize_by_n, const double *x, double *y)
{
    const int64_t dim = 1LL << n;
    const double scale = normalize_by_n ? 1.0 / (double)n : 1.0;
#pragma omp parallel for schedule(static)
    for (int64_t i = 0; i < dim; ++i) {
        const double d = (double)bko_energy(n, (uint64_t)i) * scale;
        double yr = d * x[2 * i], yi = d * x[2 * i + 1];
        for (int j = 0; j < n; ++j) {
            const int64_t k = i ^ (1LL << j);
            yr -= g * x[2 * k];
            yi -= g * x[2 * k + 1];
        }
        y[2 * i] = yr;
        y[2 * i + 1] = yi;
    }
}

/* ------------------------------------------------------------------------ */
/* Dense ED: diagonalizer.cuh:250-294 (matrix assembly), with an own         */
/* symmetric eigensolver (Householder tridiagonalisation + implicit-shift   */
/* QL, Golub & Van Loan ch. 8) in place of Eigen::SelfAdjointEigenSolver.   */
/* ------------------------------------------------------------------------ */
void bko_dense_h(int n, double g, int normalize_by_n, double *h)
{
    const int64_t dim = 1LL << n;
    const double scale = normalize_by_n ? 1.0 / (double)n : 1.0;
    memset(h, 0, sizeof(double) * (size_t)(dim * dim));
    for (int64_t a = 0; a < dim; ++a) {
        h[a * dim + a] = (double)bko_energy(n, (uint64_t)a) * scale;
        for (int bit = 0; bit < n; ++bit) {
            const int64_t b = a ^ (1LL << bit);
            h[a * dim + b] = -g;
        }
    }
}

int bko_tridiag_eig(int dim, double *d, double *e_in, double *zt)
{
    /* work on a copy of the off-diagonal extended by one zero */
    double *e = (double *)calloc((size_t)dim + 1, sizeof(double));
    if (!e) return -2;
    for (int i = 0; i + 1 < dim; ++i) e[i] = e_in[i];
    const double eps = 2.220446049250313e-16;
    for (int l = 0; l < dim; ++l) {
        int iter = 0;
        for (;;) {
            int m = l;
            for (; m + 1 < dim; ++m) {
                const double dd = fabs(d[m]) + fabs(d[m + 1]);
                if (fabs(e[m]) <= eps * dd) break;
            }
            if (m == l) break;
            if (++iter > 200) {
                free(e);
                return -3;
            }
            /* Wilkinson-type shift from the leading 2x2 block of [l, m] */
            double q = (d[l + 1] - d[l]) / (2.0 * e[l]);
            double r = hypot(q, 1.0);
            q = d[m] - d[l] + e[l] / (q + (q >= 0 ? r : -r));
            double sn = 1.0, cs = 1.0, p = 0.0;
            int i;
            int deflated = 0;
            for (i = m - 1; i >= l; --i) {
                const double f = sn * e[i];
                const double b = cs * e[i];
                r = hypot(f, q);
                e[i + 1] = r;
                if (r == 0.0) {
                    d[i + 1] -= p;
                    e[m] = 0.0;
                    deflated = 1;
                    break;
                }
                sn = f / r;
                cs = q / r;
                q = d[i + 1] - p;
                r = (d[i] - q) * sn + 2.0 * cs * b;
                p = sn * r;
                d[i + 1] = q + p;
                q = cs * r - b;
                if (zt) {
                    double *u = zt + (size_t)i * dim, *v = zt + (size_t)(i + 1) * dim;
                    for (int k = 0; k < dim; ++k) {
                        const double t = v[k];
                        v[k] = sn * u[k] + cs * t;
                        u[k] = cs * u[k] - sn * t;
                    }
                }
            }
            if (deflated) continue;
            d[l] -= p;
            e[l] = q;
            e[m] = 0.0;
        }
    }
    free(e);
    /* selection sort ascending (rows of zt follow) */
    for (int i = 0; i + 1 < dim; ++i) {
        int k = i;
        for (int j = i + 1; j < dim; ++j)
            if (d[j] < d[k]) k = j;
        if (k != i) {
            const double t = d[i];
            d[i] = d[k];
            d[k] = t;
            if (zt) {
                double *u = zt + (size_t)i * dim, *v = zt + (size_t)k * dim;
                for (int c = 0; c < dim; ++c) {
                    const double w = u[c];
                    u[c] = v[c];
                    v[c] = w;
                }
            }
        }
    }
    return 0;
}

int bko_syev(int dim, double *a, double *w)
{
    /* Householder reduction: A = Q T Q^T, Q = H_0 ... H_{dim-3} */
    double *vs = (double *)calloc((size_t)dim * dim, sizeof(double)); /* v_k in row k */
    double *beta = (double *)calloc((size_t)dim, sizeof(double));
    double *p = (double *)malloc(sizeof(double) * (size_t)dim);
    double *e = (double *)calloc((size_t)dim, sizeof(double));
    if (!vs || !beta || !p || !e) {
        free(vs); free(beta); free(p); free(e);
        return -2;
    }
    for (int k = 0; k + 2 < dim; ++k) {
        const int m = dim - k - 1;
        double *v = vs + (size_t)k * dim + (k + 1);
        double xn = 0.0;
        for (int i = 0; i < m; ++i) {
            v[i] = a[(size_t)(k + 1 + i) * dim + k];
            xn += v[i] * v[i];
        }
        xn = sqrt(xn);
        if (xn == 0.0) { beta[k] = 0.0; continue; }
        const double alpha = (v[0] >= 0.0) ? -xn : xn;
        v[0] -= alpha;
        double vn = 0.0;
        for (int i = 0; i < m; ++i) vn += v[i] * v[i];
        if (vn == 0.0) { beta[k] = 0.0; continue; }
        const double bt = 2.0 / vn;
        beta[k] = bt;
        /* p = bt * A_sub v */
#pragma omp parallel for schedule(static)
        for (int i = 0; i < m; ++i) {
            const double *row = a + (size_t)(k + 1 + i) * dim + (k + 1);
            double s = 0.0;
            for (int j = 0; j < m; ++j) s += row[j] * v[j];
            p[i] = bt * s;
        }
        double kk = 0.0;
        for (int i = 0; i < m; ++i) kk += v[i] * p[i];
        kk *= 0.5 * bt;
        for (int i = 0; i < m; ++i) p[i] -= kk * v[i]; /* p is now w */
#pragma omp parallel for schedule(static)
        for (int i = 0; i < m; ++i) {
            double *row = a + (size_t)(k + 1 + i) * dim + (k + 1);
            const double vi = v[i], wi = p[i];
            for (int j = 0; j < m; ++j) row[j] -= vi * p[j] + wi * v[j];
        }
        a[(size_t)(k + 1) * dim + k] = alpha;
        a[(size_t)k * dim + (k + 1)] = alpha;
        for (int i = 1; i < m; ++i) {
            a[(size_t)(k + 1 + i) * dim + k] = 0.0;
            a[(size_t)k * dim + (k + 1 + i)] = 0.0;
        }
    }
    for (int i = 0; i < dim; ++i) w[i] = a[(size_t)i * dim + i];
    for (int i = 0; i + 1 < dim; ++i) e[i] = a[(size_t)(i + 1) * dim + i];
    /* Q by backward accumulation, then transpose into a (rows = Q^T rows) */
    double *q = a;
    memset(q, 0, sizeof(double) * (size_t)dim * dim);
    for (int i = 0; i < dim; ++i) q[(size_t)i * dim + i] = 1.0;
    for (int k = dim - 3; k >= 0; --k) {
        if (beta[k] == 0.0) continue;
        const double *v = vs + (size_t)k * dim + (k + 1);
        const int m = dim - k - 1;
        /* rows k+1.. , columns k+1.. : Q_sub -= beta v (v^T Q_sub) */
        for (int c = 0; c < m; ++c) p[c] = 0.0;
        for (int i = 0; i < m; ++i) {
            const double *row = q + (size_t)(k + 1 + i) * dim + (k + 1);
            const double vi = v[i];
            for (int c = 0; c < m; ++c) p[c] += vi * row[c];
        }
#pragma omp parallel for schedule(static)
        for (int i = 0; i < m; ++i) {
            double *row = q + (size_t)(k + 1 + i) * dim + (k + 1);
            const double f = beta[k] * v[i];
            for (int c = 0; c < m; ++c) row[c] -= f * p[c];
        }
    }
    for (int i = 0; i < dim; ++i)
        for (int j = i + 1; j < dim; ++j) {
            const double t = q[(size_t)i * dim + j];
            q[(size_t)i * dim + j] = q[(size_t)j * dim + i];
            q[(size_t)j * dim + i] = t;
        }
    int rc = bko_tridiag_eig(dim, w, e, q);
    free(vs); free(beta); free(p); free(e);
    return rc;
}

/* connector.cuh:683-748: psi(t) = sum_j v_j[s0] exp(-i E_j t) v_j; P_< and
 * <d_H> of |psi(t)|^2.  inner_loop_trig = 1 evaluates cos/sin inside the
 * state x eigenpair loop as the reference does (O(T 4^N) trig calls);
 * 0 hoists the per-eigenpair phases (same sums, far fewer trig calls). */
int bko_ed_evolve(int n, double g, uint64_t s0, int64_t e0, const double *t, int nt,
                  double *p_less, double *dh, int inner_loop_trig, int threads)
{
#ifdef _OPENMP
    int saved = omp_get_max_threads();
    if (threads > 0) omp_set_num_threads(threads);
#else
    (void)threads;
#endif
    const int dim = 1 << n;
    s0 &= mask_n(n);
    double *h = (double *)malloc(sizeof(double) * (size_t)dim * dim);
    double *w = (double *)malloc(sizeof(double) * (size_t)dim);
    double *cre = (double *)malloc(sizeof(double) * (size_t)dim);
    double *cim = (double *)malloc(sizeof(double) * (size_t)dim);
    int32_t *en = (int32_t *)malloc(sizeof(int32_t) * (size_t)dim);
    if (!h || !w || !cre || !cim || !en) {
        free(h); free(w); free(cre); free(cim); free(en);
        return -2;
    }
    bko_dense_h(n, g, 1, h);
    int rc = bko_syev(dim, h, w);
    if (rc == 0) {
        bko_all_energies(n, en);
        for (int ti = 0; ti < nt; ++ti) {
            double pl = 0.0, dd = 0.0;
            if (!inner_loop_trig) {
                for (int j = 0; j < dim; ++j) {
                    const double ang = -w[j] * t[ti];
                    const double cj = h[(size_t)j * dim + s0];
                    cre[j] = cj * cos(ang);
                    cim[j] = cj * sin(ang);
                }
            }
#pragma omp parallel for schedule(static) reduction(+ : pl, dd)
            for (int s = 0; s < dim; ++s) {
                double re = 0.0, im = 0.0;
                if (inner_loop_trig) {
                    for (int j = 0; j < dim; ++j) {
                        const double ang = -w[j] * t[ti];
                        const double amp = h[(size_t)j * dim + s0] * h[(size_t)j * dim + s];
                        re += amp * cos(ang);
                        im += amp * sin(ang);
                    }
                } else {
                    for (int j = 0; j < dim; ++j) {
                        const double v = h[(size_t)j * dim + s];
                        re += cre[j] * v;
                        im += cim[j] * v;
                    }
                }
                const double a2 = re * re + im * im;
                if (en[s] < e0) pl += a2;
                dd += a2 * (double)__builtin_popcountll((uint64_t)s ^ s0);
            }
            p_less[ti] = pl;
            dh[ti] = dd;
        }
    }
    free(h); free(w); free(cre); free(cim); free(en);
#ifdef _OPENMP
    omp_set_num_threads(saved);
#endif
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Lanczos: diagonalizer.cuh:420-554.                                        */
/* ------------------------------------------------------------------------ */
static double dot_r(int64_t len, const double *x, const double *y)
{
    double s = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : s)
    for (int64_t i = 0; i < len; ++i) s += x[i] * y[i];
    return s;
}

int bko_lanczos(int n, double g, int normalize_by_n, int max_krylov_dim, double tol, int k_pairs,
                double *evals, double *evecs)
{
    const int64_t dim = 1LL << n;
    int m = (int)((dim < (int64_t)max_krylov_dim) ? dim : max_krylov_dim);
    double *V = (double *)malloc(sizeof(double) * (size_t)dim * (size_t)m);
    double *wv = (double *)malloc(sizeof(double) * (size_t)dim);
    double *al = (double *)calloc((size_t)m, sizeof(double));
    double *be = (double *)calloc((size_t)m, sizeof(double));
    if (!V || !wv || !al || !be) {
        free(V); free(wv); free(al); free(be);
        return -2;
    }
    bko_mt64 rng;
    bko_mt64_seed(&rng, 42);
    for (int64_t i = 0; i < dim; ++i) V[i] = bko_mt64_uniform(&rng, -0.5, 0.5);
    const double nv0 = sqrt(dot_r(dim, V, V));
    const double inv0 = 1.0 / nv0;
    for (int64_t i = 0; i < dim; ++i) V[i] *= inv0;
    int m_act = m;
    for (int j = 0; j < m; ++j) {
        double *vj = V + (size_t)j * dim;
        bko_hv_real(n, g, normalize_by_n, vj, wv);
        if (j > 0) {
            const double *vp = V + (size_t)(j - 1) * dim;
            for (int64_t i = 0; i < dim; ++i) wv[i] -= be[j - 1] * vp[i];
        }
        al[j] = dot_r(dim, vj, wv);
        for (int64_t i = 0; i < dim; ++i) wv[i] -= al[j] * vj[i];
        if (j < m - 1) {
            be[j] = sqrt(dot_r(dim, wv, wv));
            if (fabs(be[j]) < tol) {
                m_act = j + 1;
                break;
            }
            double *vn = V + (size_t)(j + 1) * dim;
            const double ib = 1.0 / be[j];
            for (int64_t i = 0; i < dim; ++i) vn[i] = wv[i] * ib;
        }
    }
    double *d = (double *)malloc(sizeof(double) * (size_t)m_act);
    double *z = (double *)calloc((size_t)m_act * m_act, sizeof(double));
    for (int i = 0; i < m_act; ++i) {
        d[i] = al[i];
        z[(size_t)i * m_act + i] = 1.0;
    }
    int rc = bko_tridiag_eig(m_act, d, be, z);
    if (rc == 0) {
        const int kp = k_pairs < m_act ? k_pairs : m_act;
        for (int k = 0; k < kp; ++k) {
            evals[k] = d[k];
            if (evecs) {
                double *x = evecs + (size_t)k * dim;
                for (int64_t i = 0; i < dim; ++i) x[i] = 0.0;
                for (int j = 0; j < m_act; ++j) {
                    const double c = z[(size_t)k * m_act + j];
                    const double *vj = V + (size_t)j * dim;
                    for (int64_t i = 0; i < dim; ++i) x[i] += c * vj[i];
                }
            }
        }
        rc = kp;
    }
    free(V); free(wv); free(al); free(be); free(d); free(z);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Krylov propagator exp(-i H tau) psi: Lanczos with full reorthogonalisation */
/* (Saad 1992; Hochbruck & Lubich 1997), T exponentiated through its        */
/* eigendecomposition.  No reference counterpart (the reference has only the */
/* eigen-solver); H is the same operator as diagonalizer.cuh:60-77.         */
/* ------------------------------------------------------------------------ */
int bko_krylov_propagate(int n, double g, int normalize_by_n, double *psi, double tau, int m)
{
    const int64_t dim = 1LL << n;
    double *V = (double *)malloc(sizeof(double) * 2 * (size_t)dim * (size_t)m);
    double *w = (double *)malloc(sizeof(double) * 2 * (size_t)dim);
    double *al = (double *)calloc((size_t)m, sizeof(double));
    double *be = (double *)calloc((size_t)m, sizeof(double));
    if (!V || !w || !al || !be) {
        free(V); free(w); free(al); free(be);
        return -2;
    }
    const double b0 = sqrt(dot_r(2 * dim, psi, psi));
    if (b0 == 0.0) {
        free(V); free(w); free(al); free(be);
        return 0;
    }
    for (int64_t i = 0; i < 2 * dim; ++i) V[i] = psi[i] / b0;
    int mm = m;
    for (int j = 0; j < m; ++j) {
        const double *vj = V + (size_t)j * 2 * dim;
        bko_hv_complex(n, g, normalize_by_n, vj, w);
        /* full reorthogonalisation (two passes of classical Gram-Schmidt) */
        for (int pass = 0; pass < 2; ++pass) {
            for (int i = 0; i <= j; ++i) {
                const double *vi = V + (size_t)i * 2 * dim;
                double cr = 0.0, ci = 0.0;
#pragma omp parallel for schedule(static) reduction(+ : cr, ci)
                for (int64_t k = 0; k < dim; ++k) {
                    cr += vi[2 * k] * w[2 * k] + vi[2 * k + 1] * w[2 * k + 1];
                    ci += vi[2 * k] * w[2 * k + 1] - vi[2 * k + 1] * w[2 * k];
                }
                /* T_jj is the projection on v_j; the second pass adds its correction */
                if (i == j) al[j] = (pass == 0) ? cr : al[j] + cr;
#pragma omp parallel for schedule(static)
                for (int64_t k = 0; k < dim; ++k) {
                    w[2 * k] -= cr * vi[2 * k] - ci * vi[2 * k + 1];
                    w[2 * k + 1] -= cr * vi[2 * k + 1] + ci * vi[2 * k];
                }
            }
        }
        if (j == m - 1) break;
        const double bj = sqrt(dot_r(2 * dim, w, w));
        if (bj < 1e-13 * (fabs(al[j]) + 1.0)) {
            mm = j + 1;
            break;
        }
        be[j] = bj;
        double *vn = V + (size_t)(j + 1) * 2 * dim;
        for (int64_t k = 0; k < 2 * dim; ++k) vn[k] = w[k] / bj;
    }
    double *d = (double *)malloc(sizeof(double) * (size_t)mm);
    double *z = (double *)calloc((size_t)mm * mm, sizeof(double));
    for (int i = 0; i < mm; ++i) {
        d[i] = al[i];
        z[(size_t)i * mm + i] = 1.0;
    }
    int rc = bko_tridiag_eig(mm, d, be, z);
    if (rc == 0) {
        /* y = S exp(-i L tau) S^T e1 ; row k of z is eigenvector k */
        double *yr = (double *)calloc((size_t)mm, sizeof(double));
        double *yi = (double *)calloc((size_t)mm, sizeof(double));
        for (int k = 0; k < mm; ++k) {
            const double s0k = z[(size_t)k * mm + 0];
            const double cr = cos(-d[k] * tau) * s0k, ci = sin(-d[k] * tau) * s0k;
            for (int j = 0; j < mm; ++j) {
                yr[j] += z[(size_t)k * mm + j] * cr;
                yi[j] += z[(size_t)k * mm + j] * ci;
            }
        }
#pragma omp parallel for schedule(static)
        for (int64_t k = 0; k < dim; ++k) {
            double ar = 0.0, ai = 0.0;
            for (int j = 0; j < mm; ++j) {
                const double *vj = V + (size_t)j * 2 * dim;
                ar += yr[j] * vj[2 * k] - yi[j] * vj[2 * k + 1];
                ai += yr[j] * vj[2 * k + 1] + yi[j] * vj[2 * k];
            }
            psi[2 * k] = b0 * ar;
            psi[2 * k + 1] = b0 * ai;
        }
        free(yr);
        free(yi);
    }
    free(V); free(w); free(al); free(be); free(d); free(z);
    return rc;
}

/* ------------------------------------------------------------------------ */
/* Projective sampling by inverse CDF (the cuStateVec sampler's published   */
/* semantics, sampler.cuh:564-573), outputs in random-number order.          */
/* ------------------------------------------------------------------------ */
void bko_sample_inverse_cdf(int n, const double *psi, const double *randnums, int shots,
                            int64_t *out)
{
    const int64_t dim = 1LL << n;
    double *cdf = (double *)malloc(sizeof(double) * (size_t)dim);
    double acc = 0.0;
    for (int64_t i = 0; i < dim; ++i) {
        acc += psi[2 * i] * psi[2 * i] + psi[2 * i + 1] * psi[2 * i + 1];
        cdf[i] = acc;
    }
    for (int s = 0; s < shots; ++s) {
        const double target = randnums[s] * acc;
        int64_t lo = 0, hi = dim - 1;
        while (lo < hi) {
            const int64_t mid = lo + (hi - lo) / 2;
            if (cdf[mid] > target) hi = mid;
            else lo = mid + 1;
        }
        out[s] = lo;
    }
    free(cdf);
}

/* ------------------------------------------------------------------------ */
/* Derived metrics: sampler.cuh:642-683 (mirrored at connector.cuh:752-793, */
/* window P_< at connector.cuh:817-830).                                     */
/* ------------------------------------------------------------------------ */
void bko_derived_metrics(const double *t, const double *p, const double *dh, int len,
                         double avg_start, double avg_end, double *rise_avg, double *rise_max,
                         double *peak_p, double *peak_t, double *window_dh, double *window_p)
{
    double ra = 0.0, rm = 0.0, pp = 0.0, pt = 0.0;
    if (len > 0) {
        int first_drop = len - 1;
        for (int i = 1; i < len; ++i)
            if (p[i] < p[i - 1]) {
                first_drop = i - 1;
                break;
            }
        ra = (t[first_drop] > 0.0) ? p[first_drop] / t[first_drop] : 0.0;
        rm = (t[0] > 1e-9) ? p[0] / t[0] : 0.0;
        for (int i = 0; i < first_drop; ++i) {
            const double d = t[i + 1] - t[i];
            if (d > 1e-9) {
                const double sl = (p[i + 1] - p[i]) / d;
                if (sl > rm) rm = sl;
            }
        }
        /* running peak, strict '>' starting from 0 (sampler.cuh:523-526) */
        for (int i = 0; i < len; ++i)
            if (p[i] > pp) {
                pp = p[i];
                pt = t[i];
            }
    }
    double sd = 0.0, sp = 0.0;
    int cnt = 0;
    for (int i = 0; i < len; ++i)
        if (t[i] >= avg_start - 1e-9 && t[i] <= avg_end + 1e-9) {
            sd += dh[i];
            sp += p[i];
            ++cnt;
        }
    *rise_avg = ra;
    *rise_max = rm;
    *peak_p = pp;
    *peak_t = pt;
    *window_dh = cnt > 0 ? sd / cnt : (len > 0 ? dh[len - 1] : 0.0);
    *window_p = cnt > 0 ? sp / cnt : 0.0;
}

/* ------------------------------------------------------------------------ */
/* Schedule: sampler.cuh:496-603.                                             */
/* ------------------------------------------------------------------------ */
int bko_schedule(double t_max, double dt, int num_sample_points, int shots_per_point,
                 double t_start, double dt_sample, int track_p_less, int record_trajectory,
                 double dt_trajectory, unsigned char *flags, int flags_len)
{
    const int total = (int)llround(t_max / dt);
    const int npts = num_sample_points > 1 ? num_sample_points : 1;
    const double sample_interval = (dt_sample > 0.0) ? dt_sample : (t_max / npts);
    double next_sample = t_start;
    const double traj_interval = (dt_trajectory > 0.0) ? dt_trajectory : dt;
    double next_traj = 0.0;
    const double sample_end = (num_sample_points > 1 && dt_sample > 0.0)
                                  ? (t_start + (num_sample_points - 1) * dt_sample)
                                  : t_max;
    int taken = 0;
    for (int step = 0; step <= total; ++step) {
        const double t = step * dt;
        unsigned char f = 0;
        if (record_trajectory) {
            if (t >= next_traj - 1e-9 && next_traj <= t_max + 1e-9) {
                f |= 1;
                next_traj += traj_interval;
            }
        } else if (track_p_less) {
            if (t >= next_sample - 1e-9 && next_sample <= t_max + 1e-9) f |= 1;
        }
        if (shots_per_point > 0 && t >= t_start - 1e-9 && t >= next_sample - 1e-9 &&
            next_sample <= sample_end + 1e-9 &&
            (num_sample_points <= 0 || taken < num_sample_points)) {
            ++taken;
            f |= 2;
            next_sample += sample_interval;
        } else if (shots_per_point == 0 && !record_trajectory && track_p_less &&
                   t >= next_sample - 1e-9 && next_sample <= t_max + 1e-9) {
            next_sample += sample_interval;
        }
        if (flags && step < flags_len) flags[step] = f;
    }
    return total;
}

/* ------------------------------------------------------------------------ */
/* Exhaustive one-flip local minima: exhaustive_minima.cpp:74-208.           */
/* ------------------------------------------------------------------------ */
/* is_one_flip_local_minimum (exhaustive_minima.cpp:74-107): s is a minimum
 * when no single flip lowers E.  With C_k the autocorrelations of s, flipping
 * spin p changes C_k by d_k = -2 s_p (s_{p+k} + s_{p-k}) over the pairs present
 * and E by sum_k d_k (2 C_k + d_k); the scan stops at the first flip whose
 * change is negative. */
static int one_flip_minimum(int n, uint64_t s, const int64_t *C)
{
    for (int p = 0; p < n; ++p) {
        const int sp = ((s >> p) & 1ULL) ? -1 : 1;
        int64_t de = 0;
        for (int k = 1; k < n; ++k) {
            int d = 0;
            if (p + k < n) d -= 2 * sp * (((s >> (p + k)) & 1ULL) ? -1 : 1);
            if (p - k >= 0) d -= 2 * (((s >> (p - k)) & 1ULL) ? -1 : 1) * sp;
            de += (int64_t)d * (2 * C[k] + d);
        }
        if (de < 0) return 0;
    }
    return 1;
}

typedef struct {
    int64_t e;
    uint64_t c;
} bko_min_rec;

static int cmp_min_rec(const void *a, const void *b)
{
    const bko_min_rec *x = (const bko_min_rec *)a, *y = (const bko_min_rec *)b;
    if (x->e != y->e) return x->e < y->e ? -1 : 1;
    if (x->c != y->c) return x->c < y->c ? -1 : 1;
    return 0;
}

/* run_exploration (exhaustive_minima.cpp:109-208): every s in [0, 2^(n-1))
 * (the other half follows by global negation) that is a one-flip minimum adds
 * its canonical image (SpinState::canonical) to the set of its energy.  The
 * lowest energy is E*; the next 'levels' energies are the trap levels, whose
 * orbits are listed in ascending canonical order (std::set order). */
int bko_exhaustive_minima(int n, int levels, int64_t *e_star, int64_t *total_minima, int64_t *total_orbits,
                          int64_t *num_energies, int64_t *level_energy, int64_t *level_orbits, uint64_t *orbits,
                          int64_t cap)
{
    if (n < 2 || n > 40 || levels < 0) return -1;
    const uint64_t half = 1ULL << (n - 1);
    int64_t count = 0, alloc = 1024;
    bko_min_rec *recs = (bko_min_rec *)malloc(sizeof(bko_min_rec) * (size_t)alloc);
    if (!recs) return -1;
    int failed = 0;
#pragma omp parallel
    {
        int64_t lc = 0, la = 256;
        bko_min_rec *loc = (bko_min_rec *)malloc(sizeof(bko_min_rec) * (size_t)la);
        int64_t C[64];
#pragma omp for schedule(dynamic, 65536)
        for (uint64_t s = 0; s < half; ++s) {
            bko_ck(n, s, C);
            if (!one_flip_minimum(n, s, C)) continue;
            int64_t e = 0;
            for (int k = 1; k < n; ++k) e += C[k] * C[k];
            if (lc == la) {
                la *= 2;
                bko_min_rec *t = loc ? (bko_min_rec *)realloc(loc, sizeof(bko_min_rec) * (size_t)la) : NULL;
                if (!t) {
#pragma omp atomic write
                    failed = 1;
                    continue;
                }
                loc = t;
            }
            if (loc) {
                loc[lc].e = e;
                loc[lc].c = bko_canonical(n, s);
                ++lc;
            }
        }
#pragma omp critical
        {
            if (count + lc > alloc) {
                while (count + lc > alloc) alloc *= 2;
                bko_min_rec *t = (bko_min_rec *)realloc(recs, sizeof(bko_min_rec) * (size_t)alloc);
                if (t) recs = t;
                else failed = 1;
            }
            if (!failed && lc > 0) {
                memcpy(recs + count, loc, sizeof(bko_min_rec) * (size_t)lc);
                count += lc;
            }
        }
        free(loc);
    }
    if (failed) {
        free(recs);
        return -1;
    }
    qsort(recs, (size_t)count, sizeof(bko_min_rec), cmp_min_rec);
    *total_minima = 2 * count;
    int64_t orb = 0, ne = 0, written = 0;
    int lvl = -1; /* -1: E*, 0.. : trap levels */
    for (int64_t i = 0; i < count; ++i) {
        const int newe = (i == 0 || recs[i].e != recs[i - 1].e);
        const int newc = newe || recs[i].c != recs[i - 1].c;
        if (newe) {
            ++ne;
            if (ne == 1) *e_star = recs[i].e;
            lvl = (int)ne - 2;
            if (lvl >= 0 && lvl < levels) {
                level_energy[lvl] = recs[i].e;
                level_orbits[lvl] = 0;
            }
        }
        if (!newc) continue;
        ++orb;
        if (lvl >= 0 && lvl < levels) {
            ++level_orbits[lvl];
            if (written < cap) orbits[written++] = recs[i].c;
        }
    }
    *total_orbits = orb;
    *num_energies = ne;
    free(recs);
    if (count == 0) *e_star = -1;
    const int64_t lv = ne - 1;
    return (int)(lv < levels ? (lv > 0 ? lv : 0) : levels);
}
