/*
 * nyxpm_oracle.c -- CPU FP64 restatement of Nyx's particle-mesh gravity step.
 *
 * TEST INFRASTRUCTURE ONLY.  This file is the correctness oracle (the checker)
 * for the HIP library in nyx_amd/.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path never calls it.
 *
 * The reference (AMReX-Astro/Nyx) cannot be compiled here: every file on the
 * path includes AMReX headers and the AMReX submodule is empty.  This file
 * therefore restates, in plain C + OpenMP, what Nyx and the AMReX routines it
 * calls compute on one periodic level:
 *
 *   deposit      NyxParticleContainer.H:90-97 -> AMReX AssignCellDensitySingleLevel
 *                (amrex_deposit_cic weights, periodic SumBoundary, then x 1/vol)
 *   mass offset  NyxParticleContainer.H:70-79, Gravity.cpp:975-1030
 *   rhs          Gravity.cpp:1207-1244 (CorrectRhsUsingOffset), sum_utils.cpp:21-99
 *                (vol_weight_sum), Gravity.cpp:283 (x Ggravity), :290 (x 1/a)
 *   solve        Gravity.cpp:1318-1416 -> AMReX MLMG V-cycles on MLPoisson
 *   gravity      Gravity.cpp:817-850 (face gradients averaged to cell centres)
 *   particles    DarkMatterParticleContainer.cpp:249-306 (update_dm_particle_single)
 *                called by moveKickDrift (:75-181) and moveKick (:183-247)
 *   wrap         NyxParticleContainer.H:114-136 (Redistribute's periodic shift)
 *   step         Nyx_advance_particles.cpp:137-268
 *
 * Parity pinning of each function (what checks it beyond restatement) is listed
 * in tests/test_oracle_*.py: closed forms (cell-centre / corner / lattice
 * deposits, single Fourier modes), an independent numpy FFT solve of the same
 * discrete operator, adjointness of deposit and interpolation, and momentum
 * conservation.  The AMReX iterate sequence inside MLMG (BiCGStab bottom,
 * agglomeration) and AMReX's periodic-wrap round-off clamp are not pinned by any
 * check available offline ("parity unpinned"); only their fixed point / intent
 * is.
 *
 * Floating-point conventions (shared with the HIP library so that iterates can be
 * compared bit for bit): no FMA contraction (built with -ffp-contract=off),
 * operations written in the order stated in each comment.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>
#include <float.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define ORC_OK 0
#define ORC_ERR_ARG 1
#define ORC_ERR_NOMEM 2
#define ORC_ERR_NOCONV 3
#define ORC_ERR_BOTTOM 4

/* ------------------------------------------------------------------------- */
/* small helpers                                                              */
/* ------------------------------------------------------------------------- */

static inline int64_t wrapi(int64_t i, int64_t n) {
    int64_t r = i % n;
    return r < 0 ? r + n : r;
}

static inline size_t cidx(int i, int j, int k, const int n[3]) {
    return ((size_t)k * (size_t)n[1] + (size_t)j) * (size_t)n[0] + (size_t)i;
}

int orc_num_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

void orc_set_num_threads(int t) {
#ifdef _OPENMP
    if (t > 0) omp_set_num_threads(t);
#else
    (void)t;
#endif
}

/* Cell size exactly as an AMReX Geometry computes it: dx = (hi-lo)/n, dxi = 1/dx. */
static void geom_dx(const int n[3], const double plo[3], const double phi[3],
                    double dx[3], double dxi[3]) {
    for (int d = 0; d < 3; ++d) {
        dx[d] = (phi[d] - plo[d]) / (double)n[d];
        dxi[d] = 1.0 / dx[d];
    }
}

/* ------------------------------------------------------------------------- */
/* CIC deposit (AssignDensitySingleLevel)                                     */
/* ------------------------------------------------------------------------- */

/*
 * rho[c] = (sum_p m_p * w_p(c)) * (1/vol), periodic.
 *
 * amrex_deposit_cic: l = (x-plo)*dxi + 0.5, i = floor(l), xint = l - i,
 * sx = {1-xint, xint}; particle mass goes to cells (i+ii-1, j+jj-1, k+kk-1) with
 * weight sx[ii]*sy[jj]*sz[kk]*m (product left to right).  SumBoundary folds the
 * ghost cells onto their periodic partners, which is the same as wrapping the
 * target index.  AssignCellDensitySingleLevel then multiplies by 1/vol with
 * vol = dx*dy*dz.
 *
 * Summation order used here (deterministic, independent of the thread count):
 * particles are stably bucketed by their z index k mod nz; target plane t takes
 * first the kk=1 contributions of bucket t, then the kk=0 contributions of
 * bucket t+1, each bucket in original particle order, each particle in (jj, ii)
 * order.  Threads own disjoint ranges of target planes.
 */
int orc_deposit_cic(const int n[3], const double plo[3], const double phi[3],
                    int64_t np, const double *x, const double *y, const double *z,
                    const double *m, double m_const, double *rho) {
    if (n[0] <= 0 || n[1] <= 0 || n[2] <= 0 || np < 0) return ORC_ERR_ARG;
    double dx[3], dxi[3];
    geom_dx(n, plo, phi, dx, dxi);
    const size_t nc = (size_t)n[0] * n[1] * n[2];
    memset(rho, 0, nc * sizeof(double));
    if (np == 0) return ORC_OK;

    const int nz = n[2];
    int nt = orc_num_threads();
    int64_t *hist = (int64_t *)calloc((size_t)nt * (size_t)(nz + 1), sizeof(int64_t));
    int32_t *kz = (int32_t *)malloc((size_t)np * sizeof(int32_t));
    int64_t *order = (int64_t *)malloc((size_t)np * sizeof(int64_t));
    int64_t *start = (int64_t *)malloc((size_t)(nz + 1) * sizeof(int64_t));
    if (!hist || !kz || !order || !start) {
        free(hist); free(kz); free(order); free(start);
        return ORC_ERR_NOMEM;
    }

    /* stable bucket sort by k mod nz: contiguous chunks per thread */
#pragma omp parallel num_threads(nt)
    {
#ifdef _OPENMP
        int t = omp_get_thread_num();
#else
        int t = 0;
#endif
        int64_t lo = np * t / nt, hi = np * (t + 1) / nt;
        int64_t *h = hist + (size_t)t * (nz + 1);
        for (int64_t p = lo; p < hi; ++p) {
            double lz = (z[p] - plo[2]) * dxi[2] + 0.5;
            int k = (int)floor(lz);
            int kk = (int)wrapi(k, nz);
            kz[p] = kk;
            h[kk]++;
        }
#pragma omp barrier
#pragma omp single
        {
            int64_t run = 0;
            for (int b = 0; b < nz; ++b) {
                start[b] = run;
                for (int tt = 0; tt < nt; ++tt) {
                    int64_t c = hist[(size_t)tt * (nz + 1) + b];
                    hist[(size_t)tt * (nz + 1) + b] = run;
                    run += c;
                }
            }
            start[nz] = run;
        }
        for (int64_t p = lo; p < hi; ++p) {
            int b = kz[p];
            order[h[b]++] = p;
        }
    }

    /* each thread owns target planes; per target plane two source buckets */
#pragma omp parallel for schedule(static) num_threads(nt)
    for (int t = 0; t < nz; ++t) {
        for (int pass = 0; pass < 2; ++pass) {
            int b = (int)wrapi(t + pass, nz); /* bucket t (kk=1) then t+1 (kk=0) */
            int kk = 1 - pass;
            for (int64_t q = start[b]; q < start[b + 1]; ++q) {
                int64_t p = order[q];
                double lx = (x[p] - plo[0]) * dxi[0] + 0.5;
                double ly = (y[p] - plo[1]) * dxi[1] + 0.5;
                double lz = (z[p] - plo[2]) * dxi[2] + 0.5;
                int i = (int)floor(lx), j = (int)floor(ly), k = (int)floor(lz);
                double xint = lx - i, yint = ly - j, zint = lz - k;
                double sx[2] = {1.0 - xint, xint};
                double sy[2] = {1.0 - yint, yint};
                double sz[2] = {1.0 - zint, zint};
                double mp = m ? m[p] : m_const;
                int kt = (int)wrapi((int64_t)k + kk - 1, nz);
                if (kt != t) continue; /* only possible for nz == 1 mismatch; never */
                for (int jj = 0; jj <= 1; ++jj) {
                    int jt = (int)wrapi((int64_t)j + jj - 1, n[1]);
                    for (int ii = 0; ii <= 1; ++ii) {
                        int it = (int)wrapi((int64_t)i + ii - 1, n[0]);
                        double w = sx[ii] * sy[jj] * sz[kk] * mp;
                        rho[cidx(it, jt, kt, n)] += w;
                    }
                }
            }
        }
    }

    const double vol = dx[0] * dx[1] * dx[2];
    const double inv_vol = 1.0 / vol;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < (int64_t)nc; ++c) rho[c] = rho[c] * inv_vol;

    free(hist); free(kz); free(order); free(start);
    return ORC_OK;
}

/* sumParticleMass (NyxParticleContainer.H:70-79).  With one mass for every
 * particle the total is taken as np * m (one rounding). */
double orc_sum_particle_mass(int64_t np, const double *m, double m_const) {
    if (!m) return (double)np * m_const;
    double s = 0.0;
    for (int64_t p = 0; p < np; ++p) s += m[p];
    return s;
}

/* mass_offset = (sum of particle mass) / ProbSize  (Gravity.cpp:1010-1014). */
double orc_mass_offset(const double plo[3], const double phi[3], double total_mass) {
    double size = (phi[0] - plo[0]) * (phi[1] - plo[1]) * (phi[2] - plo[2]);
    return total_mass / size;
}

/*
 * Sum of a cell field, in a fixed order: each z-plane summed row by row
 * (i fastest, sequentially), then planes added in increasing z.
 */
static double field_sum(const int n[3], const double *f) {
    const int nz = n[2];
    double *ps = (double *)malloc((size_t)nz * sizeof(double));
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nz; ++k) {
        double s = 0.0;
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) s += f[cidx(i, j, k, n)];
        ps[k] = s;
    }
    double s = 0.0;
    for (int k = 0; k < nz; ++k) s += ps[k];
    free(ps);
    return s;
}

/*
 * CorrectRhsUsingOffset (Gravity.cpp:1207-1244) with vol_weight_sum
 * (sum_utils.cpp:21-99, unmasked):
 *   Rhs += -mass_offset
 *   sum  = (sum_c Rhs_c) * (dx*dy*dz) / ProbSize
 *   Rhs += -sum
 * Returns sum.
 */
double orc_correct_rhs_using_offset(const int n[3], const double plo[3], const double phi[3],
                                    double *rhs, double mass_offset) {
    double dx[3], dxi[3];
    geom_dx(n, plo, phi, dx, dxi);
    const int64_t nc = (int64_t)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) rhs[c] = rhs[c] + (-mass_offset);
    double s = field_sum(n, rhs);
    double size = (phi[0] - plo[0]) * (phi[1] - plo[1]) * (phi[2] - plo[2]);
    double sum = s * (dx[0] * dx[1] * dx[2]);
    sum = sum / size;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) rhs[c] = rhs[c] + (-sum);
    return sum;
}

/* Rhs.mult(Ggravity); Rhs.mult(1/a)  (Gravity.cpp:283,290), Ggravity = -4 pi G. */
void orc_scale_rhs(int64_t nc, double *rhs, double four_pi_G, double a) {
    const double Ggravity = -four_pi_G;
    const double inv_a = 1.0 / a;
#pragma omp parallel for schedule(static)
    for (int64_t c = 0; c < nc; ++c) rhs[c] = (rhs[c] * Ggravity) * inv_a;
}

/* ------------------------------------------------------------------------- */
/* Multigrid (MLMG on MLPoisson, single periodic level)                       */
/* ------------------------------------------------------------------------- */

typedef struct {
    int n[3];
    double dh[3];
    double gamma, inv_gamma;
    size_t nc;
    double *r, *e;
} orc_level;

#define ORC_MAX_LEVELS 32
#define ORC_BOTTOM_MAX_CELLS 4096

/*
 * Coarsening chain: halve all three dimensions while every dimension is even
 * and the halved dimension stays >= 2.  For 2^p boxes this ends at 2^3 cells,
 * as AMReX's MLLinOp does with its default minimum coarse width of 2.
 */
int orc_mg_levels(const int n[3], int dims_out[][3]) {
    int m[3] = {n[0], n[1], n[2]};
    int L = 0;
    for (;;) {
        if (dims_out) { dims_out[L][0] = m[0]; dims_out[L][1] = m[1]; dims_out[L][2] = m[2]; }
        ++L;
        if (L >= ORC_MAX_LEVELS) break;
        int ok = 1;
        for (int d = 0; d < 3; ++d)
            if ((m[d] & 1) || m[d] / 2 < 2) ok = 0;
        if (!ok) break;
        for (int d = 0; d < 3; ++d) m[d] /= 2;
    }
    return L;
}

/*
 * Periodic Green's function of the 7-point operator on the bottom level:
 *   G[c] = (1/N) sum_{k != 0} cos(2 pi sum_d ((k_d c_d) mod n_d) / n_d) / lambda_k,
 *   lambda_k = sum_d dh_d * (2 cos(2 pi k_d / n_d) - 2).
 * G is the pseudo-inverse (mean mode removed), so e = G (*) r solves L e = r
 * exactly for mean-free r.  This replaces AMReX's BiCGStab bottom solve: same
 * fixed point, exact instead of 1e-4 relative.
 */
void orc_bottom_green(const int n[3], const double dh[3], double *G) {
    const double twopi = 2.0 * M_PI;
    const int N = n[0] * n[1] * n[2];
    for (int cz = 0; cz < n[2]; ++cz)
        for (int cy = 0; cy < n[1]; ++cy)
            for (int cx = 0; cx < n[0]; ++cx) {
                double acc = 0.0;
                for (int kz = 0; kz < n[2]; ++kz)
                    for (int ky = 0; ky < n[1]; ++ky)
                        for (int kx = 0; kx < n[0]; ++kx) {
                            if (kx == 0 && ky == 0 && kz == 0) continue;
                            double lam = dh[0] * (2.0 * cos(twopi * kx / n[0]) - 2.0)
                                       + dh[1] * (2.0 * cos(twopi * ky / n[1]) - 2.0)
                                       + dh[2] * (2.0 * cos(twopi * kz / n[2]) - 2.0);
                            double th = twopi * ((double)((kx * cx) % n[0]) / n[0]
                                               + (double)((ky * cy) % n[1]) / n[1]
                                               + (double)((kz * cz) % n[2]) / n[2]);
                            acc += cos(th) / lam;
                        }
                G[cidx(cx, cy, cz, n)] = acc / (double)N;
            }
}

/*
 * L phi = dhx*(phi(i-1) - 2 phi(i) + phi(i+1)) + dhy*(...) + dhz*(...)
 * (MLPoisson apply, periodic).  r = f - L phi.
 */
static inline double lap7(const double *a, int i, int j, int k, const int n[3], const double dh[3]) {
    int im = i == 0 ? n[0] - 1 : i - 1, ip = i == n[0] - 1 ? 0 : i + 1;
    int jm = j == 0 ? n[1] - 1 : j - 1, jp = j == n[1] - 1 ? 0 : j + 1;
    int km = k == 0 ? n[2] - 1 : k - 1, kp = k == n[2] - 1 ? 0 : k + 1;
    double c = a[cidx(i, j, k, n)];
    double lx = dh[0] * (a[cidx(im, j, k, n)] - 2.0 * c + a[cidx(ip, j, k, n)]);
    double ly = dh[1] * (a[cidx(i, jm, k, n)] - 2.0 * c + a[cidx(i, jp, k, n)]);
    double lz = dh[2] * (a[cidx(i, j, km, n)] - 2.0 * c + a[cidx(i, j, kp, n)]);
    return lx + ly + lz;
}

/* r = f - L phi; returns max |r|. */
double orc_residual(const int n[3], const double dh[3], const double *f, const double *phi, double *r) {
    double mx = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int k = 0; k < n[2]; ++k)
        for (int j = 0; j < n[1]; ++j)
            for (int i = 0; i < n[0]; ++i) {
                size_t c = cidx(i, j, k, n);
                double v = f[c] - lap7(phi, i, j, k, n, dh);
                r[c] = v;
                double a = fabs(v);
                if (a > mx) mx = a;
            }
    return mx;
}

static double norm_inf(size_t nc, const double *a) {
    double mx = 0.0;
#pragma omp parallel for schedule(static) reduction(max : mx)
    for (int64_t c = 0; c < (int64_t)nc; ++c) {
        double v = fabs(a[c]);
        if (v > mx) mx = v;
    }
    return mx;
}

/*
 * One red-black Gauss-Seidel half sweep of MLPoisson (colour: cells with
 * (i+j+k+colour) even):
 *   res = r - gamma*e - dhx*(e(i-1)+e(i+1)) - dhy*(...) - dhz*(...)
 *   e   = e + res * (1/gamma),       gamma = -2 (dhx+dhy+dhz)
 * Cells of one colour only read cells of the other colour, so the result does
 * not depend on traversal order.
 */
static void gsrb_half(orc_level *lv, int colour) {
    const int *n = lv->n;
    const double *dh = lv->dh;
    double *e = lv->e;
    const double *r = lv->r;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k) {
        int km = k == 0 ? n[2] - 1 : k - 1, kp = k == n[2] - 1 ? 0 : k + 1;
        for (int j = 0; j < n[1]; ++j) {
            int jm = j == 0 ? n[1] - 1 : j - 1, jp = j == n[1] - 1 ? 0 : j + 1;
            int i0 = (j + k + colour) & 1;
            for (int i = i0; i < n[0]; i += 2) {
                int im = i == 0 ? n[0] - 1 : i - 1, ip = i == n[0] - 1 ? 0 : i + 1;
                size_t c = cidx(i, j, k, n);
                double res = r[c] - lv->gamma * e[c]
                           - dh[0] * (e[cidx(im, j, k, n)] + e[cidx(ip, j, k, n)])
                           - dh[1] * (e[cidx(i, jm, k, n)] + e[cidx(i, jp, k, n)])
                           - dh[2] * (e[cidx(i, j, km, n)] + e[cidx(i, j, kp, n)]);
                e[c] = e[c] + res * lv->inv_gamma;
            }
        }
    }
}

/*
 * Residual of the correction restricted to the coarse level:
 *   rc = r - L e  on the fine level,
 *   r_c(I) = 0.125 * sum of the 8 children of I (summed k, then j, then i
 *   innermost, starting from the (0,0,0) child), as average_down does.
 */
static void resid_restrict(orc_level *f, orc_level *c) {
    const int *nf = f->n, *ncn = c->n;
#pragma omp parallel for schedule(static)
    for (int K = 0; K < ncn[2]; ++K)
        for (int J = 0; J < ncn[1]; ++J)
            for (int I = 0; I < ncn[0]; ++I) {
                double s = 0.0;
                for (int kk = 0; kk < 2; ++kk)
                    for (int jj = 0; jj < 2; ++jj)
                        for (int ii = 0; ii < 2; ++ii) {
                            int i = 2 * I + ii, j = 2 * J + jj, k = 2 * K + kk;
                            size_t q = cidx(i, j, k, nf);
                            double rc = f->r[q] - lap7(f->e, i, j, k, nf, f->dh);
                            s += rc;
                        }
                c->r[cidx(I, J, K, ncn)] = 0.125 * s;
            }
}

/* e_f += e_c(parent)  (piecewise-constant interpolation of MLCellLinOp). */
static void prolong_add(orc_level *f, orc_level *c) {
    const int *nf = f->n, *ncn = c->n;
#pragma omp parallel for schedule(static)
    for (int k = 0; k < nf[2]; ++k)
        for (int j = 0; j < nf[1]; ++j)
            for (int i = 0; i < nf[0]; ++i)
                f->e[cidx(i, j, k, nf)] = f->e[cidx(i, j, k, nf)] + c->e[cidx(i / 2, j / 2, k / 2, ncn)];
}

/* e = G (*) r on the bottom level: e[c] = sum_{c'} G[(c - c') mod n] r[c'],
 * c' in natural (z, y, x) order. */
static void bottom_solve(orc_level *b, const double *G) {
    const int *n = b->n;
    for (int cz = 0; cz < n[2]; ++cz)
        for (int cy = 0; cy < n[1]; ++cy)
            for (int cx = 0; cx < n[0]; ++cx) {
                double acc = 0.0;
                for (int qz = 0; qz < n[2]; ++qz)
                    for (int qy = 0; qy < n[1]; ++qy)
                        for (int qx = 0; qx < n[0]; ++qx) {
                            int dxx = (int)wrapi(cx - qx, n[0]);
                            int dyy = (int)wrapi(cy - qy, n[1]);
                            int dzz = (int)wrapi(cz - qz, n[2]);
                            acc += G[cidx(dxx, dyy, dzz, n)] * b->r[cidx(qx, qy, qz, n)];
                        }
                b->e[cidx(cx, cy, cz, n)] = acc;
            }
}

static void vcycle(orc_level *lv, int L, int l, const double *G, int nu1, int nu2) {
    orc_level *f = &lv[l];
    if (l == L - 1) {
        bottom_solve(f, G);
        return;
    }
    memset(f->e, 0, f->nc * sizeof(double));
    for (int s = 0; s < nu1; ++s) { gsrb_half(f, 0); gsrb_half(f, 1); }
    resid_restrict(f, &lv[l + 1]);
    vcycle(lv, L, l + 1, G, nu1, nu2);
    prolong_add(f, &lv[l + 1]);
    for (int s = 0; s < nu2; ++s) { gsrb_half(f, 0); gsrb_half(f, 1); }
}

/*
 * MLMG::solve restated for one periodic level:
 *   r = f - L phi, resnorm0 = |r|_inf, rhsnorm = |f|_inf,
 *   target = max(abs_tol, max(rel_tol, 1e-16) * max(resnorm0, rhsnorm));
 *   repeat { e = Vcycle(r) from e = 0;  phi += e;  r = f - L phi } until
 *   |r|_inf <= target.
 * V(nu1,nu2) with red-black GS, restriction by 8-child average, piecewise-
 * constant prolongation, exact bottom solve.  fixed_iters > 0 runs exactly that
 * many cycles (no convergence test).  stats[0..2] = resnorm0, rhsnorm, final.
 */
int orc_mg_solve(const int n[3], const double dx[3], const double *f, double *phi,
                 double rel_tol, double abs_tol, int max_iter, int fixed_iters,
                 int nu1, int nu2, int *iters_out, double *stats) {
    int dims[ORC_MAX_LEVELS][3];
    int L = orc_mg_levels(n, dims);
    size_t nb = (size_t)dims[L - 1][0] * dims[L - 1][1] * dims[L - 1][2];
    if (nb > ORC_BOTTOM_MAX_CELLS) return ORC_ERR_BOTTOM;
    orc_level lv[ORC_MAX_LEVELS];
    for (int l = 0; l < L; ++l) {
        orc_level *p = &lv[l];
        for (int d = 0; d < 3; ++d) {
            p->n[d] = dims[l][d];
            double dxl = dx[d] * (double)(n[d] / dims[l][d]);
            double dxil = 1.0 / dxl;
            p->dh[d] = dxil * dxil;
        }
        p->gamma = -2.0 * (p->dh[0] + p->dh[1] + p->dh[2]);
        p->inv_gamma = 1.0 / p->gamma;
        p->nc = (size_t)p->n[0] * p->n[1] * p->n[2];
        p->r = (double *)malloc(p->nc * sizeof(double));
        p->e = (double *)malloc(p->nc * sizeof(double));
        if (!p->r || !p->e) return ORC_ERR_NOMEM;
    }
    double *G = (double *)malloc(nb * sizeof(double));
    orc_bottom_green(lv[L - 1].n, lv[L - 1].dh, G);

    int rc = ORC_OK;
    double resnorm0 = orc_residual(n, lv[0].dh, f, phi, lv[0].r);
    double rhsnorm = norm_inf(lv[0].nc, f);
    double mx = resnorm0 > rhsnorm ? resnorm0 : rhsnorm;
    double rt = rel_tol > 1e-16 ? rel_tol : 1e-16;
    double target = rt * mx;
    if (abs_tol > target) target = abs_tol;
    double res = resnorm0;
    int it = 0;
    if (L == 1) {
        /* the whole problem is the bottom: one exact solve of the correction */
        if (fixed_iters > 0 || res > target) {
            bottom_solve(&lv[0], G);
            for (size_t c = 0; c < lv[0].nc; ++c) phi[c] = phi[c] + lv[0].e[c];
            res = orc_residual(n, lv[0].dh, f, phi, lv[0].r);
            it = 1;
        }
    } else if (fixed_iters > 0) {
        for (it = 0; it < fixed_iters;) {
            vcycle(lv, L, 0, G, nu1, nu2);
#pragma omp parallel for schedule(static)
            for (int64_t c = 0; c < (int64_t)lv[0].nc; ++c) phi[c] = phi[c] + lv[0].e[c];
            res = orc_residual(n, lv[0].dh, f, phi, lv[0].r);
            ++it;
        }
    } else if (res > target) {
        rc = ORC_ERR_NOCONV;
        while (it < max_iter) {
            vcycle(lv, L, 0, G, nu1, nu2);
#pragma omp parallel for schedule(static)
            for (int64_t c = 0; c < (int64_t)lv[0].nc; ++c) phi[c] = phi[c] + lv[0].e[c];
            res = orc_residual(n, lv[0].dh, f, phi, lv[0].r);
            ++it;
            if (res <= target) { rc = ORC_OK; break; }
        }
    }
    if (iters_out) *iters_out = it;
    if (stats) { stats[0] = resnorm0; stats[1] = rhsnorm; stats[2] = res; stats[3] = target; }
    for (int l = 0; l < L; ++l) { free(lv[l].r); free(lv[l].e); }
    free(G);
    return rc;
}

/* ------------------------------------------------------------------------- */
/* Gravity vector (get_new_grav_vector) and particle update                   */
/* ------------------------------------------------------------------------- */

/*
 * Face gradient G_x(i) = dxi*(phi(i) - phi(i-1)) (MLMG getGradSolution), then
 * cell-centred g_x(i) = 0.5*(G_x(i) + G_x(i+1)) (average_face_to_cellcenter),
 * periodic.  g is component-major: g[d*nc + c].
 */
void orc_grav_vector(const int n[3], const double plo[3], const double phi_hi[3],
                     const double *phi, double *g) {
    double dx[3], dxi[3];
    geom_dx(n, plo, phi_hi, dx, dxi);
    const size_t nc = (size_t)n[0] * n[1] * n[2];
#pragma omp parallel for schedule(static)
    for (int k = 0; k < n[2]; ++k) {
        int km = k == 0 ? n[2] - 1 : k - 1, kp = k == n[2] - 1 ? 0 : k + 1;
        for (int j = 0; j < n[1]; ++j) {
            int jm = j == 0 ? n[1] - 1 : j - 1, jp = j == n[1] - 1 ? 0 : j + 1;
            for (int i = 0; i < n[0]; ++i) {
                int im = i == 0 ? n[0] - 1 : i - 1, ip = i == n[0] - 1 ? 0 : i + 1;
                size_t c = cidx(i, j, k, n);
                double pc = phi[c];
                double fxl = dxi[0] * (pc - phi[cidx(im, j, k, n)]);
                double fxh = dxi[0] * (phi[cidx(ip, j, k, n)] - pc);
                double fyl = dxi[1] * (pc - phi[cidx(i, jm, k, n)]);
                double fyh = dxi[1] * (phi[cidx(i, jp, k, n)] - pc);
                double fzl = dxi[2] * (pc - phi[cidx(i, j, km, n)]);
                double fzh = dxi[2] * (phi[cidx(i, j, kp, n)] - pc);
                g[c] = 0.5 * (fxl + fxh);
                g[nc + c] = 0.5 * (fyl + fyh);
                g[2 * nc + c] = 0.5 * (fzl + fzh);
            }
        }
    }
}

/*
 * DarkMatterParticleContainer::AssignDensityAndVels (DarkMatterParticleContainer.cpp:
 * 456-460) = AssignDensity with ncomp = AMREX_SPACEDIM + 1: component 0 is the CIC
 * mass density (orc_deposit_cic), components 1..3 the CIC mass-weighted cell
 * velocity sum(w m v_d) / sum(w m) with w m = sx*sy*sz*m (left to right) and
 * (w m) * v_d the momentum contribution.  The mass and momentum sums for the
 * ratio are taken serially in particle order; a cell without mass gets velocity
 * 0.  out holds 4 fields of n cells.  AMReX's own division step (after
 * SumBoundary) is restated from its documented behaviour, not its source.
 */
int orc_deposit_cic_vels(const int n[3], const double plo[3], const double phi[3], int64_t np,
                         const double *x, const double *y, const double *z, const double *vx,
                         const double *vy, const double *vz, const double *m, double m_const,
                         double *out) {
    const size_t nc = (size_t)n[0] * n[1] * n[2];
    int rc = orc_deposit_cic(n, plo, phi, np, x, y, z, m, m_const, out);
    if (rc) return rc;
    double dx[3], dxi[3];
    geom_dx(n, plo, phi, dx, dxi);
    double *mass = (double *)calloc(nc, sizeof(double));
    if (!mass) return ORC_ERR_NOMEM;
    double *mom = out + nc;
    memset(mom, 0, 3 * nc * sizeof(double));
    const double *v[3] = {vx, vy, vz};
    for (int64_t p = 0; p < np; ++p) {
        const double mp = m ? m[p] : m_const;
        double lx = (x[p] - plo[0]) * dxi[0] + 0.5;
        double ly = (y[p] - plo[1]) * dxi[1] + 0.5;
        double lz = (z[p] - plo[2]) * dxi[2] + 0.5;
        int64_t i = (int64_t)floor(lx), j = (int64_t)floor(ly), k = (int64_t)floor(lz);
        double xint = lx - i, yint = ly - j, zint = lz - k;
        double sx[2] = {1.0 - xint, xint}, sy[2] = {1.0 - yint, yint}, sz[2] = {1.0 - zint, zint};
        for (int kk = 0; kk < 2; ++kk)
            for (int jj = 0; jj < 2; ++jj)
                for (int ii = 0; ii < 2; ++ii) {
                    const size_t c = cidx((int)wrapi(i - 1 + ii, n[0]), (int)wrapi(j - 1 + jj, n[1]),
                                          (int)wrapi(k - 1 + kk, n[2]), n);
                    const double w = sx[ii] * sy[jj] * sz[kk] * mp;
                    mass[c] += w;
                    for (int d = 0; d < 3; ++d) mom[(size_t)d * nc + c] += w * v[d][p];
                }
    }
    for (size_t c = 0; c < nc; ++c)
        for (int d = 0; d < 3; ++d) mom[(size_t)d * nc + c] = mass[c] != 0.0 ? mom[(size_t)d * nc + c] / mass[c] : 0.0;
    free(mass);
    return ORC_OK;
}

/*
 * NyxParticleContainer::estTimestep(acceleration, a, lev, cfl)
 * (NyxParticleContainer.H:327-410), called by Nyx::particle_est_time_step with the
 * new cell-centred gravity (NyxParticles.cpp:1028-1036).  Per particle:
 *   cell = floor((x - plo) * dxi)                    (no CIC shift; wrapped here,
 *                                                    the reference reads ghosts)
 *   dt_p = cfl / max_d(|v_d| * (dxi_d / a))         (1e50 if that max is 0)
 *   |g|  = sqrt(gx*gx + gy*gy + gz*gz) at the cell; if > 0,
 *   dt_p = min(dt_p, 1 / sqrt(|g| * dxi_x))
 * and the result is the minimum over particles, -1e50 without particles.  g is
 * the 3-field gravity of orc_grav_vector.
 */
double orc_est_time_step(const int n[3], const double plo[3], const double phi_hi[3], int64_t np,
                         const double *x, const double *y, const double *z, const double *vx,
                         const double *vy, const double *vz, const double *g, double a, double cfl) {
    if (np == 0) return -1e50;
    double dx[3], dxi[3];
    geom_dx(n, plo, phi_hi, dx, dxi);
    const double adxi[3] = {dxi[0] / a, dxi[1] / a, dxi[2] / a};
    const size_t nc = (size_t)n[0] * n[1] * n[2];
    double dt = 1e50;
#pragma omp parallel for schedule(static) reduction(min : dt)
    for (int64_t p = 0; p < np; ++p) {
        const int i = (int)wrapi((int64_t)floor((x[p] - plo[0]) * dxi[0]), n[0]);
        const int j = (int)wrapi((int64_t)floor((y[p] - plo[1]) * dxi[1]), n[1]);
        const int k = (int)wrapi((int64_t)floor((z[p] - plo[2]) * dxi[2]), n[2]);
        const double mx = fabs(vx[p]) * adxi[0];
        const double my = fabs(vy[p]) * adxi[1];
        const double mz = fabs(vz[p]) * adxi[2];
        double m = mx > my ? mx : my;
        m = m > mz ? m : mz;
        double dtp = (m > 0) ? (cfl / m) : 1e50;
        const size_t c = cidx(i, j, k, n);
        const double ax = g[c], ay = g[nc + c], az = g[2 * nc + c];
        const double mag = sqrt(ax * ax + ay * ay + az * az);
        if (mag > 0) {
            const double da = 1 / sqrt(mag * dxi[0]);
            dtp = dtp < da ? dtp : da;
        }
        dt = dtp < dt ? dtp : dt;
    }
    return dt;
}

/*
 * update_dm_particle_single (DarkMatterParticleContainer.cpp:249-306), with the
 * ghost-cell lookups of the reference replaced by periodic index wrapping:
 *   l = (x-plo)*dxi + 0.5, i = floor(l), sx = {1-(l-i), l-i}
 *   val_d = sum_{kk,jj,ii} sx[1-ii]*sy[1-jj]*sz[1-kk]*acc(i-ii, j-jj, k-kk, d)
 *   v_d = (a_prev*v_d + (0.5*dt)*val_d) * (1/a_cur)
 *   if do_move: x_d = x_d + (dt*(1/a_cur))*v_d
 * moveKickDrift passes (a_prev, a_cur) = (a_old, a_half), do_move = 1;
 * moveKick passes (a_half, a_new), do_move = 0.
 */
void orc_update_particles(const int n[3], const double plo[3], const double phi_hi[3],
                          int64_t np, double *x, double *y, double *z,
                          double *vx, double *vy, double *vz, const double *g,
                          double dt, double a_prev, double a_cur, int do_move) {
    double dx[3], dxi[3];
    geom_dx(n, plo, phi_hi, dx, dxi);
    const size_t nc = (size_t)n[0] * n[1] * n[2];
    const double half_dt = 0.5 * dt;
    const double a_cur_inv = 1.0 / a_cur;
    const double dt_a_cur_inv = dt * a_cur_inv;
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
        double lx = (x[p] - plo[0]) * dxi[0] + 0.5;
        double ly = (y[p] - plo[1]) * dxi[1] + 0.5;
        double lz = (z[p] - plo[2]) * dxi[2] + 0.5;
        int i = (int)floor(lx), j = (int)floor(ly), k = (int)floor(lz);
        double xint = lx - i, yint = ly - j, zint = lz - k;
        double sx[2] = {1.0 - xint, xint};
        double sy[2] = {1.0 - yint, yint};
        double sz[2] = {1.0 - zint, zint};
        double *v[3] = {vx, vy, vz};
        for (int d = 0; d < 3; ++d) {
            double val = 0.0;
            for (int kk = 0; kk <= 1; ++kk) {
                int kc = (int)wrapi((int64_t)k - kk, n[2]);
                for (int jj = 0; jj <= 1; ++jj) {
                    int jc = (int)wrapi((int64_t)j - jj, n[1]);
                    for (int ii = 0; ii <= 1; ++ii) {
                        int ic = (int)wrapi((int64_t)i - ii, n[0]);
                        val += sx[1 - ii] * sy[1 - jj] * sz[1 - kk] * g[(size_t)d * nc + cidx(ic, jc, kc, n)];
                    }
                }
            }
            double vd = a_prev * v[d][p] + half_dt * val;
            vd = vd * a_cur_inv;
            v[d][p] = vd;
        }
        if (do_move == 1) {
            x[p] = x[p] + dt_a_cur_inv * vx[p];
            y[p] = y[p] + dt_a_cur_inv * vy[p];
            z[p] = z[p] + dt_a_cur_inv * vz[p];
        }
    }
}

/*
 * Largest double below hi whose cell index (x-lo)*dxi stays < n: the upper end
 * of the particle domain used by the periodic shift.
 */
double orc_domain_rhi(double lo, double hi, int n) {
    double dxi = 1.0 / ((hi - lo) / (double)n);
    double r = nextafter(hi, -INFINITY);
    while ((int)floor((r - lo) * dxi) >= n) r = nextafter(r, -INFINITY);
    return r;
}

/*
 * Periodic shift of Redistribute: a coordinate above rhi is moved down by the
 * domain length until it is not (clamped to lo if round-off undershoots); a
 * coordinate below lo is moved up until it is not (clamped to rhi).  The clamp
 * values follow AMReX's enforcePeriodic intent; the round-off edge case is not
 * pinned by any available check.  A coordinate more than two periods outside
 * (or not finite) is clamped into the domain instead of shifted.
 */
void orc_enforce_periodic(const int n[3], const double plo[3], const double phi[3],
                          int64_t np, double *x, double *y, double *z) {
    double rhi[3], len[3];
    for (int d = 0; d < 3; ++d) { rhi[d] = orc_domain_rhi(plo[d], phi[d], n[d]); len[d] = phi[d] - plo[d]; }
    double *pos[3] = {x, y, z};
#pragma omp parallel for schedule(static)
    for (int64_t p = 0; p < np; ++p) {
        for (int d = 0; d < 3; ++d) {
            double v = pos[d][p];
            if (!(v >= plo[d] - 2.0 * len[d] && v <= rhi[d] + 2.0 * len[d])) {
                /* more than two periods away or not finite: clamp, no loop */
                pos[d][p] = (v < plo[d]) ? plo[d] : rhi[d];
                continue;
            }
            if (v > rhi[d]) {
                while (v > rhi[d]) v = v - len[d];
                if (v < plo[d]) v = plo[d];
            } else if (v < plo[d]) {
                while (v < plo[d]) v = v + len[d];
                if (v > rhi[d]) v = rhi[d];
            }
            pos[d][p] = v;
        }
    }
}

/* ------------------------------------------------------------------------- */
/* Whole PM step                                                              */
/* ------------------------------------------------------------------------- */

static double now_s(void) {
#ifdef _OPENMP
    return omp_get_wtime();
#else
    return (double)clock() / CLOCKS_PER_SEC;
#endif
}

/*
 * Build the Poisson right-hand side from particles and solve for phi
 * (AddParticlesToRhs + solve_for_phi): rho = deposit; f = CorrectRhsUsingOffset(rho);
 * f = f*Ggravity*(1/a); MG solve with phi as the initial guess.
 * stats: [0]=iters [1]=resnorm0 [2]=rhsnorm [3]=final [4]=target [5]=t_deposit
 * [6]=t_solve
 */
int orc_solve_for_phi(const int n[3], const double plo[3], const double phi_hi[3],
                      int64_t np, const double *x, const double *y, const double *z,
                      const double *m, double m_const, double four_pi_G, double a,
                      double rel_tol, double abs_tol, double *phi, double *rho_out, double *stats) {
    const size_t nc = (size_t)n[0] * n[1] * n[2];
    double dx[3], dxi[3];
    geom_dx(n, plo, phi_hi, dx, dxi);
    double *rhs = (double *)malloc(nc * sizeof(double));
    if (!rhs) return ORC_ERR_NOMEM;
    double t0 = now_s();
    int rc = orc_deposit_cic(n, plo, phi_hi, np, x, y, z, m, m_const, rhs);
    if (rc) { free(rhs); return rc; }
    if (rho_out) memcpy(rho_out, rhs, nc * sizeof(double));
    double mo = orc_mass_offset(plo, phi_hi, orc_sum_particle_mass(np, m, m_const));
    orc_correct_rhs_using_offset(n, plo, phi_hi, rhs, mo);
    orc_scale_rhs((int64_t)nc, rhs, four_pi_G, a);
    double t1 = now_s();
    int iters = 0;
    double st[4] = {0, 0, 0, 0};
    rc = orc_mg_solve(n, dx, rhs, phi, rel_tol, abs_tol, 200, 0, 2, 2, &iters, st);
    double t2 = now_s();
    if (stats) {
        stats[0] = iters; stats[1] = st[0]; stats[2] = st[1]; stats[3] = st[2]; stats[4] = st[3];
        stats[5] = t1 - t0; stats[6] = t2 - t1;
    }
    free(rhs);
    return rc;
}

/*
 * One KDK particle step on one periodic level (Nyx_advance_particles.cpp:137-268
 * followed by the post_timestep wrap, Nyx.cpp:1421-1436):
 *   [nyx_faithful: phi = solve from a zero guess at the current positions]
 *   g_old = grav_vector(phi); moveKickDrift(g_old, dt, a_old, a_half)
 *   phi   = solve at the drifted positions, previous phi as guess
 *   g_new = grav_vector(phi); moveKick(g_new, dt, a_new, a_half); wrap positions
 * a_half = 0.5*(a_old + a_new).  On entry phi must hold the solution for the
 * current positions unless nyx_faithful is set.  g_out (3*nc, optional) receives
 * g_new.  stats (optional, 12): [0]=iters [1]=final resnorm [2]=target
 * [3]=t_total [4]=t_deposit [5]=t_solve [6]=t_grav+kick [7]=iters of the
 * faithful re-solve.
 */
int orc_pm_step(const int n[3], const double plo[3], const double phi_hi[3],
                int64_t np, double *x, double *y, double *z,
                double *vx, double *vy, double *vz, const double *m, double m_const,
                double *phi, double dt, double a_old, double a_new,
                double four_pi_G, double rel_tol, double abs_tol, int nyx_faithful,
                double *g_out, double *stats) {
    const size_t nc = (size_t)n[0] * n[1] * n[2];
    double *g = (double *)malloc(3 * nc * sizeof(double));
    if (!g) return ORC_ERR_NOMEM;
    double T0 = now_s();
    double st[8];
    int rc;
    double faithful_iters = 0;
    if (nyx_faithful) {
        memset(phi, 0, nc * sizeof(double));
        rc = orc_solve_for_phi(n, plo, phi_hi, np, x, y, z, m, m_const, four_pi_G, a_old,
                               rel_tol, abs_tol, phi, NULL, st);
        if (rc) { free(g); return rc; }
        faithful_iters = st[0];
    }
    const double a_half = 0.5 * (a_old + a_new);
    double t0 = now_s();
    orc_grav_vector(n, plo, phi_hi, phi, g);
    orc_update_particles(n, plo, phi_hi, np, x, y, z, vx, vy, vz, g, dt, a_old, a_half, 1);
    double t1 = now_s();
    rc = orc_solve_for_phi(n, plo, phi_hi, np, x, y, z, m, m_const, four_pi_G, a_new,
                           rel_tol, abs_tol, phi, NULL, st);
    if (rc) { free(g); return rc; }
    double t2 = now_s();
    orc_grav_vector(n, plo, phi_hi, phi, g);
    orc_update_particles(n, plo, phi_hi, np, x, y, z, vx, vy, vz, g, dt, a_half, a_new, 0);
    orc_enforce_periodic(n, plo, phi_hi, np, x, y, z);
    double t3 = now_s();
    if (g_out) memcpy(g_out, g, 3 * nc * sizeof(double));
    if (stats) {
        stats[0] = st[0]; stats[1] = st[3]; stats[2] = st[4];
        stats[3] = t3 - T0; stats[4] = st[5]; stats[5] = st[6];
        stats[6] = (t1 - t0) + (t3 - t2); stats[7] = faithful_iters;
    }
    free(g);
    return ORC_OK;
}
