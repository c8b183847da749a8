/*
 * tcg_oracle.c -- CPU FP64 restatement of the reference's Clifford-Gaussian
 * (toroidal) integral path.  TEST INFRASTRUCTURE ONLY: the tests, the smoke
 * check and the CPU-baseline leg of the benchmark use it as the checker; the
 * engine never links or calls it.
 *
 * Each function restates one reference routine (cited file:line, paths
 * relative to the reference tree) from the mathematics it implements, in the
 * same order of operations where that order fixes the rounding:
 *
 *   orc_iv_scaled            src/module/bessel_iv_log.f90:43-328 (CUSF iv_log, 7 methods)
 *   debye_u                  src/module/u_polynomials.f90 (Debye polynomials u1..u13,
 *                            generated here from their defining recurrence)
 *   orc_bessel_vector        src/module/clifford_tools.f90:386-421 (Miller, start N+30+sqrt(80A))
 *   orc_bessel_vector_fast   src/module/clifford_tools.f90:429-461 (Miller, start at N)
 *   nmax_est                 src/module/clifford_tools.f90:566-572
 *   fourier_sin_product      src/module/clifford_tools.f90:578-607
 *   pair_vector              src/module/clifford_tools.f90:467-549 (without the memo:
 *                            a memo hit returns exactly what would be computed)
 *   orc_fast_quad_points     src/module/precompute_t.f90:906-942
 *   orc_set_nodes            (explicit 64/128-point table, :495-891, as data)
 *   setup_tables             src/module/clifford_tools.f90:620-709
 *   orc_bary                 src/Toroidal/tools.f90:1-262, src/module/heaviside.f90
 *   hermite_E                src/module/clifford_tools.f90:757-785
 *   eri_block                src/Toroidal/ERI_shell_block_toroidal.f90:28-390
 *   orc_eri_full             src/Toroidal/ERI_toroidal.f90:36-562 (regular-lattice path)
 *   orc_jk                   src/Energies/hartree_exchange_potential.f90:30-202
 *
 * Pinned by the tests against: the reference's own routines compiled from its
 * sources (oracle/_ref, where it is built), scipy.special.ive, numpy's
 * Gauss-Legendre rule, the a->0 (large torus) limit of ordinary Gaussian
 * integrals, brute-force periodic quadrature of tiny quartets, and the
 * 8-fold / translation invariances.
 */
#include <math.h>
#include <complex.h>
#include <stdlib.h>
#include <string.h>
#include <stdio.h>

#define NMC 16                 /* NMAX_CLIFF, src/module/clifford_tools.f90:82 */
#define MAXQ 256

static const double PI = 3.14159265358979323846;

/* ------------------------------------------------------------------ */
/* Debye polynomials u_k(t), k = 1..13, of the uniform asymptotic      */
/* expansion of I_v (DLMF 10.41.10):  u_0 = 1,                        */
/*   u_{k+1}(t) = t^2 (1-t^2)/2 u_k'(t) + 1/8 Int_0^t (1-5s^2) u_k(s) ds */
/* coefficients in powers of t (degree 3k).                           */
/* ------------------------------------------------------------------ */
static double debye_c[14][40];
static int debye_ready = 0;

static void debye_init(void) {
    if (debye_ready) return;
    memset(debye_c, 0, sizeof(debye_c));
    debye_c[0][0] = 1.0;
    for (int k = 0; k < 13; ++k) {
        double *u = debye_c[k], *v = debye_c[k + 1];
        for (int j = 0; j <= 3 * k; ++j) {
            if (u[j] == 0.0) continue;
            /* t^2 (1 - t^2)/2 * d/dt (u_j t^j) = j u_j/2 (t^{j+1} - t^{j+3}) */
            if (j > 0) {
                v[j + 1] += 0.5 * j * u[j];
                v[j + 3] -= 0.5 * j * u[j];
            }
            /* 1/8 Int (1 - 5 s^2) u_j s^j = u_j/8 (t^{j+1}/(j+1) - 5 t^{j+3}/(j+3)) */
            v[j + 1] += u[j] / (8.0 * (j + 1));
            v[j + 3] -= 5.0 * u[j] / (8.0 * (j + 3));
        }
    }
    debye_ready = 1;
}

static double debye_u(int k, double t) {
    double acc = 0.0;
    for (int j = 3 * k; j >= 0; --j) acc = acc * t + debye_c[k][j];
    return acc;
}

/* log of exp(-x) I_v(x), the 7 methods of iv_log (bessel_iv_log.f90:43-318) */
static double iv_log(double v, double x) {
    const double LOG_INV_SQRT_2PI = -0.91893853320467267;
    const double LOG_2 = 0.69314718055994529;
    if (x < 0.0) return NAN;
    if (x == 0.0) return v == 0.0 ? 0.0 : -HUGE_VAL;
    if (v < 0.0) v = -v;
    double log_x = log(x);
    double log_v = (v == 0.0) ? -1000.0 : log(v);

    if ((x > 1.4e3 && v < 3.05) ||
        ((0.6229 * log_x - 3.2318) > log_v && v > 3.1 && pow(v, 8) / (384.0 * pow(x, 4)) < 1e-16)) {
        /* method 1: Hankel asymptotic series, terms k = 1,3,5 */
        double mu = 4.0 * v * v, cur = 1.0, s = 1.0;
        for (int k = 1, c = 1; k <= 5; k += 2, ++c) {
            cur *= -(mu - (double)(k * k)) / ((double)c * 8.0 * x);
            s += cur;
        }
        return LOG_INV_SQRT_2PI - 0.5 * log_x + log(fabs(s));
    }
    if ((x > 30.0 && v < 15.3919) ||
        ((0.5113 * log_x + 0.7939 > log_v) && x > 59.6925 && v < sqrt(3.0 * x))) {
        /* method 2: Hankel series to k = 39 */
        double mu = 4.0 * v * v, cur = 1.0, s = 1.0;
        for (int k = 1, c = 1; k <= 39; k += 2, ++c) {
            cur *= -(mu - (double)(k * k)) / ((double)c * 8.0 * x);
            s += cur;
        }
        return LOG_INV_SQRT_2PI - 0.5 * log_x + log(fabs(s));
    }
    int order = 0;
    if ((x > 274.2377 && v > 0.3) || v > 163.6993) order = 4;
    else if ((x > 84.4153 && v > 0.46) || v > 56.9971) order = 6;
    else if ((x > 35.9074 && v > 0.6) || v > 20.1534) order = 9;
    else if ((x > 19.6931 && v > 0.7) || v > 12.6964) order = 13;
    if (order) {
        /* methods 3-6: Debye (uniform) expansion with u_1..u_order */
        double u = v / x, xp, xp2, v_eta;
        if (u < 0.5) {
            double asinh_u = asinh(u);
            double t1 = u / (sqrt(1.0 + u * u) + 1.0);
            v_eta = v * (t1 - asinh_u);
            xp = 1.0 / u;
            xp2 = 1.0 / (u * u);
        } else {
            xp = 1.0 / u;
            xp2 = xp * xp;
            double s = sqrt(1.0 + xp2);
            v_eta = v * (s - xp + log(xp / (1.0 + s)));
        }
        double t = 1.0 / sqrt(1.0 + xp2);
        debye_init();
        double s = 1.0, vp = 1.0;
        for (int k = 1; k <= order; ++k) {
            vp *= v;
            s += debye_u(k, t) / vp;
        }
        return LOG_INV_SQRT_2PI - 0.5 * log(v) + v_eta - 0.25 * log(1.0 + xp2) + log(fabs(s));
    }
    /* method 7: power series with 41 terms, summed around its peak term */
    double terms[41];
    double x2_4 = x * x / 4.0, inv = 1.0 / x2_4, vinv = v * inv;
    terms[0] = -lgamma(v + 1.0);
    for (int k = 1; k < 41; ++k) terms[k] = terms[k - 1] - log((double)k * ((double)k * inv + vinv));
    int peak = (int)floor((-v + sqrt(v * v + x * x)) / 2.0);
    if (peak < 0) peak = 0;
    if (peak > 40) peak = 40;   /* terms(peak_k) is read inside 0..40 for these arguments */
    double s = 0.0;
    for (int k = 0; k < 41; ++k) s += exp(terms[k] - terms[peak]);
    return v * (log_x - LOG_2) + terms[peak] + log(s) - x;
}

double orc_iv_scaled(int v, double x) { return exp(iv_log((double)v, x)); }

/* exp(-A) I_n(A), n = 0..N, Miller backward recurrence (clifford_tools.f90:386-421) */
void orc_bessel_vector(int N, double A, double *Iv) {
    if (A < 1e-12) {
        for (int n = 0; n <= N; ++n) Iv[n] = 0.0;
        Iv[0] = 1.0;
        return;
    }
    int n0 = N + 30 + (int)sqrt(80.0 * A);
    double *w = (double *)malloc(sizeof(double) * (n0 + 1));
    double ip1 = 0.0, ik = 1e-300;
    w[n0] = ik;
    for (int k = n0; k >= 1; --k) {
        double im1 = ip1 + (2.0 * (double)k / A) * ik;
        ip1 = ik;
        ik = im1;
        w[k - 1] = ik;
        if (fabs(ik) > 1e250) {
            for (int m = k - 1; m <= n0; ++m) w[m] *= 1e-250;
            ip1 *= 1e-250;
            ik *= 1e-250;
        }
    }
    double norm = orc_iv_scaled(0, A) / w[0];
    for (int n = 0; n <= N; ++n) Iv[n] = w[n] * norm;
    free(w);
}

/* same, recurrence started at N (clifford_tools.f90:429-461) */
void orc_bessel_vector_fast(int N, double A, double *Iv) {
    if (A < 1e-10) {
        for (int n = 0; n <= N; ++n) Iv[n] = 0.0;
        Iv[0] = 1.0;
        return;
    }
    if (N < 2) {
        for (int k = 0; k <= N; ++k) Iv[k] = orc_iv_scaled(k, A);
        return;
    }
    double inv_A2 = 2.0 / A;
    Iv[N] = 1e-280;
    Iv[N - 1] = Iv[N] * (double)N * inv_A2;
    for (int k = N - 1; k >= 1; --k) {
        Iv[k - 1] = Iv[k + 1] + (double)k * inv_A2 * Iv[k];
        if (Iv[k - 1] > 1e250)
            for (int m = k - 1; m <= N; ++m) Iv[m] *= 1e-250;
    }
    double s = orc_iv_scaled(0, A) / Iv[0];
    for (int n = 0; n <= N; ++n) Iv[n] *= s;
}

/* truncation order of exp(-A) I_n(A) (clifford_tools.f90:566-572) */
int orc_nmax_est(double A) { return 20 + (int)ceil(9.0 * sqrt(A > 0.0 ? A : 0.0)); }

/* Fourier coefficients of sin^i(u+phA) sin^l(u+phB): d[k+(i+l)], k=-(i+l)..(i+l) */
static void fourier_sin_product(int i, int l, double phA, double phB, double complex *d) {
    int m2 = i + l;
    double complex tmp[2 * 2 * NMC + 3];
    for (int k = 0; k <= 2 * m2; ++k) d[k] = 0.0;
    d[m2] = 1.0;
    int deg = 0;
    for (int m = 1; m <= m2; ++m) {
        double ph = (m <= i) ? phA : phB;
        double complex cp = cexp(I * ph) / (2.0 * I);
        double complex cm = -cexp(-I * ph) / (2.0 * I);
        for (int k = 0; k <= 2 * m2 + 2; ++k) tmp[k] = 0.0;
        /* tmp index shifted by +1 relative to d: tmp[k+m2+1] <-> k */
        for (int k = -deg; k <= deg; ++k) {
            tmp[k + 1 + m2 + 1] += cp * d[k + m2];
            tmp[k - 1 + m2 + 1] += cm * d[k + m2];
        }
        ++deg;
        for (int k = -deg; k <= deg; ++k) d[k + m2] = tmp[k + m2 + 1];
    }
}

/* pair Fourier vector V_n = sum_k d_k I~_{|n - sgn k|}(A), n = 0..N (clifford_tools.f90:467-549) */
static void pair_vector(int N, double A, int i, int l, double phA, double phB, int sgn, double complex *V) {
    double *Iv = (double *)malloc(sizeof(double) * (N + i + l + 1));
    orc_bessel_vector_fast(N + i + l, A, Iv);
    if (i + l == 0) {
        for (int m = 0; m <= N; ++m) V[m] = Iv[m];
    } else {
        double complex d[2 * 2 * NMC + 1];
        fourier_sin_product(i, l, phA, phB, d);
        for (int m = 0; m <= N; ++m) {
            double complex acc = 0.0;
            for (int kk = -(i + l); kk <= i + l; ++kk) acc += d[kk + i + l] * Iv[abs(m - sgn * kk)];
            V[m] = acc;
        }
    }
    free(Iv);
}

/* Gauss-Legendre in u on [0,1] mapped by t = 0.6 u/(1-u) (precompute_t.f90:906-942) */
void orc_fast_quad_points(int n, double *t, double *w) {
    const double s = 0.6;
    for (int i = 1; i <= n; ++i) {
        double x = cos(PI * ((double)i - 0.25) / ((double)n + 0.5)), p0, p1, p2, dp, dx;
        for (int it = 1; it <= 100; ++it) {
            p0 = 1.0; p1 = x;
            for (int j = 2; j <= n; ++j) {
                p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / (double)j;
                p0 = p1; p1 = p2;
            }
            dp = (double)n * (x * p1 - p0) / (x * x - 1.0);
            dx = p1 / dp;
            x = x - dx;
            if (fabs(dx) < 1e-15) break;
        }
        p0 = 1.0; p1 = x;
        for (int j = 2; j <= n; ++j) {
            p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / (double)j;
            p0 = p1; p1 = p2;
        }
        dp = (double)n * (x * p1 - p0) / (x * x - 1.0);
        double u = 0.5 * (1.0 - x);
        double wu = 1.0 / ((1.0 - x * x) * dp * dp);
        t[i - 1] = s * u / (1.0 - u);
        w[i - 1] = wu * s / ((1.0 - u) * (1.0 - u));
    }
}

/* ------------------------------------------------------------------ */
/* system state                                                        */
/* ------------------------------------------------------------------ */
typedef struct {
    int ndp;
    double L[3], a[3];
    int ncell[3], nfuc, nbas;
    double *xyz;     /* nbas*3 */
    int *lxyz;       /* nbas*3 */
    int *poff;       /* nbas+1 */
    double *expo, *coef;
    /* t-tables */
    int nq, ncap;
    double t[MAXQ], w[MAXQ];
    double *tbes;    /* [d][k][n] : (ncap+1) per node */
    int *tnmax;      /* [d][k] */
    /* counters */
    long long n_shell_quartets, n_prim_quartets;
} osys;

static osys S_;

static void bary_exponent(double e1, double e2, double x, double L, double a, double *p) {
    const double eps = 1e-6;
    double r;
    if (fabs(x) > 0.5 * L - eps && fabs(x) < 0.5 * L + eps) {
        r = (fabs(e1 - e2) < 1e-8) ? 0.0 : fabs(e1 - e2);
    } else {
        r = e1 * e1 + e2 * e2 + 2.0 * e1 * e2 * cos(a * x);
        r = sqrt(fabs(r));
    }
    if (r > e1 + e2) r = e1 + e2;
    if (e1 + e2 - r < 1e-14) r = e1 + e2;
    *p = r;
}

static void bary_center(double e1, double e2, double r1, double r2, double L, double a, double *rp) {
    const double eps = 1e-6;
    double d = fabs(r1 - r2);
    if (d < 0.5 * L + eps && d > 0.5 * L - eps) {
        if (fabs(e1 - e2) < 1e-10) *rp = 0.5 * (r1 + r2);
        else if (e1 > e2) *rp = r1;
        else if (e2 > e1) *rp = r2;
        else *rp = 0.5 * (r1 + r2);
    } else if (fabs(e1 - e2) < 1e-10) {
        if (d > 0.5 * L + eps) *rp = 0.5 * (r1 + r2) - 0.5 * L;
        else *rp = 0.5 * (r1 + r2);
    } else {
        double val = e1 * cos(a * r1) + e2 * cos(a * r2);
        double H = (-val >= 0.0) ? 1.0 : 0.0;
        *rp = atan((e1 * sin(a * r1) + e2 * sin(a * r2)) / val) / a + 0.5 * L * H;
    }
}

/* d is 1-based as in the reference */
void orc_bary(int d, double e1, double e2, double x1, double x2, double *pe, double *cen) {
    bary_exponent(e1, e2, x1 - x2, S_.L[d - 1], S_.a[d - 1], pe);
    bary_center(e1, e2, x1, x2, S_.L[d - 1], S_.a[d - 1], cen);
}

/* McMurchie-Davidson E_t, t = 0..i+j (clifford_tools.f90:757-785) */
static void hermite_E(int i, int j, double p, double PA, double PB, double *E) {
    double Enew[NMC + 2], Eold[NMC + 4];
    double inv2p = 0.5 / p;
    for (int t = 0; t < NMC + 4; ++t) Eold[t] = 0.0;
    double *Eo = Eold + 1;          /* Eo[-1] valid */
    Eo[0] = 1.0;
    int deg = 0;
    for (int m = 1; m <= i + j; ++m) {
        double X = (m <= i) ? PA : PB;
        for (int t = 0; t <= deg + 1; ++t) Enew[t] = inv2p * Eo[t - 1] + X * Eo[t] + (double)(t + 1) * Eo[t + 1];
        ++deg;
        for (int t = 0; t <= deg; ++t) Eo[t] = Enew[t];
        Eo[deg + 1] = 0.0;
    }
    for (int t = 0; t <= NMC; ++t) E[t] = 0.0;
    for (int t = 0; t <= deg; ++t) E[t] = Eo[t];
}

int orc_init(int ndp, const double *L, const int *ncell, int nfuc, int nq) {
    osys *s = &S_;
    s->ndp = ndp;
    for (int d = 0; d < 3; ++d) {
        s->L[d] = L[d];
        s->a[d] = L[d] > 0.0 ? 2.0 * PI / L[d] : 0.0;
        s->ncell[d] = ncell[d] > 0 ? ncell[d] : 1;
    }
    s->nfuc = nfuc;
    if (nq < 1 || nq > MAXQ) return -1;
    s->nq = nq;
    orc_fast_quad_points(nq, s->t, s->w);
    return 0;
}

/* an explicit rule (the 64/128-point tables of gauss_legendre_points,
 * src/module/precompute_t.f90:495-891, taken as data); call before
 * orc_setup_tables */
int orc_set_nodes(int nq, const double *t, const double *w) {
    osys *s = &S_;
    if (nq < 1 || nq > MAXQ) return -1;
    s->nq = nq;
    memcpy(s->t, t, sizeof(double) * nq);
    memcpy(s->w, w, sizeof(double) * nq);
    return 0;
}

int orc_set_basis(int nbas, const double *xyz, const int *lxyz, const int *poff,
                  const double *expo, const double *coef) {
    osys *s = &S_;
    free(s->xyz); free(s->lxyz); free(s->poff); free(s->expo); free(s->coef);
    s->nbas = nbas;
    int np = poff[nbas];
    s->xyz = malloc(sizeof(double) * 3 * nbas);
    s->lxyz = malloc(sizeof(int) * 3 * nbas);
    s->poff = malloc(sizeof(int) * (nbas + 1));
    s->expo = malloc(sizeof(double) * np);
    s->coef = malloc(sizeof(double) * np);
    memcpy(s->xyz, xyz, sizeof(double) * 3 * nbas);
    memcpy(s->lxyz, lxyz, sizeof(int) * 3 * nbas);
    memcpy(s->poff, poff, sizeof(int) * (nbas + 1));
    memcpy(s->expo, expo, sizeof(double) * np);
    memcpy(s->coef, coef, sizeof(double) * np);
    return 0;
}

/* t-tables for ncap (clifford_tools.f90:620-709) */
static void setup_tables(int ncap) {
    osys *s = &S_;
    free(s->tbes); free(s->tnmax);
    s->ncap = ncap;
    s->tbes = malloc(sizeof(double) * (size_t)s->ndp * s->nq * (ncap + 1));
    s->tnmax = malloc(sizeof(int) * s->ndp * s->nq);
    for (int d = 0; d < s->ndp; ++d)
        for (int k = 0; k < s->nq; ++k) {
            double *row = s->tbes + ((size_t)d * s->nq + k) * (ncap + 1);
            orc_bessel_vector(ncap, 2.0 * s->t[k] * s->t[k] / (s->a[d] * s->a[d]), row);
            int nm = ncap;
            for (int n = 0; n <= ncap; ++n)
                if (row[n] < 1e-16 * row[0]) { nm = (n + 5 < ncap) ? n + 5 : ncap; break; }
            s->tnmax[d * s->nq + k] = nm;
        }
}

/* ncap as the ERI driver sets it (ERI_toroidal.f90:94-104) */
int orc_setup_tables(void) {
    osys *s = &S_;
    int lmax = 0;
    double emax = 0.0, amin = 1e300;
    for (int i = 0; i < s->nbas; ++i) {
        for (int d = 0; d < s->ndp; ++d) if (s->lxyz[3 * i + d] > lmax) lmax = s->lxyz[3 * i + d];
        for (int p = s->poff[i]; p < s->poff[i + 1]; ++p) if (s->expo[p] > emax) emax = s->expo[p];
    }
    for (int d = 0; d < s->ndp; ++d) if (s->a[d] < amin) amin = s->a[d];
    int ncap = orc_nmax_est(4.0 * emax / (amin * amin)) + 2 * lmax + 1;
    setup_tables(ncap);
    return ncap;
}

void orc_get_tables(double *t, double *w, double *tbes, int *tnmax) {
    osys *s = &S_;
    memcpy(t, s->t, sizeof(double) * s->nq);
    memcpy(w, s->w, sizeof(double) * s->nq);
    if (tbes) memcpy(tbes, s->tbes, sizeof(double) * (size_t)s->ndp * s->nq * (s->ncap + 1));
    if (tnmax) memcpy(tnmax, s->tnmax, sizeof(int) * s->ndp * s->nq);
}


/* ------------------------------------------------------------------ */
/* the shell-quartet kernel (ERI_shell_block_toroidal.f90:28-390)       */
/* shells are given as (first AO, number of components); blk is          */
/* column-major blk(nA,nB,nC,nD).                                        */
/* ------------------------------------------------------------------ */
#define LMX 5                                   /* up to g shells */
static int lsum(int ao) { return S_.lxyz[3 * ao] + S_.lxyz[3 * ao + 1] + S_.lxyz[3 * ao + 2]; }

typedef struct {                                 /* per primitive pair */
    double fac;                                  /* prefactor, coefficient free */
    double cen[3], q;
    int N[3];                                    /* Bessel lengths, periodic d */
    double E[3][LMX][LMX][2 * NMC + 1];          /* Hermite E, real d */
} pairdat;

static void pair_setup(int ndp, double e1, double e2, const double *r1, const double *r2,
                       int l1, int l2, int ncap, pairdat *P) {
    const double *av = S_.a, *Lc = S_.L;
    double esum = e1 + e2, mu = e1 * e2 / esum, cf = 1.0;
    P->q = esum;
    for (int d = 0; d < 3; ++d) {
        if (d < ndp) {
            double pe, cen;
            bary_exponent(e1, e2, r1[d] - r2[d], Lc[d], av[d], &pe);
            bary_center(e1, e2, r1[d], r2[d], Lc[d], av[d], &cen);
            P->cen[d] = cen;
            cf *= exp(-2.0 * (esum - pe) / (av[d] * av[d]));
            int est = orc_nmax_est(2.0 * pe / (av[d] * av[d]));
            P->N[d] = est < ncap ? est : ncap;
        } else {
            double cen = (e1 * r1[d] + e2 * r2[d]) / esum;
            P->cen[d] = cen;
            cf *= exp(-mu * (r1[d] - r2[d]) * (r1[d] - r2[d]));
            for (int i = 0; i <= l1; ++i)
                for (int j = 0; j <= l2; ++j)
                    hermite_E(i, j, esum, cen - r1[d], cen - r2[d], P->E[d][i][j]);
        }
    }
    P->fac = cf;
}

/* G/H vectors of a pair along periodic d for all power pairs: V[(i*(l2+1)+j)*(N+1) + n] */
static double complex *pair_vectors(int d, double e1, double e2, const double *r1, const double *r2,
                                    int l1, int l2, int N, int sgn) {
    double pe, cen;
    bary_exponent(e1, e2, r1[d] - r2[d], S_.L[d], S_.a[d], &pe);
    bary_center(e1, e2, r1[d], r2[d], S_.L[d], S_.a[d], &cen);
    double A = 2.0 * pe / (S_.a[d] * S_.a[d]);
    double complex *V = malloc(sizeof(double complex) * (size_t)(l1 + 1) * (l2 + 1) * (N + 1));
    for (int i = 0; i <= l1; ++i)
        for (int j = 0; j <= l2; ++j)
            pair_vector(N, A, i, j, S_.a[d] * (cen - r1[d]), S_.a[d] * (cen - r2[d]), sgn,
                        V + (size_t)(i * (l2 + 1) + j) * (N + 1));
    return V;
}

typedef struct {
    int ps_on, ps_noscreen, mp;
    const double *ps_bra, *ps_ket;   /* ps(p1,p2) at p1 + mp*p2 */
    int p0[4], p1[4];                /* primitive sub-ranges (all when p1 < 0) */
    long long nprim;
} kopts;

static int eri_block_impl(int fa, int nA, int fb, int nB, int fc, int nC, int fd, int nD, double *blk, kopts *o) {
    osys *s = &S_;
    const int ndp = s->ndp, nq = s->nq, ncap = s->ncap;
    const double screen_tol = 2.22e-16;
    for (int x = 0; x < nA * nB * nC * nD; ++x) blk[x] = 0.0;
    if (!s->tbes) return -1;

    const int lsA = lsum(fa), lsB = lsum(fb), lsC = lsum(fc), lsD = lsum(fd);
    const int NN = lsA + lsB + lsC + lsD;
    const int ncomb = (lsA + 1) * (lsB + 1) * (lsC + 1) * (lsD + 1);
    const double *av = s->a, *Lc = s->L;
    const double *rA = s->xyz + 3 * fa, *rB = s->xyz + 3 * fb, *rC = s->xyz + 3 * fc, *rD = s->xyz + 3 * fd;
    double fac_const = 2.0 / sqrt(PI);
    for (int d = 0; d < ndp; ++d) fac_const *= Lc[d] * Lc[d];
    const int n1 = s->poff[fa + 1] - s->poff[fa], n2 = s->poff[fb + 1] - s->poff[fb];
    const int n3 = s->poff[fc + 1] - s->poff[fc], n4 = s->poff[fd + 1] - s->poff[fd];
    int lo[4] = {0, 0, 0, 0}, hi[4] = {n1, n2, n3, n4};
    for (int x = 0; x < 4; ++x) if (o->p1[x] >= 0) { lo[x] = o->p0[x]; hi[x] = o->p1[x]; }
    const double *exA = s->expo + s->poff[fa], *exB = s->expo + s->poff[fb];
    const double *exC = s->expo + s->poff[fc], *exD = s->expo + s->poff[fd];
#define CO(base, comp, prim) (s->coef[s->poff[(base) + (comp)] + (prim)])

    int nket = n3 * n4;
    pairdat *kp = malloc(sizeof(pairdat) * nket);
    double complex **H = calloc((size_t)nket * 3, sizeof(double complex *));
    double ket_max = 0.0;
    for (int k = 0, kl = 0; k < n3; ++k)
        for (int l = 0; l < n4; ++l, ++kl) {
            if (k < lo[2] || k >= hi[2] || l < lo[3] || l >= hi[3]) continue;
            pair_setup(ndp, exC[k], exD[l], rC, rD, lsC, lsD, ncap, &kp[kl]);
            double mc = 0.0, md = 0.0;
            for (int c = 0; c < nC; ++c) if (fabs(CO(fc, c, k)) > mc) mc = fabs(CO(fc, c, k));
            for (int c = 0; c < nD; ++c) if (fabs(CO(fd, c, l)) > md) md = fabs(CO(fd, c, l));
            double v = fabs(kp[kl].fac) * mc * md;
            if (v > ket_max) ket_max = v;
        }

    double *Idir = malloc(sizeof(double) * 3 * ncomb * nq);      /* [d][cmb][k] */
    double *F = malloc(sizeof(double) * (ncap + 2));
    double *cph = malloc(sizeof(double) * (ncap + 2)), *sph = malloc(sizeof(double) * (ncap + 2));
    double rw[MAXQ], rs[MAXQ], gt[MAXQ][4 * LMX + 1];
    int pwA[6 * 6][3], pwB[36][3], pwC[36][3], pwD[36][3];
    for (int c = 0; c < nA; ++c) for (int d = 0; d < 3; ++d) pwA[c][d] = s->lxyz[3 * (fa + c) + d];
    for (int c = 0; c < nB; ++c) for (int d = 0; d < 3; ++d) pwB[c][d] = s->lxyz[3 * (fb + c) + d];
    for (int c = 0; c < nC; ++c) for (int d = 0; d < 3; ++d) pwC[c][d] = s->lxyz[3 * (fc + c) + d];
    for (int c = 0; c < nD; ++c) for (int d = 0; d < 3; ++d) pwD[c][d] = s->lxyz[3 * (fd + c) + d];

    for (int i = lo[0]; i < hi[0]; ++i)
        for (int j = lo[1]; j < hi[1]; ++j) {
            double cAmax = 0.0, cBmax = 0.0;
            for (int c = 0; c < nA; ++c) if (fabs(CO(fa, c, i)) > cAmax) cAmax = fabs(CO(fa, c, i));
            for (int c = 0; c < nB; ++c) if (fabs(CO(fb, c, j)) > cBmax) cBmax = fabs(CO(fb, c, j));
            pairdat bp;
            pair_setup(ndp, exA[i], exB[j], rA, rB, lsA, lsB, ncap, &bp);
            double bra_fac = fac_const * bp.fac;
            if (!o->ps_noscreen && fabs(bra_fac) * cAmax * cBmax * ket_max <= screen_tol) continue;
            double p = bp.q;
            double complex *G[3] = {0, 0, 0};
            for (int d = 0; d < ndp; ++d)
                G[d] = pair_vectors(d, exA[i], exB[j], rA, rB, lsA, lsB, bp.N[d], -1);

            for (int k = 0, kl = 0; k < n3; ++k)
                for (int l = 0; l < n4; ++l, ++kl) {
                    if (k < lo[2] || k >= hi[2] || l < lo[3] || l >= hi[3]) continue;
                    double fac = bra_fac * kp[kl].fac;
                    double mc = 0.0, md = 0.0;
                    for (int c = 0; c < nC; ++c) if (fabs(CO(fc, c, k)) > mc) mc = fabs(CO(fc, c, k));
                    for (int c = 0; c < nD; ++c) if (fabs(CO(fd, c, l)) > md) md = fabs(CO(fd, c, l));
                    if (!o->ps_noscreen && fabs(fac) * cAmax * cBmax * mc * md <= screen_tol) continue;
                    if (o->ps_on && o->ps_bra[i + o->mp * j] * o->ps_ket[k + o->mp * l] < 1e-14) continue;
                    o->nprim++;
                    double q = kp[kl].q;
                    if (!H[3 * kl] && ndp > 0)
                        for (int d = 0; d < ndp; ++d)
                            H[3 * kl + d] = pair_vectors(d, exC[k], exD[l], rC, rD, lsC, lsD, kp[kl].N[d], +1);
                    if (ndp < 3)
                        for (int kq = 0; kq < nq; ++kq) {
                            double t2 = s->t[kq] * s->t[kq];
                            double D2 = 1.0 / (p * q + (p + q) * t2);
                            rw[kq] = p * q * t2 * D2;
                            rs[kq] = PI * sqrt(D2);
                        }
                    for (int d = 0; d < 3; ++d) {
                        if (d < ndp) {
                            int bN = bp.N[d], kN = kp[kl].N[d];
                            int nm = bN < kN ? bN : kN;
                            double phi = av[d] * (bp.cen[d] - kp[kl].cen[d]);
                            double c1 = cos(phi), s1 = sin(phi), cn = 1.0, sn = 0.0;
                            for (int n = 0; n <= nm; ++n) {
                                if (n > 0) {
                                    if (n % 32 == 0) { cn = cos(n * phi); sn = sin(n * phi); }
                                    else { double tmp = cn * c1 - sn * s1; sn = sn * c1 + cn * s1; cn = tmp; }
                                }
                                cph[n] = cn; sph[n] = sn;
                            }
                            const double *tb = s->tbes + (size_t)d * nq * (ncap + 1);
                            const int *tnm = s->tnmax + d * nq;
                            for (int id = 0; id <= lsD; ++id)
                                for (int ic = 0; ic <= lsC; ++ic)
                                    for (int ib = 0; ib <= lsB; ++ib)
                                        for (int ia = 0; ia <= lsA; ++ia) {
                                            int cmb = ia + (lsA + 1) * (ib + (lsB + 1) * (ic + (lsC + 1) * id));
                                            const double complex *g = G[d] + (size_t)(ia * (lsB + 1) + ib) * (bN + 1);
                                            const double complex *h = H[3 * kl + d] + (size_t)(ic * (lsD + 1) + id) * (kN + 1);
                                            double Fmx = 0.0;
                                            for (int n = 0; n <= nm; ++n) {
                                                double complex acc = g[n] * h[n];
                                                F[n] = creal(acc) * cph[n] - cimag(acc) * sph[n];
                                                if (fabs(F[n]) > Fmx) Fmx = fabs(F[n]);
                                            }
                                            int ne = nm;
                                            if (ndp > 1) {    /* F tail below 1e-17 max|F| dropped (2D/3D form) */
                                                int tmx = 0;
                                                for (int kq = 0; kq < nq; ++kq) if (tnm[kq] > tmx) tmx = tnm[kq];
                                                if (ne > tmx) ne = tmx;
                                                Fmx = 0.0;
                                                for (int n = 0; n <= ne; ++n) if (fabs(F[n]) > Fmx) Fmx = fabs(F[n]);
                                                while (ne > 0 && !(fabs(F[ne]) >= 1e-17 * Fmx)) --ne;
                                            }
                                            double sgn = pow(av[d], -(double)(ia + ib + ic + id));
                                            for (int kq = 0; kq < nq; ++kq) {
                                                int nt = ne < tnm[kq] ? ne : tnm[kq];
                                                const double *tk = tb + (size_t)kq * (ncap + 1);
                                                double val = F[0] * tk[0], dot = 0.0;
                                                for (int n = 1; n <= nt; ++n) dot += F[n] * tk[n];
                                                if (nt >= 1) val += 2.0 * dot;
                                                Idir[(d * ncomb + cmb) * nq + kq] = val * sgn;
                                            }
                                        }
                        } else {
                            double Rr = bp.cen[d] - kp[kl].cen[d];
                            for (int kq = 0; kq < nq; ++kq) {
                                double w = rw[kq];
                                gt[kq][0] = (Rr == 0.0) ? rs[kq] : rs[kq] * exp(-w * Rr * Rr);
                                double gm1 = 0.0, g0 = gt[kq][0];
                                for (int n = 0; n < NN; ++n) {
                                    double gp1 = 2.0 * w * Rr * g0 - 2.0 * (double)n * w * gm1;
                                    gt[kq][n + 1] = gp1;
                                    gm1 = g0; g0 = gp1;
                                }
                            }
                            for (int id = 0; id <= lsD; ++id)
                                for (int ic = 0; ic <= lsC; ++ic)
                                    for (int ib = 0; ib <= lsB; ++ib)
                                        for (int ia = 0; ia <= lsA; ++ia) {
                                            int cmb = ia + (lsA + 1) * (ib + (lsB + 1) * (ic + (lsC + 1) * id));
                                            double SN[4 * LMX + 1];
                                            for (int n = 0; n <= NN; ++n) SN[n] = 0.0;
                                            for (int n = 0; n <= ia + ib; ++n)
                                                for (int kk = 0; kk <= ic + id; ++kk)
                                                    SN[n + kk] += ((n % 2 == 1) ? -1.0 : 1.0) * bp.E[d][ia][ib][n] *
                                                                  kp[kl].E[d][ic][id][kk];
                                            for (int kq = 0; kq < nq; ++kq) {
                                                double v = 0.0;
                                                for (int n = 0; n <= ia + ib + ic + id; ++n) v += SN[n] * gt[kq][n];
                                                Idir[(d * ncomb + cmb) * nq + kq] = v;
                                            }
                                        }
                        }
                    }
                    for (int cmb = 0; cmb < ncomb; ++cmb)
                        for (int kq = 0; kq < nq; ++kq) Idir[cmb * nq + kq] *= s->w[kq];
                    for (int cd = 0; cd < nD; ++cd)
                        for (int cc = 0; cc < nC; ++cc)
                            for (int cb = 0; cb < nB; ++cb)
                                for (int ca = 0; ca < nA; ++ca) {
                                    int cx = pwA[ca][0] + (lsA + 1) * (pwB[cb][0] + (lsB + 1) * (pwC[cc][0] + (lsC + 1) * pwD[cd][0]));
                                    int cy = pwA[ca][1] + (lsA + 1) * (pwB[cb][1] + (lsB + 1) * (pwC[cc][1] + (lsC + 1) * pwD[cd][1]));
                                    int cz = pwA[ca][2] + (lsA + 1) * (pwB[cb][2] + (lsB + 1) * (pwC[cc][2] + (lsC + 1) * pwD[cd][2]));
                                    double val = 0.0;
                                    for (int kq = 0; kq < nq; ++kq)
                                        val += Idir[(0 * ncomb + cx) * nq + kq] * Idir[(1 * ncomb + cy) * nq + kq] *
                                               Idir[(2 * ncomb + cz) * nq + kq];
                                    blk[ca + nA * (cb + nB * (cc + nC * cd))] +=
                                        fac * CO(fa, ca, i) * CO(fb, cb, j) * CO(fc, cc, k) * CO(fd, cd, l) * val;
                                }
                }
            for (int d = 0; d < 3; ++d) free(G[d]);
        }
    for (int x = 0; x < 3 * nket; ++x) free(H[x]);
    free(H); free(kp); free(Idir); free(F); free(cph); free(sph);
    return 0;
#undef CO
}

int orc_eri_block(int fa, int nA, int fb, int nB, int fc, int nC, int fd, int nD, double *blk) {
    kopts o;
    memset(&o, 0, sizeof(o));
    for (int x = 0; x < 4; ++x) o.p1[x] = -1;
    int r = eri_block_impl(fa, nA, fb, nB, fc, nC, fd, nD, blk, &o);
    S_.n_prim_quartets += o.nprim;
    return r;
}

void orc_counters(long long *c) { c[0] = S_.n_shell_quartets; c[1] = S_.n_prim_quartets; }

/* ------------------------------------------------------------------ */
/* ERI driver, regular-lattice path (ERI_toroidal.f90:36-470,551-552):  */
/* shells, Schwarz table, primitive Schwarz table, translation x 8-fold */
/* unique shell quartets, scatter with translation to the unit cell.    */
/* eri: Fortran ERI(n1,nbas,nbas,nbas), n1 = nfuc (sao) or nbas.        */
/* Returns -2 when the shells do not repeat in every cell (the          */
/* reference then takes its general loop, not restated here).           */
/* ------------------------------------------------------------------ */
static const int IPERM8[8][4] = {{0, 1, 2, 3}, {1, 0, 2, 3}, {0, 1, 3, 2}, {1, 0, 3, 2},
                                 {2, 3, 0, 1}, {3, 2, 0, 1}, {2, 3, 1, 0}, {3, 2, 1, 0}};

typedef struct {
    int nsh, nshuc, ncell, *sh_first, *sh_n;
    int ncl[3];
    int *relc;                 /* relc[c1*ncell + c2] = cell of (c2 - c1) */
    double *Q, *Qsh;           /* Q[i + nbas*j], Qsh[SI + nsh*SJ] */
    double *psq;               /* ps(p1,p2,SI,SJ) at p1 + mp*(p2 + mp*(SI + nshuc*SJ)) */
    int mp;
} drv;

static int cell_of(const int *ncl, int cx, int cy, int cz) { return cx + ncl[0] * (cy + ncl[1] * cz); }

static int build_shells(drv *D) {
    osys *s = &S_;
    D->sh_first = malloc(sizeof(int) * s->nbas);
    D->sh_n = malloc(sizeof(int) * s->nbas);
    int nsh = 0;
    for (int i = 0; i < s->nbas;) {
        int lt = lsum(i), nc = (lt + 1) * (lt + 2) / 2;
        int same = (i + nc <= s->nbas);
        for (int j = i + 1; same && j < i + nc; ++j) {
            same = lsum(j) == lt && s->xyz[3 * j] == s->xyz[3 * i] && s->xyz[3 * j + 1] == s->xyz[3 * i + 1] &&
                   s->xyz[3 * j + 2] == s->xyz[3 * i + 2] &&
                   (s->poff[j + 1] - s->poff[j]) == (s->poff[i + 1] - s->poff[i]);
            if (same)
                for (int p = 0; p < s->poff[i + 1] - s->poff[i]; ++p)
                    if (s->expo[s->poff[j] + p] != s->expo[s->poff[i] + p]) same = 0;
        }
        if (!same) nc = 1;
        D->sh_first[nsh] = i;
        D->sh_n[nsh] = nc;
        ++nsh;
        i += nc;
    }
    D->nsh = nsh;
    for (int d = 0; d < 3; ++d) D->ncl[d] = s->ncell[d];
    D->ncell = D->ncl[0] * D->ncl[1] * D->ncl[2];
    D->nshuc = 0;
    for (int SI = 0; SI < nsh; ++SI) if (D->sh_first[SI] < s->nfuc) D->nshuc++;
    int ok = (nsh == D->nshuc * D->ncell) && (s->nbas == s->nfuc * D->ncell);
    for (int SI = 0; ok && SI < nsh - D->nshuc; ++SI)
        if (D->sh_n[SI + D->nshuc] != D->sh_n[SI] || D->sh_first[SI + D->nshuc] != D->sh_first[SI] + s->nfuc) ok = 0;
    if (!ok || D->ncell > 4096) return -2;
    int nc = D->ncell;
    D->relc = malloc(sizeof(int) * nc * nc);
    for (int c1 = 0; c1 < nc; ++c1)
        for (int c2 = 0; c2 < nc; ++c2) {
            int n0 = D->ncl[0], n1 = D->ncl[1], n2 = D->ncl[2];
            int dx = ((c2 % n0 - c1 % n0) % n0 + n0) % n0;
            int dy = (((c2 / n0) % n1 - (c1 / n0) % n1) % n1 + n1) % n1;
            int dz = ((c2 / (n0 * n1) - c1 / (n0 * n1)) % n2 + n2) % n2;
            D->relc[c1 * nc + c2] = cell_of(D->ncl, dx, dy, dz);
        }
    return 0;
}

static void free_drv(drv *D) {
    free(D->sh_first); free(D->sh_n); free(D->relc); free(D->Q); free(D->Qsh); free(D->psq);
}

static void schwarz(drv *D) {
    osys *s = &S_;
    int nb = s->nbas, nsh = D->nsh, nshuc = D->nshuc, nc = D->ncell, nf = s->nfuc;
    D->Q = calloc((size_t)nb * nb, sizeof(double));
    D->Qsh = calloc((size_t)nsh * nsh, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int SI = 0; SI < nshuc; ++SI)
        for (int SJ = 0; SJ < nsh; ++SJ) {
            int nI = D->sh_n[SI], nJ = D->sh_n[SJ];
            double *blk = malloc(sizeof(double) * nI * nJ * nI * nJ);
            kopts o;
            memset(&o, 0, sizeof(o));
            for (int x = 0; x < 4; ++x) o.p1[x] = -1;
            eri_block_impl(D->sh_first[SI], nI, D->sh_first[SJ], nJ, D->sh_first[SI], nI, D->sh_first[SJ], nJ, blk, &o);
            for (int a = 0; a < nI; ++a)
                for (int b = 0; b < nJ; ++b) {
                    int i = D->sh_first[SI] + a, j = D->sh_first[SJ] + b;
                    double v = sqrt(fabs(blk[a + nI * (b + nJ * (a + nI * b))]));
                    int ci = i / nf, li = i % nf, cj = j / nf, lj = j % nf;
                    (void)ci;
                    for (int ctr = 0; ctr < nc; ++ctr) {
                        int cc = D->relc[ctr * nc + 0];
                        int r1 = D->relc[cc * nc + 0] * nf + li;
                        int r2 = D->relc[cc * nc + cj] * nf + lj;
                        D->Q[r1 + (size_t)nb * r2] = v;
                        D->Q[r2 + (size_t)nb * r1] = v;
                    }
                }
            free(blk);
        }
    for (int SI = 0; SI < nsh; ++SI)
        for (int SJ = 0; SJ < nsh; ++SJ) {
            double m = 0.0;
            for (int a = 0; a < D->sh_n[SI]; ++a)
                for (int b = 0; b < D->sh_n[SJ]; ++b) {
                    double v = D->Q[(D->sh_first[SI] + a) + (size_t)nb * (D->sh_first[SJ] + b)];
                    if (v > m) m = v;
                }
            D->Qsh[SI + nsh * SJ] = m;
        }
}

static void prim_schwarz(drv *D) {
    osys *s = &S_;
    int nsh = D->nsh, nshuc = D->nshuc;
    int mp = 0;
    for (int SI = 0; SI < nsh; ++SI) {
        int n = s->poff[D->sh_first[SI] + 1] - s->poff[D->sh_first[SI]];
        if (n > mp) mp = n;
    }
    D->mp = mp;
    D->psq = calloc((size_t)mp * mp * nshuc * nsh, sizeof(double));
#pragma omp parallel for schedule(dynamic, 1) collapse(2)
    for (int SI = 0; SI < nshuc; ++SI)
        for (int SJ = 0; SJ < nsh; ++SJ) {
            int nI = D->sh_n[SI], nJ = D->sh_n[SJ];
            int fI = D->sh_first[SI], fJ = D->sh_first[SJ];
            double *blk = malloc(sizeof(double) * nI * nJ * nI * nJ);
            for (int p1 = 0; p1 < s->poff[fI + 1] - s->poff[fI]; ++p1)
                for (int p2 = 0; p2 < s->poff[fJ + 1] - s->poff[fJ]; ++p2) {
                    kopts o;
                    memset(&o, 0, sizeof(o));
                    o.ps_noscreen = 1;
                    o.p0[0] = o.p0[2] = p1; o.p1[0] = o.p1[2] = p1 + 1;
                    o.p0[1] = o.p0[3] = p2; o.p1[1] = o.p1[3] = p2 + 1;
                    eri_block_impl(fI, nI, fJ, nJ, fI, nI, fJ, nJ, blk, &o);
                    double m = 0.0;
                    for (int a = 0; a < nI; ++a)
                        for (int b = 0; b < nJ; ++b) {
                            double v = sqrt(fabs(blk[a + nI * (b + nJ * (a + nI * b))]));
                            if (v > m) m = v;
                        }
                    D->psq[p1 + mp * (p2 + mp * (SI + (size_t)nshuc * SJ))] = m;
                }
            free(blk);
        }
}

/* stats: [0] canonical shell quartets computed, [1] primitive quartets, [2] elements stored */
int orc_eri_full(int sao, int physicist, double *eri, long long *stats) {
    osys *s = &S_;
    const double schwarz_thresh = 1e-13;
    drv D;
    memset(&D, 0, sizeof(D));
    orc_setup_tables();
    int rc = build_shells(&D);
    if (rc) { free_drv(&D); return rc; }
    int nb = s->nbas, nf = s->nfuc, nsh = D.nsh, nshuc = D.nshuc, nc = D.ncell;
    int n1 = sao ? nf : nb;
    schwarz(&D);
    double Qmax = 0.0;
    for (size_t x = 0; x < (size_t)nsh * nsh; ++x) if (D.Qsh[x] > Qmax) Qmax = D.Qsh[x];
    prim_schwarz(&D);
    memset(eri, 0, sizeof(double) * (size_t)n1 * nb * nb * nb);
    long long nsq = 0, npq = 0;
#define SC(S) ((S) / nshuc)
#define SL(S) ((S) % nshuc)
#pragma omp parallel for schedule(dynamic, 1) collapse(2) reduction(+ : nsq, npq)
    for (int SI = 0; SI < nshuc; ++SI)
        for (int SJ = 0; SJ < nsh; ++SJ) {
            if (SL(SJ) < SI) continue;
            if (D.Qsh[SI + nsh * SJ] * Qmax < schwarz_thresh) continue;
            for (int SK = 0; SK < nsh; ++SK) {
                if (SL(SK) < SI) continue;
                for (int SLs = 0; SLs < nsh; ++SLs) {
                    if (SL(SLs) < SI) continue;
                    if (D.Qsh[SI + nsh * SJ] * D.Qsh[SK + nsh * SLs] < schwarz_thresh) continue;
                    int q4[4] = {SI, SJ, SK, SLs}, smaller = 0;
                    for (int ip = 1; ip < 8 && !smaller; ++ip) {
                        int t4[4], r4[4];
                        for (int x = 0; x < 4; ++x) t4[x] = q4[IPERM8[ip][x]];
                        for (int x = 0; x < 4; ++x) r4[x] = D.relc[SC(t4[0]) * nc + SC(t4[x])] * nshuc + SL(t4[x]);
                        for (int x = 0; x < 4; ++x) {
                            if (r4[x] < q4[x]) { smaller = 1; break; }
                            if (r4[x] > q4[x]) break;
                        }
                    }
                    if (smaller) continue;
                    ++nsq;
                    int ab0 = SL(SK), ab1 = D.relc[SC(SK) * nc + SC(SLs)] * nshuc + SL(SLs);
                    kopts o;
                    memset(&o, 0, sizeof(o));
                    for (int x = 0; x < 4; ++x) o.p1[x] = -1;
                    o.ps_on = 1;
                    o.mp = D.mp;
                    o.ps_bra = D.psq + (size_t)D.mp * D.mp * (SI + (size_t)nshuc * SJ);
                    o.ps_ket = D.psq + (size_t)D.mp * D.mp * (ab0 + (size_t)nshuc * ab1);
                    int nI = D.sh_n[SI], nJ = D.sh_n[SJ], nK = D.sh_n[SK], nL = D.sh_n[SLs];
                    double blk[6 * 6 * 6 * 6 * 4];
                    double *bp = (nI * nJ * nK * nL <= 6 * 6 * 6 * 6 * 4) ? blk : malloc(sizeof(double) * nI * nJ * nK * nL);
                    eri_block_impl(D.sh_first[SI], nI, D.sh_first[SJ], nJ, D.sh_first[SK], nK, D.sh_first[SLs], nL, bp, &o);
                    npq += o.nprim;
                    for (int d = 0; d < nL; ++d)
                        for (int c = 0; c < nK; ++c)
                            for (int b = 0; b < nJ; ++b)
                                for (int a = 0; a < nI; ++a) {
                                    int ii = D.sh_first[SI] + a, jj = D.sh_first[SJ] + b;
                                    int kk = D.sh_first[SK] + c, ll = D.sh_first[SLs] + d;
                                    if (D.Q[ii + (size_t)nb * jj] * D.Q[kk + (size_t)nb * ll] < schwarz_thresh) continue;
                                    double val = bp[a + nI * (b + nJ * (c + nK * d))];
                                    int f4[4] = {ii, jj, kk, ll};
                                    for (int ip = 0; ip < 8; ++ip) {
                                        int t4[4], pp[4];
                                        for (int x = 0; x < 4; ++x) t4[x] = f4[IPERM8[ip][x]];
                                        for (int x = 0; x < 4; ++x) pp[x] = D.relc[(t4[0] / nf) * nc + t4[x] / nf] * nf + t4[x] % nf;
                                        size_t idx = physicist
                                            ? pp[0] + (size_t)n1 * (pp[2] + (size_t)nb * (pp[1] + (size_t)nb * pp[3]))
                                            : pp[0] + (size_t)n1 * (pp[1] + (size_t)nb * (pp[2] + (size_t)nb * pp[3]));
                                        eri[idx] = val;
                                    }
                                }
                    if (bp != blk) free(bp);
                }
            }
        }
#undef SC
#undef SL
    if (!sao) {
        /* rows i > nfuc by translation (src/Tools/symmetric_matrix.f90:62-131):
           ERI(c*nf+i, p, q, r) = ERI(i, p-c, q-c, r-c) with cell arithmetic on the lattice */
        for (int c = 1; c < nc; ++c) {
            int *shift = malloc(sizeof(int) * nb);
            for (int p = 0; p < nb; ++p) shift[p] = D.relc[c * nc + (p / nf)] * nf + p % nf;   /* p - c */
            for (size_t r = 0; r < (size_t)nb; ++r)
                for (size_t q = 0; q < (size_t)nb; ++q)
                    for (size_t p = 0; p < (size_t)nb; ++p)
                        for (int i = 0; i < nf; ++i)
                            eri[(size_t)(c * nf + i) + (size_t)nb * (p + (size_t)nb * (q + (size_t)nb * r))] =
                                eri[(size_t)i + (size_t)nb * (shift[p] + (size_t)nb * (shift[q] + (size_t)nb * shift[r]))];
            free(shift);
        }
    }
    if (stats) { stats[0] = nsq; stats[1] = npq; }
    S_.n_shell_quartets += nsq;
    S_.n_prim_quartets += npq;
    free_drv(&D);
    return 0;
}

/* J and K from the stored ERI (hartree_exchange_potential.f90:30-202);   */
/* rows mu <= n1 of J, K; P, J, K column-major (nbas,nbas).              */
void orc_jk(int n1, int nbas, int physicist, const double *P, const double *eri, double *J, double *K) {
    size_t nb = nbas;
    memset(J, 0, sizeof(double) * nb * nb);
    memset(K, 0, sizeof(double) * nb * nb);
#define E4(a, b, c, d) eri[(a) + (size_t)n1 * ((b) + nb * ((c) + nb * (d)))]
    double *kcol = malloc(sizeof(double) * n1);
    for (size_t d = 0; d < nb; ++d) {
        for (int a = 0; a < n1; ++a) kcol[a] = 0.0;
        for (size_t c = 0; c < nb; ++c)
            for (size_t b = 0; b < nb; ++b) {
                if (!physicist) {
                    double pcd = P[c + nb * d], pbc = P[b + nb * c];
                    for (int a = 0; a < n1; ++a) {
                        double e = E4(a, b, c, d);
                        J[a + nb * b] += e * pcd;
                        kcol[a] += e * pbc;
                    }
                } else {
                    double pbd = P[b + nb * d], pcb = P[c + nb * b];
                    for (int a = 0; a < n1; ++a) {
                        double e = E4(a, b, c, d);
                        J[a + nb * c] += e * pbd;
                        kcol[a] += e * pcb;
                    }
                }
            }
        for (int a = 0; a < n1; ++a) K[a + nb * d] = -0.5 * kcol[a];
    }
    free(kcol);
#undef E4
}
