// Scalar special functions and primitive-pair geometry of the engine (FP64;
// the pair geometry, truncation order, Hermite and Fourier coefficients are
// shared by the host and the GPU pair set-up).  Written from the mathematics; the reference routines they
// must agree with are cited per function.
#pragma once
#include <hip/hip_runtime.h>

#include <cmath>
#include <complex>
#include <vector>

#include "tcg_engine.hpp"

namespace tcg {

constexpr double kPi = 3.14159265358979323846;

// exp(-x) I_0(x): power series below 20 (A&S 9.6.10), Hankel asymptotic
// series above (A&S 9.7.1), summed to the smallest term.  Agrees with the
// reference's iv_scaled(0,x) (src/module/bessel_iv_log.f90:320) to ~1e-13.
__host__ __device__ inline double i0_scaled(double x) {
    if (x <= 0.0) return 1.0;
    if (x <= 20.0) {
        double q = 0.25 * x * x, term = 1.0, sum = 1.0;
        for (int k = 1; k < 200; ++k) {
            term *= q / ((double)k * (double)k);
            sum += term;
            if (term < 1e-17 * sum) break;
        }
        return sum * exp(-x);
    }
    double term = 1.0, sum = 1.0, prev = 1e300;
    for (int k = 1; k < 200; ++k) {
        double t = term * (2.0 * k - 1) * (2.0 * k - 1) / (8.0 * k * x);
        if (t >= prev || t < 1e-18 * sum) break;
        prev = t;
        term = t;
        sum += term;
    }
    return sum / sqrt(2.0 * kPi * x);
}

// exp(-A) I_n(A), n = 0..N, by Miller's backward recurrence
//   I_{k-1} = I_{k+1} + (2k/A) I_k
// started at n0 with I_{n0+1} = 0 and rescaled against overflow, then
// normalised to exp(-A) I_0(A).  start = N reproduces the reference's
// bessel_scaled_vector_fast (src/module/clifford_tools.f90:429-461), start =
// N + 30 + sqrt(80A) its bessel_scaled_vector (:386-421).
inline void bessel_miller(int N, double A, int n0, double *out, std::vector<double> &w) {
    if (A < 1e-12 || (n0 == N && A < 1e-10)) {
        for (int n = 0; n <= N; ++n) out[n] = 0.0;
        out[0] = 1.0;
        return;
    }
    w.assign(n0 + 2, 0.0);
    w[n0] = 1e-280;
    double r = 2.0 / A;
    for (int k = n0; k >= 1; --k) {
        w[k - 1] = w[k + 1] + (double)k * r * w[k];
        if (w[k - 1] > 1e250)
            for (int m = k - 1; m <= n0; ++m) w[m] *= 1e-250;
    }
    double s = i0_scaled(A) / w[0];
    for (int n = 0; n <= N; ++n) out[n] = w[n] * s;
}

inline void bessel_miller(int N, double A, int n0, double *out) {
    std::vector<double> w;
    bessel_miller(N, A, n0, out, w);
}

// exp(-A) I_n(A), n = 0..N, as the reference's bessel_scaled_vector
// (src/module/clifford_tools.f90:386-421): Miller's recurrence from
// N + 30 + sqrt(80 A).  For A >= 100 (N+1)^2 that start lies millions of
// orders up, so there the Hankel expansion
//   exp(-A) I_n(A) = (2 pi A)^-1/2 sum_k (-1)^k prod_{j<=k} (4n^2 - (2j-1)^2) / (k! (8A)^k)
// is summed instead (ratio of successive terms <= 4n^2/(8A) <= 1/200, so a
// few terms reach round-off; agrees with the recurrence to its own ~1e-14).
inline void bessel_scaled_vector(int N, double A, double *out) {
    if (A >= 100.0 * (N + 1.0) * (N + 1.0) && A > 1000.0) {
        const double pre = 1.0 / std::sqrt(2.0 * kPi * A);
        for (int n = 0; n <= N; ++n) {
            const double mu = 4.0 * n * n;
            double term = 1.0, sum = 1.0;
            for (int k = 1; k < 60; ++k) {
                term *= -(mu - (2.0 * k - 1.0) * (2.0 * k - 1.0)) / (8.0 * k * A);
                sum += term;
                if (std::fabs(term) < 1e-17 * std::fabs(sum)) break;
            }
            out[n] = sum * pre;
        }
        return;
    }
    bessel_miller(N, A, N + 30 + (int)std::sqrt(80.0 * A), out);
}

// Truncation order of exp(-A) I_n(A) (src/module/clifford_tools.f90:566-572)
__host__ __device__ inline int nmax_est(double A) { return 20 + (int)ceil(9.0 * sqrt(A > 0.0 ? A : 0.0)); }

// n-point Gauss-Legendre rule on u in [0,1] mapped to t in [0,inf) by
// t = s u/(1-u), weights w_u s/(1-u)^2, s = 0.6: the reference's fast_quad
// rule (src/module/precompute_t.f90:906-942).  Nodes by Newton on P_n.
inline void mapped_gauss_legendre(int n, double *t, double *w) {
    const double s = 0.6;
    for (int i = 0; i < n; ++i) {
        double x = std::cos(kPi * (i + 0.75) / (n + 0.5)), dp = 0.0, p0, p1;
        for (int it = 0; it < 100; ++it) {
            p0 = 1.0;
            p1 = x;
            for (int j = 2; j <= n; ++j) {
                double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
                p0 = p1;
                p1 = p2;
            }
            dp = n * (x * p1 - p0) / (x * x - 1.0);
            double dx = p1 / dp;
            x -= dx;
            if (std::fabs(dx) < 1e-15) break;
        }
        p0 = 1.0;
        p1 = x;
        for (int j = 2; j <= n; ++j) {
            double p2 = ((2.0 * j - 1.0) * x * p1 - (j - 1.0) * p0) / j;
            p0 = p1;
            p1 = p2;
        }
        dp = n * (x * p1 - p0) / (x * x - 1.0);
        double u = 0.5 * (1.0 - x);
        t[i] = s * u / (1.0 - u);
        w[i] = s / ((1.0 - x * x) * dp * dp) / ((1.0 - u) * (1.0 - u));
    }
}

// Product of two Clifford primitives along a periodic direction of length
// L (a = 2 pi / L): barycentric exponent p and centre P with
// e1 e^{i a x1} + e2 e^{i a x2} = p e^{i a P}, including the reference's
// branch rules (src/Toroidal/tools.f90:1-262):
//   |x1-x2| within 1e-6 of L/2 (antipodal): p = |e1-e2| (0 below 1e-8),
//     centre of the tighter primitive (midpoint when e1 ~ e2 to 1e-10);
//   e1 ~ e2 (1e-10): midpoint, shifted by -L/2 when |x1-x2| > L/2;
//   p clamped to e1+e2, and snapped to e1+e2 within 1e-14.
__host__ __device__ inline void clifford_pair(double e1, double e2, double x1, double x2, double L, double a,
                                              double *p, double *P) {
    const double eps = 1e-6;
    double dx = x1 - x2, adx = fabs(dx);
    double r;
    if (adx > 0.5 * L - eps && adx < 0.5 * L + eps) r = (fabs(e1 - e2) < 1e-8) ? 0.0 : fabs(e1 - e2);
    else r = sqrt(fabs(e1 * e1 + e2 * e2 + 2.0 * e1 * e2 * cos(a * dx)));
    if (r > e1 + e2) r = e1 + e2;
    if (e1 + e2 - r < 1e-14) r = e1 + e2;
    *p = r;
    if (adx < 0.5 * L + eps && adx > 0.5 * L - eps) {
        if (fabs(e1 - e2) < 1e-10) *P = 0.5 * (x1 + x2);
        else *P = (e1 > e2) ? x1 : x2;
    } else if (fabs(e1 - e2) < 1e-10) {
        *P = 0.5 * (x1 + x2) - ((adx > 0.5 * L + eps) ? 0.5 * L : 0.0);
    } else {
        double c = e1 * cos(a * x1) + e2 * cos(a * x2);
        double sn = e1 * sin(a * x1) + e2 * sin(a * x2);
        *P = atan(sn / c) / a + ((c <= 0.0) ? 0.5 * L : 0.0);
    }
}

// Hermite expansion of (y-A)^i (y-B)^j exp(-p(y-P)^2) = sum_t E_t d^t/dP^t exp(-p(y-P)^2),
// t = 0..i+j (McMurchie-Davidson; src/module/clifford_tools.f90:757-785).
__host__ __device__ inline void hermite_E(int i, int j, double p, double PA, double PB, double *E /* [i+j+1] */) {
    double cur[2 * 8 + 3] = {0}, nxt[2 * 8 + 3];
    double *c = cur + 1;   // c[-1] = 0
    c[0] = 1.0;
    int deg = 0;
    double h = 0.5 / p;
    for (int m = 1; m <= i + j; ++m) {
        double X = (m <= i) ? PA : PB;
        for (int t = 0; t <= deg + 1; ++t) nxt[t] = h * c[t - 1] + X * c[t] + (t + 1) * c[t + 1];
        ++deg;
        for (int t = 0; t <= deg; ++t) c[t] = nxt[t];
        c[deg + 1] = 0.0;
    }
    for (int t = 0; t <= i + j; ++t) E[t] = c[t];
}

// Fourier coefficients d_k, k = -(i+l)..(i+l), of sin^i(u+phA) sin^l(u+phB)
// (src/module/clifford_tools.f90:578-607): repeated multiplication by
// sin(u+ph) = e^{i ph} e^{iu}/(2i) - e^{-i ph} e^{-iu}/(2i).
inline void sin_product_fourier(int i, int l, double phA, double phB, std::complex<double> *d) {
    const int K = i + l;                 // <= 4 kMaxL (callers' d holds 4 kMaxL + 1 values)
    constexpr int kCap = 4 * kMaxL + 3;
    std::complex<double> buf[2][kCap];
    std::complex<double> *cur = buf[0], *nxt = buf[1];
    for (int k = 0; k < 2 * K + 3; ++k) cur[k] = 0.0;
    cur[K + 1] = 1.0;   // index k + K + 1
    const double sA = std::sin(phA), cA = std::cos(phA), sB = std::sin(phB), cB = std::cos(phB);
    for (int m = 1; m <= K; ++m) {
        const double sn = (m <= i) ? sA : sB, cs = (m <= i) ? cA : cB;
        const std::complex<double> cp = std::complex<double>(sn, -cs) * 0.5;   // e^{i ph}/(2i)
        const std::complex<double> cm = std::complex<double>(sn, cs) * 0.5;    // -e^{-i ph}/(2i)
        for (int k = 0; k < 2 * K + 3; ++k) nxt[k] = 0.0;
        for (int k = -(m - 1); k <= m - 1; ++k) {
            nxt[k + 1 + K + 1] += cp * cur[k + K + 1];
            nxt[k - 1 + K + 1] += cm * cur[k + K + 1];
        }
        std::swap(cur, nxt);
    }
    for (int k = -K; k <= K; ++k) d[k + K] = cur[k + K + 1];
}

// The same coefficients as (re, im) pairs, for the host and the GPU pair set-up:
// d[k + K] = d_k, k = -K..K, K = i + l <= 2 kMaxL.
__host__ __device__ inline void sin_fourier_coefs(int i, int l, double phA, double phB, double2 *d) {
    const int K = i + l;
    constexpr int kCap = 4 * kMaxL + 3;
    double2 buf0[kCap], buf1[kCap];
    double2 *cur = buf0, *nxt = buf1;
    for (int k = 0; k < 2 * K + 3; ++k) cur[k] = double2{0.0, 0.0};
    cur[K + 1] = double2{1.0, 0.0};   // index k + K + 1
    const double sA = sin(phA), cA = cos(phA), sB = sin(phB), cB = cos(phB);
    for (int m = 1; m <= K; ++m) {
        const double sn = (m <= i) ? sA : sB, cs = (m <= i) ? cA : cB;
        const double pr = 0.5 * sn, pi = -0.5 * cs;        // e^{i ph}/(2i)
        const double mr = 0.5 * sn, mi = 0.5 * cs;         // -e^{-i ph}/(2i)
        for (int k = 0; k < 2 * K + 3; ++k) nxt[k] = double2{0.0, 0.0};
        for (int k = -(m - 1); k <= m - 1; ++k) {
            const double2 v = cur[k + K + 1];
            double2 &up = nxt[k + 1 + K + 1], &dn = nxt[k - 1 + K + 1];
            up.x += pr * v.x - pi * v.y;
            up.y += pr * v.y + pi * v.x;
            dn.x += mr * v.x - mi * v.y;
            dn.y += mr * v.y + mi * v.x;
        }
        double2 *tmp = cur;
        cur = nxt;
        nxt = tmp;
    }
    for (int k = -K; k <= K; ++k) d[k + K] = cur[k + K + 1];
}

}  // namespace tcg
