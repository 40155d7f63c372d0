// geometry.h - per-label shape from exact integer moment sums: the largest
// eigenvalue and axis of the covariance, and the major-axis length the
// length term prices (glazier-core/src/moments.rs:152-157, where the
// reference power-iterates an f64 covariance instead).
//
// The sums are integers about each label's anchor, so the scaled covariance
// C = n S2 - S1 S1^T is exact in int64 and the f64 arithmetic below is the
// only rounding. Shared by the sweep (length term) and division (cut axis).
#pragma once

#include "ghip_common.h"

namespace ghip {

// One Jacobi rotation in the (P, R) plane: zero a[P][R] of the symmetric
// 3x3 `a`, applying the same rotation to the eigenvector columns `v`. The
// plane is a template argument so every index is a constant and the
// matrices stay in registers.
template <int P, int R>
GHD void jacobi_rotate(double (&a)[3][3], double (&v)[3][3]) {
    const double apq = a[P][R];
    if (apq == 0.0) return;
    const double tau = (a[R][R] - a[P][P]) / (2.0 * apq);
    const double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
    const double c = 1.0 / sqrt(1.0 + t * t);
    const double s = t * c;
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double akp = a[k][P], akr = a[k][R];
        a[k][P] = c * akp - s * akr;
        a[k][R] = s * akp + c * akr;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double apk = a[P][k], ark = a[R][k];
        a[P][k] = c * apk - s * ark;
        a[R][k] = s * apk + c * ark;
    }
#pragma unroll
    for (int k = 0; k < 3; ++k) {
        const double vkp = v[k][P], vkr = v[k][R];
        v[k][P] = c * vkp - s * vkr;
        v[k][R] = s * vkp + c * vkr;
    }
}

// Moment word order after volume/surface: sx, sy, sz, sxx, syy, szz, sxy,
// sxz, syz. The largest eigenvalue and its axis, for the division plan (a
// volume's length takes the closed form length_of uses).
GHD double cov_largest(const int64_t *m, int64_t n, bool three_d, double *axis) {
    double cxx = (double)(n * m[3] - m[0] * m[0]);
    double cyy = (double)(n * m[4] - m[1] * m[1]);
    double cxy = (double)(n * m[6] - m[0] * m[1]);
    if (!three_d) {
        double half = 0.5 * (cxx - cyy);
        double r = sqrt(half * half + cxy * cxy);
        double lam = 0.5 * (cxx + cyy) + r;
        if (axis) {
            double ax, ay;
            if (r == 0.0) {
                ax = 1.0;
                ay = 0.0;
            } else if (half >= 0.0) {
                ax = half + r;
                ay = cxy;
            } else {
                ax = cxy;
                ay = r - half;
            }
            double nrm = sqrt(ax * ax + ay * ay);
            axis[0] = ax / nrm;
            axis[1] = ay / nrm;
            axis[2] = 0.0;
        }
        return lam;
    }
    // Cyclic Jacobi: three sweeps of the three plane rotations bring the
    // largest diagonal entry within about 1e-10 (relative) of the largest
    // eigenvalue of an integer covariance; lengths need far less.
    double czz = (double)(n * m[5] - m[2] * m[2]);
    double cxz = (double)(n * m[7] - m[0] * m[2]);
    double cyz = (double)(n * m[8] - m[1] * m[2]);
    double a[3][3] = {{cxx, cxy, cxz}, {cxy, cyy, cyz}, {cxz, cyz, czz}};
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
#pragma unroll
    for (int sweep = 0; sweep < 3; ++sweep) {
        jacobi_rotate<0, 1>(a, v);
        jacobi_rotate<0, 2>(a, v);
        jacobi_rotate<1, 2>(a, v);
    }
    // the largest diagonal entry, the first of equals
    double best = a[0][0];
    int col = 0;
    if (a[1][1] > best) {
        best = a[1][1];
        col = 1;
    }
    if (a[2][2] > best) {
        best = a[2][2];
        col = 2;
    }
    if (axis) {
        axis[0] = col == 0 ? v[0][0] : col == 1 ? v[0][1] : v[0][2];
        axis[1] = col == 0 ? v[1][0] : col == 1 ? v[1][1] : v[1][2];
        axis[2] = col == 0 ? v[2][0] : col == 1 ? v[2][1] : v[2][2];
    }
    return best > 0.0 ? best : 0.0;
}

// In-plane anisotropy (Moments::anisotropy, glazier-core/src/moments.rs:131-141)
// of the scaled covariance, which it does not depend on: (C_xx - C_yy)/tr and
// 2 C_xy/tr with tr = C_xx + C_yy; zero below two sites or at tr = 0 (an
// integer trace is 0 or at least 1, i.e. at least 1/n^2 unscaled).
GHD void anisotropy(const int64_t *m, int64_t n, double *a0, double *a1) {
    *a0 = 0.0;
    *a1 = 0.0;
    if (n < 2) return;
    const int64_t cxx = n * m[3] - m[0] * m[0];
    const int64_t cyy = n * m[4] - m[1] * m[1];
    const int64_t cxy = n * m[6] - m[0] * m[1];
    const int64_t tr = cxx + cyy;
    if (tr <= 0) return;
    *a0 = (double)(cxx - cyy) / (double)tr;
    *a1 = 2.0 * (double)cxy / (double)tr;
}

// h(u) = cos(2/3 acos u) for u in [0, 1], analytic there, as a degree-14
// polynomial in t = 2u - 1 (a Chebyshev fit at 200 001 nodes written in
// powers of t; largest error 3.9e-14). cos(acos(r)/3) = h(sqrt((1 + r)/2))
// for r in [-1, 1].
GHD double third_angle_cos(double u) {
    const double t = 2.0 * u - 1.0;
    double h = -1.7072390069472367e-09;
    h = h * t + 5.577076476677568e-09;
    h = h * t + -1.24175244399134e-08;
    h = h * t + 4.321603925338742e-08;
    h = h * t + -1.6036631021280584e-07;
    h = h * t + 5.662416587052154e-07;
    h = h * t + -2.035473767371784e-06;
    h = h * t + 7.540205193337502e-06;
    h = h * t + -2.892003887918587e-05;
    h = h * t + 0.00011642563899772694;
    h = h * t + -0.0005041246781468507;
    h = h * t + 0.0024663527976783333;
    h = h * t + -0.015509188437110516;
    h = h * t + 0.24740906632332305;
    h = h * t + 0.7660444431189831;
    return h;
}

// Largest eigenvalue of the symmetric 3 x 3 matrix with diagonal (a00, a11,
// a22) and off-diagonal (a01, a02, a12), by the trigonometric closed form
// (O. K. Smith, Comm. ACM 4 (1961) 168): with q = tr/3, p^2 = |A - qI|^2/6
// and r = det((A - qI)/p)/2 in [-1, 1], it is q + 2p cos(acos(r)/3). Plain
// multiplies and adds, one division and two square roots, written the same
// in the emulation.
GHD double largest_eigen3(double a00, double a11, double a22, double a01, double a02, double a12) {
    const double q = (a00 + a11 + a22) / 3.0;
    const double d0 = a00 - q, d1 = a11 - q, d2 = a22 - q;
    const double p2 = d0 * d0 + d1 * d1 + d2 * d2 + 2.0 * (a01 * a01 + a02 * a02 + a12 * a12);
    if (!(p2 > 0.0)) return q;
    const double p = sqrt(p2 / 6.0);
    const double ip = 1.0 / p;
    const double b00 = d0 * ip, b11 = d1 * ip, b22 = d2 * ip;
    const double b01 = a01 * ip, b02 = a02 * ip, b12 = a12 * ip;
    const double det = b00 * (b11 * b22 - b12 * b12) - b01 * (b01 * b22 - b12 * b02) + b02 * (b01 * b12 - b11 * b02);
    double r = 0.5 * det;
    r = r < -1.0 ? -1.0 : (r > 1.0 ? 1.0 : r);
    return q + 2.0 * p * third_angle_cos(sqrt(0.5 * (1.0 + r)));
}

// Major axis 4 sqrt(lambda_max(cov)) of n sites with sums m: the planar
// closed form in a plane, the closed form above in a volume (the division
// plan's axis keeps the Jacobi rotations, which give the eigenvector).
GHD double length_of(const int64_t *m, int64_t n, bool three_d) {
    if (n < 2) return 0.0;
    double lam;
    if (three_d) {
        lam = largest_eigen3((double)(n * m[3] - m[0] * m[0]), (double)(n * m[4] - m[1] * m[1]),
                             (double)(n * m[5] - m[2] * m[2]), (double)(n * m[6] - m[0] * m[1]),
                             (double)(n * m[7] - m[0] * m[2]), (double)(n * m[8] - m[1] * m[2]));
    } else {
        lam = cov_largest(m, n, false, nullptr);
    }
    if (lam < 0.0) lam = 0.0;
    return 4.0 * sqrt(lam) / (double)n;
}

}  // namespace ghip
