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
// sxz, syz.
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

// Major axis 4 sqrt(lambda_max(cov)) of n sites with sums m.
GHD double length_of(const int64_t *m, int64_t n, bool three_d) {
    if (n < 2) return 0.0;
    double lam = cov_largest(m, n, three_d, nullptr);
    if (lam < 0.0) lam = 0.0;
    return 4.0 * sqrt(lam) / (double)n;
}

}  // namespace ghip
