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
    // Cyclic Jacobi: five sweeps of the three plane rotations, which drive a
    // symmetric 3x3 to diagonal well past f64 resolution.
    double czz = (double)(n * m[5] - m[2] * m[2]);
    double cxz = (double)(n * m[7] - m[0] * m[2]);
    double cyz = (double)(n * m[8] - m[1] * m[2]);
    double a[3][3] = {{cxx, cxy, cxz}, {cxy, cyy, cyz}, {cxz, cyz, czz}};
    double v[3][3] = {{1, 0, 0}, {0, 1, 0}, {0, 0, 1}};
    const int pp[3] = {0, 0, 1};
    const int rr[3] = {1, 2, 2};
    for (int sweep = 0; sweep < 5; ++sweep) {
        for (int q = 0; q < 3; ++q) {
            int p = pp[q], r = rr[q];
            double apq = a[p][r];
            if (apq == 0.0) continue;
            double tau = (a[r][r] - a[p][p]) / (2.0 * apq);
            double t = (tau >= 0.0 ? 1.0 : -1.0) / (fabs(tau) + sqrt(1.0 + tau * tau));
            double c = 1.0 / sqrt(1.0 + t * t);
            double s = t * c;
            for (int k = 0; k < 3; ++k) {
                double akp = a[k][p], akr = a[k][r];
                a[k][p] = c * akp - s * akr;
                a[k][r] = s * akp + c * akr;
            }
            for (int k = 0; k < 3; ++k) {
                double apk = a[p][k], ark = a[r][k];
                a[p][k] = c * apk - s * ark;
                a[r][k] = s * apk + c * ark;
            }
            for (int k = 0; k < 3; ++k) {
                double vkp = v[k][p], vkr = v[k][r];
                v[k][p] = c * vkp - s * vkr;
                v[k][r] = s * vkp + c * vkr;
            }
        }
    }
    int best = 0;
    if (a[1][1] > a[best][best]) best = 1;
    if (a[2][2] > a[best][best]) best = 2;
    if (axis) {
        axis[0] = v[0][best];
        axis[1] = v[1][best];
        axis[2] = v[2][best];
    }
    return a[best][best] > 0.0 ? a[best][best] : 0.0;
}

// Major axis 4 sqrt(lambda_max(cov)) of n sites with sums m.
GHD double length_of(const int64_t *m, int64_t n, bool three_d) {
    if (n < 2) return 0.0;
    double lam = cov_largest(m, n, three_d, nullptr);
    if (lam < 0.0) lam = 0.0;
    return 4.0 * sqrt(lam) / (double)n;
}

}  // namespace ghip
