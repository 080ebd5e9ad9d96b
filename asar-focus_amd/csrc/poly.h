// FP64 polynomial least squares and evaluation, usable on host and device.
// Follows util/src/math_utils.cc:23-44 (Vandermonde with columns x^0..x^d, Householder QR, coefficients returned
// highest power first) and util/include/math_utils.h:55-63 (Horner).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

namespace asarf {

__host__ __device__ inline double polyval(const double* p, int n, double x) {
    double v = 0.0;
    for (int i = 0; i < n; ++i) v = v * x + p[i];
    return v;
}

// n <= 64 points, degree <= 5. Returns 0 on success.
__host__ __device__ inline int polyfit_qr(const double* x, const double* y, int n, int degree, double* out) {
    const int m = degree + 1;
    if (m < 1 || m > 6 || n < m || n > 128) return -1;
    double a[6][128];
    double b[128];
    for (int r = 0; r < n; ++r) {
        double v = 1.0;
        for (int c = 0; c < m; ++c) {
            a[c][r] = v;
            v *= x[r];
        }
        b[r] = y[r];
    }
    for (int k = 0; k < m; ++k) {
        double nrm = 0.0;
        for (int r = k; r < n; ++r) nrm += a[k][r] * a[k][r];
        nrm = sqrt(nrm);
        if (nrm == 0.0) return -2;
        const double alpha = a[k][k] > 0.0 ? -nrm : nrm;
        a[k][k] -= alpha;
        double vv = 0.0;
        for (int r = k; r < n; ++r) vv += a[k][r] * a[k][r];
        for (int c = k + 1; c < m; ++c) {
            double d = 0.0;
            for (int r = k; r < n; ++r) d += a[k][r] * a[c][r];
            const double f = 2.0 * d / vv;
            for (int r = k; r < n; ++r) a[c][r] -= f * a[k][r];
        }
        double d = 0.0;
        for (int r = k; r < n; ++r) d += a[k][r] * b[r];
        const double f = 2.0 * d / vv;
        for (int r = k; r < n; ++r) b[r] -= f * a[k][r];
        a[k][k] = alpha;
    }
    double c[6];
    for (int k = m - 1; k >= 0; --k) {
        double s = b[k];
        for (int j = k + 1; j < m; ++j) s -= a[j][k] * c[j];
        c[k] = s / a[k][k];
    }
    for (int k = 0; k < m; ++k) out[k] = c[m - 1 - k];
    return 0;
}

}  // namespace asarf
