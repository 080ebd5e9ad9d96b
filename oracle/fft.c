/*
 * FP64 mixed-radix Stockham FFT for the oracle (test infrastructure; see oracle.h).
 *
 * Stands in for cuFFT C2C FP32 (call sites range_compression.cu:80-100, range_doppler_algorithm.cu:318-390), at
 * higher precision. Algorithm: self-sorting Stockham passes, radix order 4,2,3,5,7 then any remaining prime by a
 * direct O(R^2) butterfly; twiddles are exact FP64 cos/sin of 2*pi*m/N. Pinned by brute-force DFT and numpy
 * golden vectors (tests/test_oracle_fft.py).
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "oracle.h"

struct or_fft_plan {
    int n;
    int nf;
    int radix[40];
    or_zc* tw; /* W_n^m = exp(-2 pi i m / n), m in [0, n) */
};

or_fft_plan* or_fft_plan_create(int n) {
    if (n < 1) return NULL;
    or_fft_plan* p = (or_fft_plan*)calloc(1, sizeof(or_fft_plan));
    p->n = n;
    int m = n;
    const int pref[] = {4, 2, 3, 5, 7};
    for (int k = 0; k < 5; k++) {
        while (m % pref[k] == 0) {
            p->radix[p->nf++] = pref[k];
            m /= pref[k];
        }
    }
    for (int f = 11; m > 1; f += 2) {
        while (m % f == 0) {
            p->radix[p->nf++] = f;
            m /= f;
        }
    }
    p->tw = (or_zc*)malloc(sizeof(or_zc) * (size_t)n);
    for (int i = 0; i < n; i++) {
        /* reduce the angle exactly in integers first: 2*pi*i/n with i < n */
        const double a = -2.0 * M_PI * (double)i / (double)n;
        p->tw[i] = cos(a) + I * sin(a);
    }
    return p;
}

void or_fft_plan_destroy(or_fft_plan* p) {
    if (!p) return;
    free(p->tw);
    free(p);
}

static inline or_zc tw_get(const or_fft_plan* p, long idx, int inverse) {
    const or_zc w = p->tw[idx % p->n];
    return inverse ? conj(w) : w;
}

void or_fft_exec(const or_fft_plan* p, or_zc* data, or_zc* scratch, int inverse) {
    const int n = p->n;
    if (n == 1) return;
    or_zc* own = NULL;
    if (!scratch) {
        own = (or_zc*)malloc(sizeof(or_zc) * (size_t)n);
        scratch = own;
    }
    or_zc* src = data;
    or_zc* dst = scratch;
    long ns = 1;
    or_zc v[64], o[64];
    for (int f = 0; f < p->nf; f++) {
        const int r_ = p->radix[f];
        const int m = n / r_;
        const long step_small = n / r_; /* W_R^k = W_n^{k n/R} */
        for (int j = 0; j < m; j++) {
            const long k = j % ns;
            const long tw_stride = (long)n / (ns * r_);
            for (int r = 0; r < r_; r++) {
                or_zc x = src[j + (long)r * m];
                if (r && k) x *= tw_get(p, k * r * tw_stride, inverse);
                v[r] = x;
            }
            if (r_ == 2) {
                o[0] = v[0] + v[1];
                o[1] = v[0] - v[1];
            } else if (r_ == 4) {
                const or_zc a0 = v[0] + v[2], a1 = v[0] - v[2];
                const or_zc b0 = v[1] + v[3], b1 = v[1] - v[3];
                /* forward: -i * b1 ; inverse: +i * b1 */
                const or_zc jb1 = inverse ? I * b1 : -I * b1;
                o[0] = a0 + b0;
                o[1] = a1 + jb1;
                o[2] = a0 - b0;
                o[3] = a1 - jb1;
            } else {
                for (int q = 0; q < r_; q++) {
                    or_zc acc = 0;
                    for (int r = 0; r < r_; r++) acc += v[r] * tw_get(p, (long)((q * r) % r_) * step_small, inverse);
                    o[q] = acc;
                }
            }
            const long base = (j / ns) * ns * r_ + k;
            for (int r = 0; r < r_; r++) dst[base + r * ns] = o[r];
        }
        or_zc* t = src;
        src = dst;
        dst = t;
        ns *= r_;
    }
    if (src != data) memcpy(data, src, sizeof(or_zc) * (size_t)n);
    free(own);
}

void or_dft(const or_zc* in, or_zc* out, int n, int inverse) {
    const double sgn = inverse ? 1.0 : -1.0;
    for (int k = 0; k < n; k++) {
        or_zc acc = 0;
        for (int j = 0; j < n; j++) {
            const long m = ((long)j * k) % n;
            const double a = sgn * 2.0 * M_PI * (double)m / (double)n;
            acc += in[j] * (cos(a) + I * sin(a));
        }
        out[k] = acc;
    }
}

void or_fft_rows(or_zc* data, int n, int rows, int stride, int inverse) {
    or_fft_plan* p = or_fft_plan_create(n);
#pragma omp parallel
    {
        or_zc* scratch = (or_zc*)malloc(sizeof(or_zc) * (size_t)n);
#pragma omp for schedule(static)
        for (int y = 0; y < rows; y++) or_fft_exec(p, data + (size_t)y * stride, scratch, inverse);
        free(scratch);
    }
    or_fft_plan_destroy(p);
}

void or_fft_cols(or_zc* data, int n, int cols, int stride, int inverse) {
    or_fft_plan* p = or_fft_plan_create(n);
    const int blk = 8; /* gather 8 columns at a time for cache locality */
#pragma omp parallel
    {
        or_zc* buf = (or_zc*)malloc(sizeof(or_zc) * (size_t)n * blk);
        or_zc* scratch = (or_zc*)malloc(sizeof(or_zc) * (size_t)n);
#pragma omp for schedule(static)
        for (int c0 = 0; c0 < cols; c0 += blk) {
            const int nb = (cols - c0) < blk ? (cols - c0) : blk;
            for (int y = 0; y < n; y++)
                for (int b = 0; b < nb; b++) buf[(size_t)b * n + y] = data[(size_t)y * stride + c0 + b];
            for (int b = 0; b < nb; b++) or_fft_exec(p, buf + (size_t)b * n, scratch, inverse);
            for (int y = 0; y < n; y++)
                for (int b = 0; b < nb; b++) data[(size_t)y * stride + c0 + b] = buf[(size_t)b * n + y];
        }
        free(buf);
        free(scratch);
    }
    or_fft_plan_destroy(p);
}
