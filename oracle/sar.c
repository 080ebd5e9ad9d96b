/*
 * FP64 restatement of the reference's GPU focusing stage (test infrastructure; see oracle.h).
 * Every function cites the reference file:line it follows. All arithmetic is double precision; index arithmetic
 * that the reference performs in float (multilook resampling positions) is kept in float so that window choices
 * agree exactly.
 */
#include <float.h>
#include <math.h>
#include <stdlib.h>
#include <string.h>

#ifdef _OPENMP
#include <omp.h>
#endif

#include "oracle.h"

void or_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

int or_get_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/* ------------------------------------------------------------------------------------------------------------
 * Padding rule: cuda_util/cufft_plan.cc:147-181 builds the set of 2^a 3^b 5^c 7^d <= 131072 and main.cc:217,223
 * takes std::map::upper_bound, i.e. the smallest member strictly greater than the request.
 * ------------------------------------------------------------------------------------------------------------ */
int or_next_fft_size(int n) {
    const int limit = 131072;
    int best = -1;
    for (long a = 1; a <= limit; a *= 2)
        for (long b = a; b <= limit; b *= 3)
            for (long c = b; c <= limit; c *= 5)
                for (long d = c; d <= limit; d *= 7)
                    if (d > n && (best < 0 || d < best)) best = (int)d;
    return best;
}

/* ------------------------------------------------------------------------------------------------------------
 * Chirp: sar/sar_chirp.cc:70-104 (GenerateChirpData), :27-35 (periodic Hamming, denominator N = n_samples),
 * :37-67 (unit energy: sum |FFT|^2 / N, which by Parseval is the time-domain energy).
 * Time origin t = (i - n/2)/fs with integer n/2 (sar_chirp.cc:80).
 * ------------------------------------------------------------------------------------------------------------ */
void or_generate_chirp(double kr, double fs, int n_samples, int apply_window, int padding, float* out_iq) {
    const double dt = 1.0 / fs;
    double* re = (double*)calloc((size_t)n_samples, sizeof(double));
    double* im = (double*)calloc((size_t)n_samples, sizeof(double));
    double energy = 0.0;
    for (int i = 0; i < n_samples; i++) {
        const double t = (double)(i - n_samples / 2) * dt;
        const double ph = M_PI * kr * t * t;
        double w = 1.0;
        if (apply_window) w = 0.54 - 0.46 * cos(2.0 * M_PI * (double)i / (double)n_samples);
        re[i] = cos(ph) * w;
        im[i] = sin(ph) * w;
        energy += re[i] * re[i] + im[i] * im[i];
    }
    const double scale = apply_window ? sqrt(energy) : sqrt((double)n_samples);
    for (int i = 0; i < padding; i++) {
        if (i < n_samples) {
            out_iq[2 * i] = (float)(re[i] / scale);
            out_iq[2 * i + 1] = (float)(im[i] / scale);
        } else {
            out_iq[2 * i] = 0.0f;
            out_iq[2 * i + 1] = 0.0f;
        }
    }
    free(re);
    free(im);
}

/* ------------------------------------------------------------------------------------------------------------
 * Polyfit: util/src/math_utils.cc:23-44 (Vandermonde [1, x, x^2..], Householder QR least squares, result reversed
 * to highest power first). Polyval: util/include/math_utils.h:55-63 (Horner).
 * ------------------------------------------------------------------------------------------------------------ */
int or_polyfit(const double* x, const double* y, int n, int degree, double* coeffs_out) {
    const int m = degree + 1;
    if (n < m || m < 1) return -1;
    double* a = (double*)malloc(sizeof(double) * (size_t)n * m); /* column-major n x m */
    double* b = (double*)malloc(sizeof(double) * (size_t)n);
    for (int r = 0; r < n; r++) {
        double v = 1.0;
        for (int c = 0; c < m; c++) {
            a[(size_t)c * n + r] = v;
            v *= x[r];
        }
        b[r] = y[r];
    }
    for (int k = 0; k < m; k++) {
        double* col = a + (size_t)k * n;
        double nrm = 0.0;
        for (int r = k; r < n; r++) nrm += col[r] * col[r];
        nrm = sqrt(nrm);
        if (nrm == 0.0) continue;
        const double alpha = col[k] > 0 ? -nrm : nrm;
        /* v = x - alpha e_k, stored in col[k..] */
        col[k] -= alpha;
        double vnorm2 = 0.0;
        for (int r = k; r < n; r++) vnorm2 += col[r] * col[r];
        /* apply H = I - 2 v v^T / (v^T v) to remaining columns and b */
        for (int c = k + 1; c < m; c++) {
            double* cc = a + (size_t)c * n;
            double dot = 0.0;
            for (int r = k; r < n; r++) dot += col[r] * cc[r];
            const double f = 2.0 * dot / vnorm2;
            for (int r = k; r < n; r++) cc[r] -= f * col[r];
        }
        {
            double dot = 0.0;
            for (int r = k; r < n; r++) dot += col[r] * b[r];
            const double f = 2.0 * dot / vnorm2;
            for (int r = k; r < n; r++) b[r] -= f * col[r];
        }
        col[k] = alpha; /* R diagonal */
    }
    /* back substitution R c = Q^T b (upper triangle lives above the diagonal of a) */
    double coef[16];
    for (int k = m - 1; k >= 0; k--) {
        double s = b[k];
        for (int c = k + 1; c < m; c++) s -= a[(size_t)c * n + k] * coef[c];
        coef[k] = s / a[(size_t)k * n + k];
    }
    for (int k = 0; k < m; k++) coeffs_out[k] = coef[m - 1 - k];
    free(a);
    free(b);
    return 0;
}

double or_polyval(const double* p, int n, double x) {
    double v = 0.0;
    for (int i = 0; i < n; i++) v = v * x + p[i];
    return v;
}

/* ------------------------------------------------------------------------------------------------------------
 * I/Q correction: sar/iq_correction.cu:188-273 with CalcMean/CalcStdDev (cuda_util/device_padded_image.cu:331-414)
 * and ZeroNaNs (:322-329).
 *  mean/std over the logical region, NaN pairs skipped, divided by n_total (Q3);
 *  DC subtraction, gain and phase estimation and application over x < XStride when OR_Q1 is set (Q1);
 *  cos_corr = 1 - d^2 when OR_Q2 is set (Q2); phase_error = asin(d) in degrees; NaN -> 0 over the padded image.
 * ------------------------------------------------------------------------------------------------------------ */
void or_iq_correction(or_zc* img, int xs, int ys, int xstr, int ystr, size_t ntot, int quirks, or_iq_results* res) {
    if (ntot == 0) ntot = (size_t)xs * (size_t)ys;
    double si = 0, sq = 0;
#pragma omp parallel for reduction(+ : si, sq) schedule(static)
    for (int y = 0; y < ys; y++) {
        const or_zc* row = img + (size_t)y * xstr;
        for (int x = 0; x < xs; x++) {
            const double i_ = creal(row[x]), q_ = cimag(row[x]);
            if (!isnan(i_) && !isnan(q_)) {
                si += i_;
                sq += q_;
            }
        }
    }
    const double dc_i = si / (double)ntot, dc_q = sq / (double)ntot;
    double vi = 0, vq = 0;
#pragma omp parallel for reduction(+ : vi, vq) schedule(static)
    for (int y = 0; y < ys; y++) {
        const or_zc* row = img + (size_t)y * xstr;
        for (int x = 0; x < xs; x++) {
            const double i_ = creal(row[x]), q_ = cimag(row[x]);
            if (!isnan(i_) && !isnan(q_)) {
                vi += (i_ - dc_i) * (i_ - dc_i);
                vq += (q_ - dc_q) * (q_ - dc_q);
            }
        }
    }
    res->dc_i = dc_i;
    res->dc_q = dc_q;
    res->stddev_i = sqrt(vi / (double)ntot);
    res->stddev_q = sqrt(vq / (double)ntot);

    const int w = (quirks & OR_Q1_IQ_PADDED_WIDTH) ? xstr : xs;
    double sii = 0, sqq = 0;
#pragma omp parallel for reduction(+ : sii, sqq) schedule(static)
    for (int y = 0; y < ys; y++) {
        or_zc* row = img + (size_t)y * xstr;
        for (int x = 0; x < w; x++) {
            const double i_ = creal(row[x]) - dc_i, q_ = cimag(row[x]) - dc_q;
            row[x] = i_ + I * q_;
            if (!isnan(i_) && !isnan(q_)) {
                sii += i_ * i_;
                sqq += q_ * q_;
            }
        }
    }
    const double g = sqrt(sii / sqq);
    res->iq_gain = g;
    double siq = 0, sii2 = 0;
#pragma omp parallel for reduction(+ : siq, sii2) schedule(static)
    for (int y = 0; y < ys; y++) {
        or_zc* row = img + (size_t)y * xstr;
        for (int x = 0; x < w; x++) {
            const double i_ = creal(row[x]) / g, q_ = cimag(row[x]);
            row[x] = i_ + I * q_;
            if (!isnan(i_) && !isnan(q_)) {
                siq += i_ * q_;
                sii2 += i_ * i_;
            }
        }
    }
    const double d = siq / sii2;
    const double sin_corr = d;
    const double cos_corr = (quirks & OR_Q2_COS_NO_SQRT) ? (1.0 - d * d) : sqrt(1.0 - d * d);
    res->phase_error = asin(d) * 180.0 / M_PI;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ystr; y++) {
        or_zc* row = img + (size_t)y * xstr;
        for (int x = 0; x < xstr; x++) {
            double i_ = creal(row[x]), q_ = cimag(row[x]);
            if (y < ys && x < w) q_ = (q_ - i_ * sin_corr) / cos_corr;
            if (isnan(i_)) i_ = 0.0;
            if (isnan(q_)) q_ = 0.0;
            row[x] = i_ + I * q_;
        }
    }
}

/* ------------------------------------------------------------------------------------------------------------
 * Doppler centroid (ACCC): sar/fractional_doppler_centroid.cu:98-162 with PhaseDifference :33-69.
 * step = Nr/32; blocks start at 0, step, ... while x_begin < Nr; x_end = x_begin + step is not clipped (Q4), so
 * for Nr % 32 != 0 a 33rd block reads padding columns. f_b = prf * atan2(C_b)/(2 pi) at x = (x_begin+x_end)/2
 * (integer division); degree-2 Polyfit. Columns at or beyond the row stride are never read.
 * ------------------------------------------------------------------------------------------------------------ */
int or_doppler_centroid(const or_zc* img, int xs, int ys, int xstr, double prf, int quirks, double* block_freq,
                        double* block_x, int max_blocks, double* poly3) {
    const int step = xs / 32;
    if (step <= 0) return -1;
    int nb = 0;
    double fx[64], fy[64];
    for (int xb = 0; xb < xs; xb += step) {
        int xe = xb + step;
        if (!(quirks & OR_Q4_DC_EXTRA_BLOCK)) {
            if (xb + step > xs) break;
        }
        const int xe_read = xe < xstr ? xe : xstr;
        double cr = 0, ci = 0;
#pragma omp parallel for reduction(+ : cr, ci) schedule(static)
        for (int y = 1; y < ys; y++) {
            const or_zc* prev = img + (size_t)(y - 1) * xstr;
            const or_zc* cur = img + (size_t)y * xstr;
            for (int x = xb; x < xe_read; x++) {
                const or_zc p = conj(prev[x]) * cur[x];
                cr += creal(p);
                ci += cimag(p);
            }
        }
        const double ang = atan2(ci, cr);
        const double f = prf * ang / (2.0 * M_PI);
        if (nb < 64) {
            fx[nb] = (double)((xb + xe) / 2);
            fy[nb] = f;
        }
        if (nb < max_blocks) {
            if (block_freq) block_freq[nb] = f;
            if (block_x) block_x[nb] = fx[nb];
        }
        nb++;
    }
    if (or_polyfit(fx, fy, nb, 2, poly3) != 0) return -1;
    return nb;
}

/* ------------------------------------------------------------------------------------------------------------
 * Range compression: sar/range_compression.cu:57-111. Row FFT over the padded width, multiply by conj(FFT(chirp)),
 * inverse FFT, scale 1/Nrp for x < Nr - Nc and zero x >= Nr - Nc (rows < Na only, Q9).
 * ------------------------------------------------------------------------------------------------------------ */
int or_range_compression(or_zc* img, int xs, int ys, int xstr, const float* chirp_iq, int nc) {
    const int n = xstr;
    const int nrv = xs - nc;
    or_zc* h = (or_zc*)malloc(sizeof(or_zc) * (size_t)n);
    for (int i = 0; i < n; i++) h[i] = (double)chirp_iq[2 * i] + I * (double)chirp_iq[2 * i + 1];
    or_fft_plan* p = or_fft_plan_create(n);
    or_fft_exec(p, h, NULL, 0);
    for (int i = 0; i < n; i++) h[i] = conj(h[i]);
#pragma omp parallel
    {
        or_zc* scratch = (or_zc*)malloc(sizeof(or_zc) * (size_t)n);
#pragma omp for schedule(static)
        for (int y = 0; y < ys; y++) {
            or_zc* row = img + (size_t)y * xstr;
            or_fft_exec(p, row, scratch, 0);
            for (int x = 0; x < n; x++) row[x] *= h[x];
            or_fft_exec(p, row, scratch, 1);
            for (int x = 0; x < n; x++) row[x] = (x < nrv) ? row[x] / (double)n : 0.0;
        }
        free(scratch);
    }
    or_fft_plan_destroy(p);
    free(h);
    return nrv;
}

/* ------------------------------------------------------------------------------------------------------------
 * RCMC window: numpy.blackman(16) as tabulated at sar/range_doppler_algorithm.cu:49-52, rounded to float.
 * ------------------------------------------------------------------------------------------------------------ */
void or_rcmc_window(float* w16) {
    for (int i = 0; i < 16; i++) {
        const double a = 2.0 * M_PI * (double)i / 15.0;
        w16[i] = (float)(0.42 - 0.5 * cos(a) + 0.08 * cos(2.0 * a));
    }
}

static double sinc_d(double x) {
    if (x == 0.0) return 1.0;
    return sin(M_PI * x) / (M_PI * x);
}

/* Bin frequency of azimuth FFT bin y (range_doppler_algorithm.cu:122-127, :212-218). */
static inline double bin_frequency(int y, int n_bins, double step) {
    return (y < n_bins / 2) ? (double)y * step : (double)(y - n_bins) * step;
}

/* Alias a baseband frequency around the Doppler centroid (range_doppler_algorithm.cu:131-140, :226-236). */
static inline double alias_about_dc(double f, double dc, double prf) {
    const double half = 0.5 * prf;
    if (dc < 0.0) {
        if (f > half + dc) f -= prf;
    } else if (dc > 0.0) {
        if (f < -half + dc) f += prf;
    }
    return f;
}

/* ------------------------------------------------------------------------------------------------------------
 * RCMC: sar/range_doppler_algorithm.cu:107-195 (16-tap windowed sinc, shift = lambda^2 f^2 R0 / (8 Vr^2 dr),
 * integer part std::round, zero when a tap leaves [0, Nrv), normalised by the sum of weights (Q7)).
 * ------------------------------------------------------------------------------------------------------------ */
void or_rcmc(const or_zc* in, or_zc* out, int xs, int nbins, int xstr, const or_rda_params* p) {
    float wf[16];
    or_rcmc_window(wf);
    const double step = p->prf / (double)nbins;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < nbins; y++) {
        const or_zc* src = in + (size_t)y * xstr;
        or_zc* dst = out + (size_t)y * xstr;
        for (int x = 0; x < xs; x++) {
            const double dc = or_polyval(p->dc_poly, 3, (double)x);
            const double fn = alias_about_dc(bin_frequency(y, nbins, step), dc, p->prf);
            const double r0 = p->slant_range_first + (double)x * p->range_spacing;
            const double vr = or_polyval(p->vr_poly, 3, (double)x);
            const double dr = p->wavelength * p->wavelength * fn * fn * r0 / (8.0 * vr * vr);
            const double shift = dr / p->range_spacing;
            const int pix = (int)round(shift);
            const int cx = x + pix;
            if (cx - 8 < 0 || cx + 8 >= xs) {
                dst[x] = 0.0;
                continue;
            }
            or_zc acc = 0.0;
            double norm = 0.0;
            for (int i = 0; i < 16; i++) {
                const double m = (double)wf[i] * sinc_d(shift - (double)pix + 8.0 - (double)i);
                acc += src[cx - 8 + i] * m;
                norm += m;
            }
            dst[x] = (norm == 0.0) ? 0.0 : acc / norm;
        }
    }
}

/* ------------------------------------------------------------------------------------------------------------
 * Azimuth matched filter: sar/range_doppler_algorithm.cu:197-309. Band cut to +-prf*beta/2 about the
 * bin-quantised DC offset, H = exp(-i pi fc^2/Ka) exp(-i 2 pi (dc_off/Ka) f_bin), Ka = 2 Vr^2/(lambda R0).
 * ------------------------------------------------------------------------------------------------------------ */
void or_az_matched_filter(or_zc* img, int xs, int nbins, int xstr, const or_rda_params* p) {
    const double step = p->prf / (double)nbins;
    const double half = 0.5 * p->prf;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < nbins; y++) {
        or_zc* row = img + (size_t)y * xstr;
        for (int x = 0; x < xs; x++) {
            const double dc = or_polyval(p->dc_poly, 3, (double)x);
            const double dc_off = round(dc / step) * step;
            const double fb = bin_frequency(y, nbins, step);
            const double fa = alias_about_dc(fb, dc, p->prf);
            /* the band test on bin offsets: fa and dc_off are whole multiples of step, so a bin exactly on
             * the edge is kept (the reference's strict inequality in exact arithmetic) however the edge rounds */
            const double jb = rint(fa / step) - rint(dc_off / step);
            if (fabs(jb) > 0.5 * p->az_bandwidth_fraction * (double)nbins + 1e-9) {
                row[x] = 0.0;
                continue;
            }
            double fc = fb - dc_off;
            if (fc > half)
                fc -= p->prf;
            else if (fc < -half)
                fc += p->prf;
            const double r0 = p->slant_range_first + p->range_spacing * (double)x;
            const double vr = or_polyval(p->vr_poly, 3, (double)x);
            const double ka = 2.0 * vr * vr / (p->wavelength * r0);
            const double ph = M_PI * fc * fc / ka + 2.0 * M_PI * (dc_off / ka) * fb;
            row[x] *= cos(ph) - I * sin(ph);
        }
    }
}

/* ------------------------------------------------------------------------------------------------------------
 * RangeDopplerAlgorithm: sar/range_doppler_algorithm.cu:312-407. Azimuth FFT of the Nrv valid columns over the
 * padded height (Q6), RCMC src -> zeroed out, matched filter in place, azimuth IFFT, x 1/Nap on the logical region.
 * ------------------------------------------------------------------------------------------------------------ */
void or_range_doppler(or_zc* src, or_zc* out, int xs, int ys, int xstr, int ystr, const or_rda_params* p) {
    or_fft_cols(src, ystr, xs, xstr, 0);
    memset(out, 0, sizeof(or_zc) * (size_t)xstr * ystr);
    or_rcmc(src, out, xs, ystr, xstr, p);
    or_az_matched_filter(out, xs, ystr, xstr, p);
    or_fft_cols(out, ystr, xs, xstr, 1);
    const double sc = 1.0 / (double)ystr;
#pragma omp parallel for schedule(static)
    for (int y = 0; y < ys; y++)
        for (int x = 0; x < xs; x++) out[(size_t)y * xstr + x] *= sc;
}

/* ------------------------------------------------------------------------------------------------------------
 * Azimuth multilook: sar/az_multilook.cu:25-101. ratio = float(12.5 / az_spacing); out rows = round(Na/ratio);
 * two 4-line means of |s|^2 at [trunc(y r) - 2, +4) and [trunc(y r) - 1, +4) (clipped), linearly interpolated by
 * frac(y r) (Q14). Resampling positions use float arithmetic exactly as the reference does.
 * ------------------------------------------------------------------------------------------------------------ */
int or_az_multilook_rows(int in_rows, double az_spacing) {
    const float ratio = (float)(12.5 / az_spacing);
    return (int)roundf((float)in_rows / ratio);
}

void or_az_multilook(const or_zc* in, int nr, int in_rows, int in_stride, double az_spacing, double* out) {
    const float ratio = (float)(12.5 / az_spacing);
    const int out_rows = or_az_multilook_rows(in_rows, az_spacing);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < out_rows; y++) {
        const float in_idx = (float)y * ratio;
        const float mu = in_idx - (float)(int)in_idx;
        int s0 = (int)(in_idx - 2.0f), e0 = (int)(in_idx + 2.0f);
        int s1 = (int)(in_idx - 1.0f), e1 = (int)(in_idx + 3.0f);
        if (s0 < 0) s0 = 0;
        if (e0 > in_rows) e0 = in_rows;
        if (s1 < 0) s1 = 0;
        if (e1 > in_rows) e1 = in_rows;
        for (int x = 0; x < nr; x++) {
            double a = 0, b = 0;
            int ca = 0, cb = 0;
            for (int i = s0; i < e0; i++) {
                const or_zc v = in[(size_t)i * in_stride + x];
                a += creal(v) * creal(v) + cimag(v) * cimag(v);
                ca++;
            }
            for (int i = s1; i < e1; i++) {
                const or_zc v = in[(size_t)i * in_stride + x];
                b += creal(v) * creal(v) + cimag(v) * cimag(v);
                cb++;
            }
            if (ca) a /= ca;
            if (cb) b /= cb;
            out[(size_t)y * nr + x] = (ca && cb) ? a + (b - a) * (double)mu : 0.0;
        }
    }
}

/* Range spreading loss: sar/range_spreading_loss.cu:42-69; the FP64 factor is tabulated as float (Q16). */
void or_rsl(double* img, int nr, int rows, double r_first, double dr, double r_ref) {
    float* tab = (float*)malloc(sizeof(float) * (size_t)nr);
    for (int i = 0; i < nr; i++) {
        const double r = r_first + (double)i * dr;
        tab[i] = (float)(1.0 / sqrt(pow(r_ref / r, 3.0)));
    }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; y++)
        for (int x = 0; x < nr; x++) img[(size_t)y * nr + x] *= (double)tab[x];
    free(tab);
}

/* SRGR: sar/srgr.cu:49-71 in FP64 (the reference evaluates in float, Q15); indices clamped to the row. */
void or_srgr(const double* in, int in_nr, int rows, const double* grsr5, double dr, double gs, int out_nr,
             double* out) {
#pragma omp parallel for schedule(static)
    for (int y = 0; y < rows; y++) {
        for (int x = 0; x < out_nr; x++) {
            const double rd = or_polyval(grsr5, 5, (double)x * gs);
            const double ri = rd / dr;
            int r0 = (int)ri;
            double fx = ri - (double)r0;
            if (r0 < 0) {
                r0 = 0;
                fx = 0.0;
            }
            if (r0 > in_nr - 1) {
                r0 = in_nr - 1;
                fx = 0.0;
            }
            const int r1 = (r0 + 1 < in_nr) ? r0 + 1 : in_nr - 1;
            const double v0 = in[(size_t)y * in_nr + r0], v1 = in[(size_t)y * in_nr + r1];
            out[(size_t)y * out_nr + x] = v0 + fx * (v1 - v0);
        }
    }
}

/* CalcAperturePixels: sar/sar_metadata.h:109-124 (Ka from carrier frequency = c / lambda). */
int or_aperture_pixels(double lambda, double prf, double beta, const double* vr_poly, const double* dc_poly,
                       int dc_terms, double r_first, double dr, int x) {
    const double vr = or_polyval(vr_poly, 3, (double)x);
    const double r0 = r_first + (double)x * dr;
    const double ka = 2.0 * vr * vr / (lambda * r0);
    const double fmax = or_polyval(dc_poly, dc_terms, (double)x) + prf * beta / 2.0;
    return (int)round(fmax / ka * prf);
}

/* ------------------------------------------------------------------------------------------------------------
 * Whole GPU stage: main.cc:233-301. Padded image (zeros), I/Q correction + NaN zero, chirp, DC estimate, range
 * compression, RDA; returns the logical SLC (Na x Nrv).
 * ------------------------------------------------------------------------------------------------------------ */
int or_focus_slc(const float* raw_iq, const or_focus_params* fp, or_zc* slc_out, or_focus_results* res) {
    const int na = fp->na, nr = fp->nr, nc = fp->nc;
    const int nrp = fp->nrp > 0 ? fp->nrp : or_next_fft_size(nr + nc);
    const int nap = fp->nap > 0 ? fp->nap : or_next_fft_size(na);
    if (nrp < nr + nc || nap < na) return -1;
    const size_t npad = (size_t)nrp * nap;
    or_zc* img = (or_zc*)calloc(npad, sizeof(or_zc));
    or_zc* out = (or_zc*)malloc(npad * sizeof(or_zc));
    if (!img || !out) {
        free(img);
        free(out);
        return -2;
    }
#pragma omp parallel for schedule(static)
    for (int y = 0; y < na; y++)
        for (int x = 0; x < nr; x++) {
            const size_t s = ((size_t)y * nr + x) * 2;
            img[(size_t)y * nrp + x] = (double)raw_iq[s] + I * (double)raw_iq[s + 1];
        }
    const size_t ntot = fp->n_total_samples ? fp->n_total_samples : (size_t)na * nr;
    or_iq_correction(img, nr, na, nrp, nap, ntot, fp->quirks, &res->iq);
    float* chirp = (float*)malloc(sizeof(float) * 2 * (size_t)nrp);
    or_generate_chirp(fp->chirp_kr, fp->range_sampling_rate, nc, fp->apply_range_window, nrp, chirp);
    double bf[64], bx[64];
    if (or_doppler_centroid(img, nr, na, nrp, fp->prf, fp->quirks, bf, bx, 64, res->dc_poly) < 0) {
        free(img);
        free(out);
        free(chirp);
        return -3;
    }
    const int nrv = or_range_compression(img, nr, na, nrp, chirp, nc);
    res->nrv = nrv;
    or_rda_params rp;
    rp.wavelength = fp->wavelength;
    rp.range_spacing = fp->range_spacing;
    rp.slant_range_first = fp->slant_range_first;
    rp.prf = fp->prf;
    rp.az_bandwidth_fraction = fp->az_bandwidth_fraction;
    memcpy(rp.dc_poly, res->dc_poly, sizeof(rp.dc_poly));
    memcpy(rp.vr_poly, fp->vr_poly, sizeof(rp.vr_poly));
    or_range_doppler(img, out, nrv, na, nrp, nap, &rp);
#pragma omp parallel for schedule(static)
    for (int y = 0; y < na; y++) memcpy(slc_out + (size_t)y * nrv, out + (size_t)y * nrp, sizeof(or_zc) * nrv);
    free(img);
    free(out);
    free(chirp);
    return 0;
}
