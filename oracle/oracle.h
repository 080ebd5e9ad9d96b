/*
 * FP64 CPU oracle for the asarf range-Doppler focusing path.
 *
 * TEST INFRASTRUCTURE ONLY. This library restates, in double precision, the algorithm of the reference's GPU
 * focusing stage (cgi-estonia-space/asar-focus 0.3.2, files cited per function). It is the checker the HIP path is
 * compared against and the timed CPU baseline. Only tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg
 * may load it; the product library (libasarf.so) never links or calls it.
 *
 * Conventions
 *  - Images are row-major, range (x) fastest, complex FP64 (double _Complex), with an explicit row stride.
 *    "Logical" region = x < x_size, y < y_size; "padded" = x < x_stride, y < y_stride (the reference's
 *    DevicePaddedImage model, cuda_util/device_padded_image.cuh:28-158).
 *  - FFTs are unnormalised (forward e^{-2 pi i nk/N}, inverse e^{+2 pi i nk/N}), as cuFFT C2C is.
 *  - Reference quirks (pinned numerically, not bugs of this oracle) are switched by OR_Q* flags; OR_QUIRKS_REFERENCE
 *    selects the reference's behaviour.
 *
 * Parity pins (how each function is checked, tests/test_oracle_*.py):
 *  - or_fft_*: brute-force DFT, numpy FFT golden vectors (tests/golden), Parseval, round trip.
 *  - or_next_fft_size: brute-force enumeration of 7-smooth numbers.
 *  - or_generate_chirp: closed-form chirp phase, Hamming formula, unit energy.
 *  - or_polyfit: numpy.polyfit golden vectors, exact recovery of polynomials.
 *  - or_iq_correction: recovery of a synthetic DC/gain/phase imbalance; closed form of the padding quirk.
 *  - or_doppler_centroid: pure azimuth phase ramps of known frequency; numpy polyfit of the block estimates.
 *  - or_range_compression: direct time-domain correlation; analytic Hamming/rect compressed-pulse PSLR.
 *  - or_range_doppler: identity of RCMC at zero shift, closed-form windowed sinc at fractional shift, point-target
 *    focus position; unit-modulus matched filter.
 *  - or_az_multilook / or_rsl / or_srgr: constant and ramp inputs, brute-force loops, identity polynomial.
 */
#ifndef ASARF_ORACLE_H
#define ASARF_ORACLE_H

#include <complex.h>
#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef double _Complex or_zc;

/* Quirk flags. */
enum {
    OR_Q1_IQ_PADDED_WIDTH = 1 << 0, /* I/Q DC/gain/phase run over the padded row width (iq_correction.cu:197-198) */
    OR_Q2_COS_NO_SQRT = 1 << 1,     /* cos_corr = 1 - d^2, no sqrt (iq_correction.cu:259-262) */
    OR_Q4_DC_EXTRA_BLOCK = 1 << 2,  /* 33rd DC block reading padding when Nr % 32 != 0 (fractional_doppler_centroid.cu:107-112) */
};
#define OR_QUIRKS_REFERENCE (OR_Q1_IQ_PADDED_WIDTH | OR_Q2_COS_NO_SQRT | OR_Q4_DC_EXTRA_BLOCK)

/* ---------------- FFT (FP64 mixed radix Stockham, own implementation) ---------------- */
typedef struct or_fft_plan or_fft_plan;
or_fft_plan* or_fft_plan_create(int n);
void or_fft_plan_destroy(or_fft_plan* p);
/* In-place unnormalised transform of one contiguous vector. scratch: n elements (may be NULL -> malloc). */
void or_fft_exec(const or_fft_plan* p, or_zc* data, or_zc* scratch, int inverse);
/* Convenience: batch of rows (contiguous, stride in elements), OpenMP over rows. */
void or_fft_rows(or_zc* data, int n, int rows, int stride, int inverse);
/* Convenience: batch of columns (x < cols, length n, row stride), OpenMP over columns. */
void or_fft_cols(or_zc* data, int n, int cols, int stride, int inverse);
/* O(n^2) direct DFT (for tests). */
void or_dft(const or_zc* in, or_zc* out, int n, int inverse);

/* ---------------- Padding rule (cufft_plan.cc:147-181, main.cc:217,223) ---------------- */
/* Smallest 2^a 3^b 5^c 7^d strictly greater than n (limit 131072), or -1. */
int or_next_fft_size(int n);

/* ---------------- Chirp (sar_chirp.cc:27-104) ---------------- */
/* Writes `padding` complex-float values (I,Q interleaved) of the windowed, zero-padded, unit-energy chirp. */
void or_generate_chirp(double kr, double fs, int n_samples, int apply_window, int padding, float* out_iq);

/* ---------------- Polynomials (math_utils.cc:23-44, math_utils.h:55-63) ---------------- */
/* Least squares, Householder QR; coefficients highest power first (numpy.polyfit order). Returns 0 on success. */
int or_polyfit(const double* x, const double* y, int n, int degree, double* coeffs_out);
double or_polyval(const double* p, int n, double x);

/* ---------------- I/Q correction (iq_correction.cu:140-273, device_padded_image.cu:156-257,322-414) -------- */
typedef struct {
    double dc_i, dc_q;
    double iq_gain;
    double phase_error; /* degrees */
    double stddev_i, stddev_q;
} or_iq_results;

void or_iq_correction(or_zc* img, int x_size, int y_size, int x_stride, int y_stride, size_t n_total_samples,
                      int quirks, or_iq_results* res);

/* ---------------- Doppler centroid (fractional_doppler_centroid.cu:33-162) ---------------- */
/* Writes block estimates (Hz) and centres (range index) for up to `max_blocks`, and the degree-2 fit poly[3].
 * Returns the number of blocks. */
int or_doppler_centroid(const or_zc* img, int x_size, int y_size, int x_stride, double prf, int quirks,
                        double* block_freq, double* block_x, int max_blocks, double* poly3);

/* ---------------- Range compression (range_compression.cu:25-111) ---------------- */
/* In place over rows y < y_size of a padded image with row stride x_stride (= FFT length).
 * chirp_iq: x_stride complex-float values (the padded chirp). Returns the new x_size (x_size - nc). */
int or_range_compression(or_zc* img, int x_size, int y_size, int x_stride, const float* chirp_iq, int nc);

/* ---------------- Range-Doppler algorithm (range_doppler_algorithm.cu:29-407) ---------------- */
typedef struct {
    double wavelength;
    double range_spacing;
    double slant_range_first;
    double prf;
    double az_bandwidth_fraction;
    double dc_poly[3]; /* highest power first, range index */
    double vr_poly[3]; /* highest power first, range index */
} or_rda_params;

/* src: range-compressed image (x_size = Nrv, y_size = Na, strides x_stride/y_stride = Nap rows).
 * out: same geometry; receives the focused SLC in the logical region (x < x_size, y < y_size).
 * src is used as scratch (its contents are destroyed), as the reference trades it away. */
void or_range_doppler(or_zc* src, or_zc* out, int x_size, int y_size, int x_stride, int y_stride,
                      const or_rda_params* p);

/* Individual RDA pieces (for per-stage tests). `freq_img` is in the range-Doppler domain (natural bin order). */
void or_rcmc(const or_zc* in, or_zc* out, int x_size, int n_bins, int x_stride, const or_rda_params* p);
void or_az_matched_filter(or_zc* img, int x_size, int n_bins, int x_stride, const or_rda_params* p);
/* 16-tap Blackman window used by RCMC (numpy.blackman(16) values, range_doppler_algorithm.cu:49-52). */
void or_rcmc_window(float* w16);

/* ---------------- IMP tail ---------------- */
/* az_multilook.cu:25-101. in: Na_s x Nr_s complex (row stride in_stride). Returns out rows; out: out_rows x nr. */
int or_az_multilook_rows(int in_rows, double az_spacing);
void or_az_multilook(const or_zc* in, int nr, int in_rows, int in_stride, double az_spacing, double* out);
/* range_spreading_loss.cu:29-69: img[y][x] *= float((Rref/R(x))^-1.5). */
void or_rsl(double* img, int nr, int rows, double r_first, double range_spacing, double r_ref);
/* srgr.cu:49-71: out[y][x] = lerp of in[y] at Polyval(grsr, x*gs)/dr. */
void or_srgr(const double* in, int in_nr, int rows, const double* grsr5, double slant_spacing,
             double ground_spacing, int out_nr, double* out);

/* ---------------- Results window (main_flow.cc:155-201, sar_metadata.h:109-141) ---------------- */
int or_aperture_pixels(double carrier_wavelength, double prf, double az_bw_fraction, const double* vr_poly,
                       const double* dc_poly, int dc_terms, double r_first, double range_spacing, int range_pixel);

/* ---------------- Whole path (main.cc:233-301) ---------------- */
typedef struct {
    int na, nr, nc, nrp, nap;
    double prf, wavelength, range_spacing, slant_range_first, az_bandwidth_fraction;
    double vr_poly[3];
    double chirp_kr, range_sampling_rate;
    int apply_range_window;
    int quirks;
    size_t n_total_samples; /* 0 -> na*nr */
} or_focus_params;

typedef struct {
    or_iq_results iq;
    double dc_poly[3];
    int nrv;
} or_focus_results;

/* raw_iq: na x nr complex-float samples (interleaved float I,Q), contiguous. slc_out: na x nrv complex FP64. */
int or_focus_slc(const float* raw_iq, const or_focus_params* fp, or_zc* slc_out, or_focus_results* res);

/* Threads used by OpenMP regions (0 -> default). */
void or_set_threads(int n);
int or_get_threads(void);

/* ---------------- learned RCMC interpolator network (mlp.c; no reference counterpart) ---------------- */
int or_mlp_num_params(void);
void or_mlp_forward(const double* w, double u0, double u1, double* out16);
double or_mlp_loss_grad(const double* w, const float* pairs, long n, double* grad);
void or_sinc8_weights(double d, double* out8);

#ifdef __cplusplus
}
#endif
#endif
