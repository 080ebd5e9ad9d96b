// Host-side helpers of the C ABI: padding rule, chirp generation, polynomial fit, spherical SRGR geometry.
#include <cmath>
#include <cstdio>
#include <vector>

#include "asarf.h"
#include "internal.h"
#include "poly.h"

using namespace asarf;

extern "C" int asarf_fft_padding_size(int n) {
    // cuda_util/cufft_plan.cc:147-181 enumerates 2^a 3^b 5^c 7^d <= 131072; main.cc:217 takes upper_bound(n).
    const long limit = 131072;
    long best = -1;
    for (long a = 1; a <= limit; a *= 2)
        for (long b = a; b <= limit; b *= 3)
            for (long c = b; c <= limit; c *= 5)
                for (long d = c; d <= limit; d *= 7)
                    if (d > n && (best < 0 || d < best)) best = d;
    return (int)best;
}

extern "C" int asarf_generate_chirp(const asarf_chirp_info* ci, int padding, float* out_iq) {
    if (!ci || !out_iq) return set_error(ASARF_EINVAL, "asarf_generate_chirp: null argument");
    if (ci->n_samples <= 0 || padding < ci->n_samples || ci->range_sampling_rate <= 0.0)
        return set_error(ASARF_EINVAL, "asarf_generate_chirp: n_samples=%d padding=%d fs=%g", ci->n_samples, padding,
                         ci->range_sampling_rate);
    // sar/sar_chirp.cc:70-104: t = (i - n/2)/fs (integer n/2), phase = pi Kr t^2; periodic Hamming over n;
    // unit energy (the reference sums |FFT|^2/N, equal by Parseval to the time-domain energy used here).
    const int n = ci->n_samples;
    const double dt = 1.0 / ci->range_sampling_rate;
    std::vector<double> re(n), im(n);
    double energy = 0.0;
    for (int i = 0; i < n; ++i) {
        const double t = (double)(i - n / 2) * dt;
        const double ph = M_PI * ci->kr * t * t;
        const double w = ci->apply_window ? 0.54 - 0.46 * std::cos(2.0 * M_PI * (double)i / (double)n) : 1.0;
        re[i] = std::cos(ph) * w;
        im[i] = std::sin(ph) * w;
        energy += re[i] * re[i] + im[i] * im[i];
    }
    const double scale = ci->apply_window ? std::sqrt(energy) : std::sqrt((double)n);
    for (int i = 0; i < padding; ++i) {
        out_iq[2 * i] = i < n ? (float)(re[i] / scale) : 0.0f;
        out_iq[2 * i + 1] = i < n ? (float)(im[i] / scale) : 0.0f;
    }
    return clear_error();
}

extern "C" int asarf_polyfit(const double* x, const double* y, int n, int degree, double* coeffs) {
    if (!x || !y || !coeffs) return set_error(ASARF_EINVAL, "asarf_polyfit: null argument");
    if (degree < 0 || degree > 5 || n < degree + 1 || n > 128)
        return set_error(ASARF_EINVAL, "asarf_polyfit: n=%d degree=%d unsupported", n, degree);
    if (polyfit_qr(x, y, n, degree, coeffs) != 0) return set_error(ASARF_EINVAL, "asarf_polyfit: rank deficient");
    return clear_error();
}

extern "C" double asarf_polyval(const double* p, int n, double x) { return p ? polyval(p, n, x) : 0.0; }

extern "C" int asarf_rcmc_window_size(void) { return 16; }

extern "C" int asarf_srgr_fit_spherical(double r_first, double dr, int range_size, double rs, double re, double gs,
                                        double* grsr5, int* out_range_size) {
    if (!grsr5 || !out_range_size) return set_error(ASARF_EINVAL, "asarf_srgr_fit_spherical: null argument");
    if (range_size < 100 || dr <= 0 || rs <= re || re <= 0 || gs <= 0)
        return set_error(ASARF_EINVAL, "asarf_srgr_fit_spherical: bad geometry");
    // sar/srgr.cu:73-121 with RangeDopplerGeoLocate replaced by the spherical-earth closed form: the ground point of
    // slant range R lies at central angle acos((rs^2 + re^2 - R^2) / (2 rs re)); distances are chords between
    // ground points, as the reference measures them between ECEF positions.
    double g[100], s[100];
    const double a0 = std::acos((rs * rs + re * re - r_first * r_first) / (2.0 * rs * re));
    for (int i = 0; i < 100; ++i) {
        const double r = r_first + (double)(i * (range_size / 100)) * dr;
        const double c = (rs * rs + re * re - r * r) / (2.0 * rs * re);
        if (c > 1.0 || c < -1.0) return set_error(ASARF_EINVAL, "asarf_srgr_fit_spherical: range beyond horizon");
        const double a = std::acos(c);
        g[i] = 2.0 * re * std::sin(0.5 * (a - a0));
        s[i] = r - r_first;
    }
    if (polyfit_qr(g, s, 100, 4, grsr5) != 0) return set_error(ASARF_EINVAL, "asarf_srgr_fit_spherical: fit failed");
    *out_range_size = (int)(g[99] / gs);
    return clear_error();
}

namespace asarf {

int aperture_pixels_max(double wavelength, double prf, double beta, const double* vr_poly, double r_first, double dr,
                        int range_size, int chirp_samples) {
    // sar/sar_metadata.h:109-141 with an empty Doppler polynomial (it runs before the estimate, main.cc:251).
    const int range_end = range_size - chirp_samples;
    const int step = range_end / 10;
    int best = 0;
    if (step <= 0) return 0;
    for (int x = 0; x < range_end; x += step) {
        const double vr = polyval(vr_poly, 3, (double)x);
        const double r0 = r_first + (double)x * dr;
        const double ka = 2.0 * vr * vr / (wavelength * r0);
        const double fmax = prf * beta / 2.0;
        const int px = (int)std::round(fmax / ka * prf);
        if (px > best) best = px;
    }
    return best;
}

}  // namespace asarf
