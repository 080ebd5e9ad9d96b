// FP64 host geometry for real products (the reference's orbit, geolocation and processing-velocity code): Lagrange
// orbit interpolation, range-Doppler geolocation on the WGS84 ellipsoid, incidence angle, the processing
// velocity Vr(x) fit and the slant-to-ground polynomial from the orbit. Host-only; used before the GPU stage.
#include <cmath>
#include <cstdint>
#include <cstring>

#include "asarf.h"
#include "internal.h"
#include "poly.h"

using namespace asarf;

#define ST_CHECK(expr)                 \
    do {                               \
        const int s_ = (expr);         \
        if (s_ != ASARF_OK) return s_; \
    } while (0)

namespace {

constexpr double kWgs84A = 6378137.0;
// the reference's geolocation uses b rounded to the micrometre (util/src/geo_tools.cc:28); its incidence angle and
// coordinate conversions use the full value (util/include/geo_tools.h:34-39)
constexpr double kWgs84B = 6356752.314245;
constexpr double kWgs84BFull = 6356752.3142451794975639665996337;
constexpr double kFlattening = (kWgs84A - kWgs84BFull) / kWgs84A;
constexpr double kE2 = 2.0 * kFlattening - kFlattening * kFlattening;  // first eccentricity squared
constexpr double kEp2 = kE2 / (1.0 - kE2);                            // second eccentricity squared
constexpr double kDegToRad = M_PI / 180.0;
constexpr double kSpeedOfLight = 299792458.0;

// Time of azimuth line idx (sar/sar_metadata.h:90-93): first line time + line interval * idx truncated to whole
// microseconds (the reference converts through uint32 microseconds; for idx >= 0 that is truncation).
double line_time(const asarf_geometry* g, int idx) {
    const double us = std::trunc(g->line_time_interval * (double)idx * 1e6);
    return g->first_line_time + us * 1e-6;
}

// Solves the 3x3 system m x = r by Gaussian elimination with partial pivoting; false when singular.
bool solve3(double m[3][3], double r[3], double x[3]) {
    int p[3] = {0, 1, 2};
    for (int c = 0; c < 3; ++c) {
        int best = c;
        for (int i = c + 1; i < 3; ++i)
            if (std::fabs(m[p[i]][c]) > std::fabs(m[p[best]][c])) best = i;
        std::swap(p[c], p[best]);
        const double d = m[p[c]][c];
        if (d == 0.0) return false;
        for (int i = c + 1; i < 3; ++i) {
            const double f = m[p[i]][c] / d;
            for (int k = c; k < 3; ++k) m[p[i]][k] -= f * m[p[c]][k];
            r[p[i]] -= f * r[p[c]];
        }
    }
    for (int c = 2; c >= 0; --c) {
        double s = r[p[c]];
        for (int k = c + 1; k < 3; ++k) s -= m[p[c]][k] * x[k];
        x[c] = s / m[p[c]][c];
    }
    return true;
}

int interpolate(const asarf_osv* osv, int n, double t, asarf_osv* out) {
    // InterpolateOrbit (sar/orbit_state_vector.cc:23-48): Lagrange polynomial through every state vector, in
    // seconds from the first one; times have the reference's whole-microsecond resolution
    auto rel = [&](double tt) { return std::round((tt - osv[0].t) * 1e6) * 1e-6; };
    asarf_osv r{};
    r.t = t;
    const double x = rel(t);
    for (int i = 0; i < n; ++i) {
        double mult = 1.0;
        const double xi = rel(osv[i].t);
        for (int j = 0; j < n; ++j) {
            if (i == j) continue;
            const double xj = rel(osv[j].t);
            mult *= (x - xj) / (xi - xj);
        }
        for (int k = 0; k < 3; ++k) {
            r.pos[k] += mult * osv[i].pos[k];
            r.vel[k] += mult * osv[i].vel[k];
        }
    }
    *out = r;
    return ASARF_OK;
}

int geolocate(const double* pos, const double* vel, const double* guess, double slant_range, double* out) {
    // RangeDopplerGeoLocate (util/src/geo_tools.cc:27-85): Newton on (v . (p - s) = 0, |p - s|^2 = R^2, p on the
    // WGS84 ellipsoid) from the initial guess, until the step is below 1e-6 m (at most 101 steps)
    double p[3] = {guess[0], guess[1], guess[2]};
    const double a2 = kWgs84A * kWgs84A, b2 = kWgs84B * kWgs84B;
    for (int iter = 1; iter <= 101; ++iter) {
        const double d[3] = {p[0] - pos[0], p[1] - pos[1], p[2] - pos[2]};
        double m[3][3] = {{vel[0], vel[1], vel[2]},
                          {2.0 * d[0], 2.0 * d[1], 2.0 * d[2]},
                          {2.0 * p[0] / a2, 2.0 * p[1] / a2, 2.0 * p[2] / b2}};
        double r[3] = {-(vel[0] * d[0] + vel[1] * d[1] + vel[2] * d[2]),
                       -(d[0] * d[0] + d[1] * d[1] + d[2] * d[2] - slant_range * slant_range),
                       -(p[0] * p[0] / a2 + p[1] * p[1] / a2 + p[2] * p[2] / b2 - 1.0)};
        double dx[3];
        if (!solve3(m, r, dx)) return set_error(ASARF_EINVAL, "geolocation: singular system");
        for (int k = 0; k < 3; ++k) p[k] += dx[k];
        if (std::sqrt(dx[0] * dx[0] + dx[1] * dx[1] + dx[2] * dx[2]) < 1e-6) {
            std::memcpy(out, p, sizeof(p));
            return ASARF_OK;
        }
    }
    return set_error(ASARF_EINVAL, "geolocation: no convergence in 101 steps");
}

int check_geometry(const asarf_geometry* g, const char* who) {
    if (!g || !g->osv) return set_error(ASARF_EINVAL, "%s: null argument", who);
    if (g->n_osv < 2) return set_error(ASARF_EINVAL, "%s: need at least 2 state vectors", who);
    for (int i = 1; i < g->n_osv; ++i)
        if (!(g->osv[i].t > g->osv[i - 1].t)) return set_error(ASARF_EINVAL, "%s: state vector times not increasing", who);
    if (g->azimuth_size <= 0 || g->range_size <= g->chirp_samples || !(g->range_spacing > 0) ||
        !(g->slant_range_first > 0) || !(g->line_time_interval > 0))
        return set_error(ASARF_EINVAL, "%s: bad image geometry", who);
    return ASARF_OK;
}

}  // namespace

extern "C" int asarf_interpolate_orbit(const asarf_osv* osv, int n, double t, asarf_osv* out) {
    if (!osv || !out || n < 1) return set_error(ASARF_EINVAL, "asarf_interpolate_orbit: argument");
    for (int i = 1; i < n; ++i)
        if (!(osv[i].t > osv[i - 1].t)) return set_error(ASARF_EINVAL, "asarf_interpolate_orbit: times not increasing");
    ST_CHECK(interpolate(osv, n, t, out));
    return clear_error();
}

extern "C" int asarf_geolocate(const double* sensor_pos, const double* sensor_vel, const double* init_guess,
                               double slant_range, double* out) {
    if (!sensor_pos || !sensor_vel || !init_guess || !out) return set_error(ASARF_EINVAL, "asarf_geolocate: null");
    if (!(slant_range > 0)) return set_error(ASARF_EINVAL, "asarf_geolocate: slant range must be > 0");
    ST_CHECK(geolocate(sensor_pos, sensor_vel, init_guess, slant_range, out));
    return clear_error();
}

extern "C" double asarf_incidence_angle(const double* earth_point, const double* sensor_pos) {
    // CalcIncidenceAngle (util/src/geo_tools.cc:87-106): angle between the ellipsoid normal and the direction to
    // the sensor, in degrees
    const double a2 = kWgs84A * kWgs84A, b2 = kWgs84BFull * kWgs84BFull;
    const double n[3] = {earth_point[0] / a2, earth_point[1] / a2, earth_point[2] / b2};
    const double v[3] = {sensor_pos[0] - earth_point[0], sensor_pos[1] - earth_point[1], sensor_pos[2] - earth_point[2]};
    const double dot = n[0] * v[0] + n[1] * v[1] + n[2] * v[2];
    const double nn = std::sqrt(n[0] * n[0] + n[1] * n[1] + n[2] * n[2]);
    const double vn = std::sqrt(v[0] * v[0] + v[1] * v[1] + v[2] * v[2]);
    return (360.0 / (2.0 * M_PI)) * std::acos(dot / (nn * vn));
}

extern "C" int asarf_estimate_processing_velocity(const asarf_geometry* g, double* vr_poly3, double* vr_samples,
                                                  double* range_samples, int* n_samples) {
    // EstimateProcessingVelocity (sar/processing_velocity_estimation.cc:56-121): at mid-azimuth and 10 range
    // positions, geolocate the zero-Doppler ground point, sample R^2(t) - R0^2 at 9 times across the synthetic
    // aperture and fit y = a t^2 + c (Vr = sqrt(a)); then a degree-2 polynomial of Vr over the range index
    ST_CHECK(check_geometry(g, "asarf_estimate_processing_velocity"));
    if (!vr_poly3) return set_error(ASARF_EINVAL, "asarf_estimate_processing_velocity: null output");
    const int az_idx = g->azimuth_size / 2;
    const int range_end = g->range_size - g->chirp_samples;
    const int step = range_end / 10;
    if (step <= 0) return set_error(ASARF_EINVAL, "asarf_estimate_processing_velocity: range too short");
    double xs[16], vrs[16];
    int n = 0;
    for (int x = 0; x < range_end && n < 16; x += step) {
        const double r0 = g->slant_range_first + (double)x * g->range_spacing;
        // CalcAperturePixels (sar/sar_metadata.h:117-124) with the geometry's current Vr and Doppler polynomials
        const double vr = polyval(g->vr_poly, 3, (double)x);
        const double ka = 2.0 * g->carrier_frequency * vr * vr / (kSpeedOfLight * r0);
        const double fmax = polyval(g->dc_poly, 3, (double)x) + g->prf * g->az_bandwidth_fraction / 2.0;
        const int aperture = (int)std::round(fmax / ka * g->prf);
        asarf_osv c{};
        ST_CHECK(interpolate(g->osv, g->n_osv, line_time(g, az_idx), &c));
        double ep[3];
        ST_CHECK(geolocate(c.pos, c.vel, g->center_point, r0, ep));
        const int astep = aperture / 8;
        // y = a t^2 + c over the 9 aperture samples: least squares on the columns (1, t^2), solved in centred form
        double t2s[9], ys[9], mt = 0.0, my = 0.0;
        for (int i = 0; i < 9; ++i) {
            const int ai = az_idx + (i - 4) * astep;
            asarf_osv o{};
            ST_CHECK(interpolate(g->osv, g->n_osv, line_time(g, ai), &o));
            const double dx = o.pos[0] - ep[0], dy = o.pos[1] - ep[1], dz = o.pos[2] - ep[2];
            // time difference in whole microseconds, as the reference's ptime arithmetic
            const double dt = std::round((c.t - o.t) * 1e6) * 1e-6;
            ys[i] = dx * dx + dy * dy + dz * dz - r0 * r0;
            t2s[i] = dt * dt;
            mt += t2s[i] / 9.0;
            my += ys[i] / 9.0;
        }
        double sxx = 0.0, sxy = 0.0;
        for (int i = 0; i < 9; ++i) {
            sxx += (t2s[i] - mt) * (t2s[i] - mt);
            sxy += (t2s[i] - mt) * (ys[i] - my);
        }
        if (!(sxx > 0.0)) return set_error(ASARF_EINVAL, "asarf_estimate_processing_velocity: degenerate aperture");
        const double a = sxy / sxx;
        xs[n] = (double)x;
        vrs[n] = std::sqrt(a);
        ++n;
    }
    if (polyfit_qr(xs, vrs, n, 2, vr_poly3) != 0)
        return set_error(ASARF_EINVAL, "asarf_estimate_processing_velocity: fit failed");
    if (vr_samples) std::memcpy(vr_samples, vrs, sizeof(double) * (size_t)n);
    if (range_samples) std::memcpy(range_samples, xs, sizeof(double) * (size_t)n);
    if (n_samples) *n_samples = n;
    return clear_error();
}

extern "C" int asarf_srgr_fit(const asarf_geometry* g, double ground_spacing, double* grsr5, int* out_range_size) {
    // SlantRangeToGroundRangeInterpolation host part (sar/srgr.cu:73-107): at mid-azimuth, geolocate 100 slant
    // ranges R_i = R_first + i * (range_size / 100) * dr; ground distance = chord from the first ground point;
    // degree-4 fit of slant - R_first against ground distance; output size floor(g_99 / ground_spacing)
    ST_CHECK(check_geometry(g, "asarf_srgr_fit"));
    if (!grsr5 || !out_range_size || !(ground_spacing > 0)) return set_error(ASARF_EINVAL, "asarf_srgr_fit: argument");
    if (g->range_size < 100) return set_error(ASARF_EINVAL, "asarf_srgr_fit: range size < 100");
    asarf_osv o{};
    ST_CHECK(interpolate(g->osv, g->n_osv, line_time(g, g->azimuth_size / 2), &o));
    double first[3];
    ST_CHECK(geolocate(o.pos, o.vel, g->center_point, g->slant_range_first, first));
    double gd[100], sr[100];
    for (int i = 0; i < 100; ++i) {
        const double r = g->slant_range_first + (double)(i * (g->range_size / 100)) * g->range_spacing;
        double p[3];
        ST_CHECK(geolocate(o.pos, o.vel, g->center_point, r, p));
        const double dx = p[0] - first[0], dy = p[1] - first[1], dz = p[2] - first[2];
        gd[i] = std::sqrt(dx * dx + dy * dy + dz * dz);
        sr[i] = r - g->slant_range_first;
    }
    if (polyfit_qr(gd, sr, 100, 4, grsr5) != 0) return set_error(ASARF_EINVAL, "asarf_srgr_fit: fit failed");
    *out_range_size = (int)(gd[99] / ground_spacing);
    return clear_error();
}

extern "C" int asarf_geo_to_xyz(double latitude, double longitude, double altitude, double* xyz) {
    // Geo2xyzWgs84 (util/include/geo_tools.h:52-74): geodetic degrees and height to ECEF metres
    if (!xyz) return set_error(ASARF_EINVAL, "asarf_geo_to_xyz: null");
    const double lat = latitude * kDegToRad, lon = longitude * kDegToRad;
    const double sl = std::sin(lat), cl = std::cos(lat);
    const double nr = kWgs84A / std::sqrt(1.0 - kE2 * sl * sl);  // prime-vertical radius of curvature
    xyz[0] = (nr + altitude) * cl * std::cos(lon);
    xyz[1] = (nr + altitude) * cl * std::sin(lon);
    xyz[2] = (nr * (1.0 - kE2) + altitude) * sl;
    return clear_error();
}

extern "C" int asarf_xyz_to_geo(const double* xyz, double* lat_lon) {
    // xyz2geoWGS84 (util/include/geo_tools.h:76-95): Bowring's closed form from the parametric latitude; longitude
    // from atan(y / x) moved into the half plane of y (0 at y = 0 with x < 0, as the reference)
    if (!xyz || !lat_lon) return set_error(ASARF_EINVAL, "asarf_xyz_to_geo: null");
    const double x = xyz[0], y = xyz[1], z = xyz[2];
    const double s = std::sqrt(x * x + y * y);
    const double th = std::atan(z * kWgs84A / (s * kWgs84BFull));
    double lon = std::atan(y / x) / kDegToRad;
    if (lon < 0.0 && y >= 0.0)
        lon += 180.0;
    else if (lon > 0.0 && y < 0.0)
        lon -= 180.0;
    const double st = std::sin(th), ct = std::cos(th);
    lat_lon[0] = std::atan((z + kEp2 * kWgs84BFull * st * st * st) / (s - kE2 * kWgs84A * ct * ct * ct)) / kDegToRad;
    lat_lon[1] = lon;
    return clear_error();
}

extern "C" int asarf_scene_center(const asarf_geometry* g, const double* guess_xyz, double* center_xyz,
                                  double* azimuth_spacing) {
    // Scene centre and azimuth pixel spacing (main_flow.cc:499-522): geolocate the mid-swath slant range
    // R_first + ((range_size - chirp_samples) / 2) * dr at the middle line from the guess; the spacing is the
    // distance to the same range's point one line later
    ST_CHECK(check_geometry(g, "asarf_scene_center"));
    if (!guess_xyz || !center_xyz) return set_error(ASARF_EINVAL, "asarf_scene_center: null");
    const int mid = g->azimuth_size / 2;
    const double r = g->slant_range_first + (double)((g->range_size - g->chirp_samples) / 2) * g->range_spacing;
    asarf_osv o{};
    ST_CHECK(interpolate(g->osv, g->n_osv, line_time(g, mid), &o));
    double c[3];
    ST_CHECK(geolocate(o.pos, o.vel, guess_xyz, r, c));
    if (azimuth_spacing) {
        asarf_osv o1{};
        ST_CHECK(interpolate(g->osv, g->n_osv, line_time(g, mid + 1), &o1));
        double c1[3];
        ST_CHECK(geolocate(o1.pos, o1.vel, c, r, c1));
        *azimuth_spacing = std::sqrt((c1[0] - c[0]) * (c1[0] - c[0]) + (c1[1] - c[1]) * (c1[1] - c[1]) +
                                     (c1[2] - c[2]) * (c1[2] - c[2]));
    }
    std::memcpy(center_xyz, c, sizeof(c));
    return clear_error();
}

extern "C" int asarf_aperture_pixels(const asarf_geometry* g, int* out, int capacity, int* n_out) {
    // CalcAperturePixels (sar/sar_metadata.h:117-141): round(fmax / Ka * PRF) at the range indices 0, step, ...
    // below range_size - chirp_samples (step = that / 10), with Ka = 2 f0 Vr^2 / (c R0) and fmax = fDC + PRF * B / 2
    if (!g || !out || !n_out) return set_error(ASARF_EINVAL, "asarf_aperture_pixels: null");
    const int range_end = g->range_size - g->chirp_samples;
    const int step = range_end / 10;
    if (step <= 0 || !(g->range_spacing > 0) || !(g->slant_range_first > 0))
        return set_error(ASARF_EINVAL, "asarf_aperture_pixels: bad image geometry");
    int n = 0;
    for (int x = 0; x < range_end; x += step) {
        if (n == capacity) return set_error(ASARF_EINVAL, "asarf_aperture_pixels: capacity %d too small", capacity);
        const double r0 = g->slant_range_first + (double)x * g->range_spacing;
        const double vr = polyval(g->vr_poly, 3, (double)x);
        const double ka = 2.0 * g->carrier_frequency * vr * vr / (kSpeedOfLight * r0);
        const double fmax = polyval(g->dc_poly, 3, (double)x) + g->prf * g->az_bandwidth_fraction / 2.0;
        out[n++] = (int)std::round(fmax / ka * g->prf);
    }
    *n_out = n;
    return clear_error();
}

extern "C" int asarf_results_window(double dc_constant, long lines_before_start, long lines_after_stop,
                                    long max_aperture, int rcmc_window, asarf_results_window_t* out) {
    // CalcResultsWindow (main_flow.cc:155-201): the lines read beyond the requested sensing window are removed;
    // when they are fewer than the longest aperture, the shortfall is removed inside the window too, at the start
    // for a positive Doppler centroid constant term and at the stop otherwise; half the RCMC window is removed at
    // near and at far range
    if (!out) return set_error(ASARF_EINVAL, "asarf_results_window: null");
    if (lines_before_start < 0 || lines_after_stop < 0 || max_aperture < 0 || rcmc_window < 0)
        return set_error(ASARF_EINVAL, "asarf_results_window: negative size");
    asarf_results_window_t w{};
    w.lines_before_sensing_start = lines_before_start;
    w.lines_after_sensing_stop = lines_after_stop;
    if (dc_constant > 0) {
        if (lines_before_start < max_aperture) w.lines_after_sensing_start = max_aperture - lines_before_start;
    } else if (lines_after_stop < max_aperture) {
        w.lines_before_sensing_stop = max_aperture - lines_after_stop;
    }
    w.near_range_pixels = rcmc_window / 2;
    w.far_range_pixels = rcmc_window / 2;
    *out = w;
    return clear_error();
}

extern "C" int asarf_geolocation_tie_points(const asarf_geometry* g, int az_idx, const double* grsr5,
                                            asarf_tie_points* out) {
    // FillGeoLocationAds (envisat_lvl1_writer.cc:277-324), one line: 11 tie points at range samples
    // i * range_size / 10; slant range R_first + sample * dr, or with the slant-to-ground polynomial applied
    // (grsr5 non-NULL; dr is then the ground spacing) R_first + grsr(sample * dr); the zero-Doppler point at the
    // line's time from the geometry's centre point; sample number + 1, two-way slant time [ns], incidence angle
    // [deg] and latitude / longitude in microdegrees (truncated toward zero into int32)
    ST_CHECK(check_geometry(g, "asarf_geolocation_tie_points"));
    if (!out) return set_error(ASARF_EINVAL, "asarf_geolocation_tie_points: null");
    asarf_osv o{};
    ST_CHECK(interpolate(g->osv, g->n_osv, line_time(g, az_idx), &o));
    asarf_tie_points r{};
    for (int i = 0; i < 11; ++i) {
        const int samp = i * g->range_size / 10;
        double R = g->slant_range_first + (double)samp * g->range_spacing;
        if (grsr5) R = g->slant_range_first + polyval(grsr5, 5, (double)samp * g->range_spacing);
        double xyz[3];
        ST_CHECK(geolocate(o.pos, o.vel, g->center_point, R, xyz));
        double ll[2];
        ST_CHECK(asarf_xyz_to_geo(xyz, ll));
        r.samp_numbers[i] = (unsigned)(samp + 1);
        r.slant_range_times[i] = (float)(2.0 * (R / kSpeedOfLight) * 1e9);
        r.angles[i] = (float)asarf_incidence_angle(xyz, o.pos);
        r.lats[i] = (int)(ll[0] * 1e6);
        r.longs[i] = (int)(ll[1] * 1e6);
    }
    *out = r;
    return clear_error();
}
