// Auxiliary inputs of a real product (host C++, no device work): DORIS precise orbit files, the ASAR processor
// configuration (CON) and external calibration (XCA) files, the swath parameters of the instrument characterisation
// file (INS), and the radar timing that the level-0 flow derives from them. Every file is taken as the bytes the
// caller read; nothing here opens files. Big-endian binary fields are read at the offsets of the reference's packed
// structs (cited per function), text fields through the MPH/SPH key lookup.
#include <cctype>
#include <cmath>
#include <cstdint>
#include <cstdlib>
#include <cstring>

#include "asarf.h"
#include "internal.h"

using namespace asarf;

namespace {

constexpr long kMph = 1247, kSph = 378;  // MPH and SPH of the auxiliary files (envisat_aux_file.cc:47-49)
constexpr double kC = 299792458.0;

uint32_t rd_be32(const unsigned char* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | p[3];
}
uint16_t rd_be16(const unsigned char* p) { return (uint16_t)(((unsigned)p[0] << 8) | p[1]); }
float rd_bef(const unsigned char* p) {
    const uint32_t w = rd_be32(p);
    float f;
    std::memcpy(&f, &w, 4);
    return f;
}

// Days from 2000-01-01 of a proleptic Gregorian date, by Howard Hinnant's published days_from_civil method
// (years starting in March, so the leap day ends the year; 400-year eras of 146097 days; the day of the March-based
// year from the 153-days-per-5-months rule).
long days_since_2000(long y, int m, int d) {
    y -= m <= 2;
    const long era = (y >= 0 ? y : y - 399) / 400;
    const long yoe = y - era * 400;
    const long doy = (153 * (m + (m > 2 ? -3 : 9)) + 2) / 5 + d - 1;
    const long doe = yoe * 365 + yoe / 4 - yoe / 100 + doy;
    return era * 146097 + doe - 719468 - 10957;  // 719468: 0000-03-01 to 1970-01-01; 10957: 1970 to 2000
}

bool all_digits(const char* s, int n) {
    for (int i = 0; i < n; ++i)
        if (!std::isdigit((unsigned char)s[i])) return false;
    return true;
}

int num(const char* s, int n) {
    int v = 0;
    for (int i = 0; i < n; ++i) v = 10 * v + (s[i] - '0');
    return v;
}

bool valid_date(int y, int m, int d) {
    static const int mdays[12] = {31, 29, 31, 30, 31, 30, 31, 31, 30, 31, 30, 31};
    if (m < 1 || m > 12 || d < 1 || d > mdays[m - 1]) return false;
    const bool leap = (y % 4 == 0 && y % 100 != 0) || y % 400 == 0;
    return !(m == 2 && d == 29 && !leap);
}

bool make_mjd(int y, int mo, int d, int hh, int mi, int ss, unsigned micros, asarf_mjd* out) {
    if (!valid_date(y, mo, d) || hh > 23 || mi > 59 || ss > 59) return false;
    out->days = (int)days_since_2000(y, mo, d);
    out->seconds = (unsigned)(hh * 3600 + mi * 60 + ss);
    out->micros = micros;
    return true;
}

// "DD-MMM-YYYY" and "HH:MM:SS[.ffffff]" (the "%d-%b-%Y %H:%M:%S%f" facet of doris_orbit.cc:58): month names in
// either case; fraction digits beyond microseconds are dropped.
bool parse_doris_time(const char* date, int dlen, const char* time, int tlen, asarf_mjd* out) {
    static const char* months[12] = {"JAN", "FEB", "MAR", "APR", "MAY", "JUN", "JUL", "AUG", "SEP", "OCT", "NOV", "DEC"};
    if (dlen != 11 || date[2] != '-' || date[6] != '-' || !all_digits(date, 2) || !all_digits(date + 7, 4)) return false;
    int mo = 0;
    for (int m = 0; m < 12 && !mo; ++m) {
        bool eq = true;
        for (int k = 0; k < 3; ++k) eq = eq && std::toupper((unsigned char)date[3 + k]) == months[m][k];
        if (eq) mo = m + 1;
    }
    if (!mo) return false;
    if (tlen < 8 || time[2] != ':' || time[5] != ':' || !all_digits(time, 2) || !all_digits(time + 3, 2) ||
        !all_digits(time + 6, 2))
        return false;
    unsigned micros = 0;
    if (tlen > 8) {
        if (time[8] != '.' || tlen == 9 || !all_digits(time + 9, tlen - 9)) return false;
        for (int k = 0; k < 6; ++k) micros = 10 * micros + (9 + k < tlen ? (unsigned)(time[9 + k] - '0') : 0u);
    }
    return make_mjd(num(date + 7, 4), mo, num(date, 2), num(time, 2), num(time + 3, 2), num(time + 6, 2), micros, out);
}

// "YYYYMMDD" + "HHMMSS" of a file name (the "%Y%m%d_%H%M%S" facet of doris_orbit.cc:60).
bool parse_name_time(const char* ymd, const char* hms, asarf_mjd* out) {
    if (!all_digits(ymd, 8) || !all_digits(hms, 6)) return false;
    return make_mjd(num(ymd, 4), num(ymd + 4, 2), num(ymd + 6, 2), num(hms, 2), num(hms + 2, 2), num(hms + 4, 2), 0,
                    out);
}

long long micros_of(const asarf_mjd& m) {
    return ((long long)m.days * 86400 + m.seconds) * 1000000LL + m.micros;
}

bool header_value(const unsigned char* text, long n, const char* key, char* value, int capacity) {
    if (asarf_header_get(reinterpret_cast<const char*>(text), n, key, value, capacity) != ASARF_OK) {
        clear_error();
        return false;
    }
    return true;
}

}  // namespace

extern "C" int asarf_doris_orbit(const unsigned char* file, long n_bytes, const asarf_mjd* start, const asarf_mjd* stop,
                                 asarf_osv* osv, asarf_mjd* times, int capacity, int* n_out, asarf_doris_meta* meta) {
    // Parsable::CreateFromFile + CreateOrbitInfo + UpdateMetadataForL1Product (doris_orbit.cc:82-150, 152-176,
    // 260-304): a DOR_VOR_AX file is a 1247-byte MPH, a 378-byte SPH and text records of 129 bytes (1589 for
    // ENVISAT, 1801 for ERS); each record line holds 11 space-separated items: date, time, UT1 - UTC, absolute orbit,
    // x, y, z [m], vx, vy, vz [m/s], quality. The vectors from 5 minutes before start to 5 minutes after stop are
    // kept (at least 10, else the file is taken as corrupt); the L1 metadata comes from the first kept vector later
    // than one minute before start.
    constexpr long kEnvisatSize = 206606, kErsSize = 233954;
    if (!file || !start || !stop || !n_out || (capacity > 0 && !osv))
        return set_error(ASARF_EINVAL, "asarf_doris_orbit: null argument");
    *n_out = 0;
    if (n_bytes != kEnvisatSize && n_bytes != kErsSize)
        return set_error(ASARF_EINVAL,
                         "asarf_doris_orbit: a DORIS orbit file has %ld (ENVISAT) or %ld (ERS) bytes, not %ld",
                         kEnvisatSize, kErsSize, n_bytes);
    const long long t0 = micros_of(*start) - 5LL * 60 * 1000000, t1 = micros_of(*stop) + 5LL * 60 * 1000000;
    const long long t_meta = micros_of(*start) - 60LL * 1000000;
    const char* p = reinterpret_cast<const char*>(file) + kMph + kSph;
    const char* end = reinterpret_cast<const char*>(file) + n_bytes;
    int kept = 0;
    bool have_meta = false;
    asarf_doris_meta m{};
    int line_no = 0;
    while (p < end) {
        const char* eol = p;
        while (eol < end && *eol != '\n') ++eol;
        ++line_no;
        // items separated by runs of spaces; a leading or trailing separator gives an empty first or last item, as a
        // compressing split does
        const char* item[12];
        int ilen[12];
        int n_items = 0;
        const char* q = p;
        bool leading = q < eol && *q == ' ';
        if (leading) {
            item[n_items] = q;
            ilen[n_items++] = 0;
        }
        while (q < eol) {
            while (q < eol && *q == ' ') ++q;
            if (q >= eol) break;
            const char* b = q;
            while (q < eol && *q != ' ') ++q;
            if (n_items < 12) {
                item[n_items] = b;
                ilen[n_items] = (int)(q - b);
            }
            ++n_items;
        }
        if (eol > p && eol[-1] == ' ') ++n_items;
        if (n_items != 11)
            return set_error(ASARF_EINVAL, "asarf_doris_orbit: record %d has %d items, not 11", line_no, n_items);
        asarf_mjd t{};
        if (!parse_doris_time(item[0], ilen[0], item[1], ilen[1], &t))
            return set_error(ASARF_EINVAL, "asarf_doris_orbit: record %d: unparseable date", line_no);
        const long long tu = micros_of(t);
        p = eol + 1;
        if (tu < t0 || tu > t1) continue;
        char buf[12][48];
        for (int k = 0; k < 11; ++k) {
            const int l = ilen[k] < 47 ? ilen[k] : 47;
            std::memcpy(buf[k], item[k], (size_t)l);
            buf[k][l] = '\0';
        }
        asarf_osv v{};
        v.t = (double)t.days * 86400.0 + (double)t.seconds + 1e-6 * (double)t.micros;
        for (int k = 0; k < 3; ++k) {
            v.pos[k] = std::strtod(buf[4 + k], nullptr);
            v.vel[k] = std::strtod(buf[7 + k], nullptr);
        }
        if (kept < capacity) {
            osv[kept] = v;
            if (times) times[kept] = t;
        }
        if (!have_meta && tu > t_meta) {
            have_meta = true;
            m.delta_ut1 = std::strtod(buf[2], nullptr);
            m.abs_orbit = (unsigned)std::strtoul(buf[3], nullptr, 10);
            m.quality = (unsigned)std::strtoul(buf[10], nullptr, 10);
            m.state_vector_time = t;
            for (int k = 0; k < 3; ++k) {
                m.pos[k] = v.pos[k];
                m.vel[k] = v.vel[k];
            }
        }
        ++kept;
    }
    if (kept < 10)
        return set_error(ASARF_EINVAL, "asarf_doris_orbit: %d state vectors within 5 minutes of the window, 10 needed",
                         kept);
    if (!have_meta) return set_error(ASARF_EINVAL, "asarf_doris_orbit: no state vector for the L1 metadata");
    if (kept > capacity)
        return set_error(ASARF_EWORKSPACE, "asarf_doris_orbit: %d state vectors, capacity %d", kept, capacity);
    *n_out = kept;
    if (meta) {
        char value[256];
        if (header_value(file, kMph, "PRODUCT", value, sizeof(value)))
            std::snprintf(m.orbit_name, sizeof(m.orbit_name), "%s", value);
        std::strcpy(m.vector_source, "??");
        if (header_value(file + kMph, kSph, "DS_NAME", value, sizeof(value))) {
            if (std::strstr(value, "DORIS PRECISE")) std::strcpy(m.vector_source, "DP");
            else if (std::strstr(value, "REAPER ORBIT")) std::strcpy(m.vector_source, "RO");
        }
        if (header_value(file, kMph, "PHASE", value, sizeof(value))) m.phase = value[0];
        if (header_value(file, kMph, "CYCLE", value, sizeof(value))) m.cycle = (unsigned)std::strtoul(value, nullptr, 10);
        *meta = m;
    }
    return clear_error();
}

extern "C" int asarf_doris_filename_window(const char* name, asarf_mjd* valid_start, asarf_mjd* valid_stop) {
    // Parsable::CreateFromDirectoryListing (doris_orbit.cc:178-246): DOR_VOR_AXVF-P20120424_120200_20040101_215528_
    // 20040103_002328 (61 characters, 8 '_'-separated items) or the ERS REAPER form with one more HHMMSS item (68
    // characters, 9 items); the validity window is items 4-5 to items 6-7 (7-8 in the longer form).
    if (!name || !valid_start || !valid_stop) return set_error(ASARF_EINVAL, "asarf_doris_filename_window: null");
    const size_t len = std::strlen(name);
    if (len != 61 && len != 68) return set_error(ASARF_EINVAL, "asarf_doris_filename_window: %zu characters", len);
    const char* it[10];
    int il[10];
    int n = 0;
    const char* s = name;
    while (*s && n < 10) {
        while (*s == '_') ++s;  // runs of '_' separate one item
        if (!*s) break;
        const char* b = s;
        while (*s && *s != '_') ++s;
        it[n] = b;
        il[n++] = (int)(s - b);
    }
    if (*s || (n != 8 && n != 9)) return set_error(ASARF_EINVAL, "asarf_doris_filename_window: %d items", n);
    const int want8[8] = {3, 3, 14, 6, 8, 6, 8, 6};
    const int want9[9] = {3, 3, 14, 6, 8, 6, 6, 8, 6};
    for (int k = 0; k < n; ++k)
        if (il[k] != (n == 8 ? want8[k] : want9[k]))
            return set_error(ASARF_EINVAL, "asarf_doris_filename_window: item %d has %d characters", k, il[k]);
    if (std::strncmp(it[0], "DOR", 3) != 0 || std::strncmp(it[1], "VOR", 3) != 0)
        return set_error(ASARF_EINVAL, "asarf_doris_filename_window: not a DOR_VOR file name");
    const int e = n == 8 ? 6 : 7;
    if (!parse_name_time(it[4], it[5], valid_start) || !parse_name_time(it[e], it[e + 1], valid_stop))
        return set_error(ASARF_EINVAL, "asarf_doris_filename_window: unparseable date");
    return clear_error();
}

extern "C" int asarf_con_process_gains(const unsigned char* con, long n_bytes, float* gain_ims, float* gain_imp) {
    // FindCONFile (envisat_aux_file.cc:41-65): a 5721-byte processor configuration file = MPH + SPH + the
    // 4096-byte configuration record, whose big-endian process_gain_imp[7] and process_gain_ims[7] (swaths IS1-IS7)
    // start at record offsets 1020 and 1048 (envisat_aux_file.h:331-332, :710).
    if (!con || !gain_ims || !gain_imp) return set_error(ASARF_EINVAL, "asarf_con_process_gains: null");
    if (n_bytes != 5721)
        return set_error(ASARF_ESHAPE, "asarf_con_process_gains: %ld bytes, a configuration file has 5721", n_bytes);
    const unsigned char* rec = con + kMph + kSph;
    for (int s = 0; s < 7; ++s) {
        gain_imp[s] = rd_bef(rec + 1020 + 4 * s);
        gain_ims[s] = rd_bef(rec + 1048 + 4 * s);
    }
    return clear_error();
}

extern "C" int asarf_xca_scaling_factors(const unsigned char* xca, long n_bytes, float* slc_vv, float* precision_vv) {
    // GetXca (envisat_aux_file.cc:99-131): a 28177-byte external calibration file = MPH + SPH + one 26552-byte GADS
    // (PO-RS-MDA-GS-2009 table 8.6.2.1-1, envisat_aux_xca.h:25-40): creation time (12 bytes), DSR length (4), then
    // big-endian scaling factors of 7 swaths each: IM SLC HH, IM SLC VV, IM precision HH, IM precision VV, ...
    // A DSR length other than 26552 is an error, except 0 (as ERS files have it).
    if (!xca || !slc_vv || !precision_vv) return set_error(ASARF_EINVAL, "asarf_xca_scaling_factors: null");
    if (n_bytes != 28177)
        return set_error(ASARF_ESHAPE, "asarf_xca_scaling_factors: %ld bytes, an XCA file has 28177", n_bytes);
    const unsigned char* g = xca + kMph + kSph;
    const uint32_t dsr = rd_be32(g + 12);
    if (dsr != 26552 && dsr != 0)
        return set_error(ASARF_EINVAL, "asarf_xca_scaling_factors: GADS DSR length %u, not 26552", dsr);
    for (int s = 0; s < 7; ++s) {
        slc_vv[s] = rd_bef(g + 16 + 28 + 4 * s);
        precision_vv[s] = rd_bef(g + 16 + 3 * 28 + 4 * s);
    }
    return clear_error();
}

extern "C" int asarf_calibration_constant(int detected, int ers, float process_gain, float scaling_factor,
                                          double* out) {
    // The product scaling of main.cc:336-337 (IMP: process_gain_imp / sqrt(scaling_factor_im_precision_vv)) and
    // main.cc:384-388 (IMS: process_gain_ims / sqrt(scaling_factor_im_slc_vv), divided by 2 * 4 for the ERS SAR
    // instrument), evaluated in FP64; it multiplies the statistics and the conditioned MDS records.
    if (!out) return set_error(ASARF_EINVAL, "asarf_calibration_constant: null");
    if (!(scaling_factor > 0.f))
        return set_error(ASARF_EINVAL, "asarf_calibration_constant: scaling factor %g is not positive",
                         (double)scaling_factor);
    double k = (double)process_gain / std::sqrt((double)scaling_factor);
    if (!detected && ers) k /= 2.0 * 4.0;
    *out = k;
    return clear_error();
}

extern "C" int asarf_ins_swath_params(const unsigned char* ins, long n_bytes, int swath_idx, asarf_ins_swath* out) {
    // FindINSFile (envisat_aux_file.cc:67-95) and the structs of envisat_aux_file.h:33-168 (big-endian, after the
    // 1625-byte MPH + SPH): FixedLifetimeParameters: radar frequency at 16, sampling rate at 20, the image-mode
    // nominal chirps im_chirp[7] (amplitude[4], phase[4], duration: 36 bytes each) at 65564, range gate bias at
    // 69912; FixedBaselineParameters (at 72608) holds mode_timelines[5] (swath_nums, m_values, r_values, g_values:
    // 7 uint16 each) at 88986, of which image mode is entry 0 (main_flow.cc:407-408).
    constexpr long kInsSize = 173273, kFlp = kMph + kSph, kFbp = kFlp + 72608;
    if (!ins || !out) return set_error(ASARF_EINVAL, "asarf_ins_swath_params: null");
    if (n_bytes != kInsSize)
        return set_error(ASARF_ESHAPE, "asarf_ins_swath_params: %ld bytes, an instrument file has 173273", n_bytes);
    if (swath_idx < 0 || swath_idx > 6) return set_error(ASARF_EINVAL, "asarf_ins_swath_params: swath %d", swath_idx);
    asarf_ins_swath r{};
    r.radar_frequency = rd_bef(ins + kFlp + 16);
    r.radar_sampling_rate = rd_bef(ins + kFlp + 20);
    const unsigned char* ch = ins + kFlp + 65564 + 36 * swath_idx;
    for (int k = 0; k < 4; ++k) {
        r.chirp_amplitude[k] = rd_bef(ch + 4 * k);
        r.chirp_phase[k] = rd_bef(ch + 16 + 4 * k);
    }
    r.chirp_duration = rd_bef(ch + 32);
    r.range_gate_bias = rd_bef(ins + kFlp + 69912);
    r.swst_rank = rd_be16(ins + kFbp + 88986 + 28 + 2 * swath_idx);
    *out = r;
    return clear_error();
}

extern "C" int asarf_radar_timing(const asarf_ins_swath* ins, int ers, int pri_code, int min_swst, int swst_mult,
                                  asarf_timing* out) {
    // The radar timing of the level-0 flow (main_flow.cc:389-460), for a window with no range recalibration:
    // chirp rate 2 x the instrument file's phase[2], duration and samples round(duration x fs); ASAR PRF = fs /
    // PRI code and two-way slant range time (min SWST + rank x PRI code) / fs - range gate bias; ERS PRI =
    // (PRI code + 2) x 210.94 ns and two-way time rank x PRI + min SWST x swst_mult / fs - range gate bias; slant
    // range to the first sample, wavelength and range spacing from c.
    if (!ins || !out) return set_error(ASARF_EINVAL, "asarf_radar_timing: null");
    if (!(ins->radar_sampling_rate > 0.f) || !(ins->radar_frequency > 0.f) || !std::isfinite(ins->radar_sampling_rate) ||
        !std::isfinite(ins->radar_frequency))
        return set_error(ASARF_EINVAL, "asarf_radar_timing: radar frequency or sampling rate not positive and finite");
    if (!std::isfinite(ins->chirp_phase[2]) || !std::isfinite(ins->range_gate_bias) ||
        !(ins->chirp_duration >= 0.f) || !((double)ins->chirp_duration * ins->radar_sampling_rate < 1e8))
        return set_error(ASARF_EINVAL, "asarf_radar_timing: chirp or range gate bias out of range");
    if (pri_code <= 0 || min_swst < 0 || swst_mult < 0)
        return set_error(ASARF_EINVAL, "asarf_radar_timing: PRI code %d, min SWST %d, SWST multiplier %d", pri_code,
                         min_swst, swst_mult);
    asarf_timing t{};
    const double fs = ins->radar_sampling_rate;
    t.chirp_kr = (double)(ins->chirp_phase[2] * 2.0f);
    t.chirp_duration = ins->chirp_duration;
    t.chirp_samples = (int)std::round(t.chirp_duration / (1.0 / fs));
    t.chirp_bandwidth = t.chirp_kr * t.chirp_duration;
    const double delay = ins->range_gate_bias;
    if (ers) {
        const double pri = (pri_code + 2.0) * 210.94e-9;
        t.prf = 1.0 / pri;
        t.two_way_slant_range_time = (double)ins->swst_rank * pri + (double)(min_swst * swst_mult) / fs - delay;
    } else {
        t.prf = fs / pri_code;
        t.two_way_slant_range_time =
            (double)((unsigned)min_swst + (unsigned)ins->swst_rank * (unsigned)pri_code) * (1.0 / fs) - delay;
    }
    t.line_time_interval = 1.0 / t.prf;
    t.slant_range_first = t.two_way_slant_range_time * kC / 2.0;
    t.wavelength = kC / (double)ins->radar_frequency;
    t.range_spacing = kC / (2.0 * fs);
    *out = t;
    return clear_error();
}
