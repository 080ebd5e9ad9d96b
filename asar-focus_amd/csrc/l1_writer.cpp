// Level-1 product header (host C++): the MPH, SPH with its 18 data set descriptors, and the annotation data sets
// (summary quality, main processing parameters, Doppler centroid, slant-to-ground, geolocation grid) that precede
// the measurement records of an ENVISAT-format IMS (SLC) or IMP (detected) file, as ConstructEnvisatFileHeader
// builds them (envisat_format/src/envisat_lvl1_writer.cc:326-570). The layouts are those of the ENVISAT product
// specification (PO-RS-MDA-GS-2009, ANNEX A data conventions): ASCII header fields "KEY=value<unit>\n" of fixed width,
// and big-endian binary annotation records. Widths and offsets are written out below, each block checked against its
// specified size; a field that does not fit its width is an error, not a silent truncation.
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

#include "asarf.h"
#include "internal.h"
#include "poly.h"

using namespace asarf;

#define ST_CHECK(expr)                     \
    do {                                   \
        const int st_ = (expr);            \
        if (st_ != ASARF_OK) return st_;   \
    } while (0)

namespace {

constexpr double kC = 299792458.0;
constexpr long kMphBytes = 1247, kSphBytes = 6099, kDsdBytes = 280;
constexpr long kSqBytes = 170, kMppBytes = 2009, kDcBytes = 55, kSrgrBytes = 55, kGeoBytes = 521;
constexpr int kGeoRecords = 12, kNumDsd = 18;

// ------------------------------------------------------------------------------------ ASCII header fields
struct Text {
    std::string buf;
    bool ok = true;
    std::string bad;  // first field that did not fit
    // one field of exactly `width` characters
    void put(const std::string& field, int width, const char* key) {
        if ((int)field.size() != width) {
            if (ok) bad = key;
            ok = false;
            buf += std::string((size_t)width, '?');
            return;
        }
        buf += field;
    }
    // quoted string value padded with spaces (or cut) to the width
    void str(int width, const char* key, const std::string& value) {
        std::string f = std::string(key) + "=\"";
        const int vw = width - (int)f.size() - 2;
        std::string v = value;
        v.resize(vw > 0 ? (size_t)vw : 0, ' ');
        put(f + v + "\"\n", width, key);
    }
    void chr(int width, const char* key, char value) { put(std::string(key) + "=" + value + "\n", width, key); }
    void num(int width, const char* key, const char* fmt, long long value, const char* unit = "") {
        char b[64];
        std::snprintf(b, sizeof(b), fmt, value);
        put(std::string(key) + "=" + b + unit + "\n", width, key);
    }
    void dbl(int width, const char* key, const char* fmt, double value, const char* unit = "") {
        char b[64];
        std::snprintf(b, sizeof(b), fmt, value);
        put(std::string(key) + "=" + b + unit + "\n", width, key);
    }
    // signed fraction without its integer digit, "+.123456" (the Ado06 type)
    void frac6(int width, const char* key, double value, const char* unit) {
        char b[64];
        std::snprintf(b, sizeof(b), "%+.6f", value);
        std::string s(b);
        if (s.size() >= 2) s.erase(1, 1);
        put(std::string(key) + "=" + s + unit + "\n", width, key);
    }
    void spare(int width) { put(std::string((size_t)(width - 1), ' ') + "\n", width, "spare"); }
};

// "%d-%b-%Y %H:%M:%S.%f" in upper case: 01-JAN-2004 21:55:28.000000
std::string mjd_text(const asarf_mjd& m) {
    static const char* months[12] = {"JAN", "FEB", "MAR", "APR", "MAY", "JUN", "JUL", "AUG", "SEP", "OCT", "NOV", "DEC"};
    // civil date from days since 2000-01-01 (Howard Hinnant's published civil_from_days method)
    long z = (long)m.days + 10957 + 719468;
    const long era = (z >= 0 ? z : z - 146096) / 146097;
    const long doe = z - era * 146097;
    const long yoe = (doe - doe / 1460 + doe / 36524 - doe / 146096) / 365;
    long y = yoe + era * 400;
    const long doy = doe - (365 * yoe + yoe / 4 - yoe / 100);
    const long mp = (5 * doy + 2) / 153;
    const long d = doy - (153 * mp + 2) / 5 + 1;
    const long mo = mp < 10 ? mp + 3 : mp - 9;
    y += mo <= 2;
    char b[64];
    std::snprintf(b, sizeof(b), "%02ld-%s-%04ld %02u:%02u:%02u.%06u", d, months[mo - 1], y, m.seconds / 3600,
                  (m.seconds / 60) % 60, m.seconds % 60, m.micros);
    return b;
}

// seconds since 2000-01-01 to mjd, truncated to whole microseconds
asarf_mjd mjd_of_seconds(double t) {
    const double us = std::floor(t * 1e6 + 0.5);  // the geometry keeps whole microseconds; round away the FP64 noise
    const long long u = (long long)us;
    long long days = u / 86400000000LL;
    long long rem = u % 86400000000LL;
    if (rem < 0) {
        rem += 86400000000LL;
        --days;
    }
    asarf_mjd m{};
    m.days = (int)days;
    m.seconds = (unsigned)(rem / 1000000);
    m.micros = (unsigned)(rem % 1000000);
    return m;
}

// ------------------------------------------------------------------------------------ binary records
struct Bin {
    std::vector<unsigned char> b;
    explicit Bin(size_t n) : b(n, 0) {}
    void u8(size_t o, unsigned v) { b[o] = (unsigned char)v; }
    void u16(size_t o, unsigned v) {
        b[o] = (unsigned char)(v >> 8);
        b[o + 1] = (unsigned char)v;
    }
    void u32(size_t o, uint32_t v) {
        for (int k = 0; k < 4; ++k) b[o + k] = (unsigned char)(v >> (24 - 8 * k));
    }
    void i32(size_t o, int32_t v) { u32(o, (uint32_t)v); }
    void f32(size_t o, float v) {
        uint32_t w;
        std::memcpy(&w, &v, 4);
        u32(o, w);
    }
    void mjd(size_t o, const asarf_mjd& m) {
        i32(o, m.days);
        u32(o + 4, m.seconds);
        u32(o + 8, m.micros);
    }
    // characters padded with spaces to n (CopyStrPad)
    void chars(size_t o, size_t n, const char* s) {
        const size_t l = s ? std::strlen(s) : 0;
        for (size_t k = 0; k < n; ++k) b[o + k] = (unsigned char)(k < l ? s[k] : ' ');
    }
};

// the value of a double assigned to a 32-bit integer field (truncation toward zero, saturated to the range)
int32_t to_i32(double v) {
    if (!(v == v)) return 0;
    if (v >= 2147483647.0) return 2147483647;
    if (v <= -2147483648.0) return (-2147483647 - 1);
    return (int32_t)v;
}

double line_time(const asarf_geometry* g, int idx) {
    const double us = std::trunc(g->line_time_interval * (double)idx * 1e6);
    return g->first_line_time + us * 1e-6;
}

bool fits(const char* s, size_t n) { return !s || std::strlen(s) <= n; }

}  // namespace

extern "C" int asarf_con_summary_quality(const unsigned char* con, long n_bytes, int detected,
                                         asarf_sq_thresholds* out) {
    // The summary-quality thresholds of the processor configuration file (envisat_lvl0_parser.cc:211-233 from
    // ConfigurationFile, envisat_aux_file.h:255-296): big-endian at record offsets 16-56 (chirp broadening,
    // sidelobe, ISLR, input mean, input std dev, Doppler centroid, Doppler ambiguity, output mean, output std dev,
    // missing lines, gaps), lines per gap at 124, expected IMS mean / std dev at 128 / 132, IMP at 144 / 148, expected
    // input mean / std dev at 192 / 196.
    if (!con || !out) return set_error(ASARF_EINVAL, "asarf_con_summary_quality: null");
    if (n_bytes != 5721)
        return set_error(ASARF_ESHAPE, "asarf_con_summary_quality: %ld bytes, a configuration file has 5721", n_bytes);
    const unsigned char* r = con + 1247 + 378;
    auto f = [&](long o) {
        const uint32_t w = ((uint32_t)r[o] << 24) | ((uint32_t)r[o + 1] << 16) | ((uint32_t)r[o + 2] << 8) | r[o + 3];
        float v;
        std::memcpy(&v, &w, 4);
        return v;
    };
    asarf_sq_thresholds t{};
    t.thresh_chirp_broadening = f(16);
    t.thresh_chirp_sidelobe = f(20);
    t.thresh_chirp_islr = f(24);
    t.thresh_input_mean = f(28);
    t.thresh_input_std_dev = f(32);
    t.thresh_dop_cen = f(36);
    t.thresh_dop_amb = f(40);
    t.thresh_output_mean = f(44);
    t.thresh_output_std_dev = f(48);
    t.thresh_input_missing_lines = f(52);
    t.thresh_input_gaps = f(56);
    t.lines_per_gap = ((uint32_t)r[124] << 24) | ((uint32_t)r[125] << 16) | ((uint32_t)r[126] << 8) | r[127];
    t.exp_output_mean = detected ? f(144) : f(128);
    t.exp_output_std_dev = detected ? f(148) : f(132);
    t.exp_input_mean = f(192);
    t.exp_input_std_dev = f(196);
    *out = t;
    return clear_error();
}

extern "C" long asarf_l1_header_size(int detected) {
    return kMphBytes + kSphBytes + kSqBytes + kMppBytes + kDcBytes + (detected ? kSrgrBytes : 0) +
           (long)kGeoRecords * kGeoBytes;
}

extern "C" int asarf_l1_header(const asarf_l1_inputs* in, unsigned char* out, long capacity, long* n_out) {
    if (!in || !out || !n_out || !in->geometry || !in->l0_product_name)
        return set_error(ASARF_EINVAL, "asarf_l1_header: null argument");
    *n_out = 0;
    const asarf_geometry* g = in->geometry;
    const bool detected = in->detected != 0, slc = !detected;
    const long need = asarf_l1_header_size(in->detected);
    if (capacity < need) return set_error(ASARF_EWORKSPACE, "asarf_l1_header: %ld bytes needed, %ld given", need, capacity);
    const std::string l0(in->l0_product_name);
    if (l0.size() < 10) return set_error(ASARF_EINVAL, "asarf_l1_header: product name '%s' too short", l0.c_str());
    if (!in->swath || std::strlen(in->swath) != 3) return set_error(ASARF_EINVAL, "asarf_l1_header: swath must be 3 characters");
    if (g->azimuth_size < 11 || g->range_size < 11 || g->n_osv < 2 || !g->osv)
        return set_error(ASARF_ESHAPE, "asarf_l1_header: azimuth %d, range %d, %d state vectors", g->azimuth_size,
                         g->range_size, g->n_osv);
    const int pix = slc ? 4 : 2;
    if ((long)g->range_size * pix != (long)in->mds_record_size - 17)
        return set_error(ASARF_ESHAPE, "asarf_l1_header: range size %d x %d bytes does not match MDS record size %d - 17",
                         g->range_size, pix, in->mds_record_size);
    if (detected && !in->grsr5) return set_error(ASARF_EINVAL, "asarf_l1_header: a detected product needs grsr5");
    if (!(in->range_sampling_rate > 0)) return set_error(ASARF_EINVAL, "asarf_l1_header: range sampling rate");
    if (in->up_code > 15) return set_error(ASARF_EINVAL, "asarf_l1_header: upconverter code %u > 15", in->up_code);
    const char* comp[8] = {in->echo_comp_method, in->echo_comp_ratio, in->init_cal_method, in->init_cal_ratio,
                           in->per_cal_method,   in->per_cal_ratio,   in->noise_method,    in->noise_ratio};
    const size_t comp_w[8] = {4, 3, 4, 3, 4, 3, 4, 3};
    for (int k = 0; k < 8; ++k)
        if (!fits(comp[k], comp_w[k])) return set_error(ASARF_EINVAL, "asarf_l1_header: compression field %d too long", k);

    // ---- MPH (envisat_lvl1_mph.h:33-146)
    std::string name = l0;
    name[6] = slc ? 'S' : 'P';
    name[8] = '1';
    const long n_records = in->mds_records, rec = in->mds_record_size;
    const int n_datasets = detected ? 6 : 5;
    const long long total = (long long)need + (long long)n_records * rec;
    const char* ref_doc = l0.compare(0, 3, "ASA") == 0   ? "PO-RS-MDA-GS-2009_4/C"
                          : l0.compare(0, 3, "SAR") == 0 ? "PX-SP-50-9105_3/1"
                                                         : "TBD.pdf";
    const asarf_doris_meta* o = in->orbit;
    asarf_doris_meta no_orbit{};
    if (!o) {
        std::strcpy(no_orbit.vector_source, "  ");
        no_orbit.phase = ' ';
        o = &no_orbit;
    }
    Text mph;
    mph.str(73, "PRODUCT", name);
    mph.chr(13, "PROC_STAGE", 'N');
    mph.str(34, "REF_DOC", ref_doc);
    mph.spare(41);
    mph.str(43, "ACQUISITION_STATION", in->acquisition_station ? in->acquisition_station : "");
    mph.str(21, "PROC_CENTER", in->processing_station ? in->processing_station : "");
    mph.str(40, "PROC_TIME", mjd_text(in->processing_time));
    mph.str(30, "SOFTWARE_VER", in->software_version ? in->software_version : "");
    mph.spare(41);
    mph.str(44, "SENSING_START", mjd_text(in->sensing_start));
    mph.str(43, "SENSING_STOP", mjd_text(in->sensing_stop));
    mph.spare(41);
    mph.chr(8, "PHASE", o->phase ? o->phase : ' ');
    mph.num(11, "CYCLE", "%+04lld", (long long)o->cycle);
    mph.num(17, "REL_ORBIT", "%+06lld", (long long)in->rel_orbit);
    mph.num(17, "ABS_ORBIT", "%+06lld", (long long)o->abs_orbit);
    mph.str(48, "STATE_VECTOR_TIME", mjd_text(o->state_vector_time));
    mph.frac6(22, "DELTA_UT1", o->delta_ut1, "<s>");
    mph.dbl(27, "X_POSITION", "%+012.3f", o->pos[0], "<m>");
    mph.dbl(27, "Y_POSITION", "%+012.3f", o->pos[1], "<m>");
    mph.dbl(27, "Z_POSITION", "%+012.3f", o->pos[2], "<m>");
    mph.dbl(29, "X_VELOCITY", "%+012.6f", o->vel[0], "<m/s>");
    mph.dbl(29, "Y_VELOCITY", "%+012.6f", o->vel[1], "<m/s>");
    mph.dbl(29, "Z_VELOCITY", "%+012.6f", o->vel[2], "<m/s>");
    mph.str(19, "VECTOR_SOURCE", o->vector_source);
    mph.spare(41);
    mph.str(43, "UTC_SBT_TIME", mjd_text(in->utc_sbt_time));
    mph.num(28, "SAT_BINARY_TIME", "%+011lld", (long long)in->sat_binary_time);
    mph.num(27, "CLOCK_STEP", "%+011lld", (long long)in->clock_step, "<ps>");
    mph.spare(33);
    mph.str(39, "LEAP_UTC", mjd_text(in->leap_utc));
    mph.num(15, "LEAP_SIGN", "%+04lld", (long long)in->leap_sign);
    mph.chr(11, "LEAP_ERR", in->leap_err ? in->leap_err : '0');
    mph.spare(41);
    mph.chr(14, "PRODUCT_ERR", in->product_err ? '1' : '0');
    mph.num(38, "TOT_SIZE", "%+021lld", total, "<bytes>");
    mph.num(28, "SPH_SIZE", "%+011lld", (long long)kSphBytes, "<bytes>");
    mph.num(20, "NUM_DSD", "%+011lld", (long long)kNumDsd);
    mph.num(28, "DSD_SIZE", "%+011lld", (long long)kDsdBytes, "<bytes>");
    mph.num(26, "NUM_DATA_SETS", "%+011lld", (long long)n_datasets);
    mph.spare(41);
    if (!mph.ok || (long)mph.buf.size() != kMphBytes)
        return set_error(ASARF_EINVAL, "asarf_l1_header: MPH field %s does not fit (%zu bytes)", mph.bad.c_str(),
                         mph.buf.size());

    // ---- SPH (envisat_lvl1_sph.h:89-209) with the scene corners from the first and last lines
    double ll_first[3][2], ll_last[3][2];
    {
        const double r_near = g->slant_range_first;
        const double r_mid = r_near + g->range_spacing * g->range_size / 2.0;
        const double r_far = r_near + g->range_spacing * g->range_size;
        const double rs[3] = {r_near, r_mid, r_far};
        const double t_last = (double)in->last_line_time.days * 86400.0 + in->last_line_time.seconds +
                              1e-6 * in->last_line_time.micros;
        const double times[2] = {g->first_line_time, t_last};
        for (int w = 0; w < 2; ++w) {
            asarf_osv s{};
            ST_CHECK(asarf_interpolate_orbit(g->osv, g->n_osv, times[w], &s));
            for (int k = 0; k < 3; ++k) {
                double xyz[3];
                ST_CHECK(asarf_geolocate(s.pos, s.vel, g->center_point, rs[k], xyz));
                ST_CHECK(asarf_xyz_to_geo(xyz, w == 0 ? ll_first[k] : ll_last[k]));
            }
        }
    }
    Text sph;
    sph.str(46, "SPH_DESCRIPTOR", slc ? "Image Mode SLC Image" : "Image Mode Precision Image");
    sph.num(36, "STRIPLINE_CONTINUITY_INDICATOR", "%+04lld", 0);
    sph.num(20, "SLICE_POSITION", "%+04lld", 1);
    sph.num(16, "NUM_SLICES", "%+04lld", 1);
    sph.str(46, "FIRST_LINE_TIME", mjd_text(in->first_line_time));
    sph.str(45, "LAST_LINE_TIME", mjd_text(in->last_line_time));
    {
        const char* keys[2][6] = {{"FIRST_NEAR_LAT", "FIRST_NEAR_LONG", "FIRST_MID_LAT", "FIRST_MID_LONG", "FIRST_FAR_LAT",
                                   "FIRST_FAR_LONG"},
                                  {"LAST_NEAR_LAT", "LAST_NEAR_LONG", "LAST_MID_LAT", "LAST_MID_LONG", "LAST_FAR_LAT",
                                   "LAST_FAR_LONG"}};
        const int widths[2][6] = {{37, 38, 36, 37, 36, 37}, {36, 37, 35, 36, 35, 36}};
        for (int w = 0; w < 2; ++w)
            for (int k = 0; k < 3; ++k) {
                const double(*ll)[2] = w == 0 ? ll_first : ll_last;
                sph.num(widths[w][2 * k], keys[w][2 * k], "%+011lld", (long long)to_i32(1e6 * ll[k][0]), "<10-6degN>");
                sph.num(widths[w][2 * k + 1], keys[w][2 * k + 1], "%+011lld", (long long)to_i32(1e6 * ll[k][1]),
                        "<10-6degE>");
            }
    }
    sph.spare(36);
    sph.str(12, "SWATH", in->swath);
    sph.str(18, "PASS", in->ascending ? "ASCENDING" : "DESCENDING");
    sph.str(23, "SAMPLE_TYPE", slc ? "COMPLEX" : "DETECTED");
    sph.str(20, "ALGORITHM", "RAN/DOP");
    sph.str(23, "MDS1_TX_RX_POLAR", in->polarization ? in->polarization : "");
    sph.str(23, "MDS2_TX_RX_POLAR", "");
    {
        std::string c = in->echo_comp_method ? in->echo_comp_method : "";
        if (in->echo_comp_ratio && in->echo_comp_ratio[0]) c += in->echo_comp_ratio[0];
        sph.str(20, "COMPRESSION", c);
    }
    sph.num(19, "AZIMUTH_LOOKS", "%+04lld", slc ? 1 : 4);
    sph.num(17, "RANGE_LOOKS", "%+04lld", 1);
    sph.dbl(33, "RANGE_SPACING", "%+.8E", (double)(float)g->range_spacing, "<m>");
    sph.dbl(35, "AZIMUTH_SPACING", "%+.8E", (double)(float)in->azimuth_spacing, "<m>");
    sph.dbl(38, "LINE_TIME_INTERVAL", "%+.8E", (double)(float)g->line_time_interval, "<s>");
    sph.num(28, "LINE_LENGTH", "%+06lld", (long long)g->range_size, "<samples>");
    sph.str(18, "DATA_TYPE", slc ? "SWORD" : "UWORD");
    sph.spare(51);
    // data set descriptors (envisat_lvl1_sph.h:27-85) and the annotation offsets they give
    long off = kMphBytes + kSphBytes;
    auto dsd = [&](const char* ds_name, char type, const char* file, long offset, long n, long size) {
        sph.str(39, "DS_NAME", ds_name);
        sph.chr(10, "DS_TYPE", type);
        sph.str(74, "FILENAME", file ? file : "");
        sph.num(39, "DS_OFFSET", "%+021lld", (long long)offset, "<bytes>");
        sph.num(37, "DS_SIZE", "%+021lld", (long long)n * size, "<bytes>");
        sph.num(20, "NUM_DSR", "%+011lld", (long long)n);
        sph.num(28, "DSR_SIZE", "%+011lld", (long long)size, "<bytes>");
        sph.spare(33);
    };
    auto internal = [&](const char* ds_name, char type, long n, long size) {
        dsd(ds_name, type, "", off, n, size);
        const long at = off;
        off += n * size;
        return at;
    };
    auto empty = [&](const char* ds_name, char type) { dsd(ds_name, type, "NOT USED", 0, 0, 0); };
    auto reference = [&](const char* ds_name, const char* file) { dsd(ds_name, 'R', file, 0, 0, 0); };
    const long off_sq = internal("MDS1 SQ ADS", 'A', 1, kSqBytes);
    empty("MDS2 SQ ADS", 'A');
    const long off_mpp = internal("MAIN PROCESSING PARAMS ADS", 'A', 1, kMppBytes);
    const long off_dc = internal("DOP CENTROID COEFFS ADS", 'A', 1, kDcBytes);
    long off_srgr = -1;
    if (slc) empty("SR GR ADS", 'A');
    else off_srgr = internal("SR GR ADS", 'A', 1, kSrgrBytes);
    empty("CHIRP PARAMS ADS", 'A');
    empty("MDS1 ANTENNA ELEV PATT ADS", 'A');
    empty("MDS2 ANTENNA ELEV PATT ADS", 'A');
    const long off_geo = internal("GEOLOCATION GRID ADS", 'A', kGeoRecords, kGeoBytes);
    empty("MAP PROJECTION GADS", 'G');
    internal("MDS1", 'M', n_records, rec);
    empty("MDS2", 'M');
    reference("LEVEL 0 PRODUCT", in->l0_product_name);
    reference("ASAR PROCESSOR CONFIG", in->configuration_file);
    reference("INSTRUMENT CHARACTERIZATION", in->instrument_file);
    if (!in->external_characterization_file || !in->external_characterization_file[0])
        empty("EXTERNAL CHARACTERIZATION", 'R');
    else
        reference("EXTERNAL_CHARACTERIZATION", in->external_characterization_file);
    reference("EXTERNAL CALIBRATION", in->external_calibration_file);
    reference("ORBIT STATE VECTOR 1", in->orbit_dataset_name);
    if (!sph.ok || (long)sph.buf.size() != kSphBytes)
        return set_error(ASARF_EINVAL, "asarf_l1_header: SPH field %s does not fit (%zu bytes)", sph.bad.c_str(),
                         sph.buf.size());

    const double dt = 1.0 / in->range_sampling_rate;
    const asarf_mjd first_zd = mjd_of_seconds(g->first_line_time);

    // ---- summary quality ADS (envisat_lvl1_ads.h:485-558, envisat_lvl1_writer.cc:204-270)
    Bin sq(kSqBytes);
    {
        const asarf_sq_thresholds& t = in->sq;
        sq.mjd(0, first_zd);
        const float in_mean[2] = {(float)in->iq.dc_i, (float)in->iq.dc_q};
        const float in_std[2] = {(float)in->iq.stddev_i, (float)in->iq.stddev_q};
        const float out_mean[2] = {(float)in->statistics[0], (float)in->statistics[1]};
        const float out_std[2] = {(float)in->statistics[2], (float)in->statistics[3]};
        auto flag = [&](const float* v, float expected, float thresh) {
            bool f = std::fabs(v[0] - expected) > thresh;
            if (slc) f = f || std::fabs(v[1] - expected) > thresh;
            return f ? 1u : 0u;
        };
        sq.u8(13, flag(in_mean, t.exp_input_mean, t.thresh_input_mean));
        sq.u8(14, flag(in_std, t.exp_input_std_dev, t.thresh_input_std_dev));
        sq.u8(19, flag(out_mean, t.exp_output_mean, t.thresh_output_mean));
        sq.u8(20, flag(out_std, t.exp_output_std_dev, t.thresh_output_std_dev));
        const float th[16] = {t.thresh_chirp_broadening, t.thresh_chirp_sidelobe, t.thresh_chirp_islr,
                              t.thresh_input_mean,       t.exp_input_mean,        t.thresh_input_std_dev,
                              t.exp_input_std_dev,       t.thresh_dop_cen,        t.thresh_dop_amb,
                              t.thresh_output_mean,      t.exp_output_mean,       t.thresh_output_std_dev,
                              t.exp_output_std_dev,      t.thresh_input_missing_lines, t.thresh_input_gaps, 0.f};
        for (int k = 0; k < 15; ++k) sq.f32(31 + 4 * k, th[k]);
        sq.u32(91, t.lines_per_gap);
        for (int k = 0; k < 2; ++k) {
            sq.f32(110 + 4 * k, in_mean[k]);
            sq.f32(118 + 4 * k, in_std[k]);
            sq.f32(134 + 4 * k, out_mean[k]);
            sq.f32(142 + 4 * k, out_std[k]);
        }
        sq.chars(154, 3, in->swath);
    }

    // ---- main processing parameters ADS (envisat_lvl1_ads.h:31-481, envisat_lvl1_writer.cc:36-203)
    Bin mp(kMppBytes);
    {
        // general summary (0): defaults, then the product's values
        mp.mjd(0, in->sensing_start);
        mp.mjd(13, in->sensing_stop);
        mp.chars(25, 12, "123");
        mp.chars(41, 3, in->swath);
        mp.f32(44, (float)g->range_spacing);
        mp.f32(48, (float)in->azimuth_spacing);
        mp.f32(52, (float)g->line_time_interval);
        mp.u32(56, (uint32_t)g->azimuth_size);
        mp.u32(60, (uint32_t)g->range_size);
        mp.chars(64, 5, slc ? "SWORD" : "UWORD");
        // image processing summary (120): Doppler centroid, SRGR, look summing, detection, spreading loss flags
        mp.u8(120 + 4, 1);
        mp.u8(120 + 3, in->grsr5 ? 1 : 0);
        mp.u8(120 + 6, slc ? 0 : 1);
        mp.u8(120 + 7, slc ? 0 : 1);
        mp.u8(120 + 8, slc ? 0 : 1);
        // raw data analysis [0] (141): calculated biases, gain and quadrature departure
        mp.f32(141 + 16, (float)in->iq.dc_i);
        mp.f32(141 + 20, (float)in->iq.dc_q);
        mp.f32(141 + 32, (float)in->iq.iq_gain);
        mp.f32(141 + 36, (float)in->iq.phase_error);
        // information from the downlink header (357)
        const size_t dl = 357;
        mp.u32(dl + 0, in->first_sbt >> 8);
        mp.u32(dl + 4, in->first_sbt & 0xFF);
        mp.mjd(dl + 8, in->first_mjd);
        mp.u16(dl + 40, in->first_swst_code);
        mp.u16(dl + 50, in->last_swst_code);
        mp.u16(dl + 60, in->pri_code);
        mp.u16(dl + 70, in->tx_pulse_len_code);
        mp.u16(dl + 80, in->tx_bw_code);
        mp.u16(dl + 90, in->echo_win_len_code);
        mp.u16(dl + 100, in->up_code);
        mp.u16(dl + 110, in->down_code);
        mp.u16(dl + 120, in->resamp_code);
        mp.u16(dl + 130, in->beam_adj_code);
        mp.u16(dl + 140, in->beam_set_num_code);
        mp.u16(dl + 150, in->tx_monitor_code);
        // values: the last SWST value repeats the first SWST code, as envisat_lvl1_writer.cc:90-91 writes it
        mp.f32(dl + 286, (float)(in->first_swst_code * dt));
        mp.f32(dl + 306, (float)(in->first_swst_code * dt));
        mp.u32(dl + 326, in->swst_changes);
        mp.f32(dl + 346, (float)g->prf);
        mp.f32(dl + 366, (float)(in->tx_pulse_len_code * dt));
        const float tx_bw = (float)((double)in->tx_bw_code / 255 * 16e6);
        mp.f32(dl + 386, tx_bw);
        mp.f32(dl + 406, (float)(in->echo_win_len_code * dt));
        // upconverter level (PO-TN-MMS-SR-0248: 25.1625 dB + 0.52 dB per code step, codes 0-15)
        mp.f32(dl + 426, (float)(25.1625 + 0.52 * in->up_code));
        mp.f32(dl + 516, in->tx_monitor_value);
        mp.u32(dl + 536, in->swst_rank);
        // range processing (975)
        const size_t rp = 975;
        mp.f32(rp + 4, 800000.0f);
        mp.f32(rp + 8, (float)in->range_sampling_rate);
        mp.f32(rp + 12, (float)g->carrier_frequency);
        mp.u16(rp + 16, 1);
        mp.chars(rp + 18, 7, in->chirp_window ? "HAMMING" : "NONE");
        mp.f32(rp + 25, in->chirp_window ? (float)in->chirp_alpha : 0.f);
        mp.f32(rp + 29, tx_bw);
        mp.f32(rp + 49, tx_bw);
        mp.f32(rp + 69, 1.0f);                          // nominal chirp [0] amplitude [0]
        mp.f32(rp + 69 + 16 + 8, (float)(in->chirp_kr / 2.0));  // nominal chirp [0] phase [2]
        // azimuth processing (1264): looks, bandwidth, window, FM rate polynomial over two-way slant time
        const size_t ap = 1264;
        mp.u16(ap + 4, 1);
        mp.f32(ap + 10, (float)(g->prf * g->az_bandwidth_fraction));
        mp.chars(ap + 14, 7, in->chirp_window ? "HAMMING" : "NONE");
        mp.f32(ap + 21, in->chirp_window ? (float)in->azimuth_window_alpha : 0.f);
        {
            double ka[10], tt[10], c3[3];
            const double lambda = kC / g->carrier_frequency;
            for (int i = 0; i < 10; ++i) {
                const int idx = (g->range_size * i) / 10;
                const double vr = polyval(g->vr_poly, 3, (double)idx);
                const double r0 = g->slant_range_first + (double)idx * g->range_spacing;
                ka[i] = -2.0 * (vr * vr) / (lambda * r0);
                tt[i] = 2.0 * (r0 / kC) - in->two_way_slant_range_time;
            }
            if (polyfit_qr(tt, ka, 10, 2, c3) != 0)
                return set_error(ASARF_EINVAL, "asarf_l1_header: azimuth FM rate fit failed");
            mp.f32(ap + 25, (float)c3[2]);
            mp.f32(ap + 29, (float)c3[1]);
            mp.f32(ap + 33, (float)c3[0]);
            mp.f32(ap + 37, (float)(in->two_way_slant_range_time * 1e9));
        }
        // data compression (1593)
        const size_t dc = 1593;
        const size_t comp_off[8] = {0, 4, 7, 11, 14, 18, 21, 25};
        const char* comp_def[8] = {"NONE", "", "NONE", "", "NONE", "", "NONE", ""};
        for (int k = 0; k < 8; ++k) mp.chars(dc + comp_off[k], comp_w[k], comp[k] ? comp[k] : comp_def[k]);
        // orbit state vectors (1765): at lines 0, 1/4, ... of the image; cm and 1e-5 m/s
        for (int i = 0; i < 5; ++i) {
            const int idx = (g->azimuth_size * i) / 4;
            const double t = line_time(g, idx);
            asarf_osv s{};
            ST_CHECK(asarf_interpolate_orbit(g->osv, g->n_osv, t, &s));
            const size_t ov = 1765 + 36 * (size_t)i;
            mp.mjd(ov, mjd_of_seconds(t));
            for (int k = 0; k < 3; ++k) {
                mp.i32(ov + 12 + 4 * k, to_i32(s.pos[k] * 1e2));
                mp.i32(ov + 24 + 4 * k, to_i32(s.vel[k] * 1e5));
            }
        }
    }

    // ---- Doppler centroid ADS (envisat_lvl1_ads.h:562-581, envisat_lvl1_writer.cc:469-489): the DC polynomial
    // over the range index converted to two-way slant range time from the first sample
    Bin dcb(kDcBytes);
    dcb.mjd(0, first_zd);
    dcb.f32(13, (float)(in->two_way_slant_range_time * 1e9));
    dcb.f32(17, (float)g->dc_poly[2]);
    dcb.f32(21, (float)(g->dc_poly[1] / dt));
    dcb.f32(25, (float)(g->dc_poly[0] / (dt * dt)));

    // ---- slant-to-ground ADS (envisat_lvl1_ads.h:690-706, envisat_lvl1_writer.cc:491-508), detected products
    Bin sr(kSrgrBytes);
    if (detected) {
        sr.mjd(0, first_zd);
        sr.f32(13, (float)(in->two_way_slant_range_time * 1e9));
        sr.f32(17, 0.f);
        sr.f32(21, (float)(in->grsr5[4] + g->slant_range_first));
        sr.f32(25, (float)in->grsr5[3]);
        sr.f32(29, (float)in->grsr5[2]);
        sr.f32(33, (float)in->grsr5[1]);
        sr.f32(37, (float)in->grsr5[0]);
    }

    // ---- geolocation grid ADS (envisat_lvl1_ads.h:640-686, envisat_lvl1_writer.cc:523-536, 277-324)
    Bin geo((size_t)kGeoRecords * kGeoBytes);
    {
        const int step = g->azimuth_size / (kGeoRecords - 1);
        for (int i = 0; i < kGeoRecords; ++i) {
            const int first = i * g->azimuth_size / (kGeoRecords - 1);
            const int last = first + step - 1;
            asarf_tie_points tp[2];
            ST_CHECK(asarf_geolocation_tie_points(g, first, in->grsr5, &tp[0]));
            ST_CHECK(asarf_geolocation_tie_points(g, last, in->grsr5, &tp[1]));
            const size_t b = (size_t)i * kGeoBytes;
            geo.mjd(b, mjd_of_seconds(line_time(g, first)));
            geo.u32(b + 13, (uint32_t)first);
            geo.u32(b + 17, (uint32_t)(last - first));
            geo.f32(b + 21, 0.f);
            for (int w = 0; w < 2; ++w) {
                const size_t t0 = b + (w == 0 ? 25 : 279);
                if (w == 1) geo.mjd(b + 267, mjd_of_seconds(line_time(g, last)));
                for (int k = 0; k < 11; ++k) {
                    geo.u32(t0 + 4 * k, tp[w].samp_numbers[k]);
                    geo.f32(t0 + 44 + 4 * k, tp[w].slant_range_times[k]);
                    geo.f32(t0 + 88 + 4 * k, tp[w].angles[k]);
                    geo.i32(t0 + 132 + 4 * k, tp[w].lats[k]);
                    geo.i32(t0 + 176 + 4 * k, tp[w].longs[k]);
                }
            }
        }
    }

    // ---- assemble: MPH, SPH, SQ, MPP, DC, [SRGR], geolocation grid
    long pos = 0;
    auto append = [&](const void* p, size_t n) {
        std::memcpy(out + pos, p, n);
        pos += (long)n;
    };
    append(mph.buf.data(), mph.buf.size());
    append(sph.buf.data(), sph.buf.size());
    if (pos != off_sq) return set_error(ASARF_EINVAL, "asarf_l1_header: SQ offset");
    append(sq.b.data(), sq.b.size());
    if (pos != off_mpp) return set_error(ASARF_EINVAL, "asarf_l1_header: MPP offset");
    append(mp.b.data(), mp.b.size());
    if (pos != off_dc) return set_error(ASARF_EINVAL, "asarf_l1_header: DC offset");
    append(dcb.b.data(), dcb.b.size());
    if (detected) {
        if (pos != off_srgr) return set_error(ASARF_EINVAL, "asarf_l1_header: SRGR offset");
        append(sr.b.data(), sr.b.size());
    }
    if (pos != off_geo) return set_error(ASARF_EINVAL, "asarf_l1_header: geolocation offset");
    append(geo.b.data(), geo.b.size());
    if (pos != need) return set_error(ASARF_EINVAL, "asarf_l1_header: %ld bytes written, %ld expected", pos, need);
    *n_out = pos;
    return clear_error();
}
