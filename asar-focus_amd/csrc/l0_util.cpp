// Level-0 host utilities (the reference's envisat_format helpers that its unit tests pin): MJD time arithmetic,
// packet-counter gaps, the ERS satellite-binary-time (SBT) correction from a PATC record, and product-type
// identification. Host-only integer work; the packets they describe are decoded on the device by k_ingest.
#include <cctype>
#include <cstdint>
#include <algorithm>
#include <cstring>
#include <vector>

#include "asarf.h"
#include "internal.h"

using namespace asarf;

namespace {

constexpr int64_t kMicrosPerDay = 86400LL * 1000000LL;
// days from 1950-01-01 to 2000-01-01 (the MJD2000 epoch): 50 years with the 12 leap days 1952 ... 1996
constexpr int64_t kDays1950To2000 = 18262;
// days from 1950-01-01 to 1991-01-01; a PATC epoch before that is rejected (ers_sbt.cc:48-54)
constexpr int32_t kMinPatcDays1950 = 14975;

// Day counts beyond +-1e8 (about 274,000 years) are not times of any product: entry points that take an mjd from
// the caller reject them, and days decoded from packet bytes are clamped here, so the microsecond arithmetic
// below cannot overflow.
constexpr int64_t kMaxMjdDays = 100000000;

int64_t mjd_to_micros(const asarf_mjd& m) {
    const int64_t days = std::min<int64_t>(std::max<int64_t>(m.days, -kMaxMjdDays), kMaxMjdDays);
    return days * kMicrosPerDay + (int64_t)m.seconds * 1000000LL + (int64_t)m.micros;
}

// PtimeToMjd (envisat_utils.h:50-59): days = floor(t / day) from 2000-01-01, then whole seconds of the day and the
// microseconds within the second
asarf_mjd micros_to_mjd(int64_t us) {
    int64_t days = us / kMicrosPerDay;
    int64_t rem = us % kMicrosPerDay;
    if (rem < 0) {
        rem += kMicrosPerDay;
        --days;
    }
    asarf_mjd r;
    r.days = (int32_t)days;
    r.seconds = (uint32_t)(rem / 1000000LL);
    r.micros = (uint32_t)(rem % 1000000LL);
    return r;
}

// std::stoul on a fixed-width ASCII field: optional leading blanks, then at least one decimal digit
bool ascii_field(const unsigned char* p, int n, uint32_t* out) {
    int i = 0;
    while (i < n && std::isspace(p[i])) ++i;
    if (i == n || !std::isdigit(p[i])) return false;
    uint64_t v = 0;
    for (; i < n && std::isdigit(p[i]); ++i) v = v * 10 + (uint64_t)(p[i] - '0');
    *out = (uint32_t)v;
    return true;
}

bool iequals(const char* a, const char* b) {
    for (; *a && *b; ++a, ++b)
        if (std::tolower((unsigned char)*a) != std::tolower((unsigned char)*b)) return false;
    return *a == *b;
}

}  // namespace

extern "C" int asarf_mjd_compare(const asarf_mjd* a, const asarf_mjd* b) {
    // mjd ordering (envisat_types.h:42-73): lexicographic on (days, seconds, micros)
    if (!a || !b) {
        set_error(ASARF_EINVAL, "asarf_mjd_compare: null");
        return 0;
    }
    if (a->days != b->days) return a->days < b->days ? -1 : 1;
    if (a->seconds != b->seconds) return a->seconds < b->seconds ? -1 : 1;
    if (a->micros != b->micros) return a->micros < b->micros ? -1 : 1;
    return 0;
}

extern "C" int asarf_mjd_add_micros(const asarf_mjd* in, long long micros, asarf_mjd* out) {
    // MjdAddMicros (envisat_utils.h:74-78)
    if (!in || !out) return set_error(ASARF_EINVAL, "asarf_mjd_add_micros: null");
    if (in->days > kMaxMjdDays || in->days < -kMaxMjdDays)
        return set_error(ASARF_EINVAL, "asarf_mjd_add_micros: %d days is out of range", in->days);
    int64_t us = 0;
    if (__builtin_add_overflow(mjd_to_micros(*in), (int64_t)micros, &us) || us / kMicrosPerDay > kMaxMjdDays ||
        us / kMicrosPerDay < -kMaxMjdDays)
        return set_error(ASARF_EINVAL, "asarf_mjd_add_micros: result out of range");
    *out = micros_to_mjd(us);
    return clear_error();
}

extern "C" unsigned long long asarf_counter_gap(unsigned long long last, unsigned long long current,
                                                unsigned long long limit) {
    // CounterGap (parse_util.h:33-45): steps from last to current of a counter that wraps after limit
    // (limit = 2^bits - 1 for a bits-wide field)
    if (last + 1 == current) return 1;
    if (current < last) return limit - last + current + 1;
    return current - last;
}

extern "C" int asarf_parse_patc(const unsigned char* bytes, long n_bytes, asarf_patc* out) {
    // ParsePatc / ConstructPatc (patc.cc:25-57): a 54-byte record; ASCII cyclic counter at 15 (4 characters),
    // mission at 20-21, ASCII orbit number at 30 (5 characters), then little-endian int32 days since 1950 at 38,
    // int32 milliseconds at 42, uint32 SBT counter at 46 and uint32 SBT period [ns] at 50
    if (!bytes || !out) return set_error(ASARF_EINVAL, "asarf_parse_patc: null");
    if (n_bytes != 54) return set_error(ASARF_EINVAL, "asarf_parse_patc: record is %ld bytes, not 54", n_bytes);
    asarf_patc p{};
    uint32_t cc = 0, orbit = 0;
    if (!ascii_field(bytes + 15, 4, &cc) || !ascii_field(bytes + 30, 5, &orbit))
        return set_error(ASARF_EINVAL, "asarf_parse_patc: counter or orbit field is not a number");
    p.cyclic_counter = (uint16_t)cc;
    p.mission[0] = (char)bytes[20];
    p.mission[1] = (char)bytes[21];
    p.orbit_number = orbit;
    auto le32 = [&](int at) {
        return (uint32_t)bytes[at] | ((uint32_t)bytes[at + 1] << 8) | ((uint32_t)bytes[at + 2] << 16) |
               ((uint32_t)bytes[at + 3] << 24);
    };
    p.epoch_1950 = (int32_t)le32(38);
    p.milliseconds = (int32_t)le32(42);
    p.sbt_counter = le32(46);
    p.sbt_period = le32(50);
    *out = p;
    return clear_error();
}

extern "C" int asarf_patc_register(const asarf_patc* patc, asarf_patc_reference* out) {
    // RegisterPatc (ers_sbt.cc:40-63): validates the PATC epoch and keeps its time point (days since 1950 plus
    // milliseconds) with the SBT counter and period. The reference keeps this in a process-wide variable; here the
    // caller holds it.
    if (!patc || !out) return set_error(ASARF_EINVAL, "asarf_patc_register: null");
    if (patc->epoch_1950 < 0) return set_error(ASARF_EINVAL, "asarf_patc_register: negative epoch days since 1950");
    if (patc->milliseconds < 0) return set_error(ASARF_EINVAL, "asarf_patc_register: negative epoch milliseconds");
    if (patc->epoch_1950 < kMinPatcDays1950)
        return set_error(ASARF_EINVAL, "asarf_patc_register: epoch %d days after 1950-01-01 is before 1991",
                         patc->epoch_1950);
    if (patc->epoch_1950 > kMaxMjdDays)
        return set_error(ASARF_EINVAL, "asarf_patc_register: epoch %d days after 1950-01-01 is out of range",
                         patc->epoch_1950);
    out->time_us = ((int64_t)patc->epoch_1950 - kDays1950To2000) * kMicrosPerDay + (int64_t)patc->milliseconds * 1000;
    out->sbt_counter = patc->sbt_counter;
    out->sbt_period_ns = patc->sbt_period;
    return clear_error();
}

extern "C" int asarf_adjust_isp_sensing_time(const asarf_patc_reference* ref, asarf_mjd* isp_time, unsigned sbt,
                                             int overflow) {
    // AdjustIspSensingTime (ers_sbt.cc:65-86): without a registered PATC the time is left as it is; otherwise it
    // becomes the PATC time plus (SBT ticks since the PATC counter) x period, in whole microseconds truncated
    // toward zero. A counter below the PATC's counts backwards unless overflow is set, in which case it is taken
    // as wrapped past 2^32 - 1
    if (!isp_time) return set_error(ASARF_EINVAL, "asarf_adjust_isp_sensing_time: null");
    if (!ref) return clear_error();
    int64_t delta;
    if (sbt < ref->sbt_counter && !overflow)
        delta = (int64_t)sbt - (int64_t)ref->sbt_counter;
    else
        delta = (int64_t)asarf_counter_gap(ref->sbt_counter, sbt, 0xFFFFFFFFull);
    int64_t ns = 0, us = 0;
    if (__builtin_mul_overflow((int64_t)ref->sbt_period_ns, delta, &ns) ||
        __builtin_add_overflow((int64_t)ref->time_us, ns / 1000, &us) || us / kMicrosPerDay > kMaxMjdDays ||
        us / kMicrosPerDay < -kMaxMjdDays)
        return set_error(ASARF_EINVAL, "asarf_adjust_isp_sensing_time: time out of range");
    *isp_time = micros_to_mjd(us);
    return clear_error();
}

extern "C" int asarf_product_type_from_name(const char* product_name) {
    // GetProductTypeFrom (asar_constants.cc:58-66): by the name's prefix
    if (!product_name) return ASARF_PRODUCT_UNIDENTIFIED;
    if (std::strncmp(product_name, "ASA_IM__0", 9) == 0) return ASARF_PRODUCT_ASA_IM0;
    if (std::strncmp(product_name, "SAR_IM__0", 9) == 0) return ASARF_PRODUCT_SAR_IM0;
    return ASARF_PRODUCT_UNIDENTIFIED;
}

extern "C" int asarf_target_product(int in_product, const char* target, int* out_product) {
    // TryDetermineTargetProductFrom (asar_constants.cc:68-90): "IMS" or "IMP", any case, from an IM level-0 input
    if (!target || !out_product) return set_error(ASARF_EINVAL, "asarf_target_product: null");
    const bool ims = iequals(target, "IMS"), imp = iequals(target, "IMP");
    if (in_product == ASARF_PRODUCT_SAR_IM0 && (ims || imp)) {
        *out_product = ims ? ASARF_PRODUCT_SAR_IMS : ASARF_PRODUCT_SAR_IMP;
        return clear_error();
    }
    if (in_product == ASARF_PRODUCT_ASA_IM0 && (ims || imp)) {
        *out_product = ims ? ASARF_PRODUCT_ASA_IMS : ASARF_PRODUCT_ASA_IMP;
        return clear_error();
    }
    return set_error(ASARF_EINVAL, "asarf_target_product: focussing '%s' from this input is not supported", target);
}

extern "C" int asarf_instrument(int product, int* out_instrument) {
    // GetInstrumentFrom (asar_constants.cc:92-102): SAR for SAR_IM0 / SAR_IMS, ASAR for ASA_IM0 / ASA_IMS, an error
    // for every other type (the detected products included, as the reference)
    if (!out_instrument) return set_error(ASARF_EINVAL, "asarf_instrument: null");
    if (product == ASARF_PRODUCT_SAR_IM0 || product == ASARF_PRODUCT_SAR_IMS) {
        *out_instrument = ASARF_INSTRUMENT_SAR;
        return clear_error();
    }
    if (product == ASARF_PRODUCT_ASA_IM0 || product == ASARF_PRODUCT_ASA_IMS) {
        *out_instrument = ASARF_INSTRUMENT_ASAR;
        return clear_error();
    }
    return set_error(ASARF_EINVAL, "asarf_instrument: unsupported product type %d", product);
}

namespace {
uint16_t be16(const unsigned char* p) { return (uint16_t)(((unsigned)p[0] << 8) | p[1]); }
uint32_t be32(const unsigned char* p) {
    return ((uint32_t)p[0] << 24) | ((uint32_t)p[1] << 16) | ((uint32_t)p[2] << 8) | (uint32_t)p[3];
}
asarf_mjd be_mjd(const unsigned char* p) {
    asarf_mjd m;
    m.days = (int32_t)be32(p);
    m.seconds = be32(p + 4);
    m.micros = be32(p + 8);
    return m;
}
}  // namespace

extern "C" int asarf_parse_asar_im_packet(const unsigned char* packet, long n_bytes, asarf_asar_echo* meta,
                                          unsigned char* record, long record_bytes) {
    // One ASAR IM level-0 data set record (envisat_im_parse.cc:433-600; PO-TN-MMS-SR-0248), all fields big-endian:
    //   0 ISP sensing time (MJD: int32 days, uint32 seconds, uint32 microseconds), 12 FEP annotations (MJD, ISP
    //   length, CRC error count, correction count, spare), 32 packet id, 34 sequence control (14-bit count),
    //   36 packet length, 38 data field header length (29), 41 mode id (0x54 for IM), 42 40-bit on-board time,
    //   48 24-bit mode packet count, 51 antenna beam set (6 bits) | compression ratio (2 bits), 52 echo / noise /
    //   calibration / calibration-type flags and the 12-bit cycle packet count, 54 PRI code, 56 SWST code,
    //   58 echo window code, 60 up/down-converter levels, polarisations and calibration row, 62 Tx pulse code and
    //   beam adjustment, 64 chirp bandwidth code, 65 auxiliary Tx monitor level, 66 resampling factor,
    //   68 sample blocks (packet length - 29 bytes).
    // For an echo packet, record (when given) receives the sample blocks as the decoder's record: a 2-byte
    // little-endian length, the bytes, and zeros up to record_bytes.
    if (!packet || !meta) return set_error(ASARF_EINVAL, "asarf_parse_asar_im_packet: null");
    if (n_bytes < 68) return set_error(ASARF_ESHAPE, "asarf_parse_asar_im_packet: %ld bytes < 68", n_bytes);
    asarf_asar_echo m{};
    m.isp_sensing_time = be_mjd(packet);
    m.fep_time = be_mjd(packet + 12);
    m.isp_length = be16(packet + 24);
    m.packet_id = be16(packet + 32);
    const uint16_t seq = be16(packet + 34);
    m.segmentation_flags = (unsigned char)(seq >> 14);
    m.sequence_count = (unsigned short)(seq & 0x3FFF);
    m.packet_length = be16(packet + 36);
    const uint16_t dfh = be16(packet + 38);
    if (dfh != 29)
        return set_error(ASARF_EINVAL, "asarf_parse_asar_im_packet: data field header length %u, not 29", dfh);
    m.mode_id = packet[41];
    if (m.mode_id != 0x54)
        return set_error(ASARF_EINVAL, "asarf_parse_asar_im_packet: mode id 0x%x, not IM (0x54)", m.mode_id);
    m.onboard_time = ((unsigned long long)packet[42] << 32) | ((unsigned long long)packet[43] << 24) |
                     ((unsigned long long)packet[44] << 16) | ((unsigned long long)packet[45] << 8) | packet[46];
    m.mode_count = ((unsigned)packet[48] << 16) | ((unsigned)packet[49] << 8) | packet[50];
    m.antenna_beam_set_no = (unsigned char)(packet[51] >> 2);
    m.comp_ratio = (unsigned char)(packet[51] & 0x3);
    m.echo_flag = (packet[52] & 0x80) != 0;
    m.noise_flag = (packet[52] & 0x40) != 0;
    m.cal_flag = (packet[52] & 0x20) != 0;
    m.cal_type = (packet[52] & 0x10) != 0;
    m.cycle_count = (unsigned short)(((packet[52] & 0x0F) << 8) | packet[53]);
    m.pri_code = be16(packet + 54);
    m.swst_code = be16(packet + 56);
    m.echo_window_code = be16(packet + 58);
    const uint16_t conv = be16(packet + 60);
    m.upconverter_raw = (unsigned char)(conv >> 12);
    m.downconverter_raw = (unsigned char)((conv >> 7) & 0x1F);
    m.tx_pol = (conv & 0x40) != 0;
    m.rx_pol = (conv & 0x20) != 0;
    m.cal_row_number = (unsigned char)(conv & 0x1F);
    const uint16_t pulse = be16(packet + 62);
    m.tx_pulse_code = (unsigned short)(pulse >> 6);
    m.beam_adj_delta = (unsigned char)(pulse & 0x3F);
    m.chirp_pulse_bw_code = packet[64];
    m.aux_tx_monitor_level = packet[65];
    m.resampling_factor = be16(packet + 66);
    if (m.packet_length < 29) return set_error(ASARF_EINVAL, "asarf_parse_asar_im_packet: packet length < 29");
    const long data_len = (long)m.packet_length - 29;
    if (68 + data_len > n_bytes)
        return set_error(ASARF_ESHAPE, "asarf_parse_asar_im_packet: %ld sample bytes beyond the %ld-byte record",
                         data_len, n_bytes);
    m.sample_bytes = (int)data_len;
    m.dsr_bytes = 39 + (int)m.isp_length;  // the next record's offset, as the reference steps (envisat_im_parse.cc:188)
    // whole 64-byte blocks hold 63 samples; a partial block holds its length minus the block-id byte
    // (envisat_im_parse.cc:588-590)
    m.samples = (int)((data_len / 64) * 63 + data_len % 64 - (data_len % 64 > 0 ? 1 : 0));
    if (record && m.echo_flag) {
        if (record_bytes < 2 + data_len)
            return set_error(ASARF_EWORKSPACE, "asarf_parse_asar_im_packet: record of %ld bytes < %ld", record_bytes,
                             2 + data_len);
        record[0] = (unsigned char)(data_len & 0xFF);
        record[1] = (unsigned char)(data_len >> 8);
        std::memcpy(record + 2, packet + 68, (size_t)data_len);
        std::memset(record + 2 + data_len, 0, (size_t)(record_bytes - 2 - data_len));
    }
    *meta = m;
    return clear_error();
}

extern "C" int asarf_parse_ers_im_packet(const unsigned char* packet, long n_bytes, asarf_ers_echo* meta,
                                         unsigned char* line) {
    // One ERS IM level-0 data set record of 11498 bytes (ers_im_parse.cc:366-548, ers_env_format.h:26-39), fields
    // big-endian: 0 ISP sensing time (MJD), 24 ISP length (11465), 32 data record number, 36 packet counter,
    // 37 subcommutation counter, 46 auxiliary section marker (0xAA), 47 OBRC flag / orbit bits, 48 satellite
    // binary time, 52 activity task, 53 sample flags, 54 image format counter, 58 SWST code, 60 PRI code, and from
    // 266 the 5616 (I, Q) byte pairs of the echo, which line (when given, 11232 bytes) receives as the ERS
    // decoder's line. The ISP sensing time is returned as recorded: the caller applies the SBT correction
    // (asarf_adjust_isp_sensing_time with sbt and overflow = sbt < the previous packet's), as the reference does.
    if (!packet || !meta) return set_error(ASARF_EINVAL, "asarf_parse_ers_im_packet: null");
    if (n_bytes < 11498) return set_error(ASARF_ESHAPE, "asarf_parse_ers_im_packet: %ld bytes < 11498", n_bytes);
    asarf_ers_echo m{};
    m.isp_sensing_time = be_mjd(packet);
    m.isp_length = be16(packet + 24);
    if (m.isp_length != 11465)
        return set_error(ASARF_EINVAL, "asarf_parse_ers_im_packet: ISP length %u, not 11465", m.isp_length);
    m.data_record_number = be32(packet + 32);
    m.packet_counter = packet[36];
    m.subcommutation_counter = packet[37];
    if (packet[46] != 0xAA)
        return set_error(ASARF_EINVAL, "asarf_parse_ers_im_packet: auxiliary section starts with 0x%x, not 0xAA",
                         packet[46]);
    m.obrc_orbit = packet[47];
    m.sbt = be32(packet + 48);
    m.activity_task = packet[52];
    m.sample_flags = packet[53];
    m.image_format_counter = be32(packet + 54);
    m.swst_code = be16(packet + 58);
    m.pri_code = be16(packet + 60);
    if (line) std::memcpy(line, packet + 266, 11232);
    *meta = m;
    return clear_error();
}

namespace {
constexpr long kMphSize = 1247, kSphSize = 1956, kDsdSize = 280;

// extract_bytes / extract_num (envisat_mph_sph_parser.h:86-107): from the first digit 1-9 after the '=' (signs and
// leading zeros are skipped, so "-0000000001" reads as 1 and an all-zero field as 0), the decimal number. False
// when it does not fit 64 bits (the reference's std::stoul throws std::out_of_range there).
bool dsd_number(const char* p, long n, unsigned long long* out) {
    long i = 0;
    while (i < n && p[i] != '=') ++i;
    for (; i < n; ++i)
        if (p[i] >= '1' && p[i] <= '9') break;
    unsigned long long v = 0;
    for (; i < n && p[i] >= '0' && p[i] <= '9'; ++i) {
        const unsigned d = (unsigned)(p[i] - '0');
        if (v > (~0ull - d) / 10) return false;
        v = v * 10 + d;
    }
    *out = v;
    return true;
}

// the idx-th '\n'-separated token of [p, p + n): its start and length (false when there are fewer tokens)
bool nth_line(const char* p, long n, int idx, const char** s, long* len) {
    long start = 0;
    for (int k = 0; k < idx; ++k) {
        while (start < n && p[start] != '\n') ++start;
        if (start >= n) return false;
        ++start;
    }
    long end = start;
    while (end < n && p[end] != '\n') ++end;
    *s = p + start;
    *len = end - start;
    return true;
}

void copy_text(char* dst, int cap, const char* src, long len) {
    if (!dst || cap <= 0) return;
    const long k = len < cap - 1 ? len : cap - 1;
    std::memcpy(dst, src, (size_t)k);
    dst[k] = '\0';
}
}  // namespace

extern "C" int asarf_header_get(const char* text, long n_bytes, const char* key, char* value, int capacity) {
    // ProductHeader::Load / Get (envisat_mph_sph_parser.h:41-77): '\n'-separated KEY=value lines; the value of the
    // first line whose key matches, without one surrounding pair of quotes
    if (!text || !key || !value || capacity <= 0) return set_error(ASARF_EINVAL, "asarf_header_get: argument");
    const long klen = (long)std::strlen(key);
    long start = 0;
    while (start < n_bytes) {
        long end = start;
        while (end < n_bytes && text[end] != '\n') ++end;
        long eq = start;
        while (eq < end && text[eq] != '=') ++eq;
        if (eq < end && eq - start == klen && std::memcmp(text + start, key, (size_t)klen) == 0) {
            long vs = eq + 1, ve = end;
            if (ve > vs && text[ve - 1] == '"') --ve;
            if (ve > vs && text[vs] == '"') ++vs;
            copy_text(value, capacity, text + vs, ve - vs);
            return clear_error();
        }
        start = end + 1;
    }
    return set_error(ASARF_EINVAL, "asarf_header_get: %s not found", key);
}

extern "C" int asarf_level0_mdsr(const unsigned char* file, long n_bytes, asarf_dsd* out) {
    // ParseMdsr / ExtractDSDs (envisat_lvl0_parser.cc:98-102, envisat_mph_sph_parser.h:110-133): the first of the
    // four 280-byte data set descriptors at the end of the 1956-byte SPH that follows the 1247-byte MPH; lines 0-2
    // as they stand (name, type, filename), 3 offset, 4 size, 5 record count, 6 record size
    if (!file || !out) return set_error(ASARF_EINVAL, "asarf_level0_mdsr: null");
    if (n_bytes < kMphSize + kSphSize)
        return set_error(ASARF_ESHAPE, "asarf_level0_mdsr: %ld bytes < MPH + SPH (%ld)", n_bytes, kMphSize + kSphSize);
    const char* sph = reinterpret_cast<const char*>(file) + kMphSize;
    for (int d = 0; d < 4; ++d) {
        const char* dsd = sph + kSphSize - 4 * kDsdSize + d * kDsdSize;
        const char* ln[7];
        long len[7];
        bool ok = true;
        for (int i = 0; i < 7 && ok; ++i) ok = nth_line(dsd, kDsdSize, i, &ln[i], &len[i]);
        if (!ok) continue;  // fewer tokens than a descriptor has: not a descriptor
        asarf_dsd r{};
        copy_text(r.ds_name, sizeof(r.ds_name), ln[0], len[0]);
        copy_text(r.ds_type, sizeof(r.ds_type), ln[1], len[1]);
        copy_text(r.filename, sizeof(r.filename), ln[2], len[2]);
        unsigned long long v[4];
        for (int i = 0; i < 4; ++i)
            if (!dsd_number(ln[3 + i], len[3 + i], &v[i]))
                return set_error(ASARF_EINVAL, "asarf_level0_mdsr: descriptor %d field %d does not fit 64 bits", d,
                                 3 + i);
        r.ds_offset = v[0];
        r.ds_size = v[1];
        r.num_dsr = v[2];
        r.dsr_size = v[3];
        *out = r;
        return clear_error();
    }
    return set_error(ASARF_EINVAL, "asarf_level0_mdsr: no data set descriptor");
}

extern "C" int asarf_asar_forecast(const unsigned char* mds, long mds_bytes, long num_dsr, const asarf_mjd* start,
                                   const asarf_mjd* stop, long* before_start, long* after_stop, long* offsets,
                                   asarf_mjd* times, long capacity, long* n_out) {
    // FetchEnvisatL0ImForecastMeta (envisat_im_parse.cc:297-384): walk the num_dsr records of the measurement data
    // set; skip packets before the first echo and packets that are neither echo nor calibration; count those
    // sensed before start and after stop; stop once more than *after_stop (the lines the caller may keep after the
    // stop) lie after it; then drop the surplus before the start beyond *before_start, and leading calibration
    // packets, so the list starts with an echo. On return *before_start / *after_stop hold the counts kept.
    if (!mds || !start || !stop || !before_start || !after_stop || !offsets || !times || !n_out)
        return set_error(ASARF_EINVAL, "asarf_asar_forecast: null");
    std::vector<long> off;
    std::vector<asarf_mjd> tim;
    std::vector<char> is_echo;
    long before = 0, after = 0, pos = 0;
    for (long i = 0; i < num_dsr; ++i) {
        if (pos + 68 > mds_bytes)
            return set_error(ASARF_ESHAPE, "asarf_asar_forecast: record %ld beyond the data set", i);
        const unsigned char* p = mds + pos;
        const asarf_mjd t = be_mjd(p);
        const uint16_t isp_length = be16(p + 24);
        if (be16(p + 38) != 29)
            return set_error(ASARF_EINVAL, "asarf_asar_forecast: record %ld data field header length is not 29", i);
        const bool echo = (p[52] & 0x80) != 0, cal = (p[52] & 0x20) != 0;
        const long here = pos;
        pos += 39 + (long)isp_length;
        if (off.empty() && !echo) continue;
        if (!echo && !cal) continue;
        if (asarf_mjd_compare(&t, start) < 0) ++before;
        if (asarf_mjd_compare(&t, stop) > 0) ++after;
        if (after > *after_stop) {
            --after;
            break;
        }
        off.push_back(here);
        tim.push_back(t);
        is_echo.push_back(echo ? 1 : 0);
    }
    size_t first = 0;
    if (before > *before_start) {
        first = (size_t)(before - *before_start);
        before = *before_start;
        // the list must start with an echo: drop leading calibration packets (the reference tests the calibration
        // flag of the last record it read instead of the erased entry's; every kept non-echo entry is one). A
        // dropped packet leaves the count before the start only if it was sensed before the start: the reference
        // decrements its unsigned count for every dropped packet, which wraps below zero when the allowance is 0.
        while (first < off.size() && !is_echo[first]) {
            if (asarf_mjd_compare(&tim[first], start) < 0 && before > 0) --before;
            ++first;
        }
    }
    const long n = (long)(off.size() - std::min(first, off.size()));
    if (n > capacity) return set_error(ASARF_EWORKSPACE, "asarf_asar_forecast: %ld entries > capacity %ld", n, capacity);
    for (long k = 0; k < n; ++k) {
        offsets[k] = off[first + k];
        times[k] = tim[first + k];
    }
    *n_out = n;
    *before_start = before;
    *after_stop = after;
    return clear_error();
}

extern "C" int asarf_ers_forecast(const unsigned char* mds, long mds_bytes, long num_dsr, const asarf_mjd* start,
                                  const asarf_mjd* stop, const asarf_patc_reference* patc, long* before_start,
                                  long* after_stop, long* offsets, asarf_mjd* times, long capacity, long* n_out) {
    // FetchErsL0ImForecastMeta (ers_im_parse.cc:228-331): walk the num_dsr 11498-byte records; each ISP time is
    // SBT-corrected (asarf_adjust_isp_sensing_time; overflow = SBT below the first record's, as the reference
    // compares with the SBT it read before the loop); a data-record-number gap of g > 1 repeats the previous entry
    // g - 1 times (at most 900 missing records), counted before / after the window by the previous entry's time;
    // the walk stops once more than *after_stop records lie after the stop; the surplus before the start beyond
    // *before_start is dropped, and with no allowance (*before_start == 0) every leading entry sensed before the
    // start too. On return *before_start / *after_stop hold the counts kept.
    constexpr long kDsr = 11498;
    if (!mds || !start || !stop || !before_start || !after_stop || !offsets || !times || !n_out)
        return set_error(ASARF_EINVAL, "asarf_ers_forecast: null");
    if (num_dsr <= 0 || mds_bytes < 0 || num_dsr > mds_bytes / kDsr)
        return set_error(ASARF_ESHAPE, "asarf_ers_forecast: %ld records of 11498 bytes exceed %ld bytes", num_dsr,
                         mds_bytes);
    const long before_in = *before_start, after_in = *after_stop;
    const uint32_t dr_first = be32(mds + 32);
    uint32_t last_dr = dr_first != 0 ? dr_first - 1 : 0xFFFFFFFFu;
    const uint32_t first_sbt = be32(mds + 48);
    std::vector<long> off;
    std::vector<asarf_mjd> tim;
    long before = 0, after = 0;
    for (long i = 0; i < num_dsr; ++i) {
        const unsigned char* p = mds + i * kDsr;
        if (be16(p + 24) != 11465)
            return set_error(ASARF_EINVAL, "asarf_ers_forecast: record %ld ISP length is not 11465", i);
        asarf_mjd t = be_mjd(p);
        const uint32_t dr = be32(p + 32), sbt = be32(p + 48);
        const int st = asarf_adjust_isp_sensing_time(patc, &t, sbt, sbt < first_sbt);
        if (st != ASARF_OK) return st;
        const unsigned long long gap = asarf_counter_gap(last_dr, dr, 0xFFFFFFFFull);
        if (gap > 1) {
            if (off.empty())
                return set_error(ASARF_EINVAL, "asarf_ers_forecast: record number gap before any record was kept");
            const unsigned long long fill = gap - 1;
            if (fill > 900)
                return set_error(ASARF_EINVAL, "asarf_ers_forecast: %llu records missing before record %ld (over 900)",
                                 fill, i);
            const long prev_off = off.back();
            const asarf_mjd prev_t = tim.back();
            for (unsigned long long k = 0; k < fill; ++k) {
                off.push_back(prev_off);
                tim.push_back(prev_t);
            }
            if (asarf_mjd_compare(&prev_t, start) < 0) before += (long)fill;
            if (asarf_mjd_compare(&prev_t, stop) > 0) after += (long)fill;
        } else {
            if (asarf_mjd_compare(&t, start) < 0) ++before;
            if (asarf_mjd_compare(&t, stop) > 0) ++after;
        }
        if (after > after_in) {
            --after;
            break;
        }
        last_dr = dr;
        off.push_back(i * kDsr);
        tim.push_back(t);
    }
    size_t first = 0;
    if (before > before_in) {
        first = (size_t)(before - before_in);
        before = before_in;
        if (before_in == 0)
            while (first < off.size() && asarf_mjd_compare(&tim[first], start) < 0) ++first;
    }
    const long n = (long)(off.size() - std::min(first, off.size()));
    if (n > capacity) return set_error(ASARF_EWORKSPACE, "asarf_ers_forecast: %ld entries > capacity %ld", n, capacity);
    for (long k = 0; k < n; ++k) {
        offsets[k] = off[first + k];
        times[k] = tim[first + k];
    }
    *n_out = n;
    *before_start = before;
    *after_stop = after;
    return clear_error();
}

extern "C" int asarf_ins_fbaq4_tables(const unsigned char* ins, long n_bytes, float* i_lut, float* q_lut) {
    // FindINSFile (envisat_aux_file.cc:67-95) and FixedBaselineParameters (envisat_aux_file.h:143-149): a 173273-byte
    // ASAR instrument characterisation file = 1247-byte MPH + 378-byte SPH + InstrumentFile, whose 89408-byte
    // FixedBaselineParameters follow the 72608-byte FixedLifetimeParameters and hold, big-endian, the 8-bit I and
    // Q tables (256 floats each), then i_LUT_fbaq4 (4096), i_LUT_fbaq3 (2048), i_LUT_fbaq2 (1024), q_LUT_fbaq4 (4096)
    constexpr long kInsSize = 173273, kFbp = 1247 + 378 + 72608;
    constexpr long kILut = kFbp + 4 * (256 + 256), kQLut = kILut + 4 * (4096 + 2048 + 1024);
    if (!ins || !i_lut || !q_lut) return set_error(ASARF_EINVAL, "asarf_ins_fbaq4_tables: null");
    if (n_bytes != kInsSize)
        return set_error(ASARF_ESHAPE, "asarf_ins_fbaq4_tables: %ld bytes, an instrument file has 173273", n_bytes);
    for (int k = 0; k < 4096; ++k) {
        const uint32_t wi = be32(ins + kILut + 4 * k), wq = be32(ins + kQLut + 4 * k);
        std::memcpy(&i_lut[k], &wi, 4);
        std::memcpy(&q_lut[k], &wq, 4);
    }
    return clear_error();
}
