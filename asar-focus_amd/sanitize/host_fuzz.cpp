// Host-side robustness check of the level-0 parsers and the FP64 geometry under AddressSanitizer and
// UndefinedBehaviorSanitizer: every entry point that reads caller bytes (packets, product headers, data sets,
// PATC and auxiliary files) is fed truncated, random and adversarial inputs from a seeded generator, and the
// geometry functions degenerate orbits and sizes. An entry point may reject an input with an error status; a read
// or write outside a buffer, an overflow the sanitizers catch, or a crash ends the run with a report.
#include <algorithm>
#include <cstdarg>
#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "asarf.h"

namespace asarf {
// The library keeps its error text in api.cpp, which this host-only build does not link.
static thread_local char t_err[512] = "";
int set_error(int status, const char* fmt, ...) {
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(t_err, sizeof(t_err), fmt, ap);
    va_end(ap);
    return status;
}
int clear_error() {
    t_err[0] = 0;
    return ASARF_OK;
}
}  // namespace asarf

namespace {

std::mt19937_64 rng(0x5EED0042);

std::vector<unsigned char> random_bytes(size_t n) {
    std::vector<unsigned char> v(n);
    for (auto& b : v) b = (unsigned char)(rng() & 0xFF);
    return v;
}

// A buffer of exactly n bytes on the heap, so that ASan sees any read past its end.
struct Exact {
    unsigned char* p;
    size_t n;
    explicit Exact(const std::vector<unsigned char>& v) : p(new unsigned char[v.empty() ? 1 : v.size()]), n(v.size()) {
        if (!v.empty()) std::memcpy(p, v.data(), v.size());
    }
    ~Exact() { delete[] p; }
};

void put_be16(std::vector<unsigned char>& v, size_t at, unsigned x) {
    if (at + 1 < v.size()) {
        v[at] = (unsigned char)(x >> 8);
        v[at + 1] = (unsigned char)x;
    }
}

void craft_asar(std::vector<unsigned char>& v, size_t at = 0) {
    if (at + 68 > v.size()) return;
    const size_t room = v.size() - at - 68;
    unsigned pl;
    switch (rng() % 4) {
        case 0: pl = 29 + (unsigned)(room ? rng() % room : 0); break;  // samples inside the buffer
        case 1: pl = (unsigned)(rng() & 0xFFFF); break;                 // any length
        case 2: pl = (unsigned)(rng() % 29); break;                      // below the header
        default: pl = 29 + (unsigned)room + (unsigned)(rng() % 8); break;  // just past the end
    }
    put_be16(v, at + 36, pl);
    put_be16(v, at + 24, (rng() % 3) ? pl + 29 : (unsigned)(rng() & 0xFFFF));  // ISP length: the record step
    put_be16(v, at + 38, (rng() % 8) ? 29 : (unsigned)(rng() & 0xFF));
    v[at + 41] = (rng() % 8) ? 0x54 : (unsigned char)rng();
}

long pick_len(long typical) {
    switch (rng() % 6) {
        case 0: return 0;
        case 1: return (long)(rng() % 64);
        case 2: return typical;
        case 3: return typical - 1 - (long)(rng() % 16);
        case 4: return typical + (long)(rng() % 64);
        default: return (long)(rng() % (2 * typical + 1));
    }
}

void fuzz_asar_packets(int iters) {
    for (int it = 0; it < iters; ++it) {
        const long n = pick_len(39 + 29 + 64 * 90);
        auto v = random_bytes((size_t)std::max(0L, n));
        // headers that pass the format checks most of the time (data field header length 29 at byte 38, IM mode
        // 0x54 at 41), with any packet length (36) and ISP length (24), so the length arithmetic is what is tried
        if ((rng() % 4) && v.size() >= 68) craft_asar(v);
        Exact pk(v);
        asarf_asar_echo meta;
        const long rec_bytes = pick_len(2 + 64 * 90);
        std::vector<unsigned char> rec((size_t)std::max(1L, rec_bytes));
        (void)asarf_parse_asar_im_packet(pk.p, (long)pk.n, &meta, rec.data(), rec_bytes);
        (void)asarf_parse_asar_im_packet(pk.p, (long)pk.n, &meta, nullptr, 0);
    }
}

void fuzz_ers_packets(int iters) {
    for (int it = 0; it < iters; ++it) {
        const long n = pick_len(11498);
        auto v = random_bytes((size_t)std::max(0L, n));
        // the ISP length (11465 at byte 24) and the auxiliary marker (0xAA at 46) that the parser checks
        if ((rng() % 4) && v.size() >= 64) {
            put_be16(v, 24, (rng() % 8) ? 11465 : (unsigned)(rng() & 0xFFFF));
            v[46] = (rng() % 8) ? 0xAA : (unsigned char)rng();
        }
        Exact pk(v);
        asarf_ers_echo meta;
        std::vector<unsigned char> line(11232);
        (void)asarf_parse_ers_im_packet(pk.p, (long)pk.n, &meta, line.data());
        (void)asarf_parse_ers_im_packet(pk.p, (long)pk.n, &meta, nullptr);
    }
}

std::string random_header_text() {
    static const char* keys[] = {"PRODUCT", "DS_NAME", "DS_TYPE", "FILENAME", "DS_OFFSET", "DS_SIZE", "NUM_DSR",
                                 "DSR_SIZE", "SENSING_START", "X"};
    std::string t;
    const int lines = (int)(rng() % 40);
    for (int i = 0; i < lines; ++i) {
        t += keys[rng() % 10];
        if (rng() % 4) t += "=";
        switch (rng() % 5) {
            case 0: t += "\"" + std::string((size_t)(rng() % 90), 'A') + "\""; break;
            case 1: t += std::string((size_t)(rng() % 40), '9'); break;  // long digit strings
            case 2: t += "+" + std::to_string(rng() % 100000) + "<bytes>"; break;
            case 3: t += "-0000000001"; break;
            default: t += "\""; break;  // unbalanced quote
        }
        if (rng() % 8) t += "\n";
    }
    return t;
}

void fuzz_headers(int iters) {
    for (int it = 0; it < iters; ++it) {
        const std::string t = random_header_text();
        std::vector<unsigned char> v(t.begin(), t.end());
        if (rng() % 3 == 0 && !v.empty()) v.resize(rng() % v.size());  // cut mid-line
        Exact buf(v);
        char value[64];
        const int cap = (int)(rng() % 65);
        static const char* keys[] = {"PRODUCT", "DS_NAME", "NUM_DSR", "X", "", "DS_SIZE"};
        (void)asarf_header_get(reinterpret_cast<const char*>(buf.p), (long)buf.n, keys[rng() % 6], value, cap);
        asarf_dsd dsd;
        (void)asarf_level0_mdsr(buf.p, (long)buf.n, &dsd);
    }
}

void fuzz_forecasts(int iters) {
    for (int it = 0; it < iters; ++it) {
        const bool ers = rng() & 1;
        const long rec = ers ? 11498 : 39 + 29 + 64 * 20;
        const long n = pick_len(rec * (long)(1 + rng() % 6));
        auto v = random_bytes((size_t)std::max(0L, n));
        // records with passing headers at their nominal steps (ERS: ISP length and marker; ASAR: crafted headers
        // whose ISP lengths set the step to the next record)
        for (size_t at = 0; at + 68 <= v.size(); at += (size_t)rec) {
            if (ers) {
                if (at + 64 > v.size()) break;
                put_be16(v, at + 24, 11465);
                v[at + 46] = 0xAA;
            } else {
                craft_asar(v, at);
            }
        }
        Exact mds(v);
        long num_dsr;
        switch (rng() % 4) {
            case 0: num_dsr = (long)(rng() % 8); break;
            case 1: num_dsr = (long)(n / rec) + (long)(rng() % 3); break;
            case 2: num_dsr = (long)(rng() >> 1); break;  // huge
            default: num_dsr = -(long)(rng() % 5); break;
        }
        asarf_mjd start{(int)(rng() % 9000) - 100, (unsigned)(rng() % 86400), (unsigned)(rng() % 1000000)};
        asarf_mjd stop = start;
        (void)asarf_mjd_add_micros(&start, (long long)(rng() % 100000000), &stop);
        long before = (long)(rng() % 50), after = (long)(rng() % 50);
        const long cap = (long)(rng() % 32);
        std::vector<long> offs((size_t)std::max(1L, cap));
        std::vector<asarf_mjd> times((size_t)std::max(1L, cap));
        long n_out = 0;
        (void)asarf_asar_forecast(mds.p, (long)mds.n, num_dsr, &start, &stop, &before, &after, offs.data(),
                                  times.data(), cap, &n_out);
        before = (long)(rng() % 50);
        after = (long)(rng() % 50);
        asarf_patc_reference ref{(long long)(rng() % 1000000000000LL), (unsigned)rng(), (unsigned)(rng() % 5000000)};
        (void)asarf_ers_forecast(mds.p, (long)mds.n, num_dsr, &start, &stop, (rng() & 1) ? &ref : nullptr, &before,
                                 &after, offs.data(), times.data(), cap, &n_out);
    }
}

void fuzz_small_files(int iters) {
    for (int it = 0; it < iters; ++it) {
        auto v = random_bytes((size_t)pick_len(54));
        Exact p(v);
        asarf_patc patc;
        if (asarf_parse_patc(p.p, (long)p.n, &patc) == ASARF_OK) {
            asarf_patc_reference ref;
            if (asarf_patc_register(&patc, &ref) == ASARF_OK) {
                asarf_mjd t{(int)(rng() % 10000), (unsigned)(rng() % 86400), (unsigned)(rng() % 1000000)};
                (void)asarf_adjust_isp_sensing_time(&ref, &t, (unsigned)rng(), (int)(rng() % 3));
            }
        }
        auto ins = random_bytes((size_t)pick_len(173273));
        Exact q(ins);
        std::vector<float> il(4096), ql(4096);
        (void)asarf_ins_fbaq4_tables(q.p, (long)q.n, il.data(), ql.data());
        asarf_mjd a{(int)rng(), (unsigned)rng(), (unsigned)rng()}, b{};
        (void)asarf_mjd_add_micros(&a, (long long)rng(), &b);
        (void)asarf_mjd_compare(&a, &b);
        (void)asarf_counter_gap(rng(), rng(), rng() % 0xFFFFFFFFull);
        (void)asarf_product_type_from_name((rng() & 1) ? "ASA_IM__0P" : "");
    }
}

// DORIS orbit files of the ENVISAT size: well-formed records (one minute apart) with random bytes flipped, or
// random text; CON, XCA and INS files of the right and wrong sizes; DORIS file names; radar timing parameters.
void fuzz_aux_files(int iters) {
    static const char* months[12] = {"JAN", "FEB", "MAR", "APR", "MAY", "JUN", "JUL", "AUG", "SEP", "OCT", "NOV", "DEC"};
    for (int it = 0; it < iters; ++it) {
        std::vector<unsigned char> v(206606, ' ');
        const size_t base = 1247 + 378;
        for (int k = 0; k < 1589; ++k) {
            char line[160];
            const int h = (k / 60) % 24, m = k % 60;
            int n = std::snprintf(line, sizeof(line), "%02d-%s-2004 %02d:%02d:00.000000 -0.4 %05d %.3f %.3f %.3f %.6f "
                                  "%.6f %.6f 0", 1 + (k / 1440) % 28, months[(it + k / 40000) % 12], h, m, 9000 + k,
                                  7.1e6, 1.0e3 * k, -2.0e3, 7450.0, -1.0, 0.5);
            std::string l(line, (size_t)n);
            if (l.size() < 128) l.insert(11, 128 - l.size(), ' ');
            l.resize(128);
            std::memcpy(v.data() + base + 129 * (size_t)k, l.data(), 128);
            v[base + 129 * (size_t)k + 128] = '\n';
        }
        const int flips = (int)(rng() % 40);
        for (int f = 0; f < flips; ++f) v[rng() % v.size()] = (unsigned char)(rng() & 0xFF);
        if (rng() % 8 == 0) v = random_bytes(rng() % 2 ? 206606 : 233954);
        Exact p(v);
        asarf_mjd a{1461 + (int)(rng() % 3), (unsigned)(rng() % 86400), 0}, b = a;
        b.seconds = (unsigned)std::min<unsigned long long>(86399, a.seconds + rng() % 600);
        std::vector<asarf_osv> osv(64);
        std::vector<asarf_mjd> times(64);
        int n_out = 0;
        asarf_doris_meta meta;
        (void)asarf_doris_orbit(p.p, (long)p.n, &a, &b, osv.data(), times.data(), (int)(rng() % 65), &n_out, &meta);
        std::string name = "DOR_VOR_AXVF-P20120424_120200_20040101_215528_20040103_002328";
        for (int f = 0; f < (int)(rng() % 4); ++f) name[rng() % name.size()] = (char)(32 + rng() % 95);
        if (rng() % 4 == 0) name.resize(rng() % 70);
        asarf_mjd w0, w1;
        (void)asarf_doris_filename_window(name.c_str(), &w0, &w1);
        float g1[7], g2[7];
        auto con = random_bytes(rng() % 2 ? 5721 : (size_t)pick_len(5721));
        Exact c(con);
        (void)asarf_con_process_gains(c.p, (long)c.n, g1, g2);
        auto xca = random_bytes(rng() % 2 ? 28177 : (size_t)pick_len(28177));
        Exact x(xca);
        (void)asarf_xca_scaling_factors(x.p, (long)x.n, g1, g2);
        double k = 0;
        (void)asarf_calibration_constant((int)(rng() & 1), (int)(rng() & 1), g1[0], g2[0], &k);
        auto ins = random_bytes(rng() % 2 ? 173273 : (size_t)pick_len(173273));
        Exact q(ins);
        asarf_ins_swath sw{};
        if (asarf_ins_swath_params(q.p, (long)q.n, (int)(rng() % 9) - 1, &sw) == ASARF_OK) {
            asarf_timing t;
            (void)asarf_radar_timing(&sw, (int)(rng() & 1), (int)(rng() % 5000) - 10, (int)(rng() % 2000) - 5,
                                     (int)(rng() % 6), &t);
        }
    }
}

// The level-1 header writer on random metadata: product names, strings of every length, sizes, times and orbits
// (values that do not fit their header fields must come back as errors).
void fuzz_l1_header(int iters) {
    std::vector<unsigned char> out((size_t)asarf_l1_header_size(1) + 64);
    for (int it = 0; it < iters; ++it) {
        const int n_osv = 2 + (int)(rng() % 10);
        std::vector<asarf_osv> osv((size_t)n_osv);
        for (int k = 0; k < n_osv; ++k) {
            const double a = 1e-3 * k;
            osv[(size_t)k] = asarf_osv{1.4e8 + 10.0 * k, {7.1e6 * std::cos(a), 7.1e6 * std::sin(a), 1e5},
                                       {-7.4e3 * std::sin(a), 7.4e3 * std::cos(a), 10.0}};
        }
        asarf_geometry g{};
        g.osv = osv.data();
        g.n_osv = n_osv;
        g.first_line_time = 1.4e8 + (double)(rng() % 30);
        g.line_time_interval = 1.0 / (1000.0 + (double)(rng() % 1000));
        g.azimuth_size = (int)(rng() % 5000);
        g.range_size = (int)(rng() % 6000);
        g.slant_range_first = 8e5 + (double)(rng() % 100000);
        g.range_spacing = (rng() % 2) ? 7.8 : 12.5;
        g.carrier_frequency = 5.3e9;
        g.prf = 1652.0;
        g.az_bandwidth_fraction = 0.8;
        g.vr_poly[2] = 7000.0;
        g.center_point[0] = 6.37e6;
        auto rand_str = [&](size_t maxlen) {
            std::string s((size_t)(rng() % (maxlen + 1)), 'A');
            for (auto& c : s) c = (char)(32 + rng() % 95);
            return s;
        };
        const std::string name = rand_str(80), station = rand_str(40), swath = rand_str(4), comp = rand_str(6);
        asarf_doris_meta meta{};
        meta.phase = 'A';
        std::strcpy(meta.vector_source, "DP");
        meta.pos[0] = (double)(int64_t)(rng() % 200000000) - 1e8;
        meta.delta_ut1 = ((double)(rng() % 2000) - 1000.0) * 1e-3;
        double grsr[5] = {0, 0, 0, 1.0, 0};
        asarf_l1_inputs in{};
        in.l0_product_name = name.c_str();
        in.detected = (int)(rng() & 1);
        in.acquisition_station = station.c_str();
        in.orbit = (rng() % 4) ? &meta : nullptr;
        in.swath = swath.c_str();
        in.echo_comp_method = comp.c_str();
        in.geometry = &g;
        in.range_sampling_rate = (rng() % 8) ? 19.2e6 : 0.0;
        in.grsr5 = (rng() % 4) ? grsr : nullptr;
        in.up_code = (unsigned short)(rng() % 20);
        in.mds_records = g.azimuth_size;
        in.mds_record_size = 17 + g.range_size * (in.detected ? 2 : 4) + (int)(rng() % 2);
        long n = 0;
        (void)asarf_l1_header(&in, out.data(), (long)out.size() - (long)(rng() % 64), &n);
    }
}

void fuzz_geometry(int iters) {
    for (int it = 0; it < iters; ++it) {
        const int n = (int)(rng() % 7);
        std::vector<asarf_osv> osv((size_t)std::max(1, n));
        for (int i = 0; i < n; ++i) {
            osv[(size_t)i].t = (rng() % 4 == 0) ? 0.0 : (double)i * 10.0;  // sometimes repeated times
            for (int k = 0; k < 3; ++k) {
                osv[(size_t)i].pos[k] = (double)((int64_t)(rng() % 14000000) - 7000000);
                osv[(size_t)i].vel[k] = (double)((int64_t)(rng() % 15000) - 7500);
            }
        }
        asarf_osv out;
        (void)asarf_interpolate_orbit(osv.data(), n, 5.0, &out);
        asarf_geometry g{};
        g.osv = osv.data();
        g.n_osv = n;
        g.first_line_time = 0.0;
        g.line_time_interval = (rng() & 1) ? 6e-4 : 0.0;
        g.azimuth_size = (int)(rng() % 3000);
        g.range_size = (int)(rng() % 6000);
        g.chirp_samples = (int)(rng() % 800);
        g.slant_range_first = 826e3;
        g.range_spacing = (rng() & 1) ? 7.8 : 0.0;
        g.carrier_frequency = 5.331e9;
        g.prf = 1652.4;
        g.az_bandwidth_fraction = 0.8;
        g.vr_poly[2] = 7100.0;
        g.center_point[0] = 6378137.0;
        double vr[3], samples[16], xs[16], grsr[5];
        int ns = 0, width = 0;
        (void)asarf_estimate_processing_velocity(&g, vr, samples, xs, &ns);
        (void)asarf_srgr_fit(&g, 12.5, grsr, &width);
        int ap[16];
        (void)asarf_aperture_pixels(&g, ap, (int)(rng() % 17), &ns);
    }
}

}  // namespace

int main(int argc, char** argv) {
    const int scale = argc > 1 ? std::atoi(argv[1]) : 1;
    fuzz_asar_packets(400 * scale);
    fuzz_ers_packets(200 * scale);
    fuzz_headers(2000 * scale);
    fuzz_forecasts(300 * scale);
    fuzz_small_files(100 * scale);
    fuzz_aux_files(40 * scale);
    fuzz_l1_header(200 * scale);
    fuzz_geometry(60 * scale);
    std::printf("host_fuzz: done\n");
    return 0;
}
