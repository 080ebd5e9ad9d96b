// Autotuner: measures candidate transform schedules on the device and records the fastest per length in the
// context, where the planner (get_plan, get_col_split) and the plans created afterwards pick them up.
//   rows (range compression): the compile-time variants where the length has them (radix sequences and lanes
//     per line), and runtime-radix plans with the factors of the length in different orders and groupings;
//   columns (azimuth): every two-pass split n1 x n2 (and the single pass where it fits LDS).
// Timing: one warm-up launch, then the minimum of `reps` launches bracketed by HIP events.
#include <algorithm>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <set>
#include <string>
#include <vector>

#include "asarf.h"
#include "internal.h"

using namespace asarf;

#define HIP_CHECK(expr)                                                                                    \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return set_error(e_ == hipErrorOutOfMemory ? ASARF_ENOMEM : ASARF_EHIP, "%s: %s (%s:%d)", #expr, \
                             hipGetErrorString(e_), __FILE__, __LINE__);                                   \
    } while (0)

namespace {

// Orderings and groupings of the prime-power factors of n that the runtime kernels accept.
std::vector<std::vector<int>> row_candidates(int n, int lanes) {
    std::vector<int> base;
    std::vector<std::vector<int>> out;
    if (!radix_plan_relaxed(n, lanes, base)) return out;
    std::set<std::vector<int>> seen;
    auto add = [&](std::vector<int> v) {
        if (radix_plan_fits(n, lanes, v) && seen.insert(v).second) out.push_back(v);
    };
    std::vector<std::vector<int>> groupings = {base};
    // regroup powers of two: 16 -> 4 x 4, 8 x 2 -> 16, 4 x 4 -> 16
    {
        std::vector<int> g;
        int p2 = 1;
        for (int r : base)
            if ((r & (r - 1)) == 0) p2 *= r;
            else g.push_back(r);
        for (int big : {16, 8, 4}) {
            std::vector<int> v = g;
            int m = p2;
            while (m % big == 0 && m > 1) {
                v.push_back(big);
                m /= big;
            }
            for (int r : {8, 4, 2})
                while (m % r == 0 && m > 1) {
                    v.push_back(r);
                    m /= r;
                }
            if (m == 1) groupings.push_back(v);
        }
    }
    for (auto g : groupings) {
        add(g);
        std::vector<int> v = g;
        std::reverse(v.begin(), v.end());
        add(v);
        std::sort(v.begin(), v.end());
        add(v);
        std::reverse(v.begin(), v.end());
        add(v);
    }
    return out;
}

struct Timer {
    hipEvent_t a = nullptr, b = nullptr;
    ~Timer() {
        if (a) (void)hipEventDestroy(a);
        if (b) (void)hipEventDestroy(b);
    }
};

}  // namespace

extern "C" int asarf_autotune_fft(asarf_ctx* c, int n, int layout, int batch, int reps, int* choice,
                                  float* best_ms, int* n_candidates) {
    if (!c || n < 2 || batch <= 0 || reps <= 0) return set_error(ASARF_EINVAL, "asarf_autotune_fft: argument");
    if (layout != ASARF_FFT_ROWS && layout != ASARF_FFT_COLUMNS)
        return set_error(ASARF_EINVAL, "asarf_autotune_fft: layout %d", layout);
    int st = ctx_init(c);
    if (st != ASARF_OK) return st;
    hipStream_t s = c->stream;
    Timer t;
    HIP_CHECK(hipEventCreate(&t.a));
    HIP_CHECK(hipEventCreate(&t.b));
    float best = INFINITY;
    int best_idx = -1, ncand = 0;
    bool found = false;

    if (layout == ASARF_FFT_ROWS) {
        // range compression of `batch` lines of zero-valued raw bytes into an n-wide image
        const int nr_in = n / 2;
        void* raw = nullptr;
        float2 *img = nullptr, *h = nullptr;
        HIP_CHECK(hipMalloc(&raw, (size_t)batch * nr_in * 2));
        auto cleanup = [&]() {
            (void)hipFree(raw);
            if (img) (void)hipFree(img);
            if (h) (void)hipFree(h);
        };
        if (hipMalloc(&img, sizeof(float2) * (size_t)batch * n) != hipSuccess ||
            hipMalloc(&h, sizeof(float2) * (size_t)n) != hipSuccess) {
            cleanup();
            return set_error(ASARF_ENOMEM, "asarf_autotune_fft: buffers");
        }
        (void)hipMemsetAsync(raw, 0, (size_t)batch * nr_in * 2, s);
        (void)hipMemsetAsync(h, 0, sizeof(float2) * (size_t)n, s);
        std::vector<std::vector<int>> cands = row_candidates(n, kRowThreads);
        std::vector<const FftPlanHost*> plans;
        float2* tw = get_twiddles(c, n, &st);
        if (!tw) {
            cleanup();
            return st;
        }
        for (auto& r : cands) {
            auto p = std::make_unique<FftPlanHost>();
            p->n = n;
            p->radix = r;
            p->d_tw = tw;
            plans.push_back(p.get());
            c->plans[std::make_pair(n, -1000 - c->tune_key++)] = std::move(p);
        }
        // candidates -1 .. -nv: the compile-time variants of the length (if any)
        const int nv = range_compress_static_variants(n);
        for (int ci = -nv; ci < (int)plans.size(); ++ci) {
            RcArgs a{};
            a.in = raw;
            a.in_stride = nr_in;
            a.nr_in = nr_in;
            a.out = img;
            a.out_stride = n;
            a.rows = batch;
            a.nrv = n - 1;
            a.zero_to = n - 1;
            a.coef = nullptr;
            a.hconj = h;
            FftPlanDev pd{};
            if (ci >= 0) {
                pd = plans[ci]->dev();
            } else if (!plans.empty()) {
                pd = plans[0]->dev();  // the compile-time kernel only reads n and the twiddle table
            } else {
                pd.n = n;
                pd.tw = tw;
            }
            a.plan = pd;
            a.no_static = ci >= 0;
            a.static_variant = ci < 0 ? -1 - ci : 0;
            hipError_t e = launch_range_compress(a, RAW_U8, s);
            if (e != hipSuccess) continue;  // candidate not launchable (e.g. LDS): skip
            float ms_min = INFINITY;
            for (int k = 0; k < reps; ++k) {
                (void)hipEventRecord(t.a, s);
                (void)launch_range_compress(a, RAW_U8, s);
                (void)hipEventRecord(t.b, s);
                (void)hipEventSynchronize(t.b);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, t.a, t.b);
                ms_min = std::min(ms_min, ms);
            }
            ++ncand;
            if (ms_min < best) {
                best = ms_min;
                best_idx = ci;
                found = true;
            }
        }
        HIP_CHECK(hipStreamSynchronize(s));
        cleanup();
        if (!found) return set_error(ASARF_EUNSUPPORTED, "asarf_autotune_fft: no row candidate ran");
        asarf_ctx::RowTune rt;
        rt.use_static = best_idx < 0;
        rt.static_variant = best_idx < 0 ? -1 - best_idx : 0;
        rt.plan = best_idx >= 0 ? plans[best_idx] : nullptr;
        c->tuned_rows[n] = rt;
        if (choice) *choice = best_idx;
    } else {
        // forward passes A and B of an n x batch column transform
        const size_t lds_cap = c->max_lds / (sizeof(float2) * (kColTile + 1));
        const int lanes = kThreads / kColTile;
        std::vector<int> n1s;
        std::vector<int> tmp;
        if ((size_t)n <= lds_cap && radix_plan_relaxed(n, lanes, tmp)) n1s.push_back(n);
        for (int n1 = 2; n1 <= n / 2; ++n1) {
            if (n % n1) continue;
            const int n2 = n / n1;
            if ((size_t)n1 > lds_cap || (size_t)n2 > lds_cap) continue;
            if (radix_plan_relaxed(n1, lanes, tmp) && radix_plan_relaxed(n2, lanes, tmp)) n1s.push_back(n1);
        }
        if (n1s.empty()) return set_error(ASARF_EUNSUPPORTED, "asarf_autotune_fft: no column split for %d", n);
        const long stride = (batch + kColTile - 1) / kColTile * kColTile;
        float2* img = nullptr;
        HIP_CHECK(hipMalloc(&img, sizeof(float2) * (size_t)n * stride));
        (void)hipMemsetAsync(img, 0, sizeof(float2) * (size_t)n * stride, s);
        const auto saved = c->tuned_col_n1.find(n) != c->tuned_col_n1.end() ? c->tuned_col_n1[n] : 0;
        for (size_t i = 0; i < n1s.size(); ++i) {
            c->tuned_col_n1[n] = n1s[i];
            ColSplit cs;
            if (get_col_split(c, n, &cs) != ASARF_OK) continue;
            auto run = [&]() {
                int r = col_forward_pass(c, cs, 0, img, stride, n, img, stride, batch);
                if (r == ASARF_OK && cs.n2 > 1) r = col_forward_pass(c, cs, 1, img, stride, n, img, stride, batch);
                return r;
            };
            if (run() != ASARF_OK) continue;
            float ms_min = INFINITY;
            for (int k = 0; k < reps; ++k) {
                (void)hipEventRecord(t.a, s);
                (void)run();
                (void)hipEventRecord(t.b, s);
                (void)hipEventSynchronize(t.b);
                float ms = 0.f;
                (void)hipEventElapsedTime(&ms, t.a, t.b);
                ms_min = std::min(ms_min, ms);
            }
            ++ncand;
            if (ms_min < best) {
                best = ms_min;
                best_idx = n1s[i];
                found = true;
            }
        }
        (void)hipStreamSynchronize(s);
        (void)hipFree(img);
        if (!found) {
            if (saved) c->tuned_col_n1[n] = saved;
            else c->tuned_col_n1.erase(n);
            return set_error(ASARF_EUNSUPPORTED, "asarf_autotune_fft: no column candidate ran");
        }
        c->tuned_col_n1[n] = best_idx;
        if (choice) *choice = best_idx;
    }
    if (best_ms) *best_ms = best;
    if (n_candidates) *n_candidates = ncand;
    clear_error();
    return ASARF_OK;
}

// ------------------------------------------------------------------------------------------ autotuner cache file
// Text, one choice per line after a header naming the format version and the device architecture:
//   asarf-tune 1 <gcnArchName>
//   row <n> static <variant>          compile-time range-compression variant of length n
//   row <n> radix <r1> <r2> ...       runtime radix order of length n
//   col <n> <n1>                      column split n = n1 x (n / n1) (n1 = n: single pass)
// Lines starting with '#' and blank lines are skipped.

namespace {

int device_arch(asarf_ctx* c, std::string* out) {
    hipDeviceProp_t prop;
    HIP_CHECK(hipGetDeviceProperties(&prop, c->device));
    *out = prop.gcnArchName;
    if (out->empty()) *out = "unknown";
    for (char& ch : *out)
        if (ch == ' ' || ch == '\t' || ch == '\n') ch = '_';
    return ASARF_OK;
}

}  // namespace

extern "C" int asarf_tune_save(asarf_ctx* c, const char* path) {
    if (!c || !path) return set_error(ASARF_EINVAL, "asarf_tune_save: null argument");
    int st = ctx_init(c);
    if (st != ASARF_OK) return st;
    std::string arch;
    if ((st = device_arch(c, &arch)) != ASARF_OK) return st;
    FILE* f = std::fopen(path, "w");
    if (!f) return set_error(ASARF_EINVAL, "asarf_tune_save: cannot open %s for writing", path);
    bool ok = std::fprintf(f, "asarf-tune 1 %s\n", arch.c_str()) > 0;
    for (const auto& kv : c->tuned_rows) {
        const auto& rt = kv.second;
        if (rt.use_static || !rt.plan) {
            ok = ok && std::fprintf(f, "row %d static %d\n", kv.first, rt.static_variant) > 0;
        } else {
            ok = ok && std::fprintf(f, "row %d radix", kv.first) > 0;
            for (int r : rt.plan->radix) ok = ok && std::fprintf(f, " %d", r) > 0;
            ok = ok && std::fprintf(f, "\n") > 0;
        }
    }
    for (const auto& kv : c->tuned_col_n1) ok = ok && std::fprintf(f, "col %d %d\n", kv.first, kv.second) > 0;
    ok = (std::fclose(f) == 0) && ok;
    if (!ok) return set_error(ASARF_EINVAL, "asarf_tune_save: write to %s failed", path);
    return clear_error();
}

extern "C" int asarf_tune_load(asarf_ctx* c, const char* path, int* n_loaded) {
    if (!c || !path) return set_error(ASARF_EINVAL, "asarf_tune_load: null argument");
    if (n_loaded) *n_loaded = 0;
    FILE* f = std::fopen(path, "r");
    if (!f) return set_error(ASARF_EINVAL, "asarf_tune_load: cannot open %s", path);
    // parse everything first (no device call): a malformed file changes nothing in the context
    struct RowEntry {
        int n;
        bool is_static;
        int variant;
        std::vector<int> radix;
    };
    std::vector<RowEntry> rows;
    std::vector<std::pair<int, int>> cols;
    char line[1024];
    int lineno = 0;
    bool header = false;
    std::string file_arch;
    auto bad = [&](const char* what) {
        std::fclose(f);
        return set_error(ASARF_EINVAL, "asarf_tune_load: %s:%d: %s", path, lineno, what);
    };
    while (std::fgets(line, sizeof line, f)) {
        ++lineno;
        if (!std::strchr(line, '\n') && !std::feof(f)) return bad("line too long");
        char* save = nullptr;
        char* tok = strtok_r(line, " \t\r\n", &save);
        if (!tok || tok[0] == '#') continue;
        auto next_int = [&](int* v) {
            char* t = strtok_r(nullptr, " \t\r\n", &save);
            if (!t) return false;
            char* end = nullptr;
            const long x = std::strtol(t, &end, 10);
            if (*end != '\0' || x < 0 || x > (1L << 26)) return false;
            *v = (int)x;
            return true;
        };
        if (!header) {
            int version = 0;
            if (std::strcmp(tok, "asarf-tune") != 0 || !next_int(&version) || version != 1)
                return bad("not an asarf-tune version 1 file");
            const char* a = strtok_r(nullptr, " \t\r\n", &save);
            if (!a) return bad("header: no device architecture");
            file_arch = a;
            header = true;
            continue;
        }
        if (std::strcmp(tok, "row") == 0) {
            RowEntry e{};
            if (!next_int(&e.n) || e.n < 2) return bad("row: length");
            const char* kind = strtok_r(nullptr, " \t\r\n", &save);
            if (kind && std::strcmp(kind, "static") == 0) {
                e.is_static = true;
                if (!next_int(&e.variant) || e.variant >= std::max(1, range_compress_static_variants(e.n)))
                    return bad("row: compile-time variant out of range");
            } else if (kind && std::strcmp(kind, "radix") == 0) {
                int r = 0;
                long prod = 1;
                while (next_int(&r)) {
                    if (r < 2 || (int)e.radix.size() >= kMaxPasses) return bad("row: radix");
                    e.radix.push_back(r);
                    prod *= r;
                }
                if (prod != e.n || !radix_plan_fits(e.n, kRowThreads, e.radix))
                    return bad("row: radix order does not factor the length on the row kernel");
            } else {
                return bad("row: expected 'static' or 'radix'");
            }
            rows.push_back(std::move(e));
        } else if (std::strcmp(tok, "col") == 0) {
            int n = 0, n1 = 0;
            if (!next_int(&n) || !next_int(&n1) || n < 2 || n1 < 2 || n1 > n || n % n1 != 0)
                return bad("col: length and first-pass length");
            cols.emplace_back(n, n1);
        } else {
            return bad("unknown keyword");
        }
    }
    const bool read_err = std::ferror(f) != 0;
    std::fclose(f);
    if (read_err) return set_error(ASARF_EINVAL, "asarf_tune_load: read error on %s", path);
    if (!header) return set_error(ASARF_EINVAL, "asarf_tune_load: %s is empty", path);
    int st = ctx_init(c);
    if (st != ASARF_OK) return st;
    std::string arch;
    if ((st = device_arch(c, &arch)) != ASARF_OK) return st;
    if (arch != file_arch)
        return set_error(ASARF_EINVAL, "asarf_tune_load: %s was tuned on %s, this device is %s", path,
                         file_arch.c_str(), arch.c_str());
    for (const auto& e : rows) {
        asarf_ctx::RowTune rt;
        rt.use_static = e.is_static;
        rt.static_variant = e.is_static ? e.variant : 0;
        if (!e.is_static) {
            auto p = std::make_unique<FftPlanHost>();
            p->n = e.n;
            p->radix = e.radix;
            p->d_tw = get_twiddles(c, e.n, &st);
            if (!p->d_tw) return st;
            rt.plan = p.get();
            c->plans[std::make_pair(e.n, -1000 - c->tune_key++)] = std::move(p);
        }
        c->tuned_rows[e.n] = rt;
    }
    for (const auto& kv : cols) c->tuned_col_n1[kv.first] = kv.second;
    if (n_loaded) *n_loaded = (int)(rows.size() + cols.size());
    return clear_error();
}

extern "C" int asarf_fft_static_column_length(int n) { return col_static_length(n) ? 1 : 0; }

extern "C" int asarf_fft_fast_range_size(int n) {
    const int pad = asarf_fft_padding_size(n);
    if (pad <= 0) return pad;
    for (int len : {6144, 6400, 8192})
        if (len >= pad && 2L * len <= 3L * pad && range_compress_static_variants(len) > 0) return len;
    return pad;
}
