// Autotuner: measures candidate transform schedules on the device and records the fastest per length in the
// context, where the planner (get_plan, get_col_split) and the plans created afterwards pick them up.
//   rows (range compression): the compile-time variants where the length has them (radix sequences and lanes
//     per line), and runtime-radix plans with the factors of the length in different orders and groupings;
//   columns (azimuth): every two-pass split n1 x n2 (and the single pass where it fits LDS).
// Timing: one warm-up launch, then the minimum of `reps` launches bracketed by HIP events.
#include <algorithm>
#include <cmath>
#include <set>
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
