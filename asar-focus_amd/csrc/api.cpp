// C ABI of libasarf.so: context, memory model, FFT planning, per-stage entry points and the fused plan.
// Every entry point validates its arguments before touching the device.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <vector>

#include "asarf.h"
#include "internal.h"
#include "poly.h"

using namespace asarf;

// ---------------------------------------------------------------------------------------------- errors ----
namespace asarf {
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

#define HIP_CHECK(expr)                                                                                   \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return set_error(e_ == hipErrorOutOfMemory ? ASARF_ENOMEM : ASARF_EHIP, "%s: %s (%s:%d)", #expr, \
                             hipGetErrorString(e_), __FILE__, __LINE__);                                  \
    } while (0)

#define ST_CHECK(expr)             \
    do {                           \
        int s_ = (expr);           \
        if (s_ != ASARF_OK) return s_; \
    } while (0)

extern "C" const char* asarf_last_error(void) { return t_err; }
extern "C" const char* asarf_version(void) { return "asarf 0.1.0 (gfx950)"; }

// ---------------------------------------------------------------------------------------------- context ----
extern "C" int asarf_ctx_create(int device, void* stream, asarf_ctx** out) {
    if (!out) return set_error(ASARF_EINVAL, "asarf_ctx_create: out is null");
    if (device < 0) return set_error(ASARF_EINVAL, "asarf_ctx_create: device %d", device);
    auto* c = new asarf_ctx();
    c->device = device;
    c->stream = static_cast<hipStream_t>(stream);
    c->own_stream = false;
    *out = c;
    return clear_error();
}

namespace asarf {
int ctx_init(asarf_ctx* c) {
    if (c->initialized) return ASARF_OK;
    int n = 0;
    hipError_t e = hipGetDeviceCount(&n);
    if (e != hipSuccess || n <= c->device)
        return set_error(ASARF_EHIP, "no HIP device %d (%d visible): %s", c->device, n, hipGetErrorString(e));
    HIP_CHECK(hipSetDevice(c->device));
    int lds = 0;
    HIP_CHECK(hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, c->device));
    c->max_lds = (size_t)lds;
    // ASARF_COL_STATIC=0 keeps the runtime-radix column kernels for every length (comparison runs)
    const char* cs_env = std::getenv("ASARF_COL_STATIC");
    c->col_static = !(cs_env && cs_env[0] == '0');
    // ASARF_COL_WIDTH=16|32 forces the column tile width of the compile-time column kernels (default: by pass)
    const char* cw_env = std::getenv("ASARF_COL_WIDTH");
    const int cw = cw_env ? std::atoi(cw_env) : 0;
    c->col_width = (cw == 16 || cw == 32) ? cw : 0;
    c->initialized = true;
    return ASARF_OK;
}

int ctx_scratch(asarf_ctx* c, size_t bytes, void** out) {
    if (bytes > c->scratch_bytes) {
        if (c->d_scratch) {
            HIP_CHECK(hipStreamSynchronize(c->stream));
            HIP_CHECK(hipFree(c->d_scratch));
            c->d_scratch = nullptr;
            c->scratch_bytes = 0;
        }
        const size_t want = std::max(bytes, (size_t)1 << 20);
        HIP_CHECK(hipMalloc(&c->d_scratch, want));
        c->scratch_bytes = want;
    }
    *out = c->d_scratch;
    return ASARF_OK;
}

int ctx_pinned(asarf_ctx* c, size_t bytes, void** out) {
    if (bytes > c->pinned_bytes) {
        if (c->h_pinned) {
            HIP_CHECK(hipStreamSynchronize(c->stream));
            HIP_CHECK(hipHostFree(c->h_pinned));
            c->h_pinned = nullptr;
            c->pinned_bytes = 0;
        }
        const size_t want = std::max(bytes, (size_t)1 << 16);
        HIP_CHECK(hipHostMalloc(&c->h_pinned, want, hipHostMallocDefault));
        c->pinned_bytes = want;
    }
    *out = c->h_pinned;
    return ASARF_OK;
}

// Radix order: 16s first, then one 8/4/2 for the remaining power of two, then 3, 5, 7.
bool radix_plan(int n, int lanes, std::vector<int>& out) {
    out.clear();
    if (n < 2) return false;
    int m = n;
    while (m % 16 == 0) {
        out.push_back(16);
        m /= 16;
    }
    if (m % 8 == 0) {
        out.push_back(8);
        m /= 8;
    } else if (m % 4 == 0) {
        out.push_back(4);
        m /= 4;
    } else if (m % 2 == 0) {
        out.push_back(2);
        m /= 2;
    }
    for (int f : {3, 5, 7})
        while (m % f == 0) {
            out.push_back(f);
            m /= f;
        }
    if (m != 1 || (int)out.size() > kMaxPasses) return false;
    // every pass must fit the per-lane register budget: ceil((n/R)/lanes) butterflies of R elements <= kMaxE
    for (int r : out) {
        const int nb = (n / r + lanes - 1) / lanes;
        if (nb > kMaxE / r) {
            // split a 16 into 4x4 or an 8 into 4x2 and retry once
            return false;
        }
    }
    return true;
}

bool radix_plan_fits(int n, int lanes, const std::vector<int>& radix) {
    if (radix.empty() || (int)radix.size() > kMaxPasses) return false;
    long prod = 1;
    for (int r : radix) {
        if (r != 2 && r != 3 && r != 4 && r != 5 && r != 7 && r != 8 && r != 16) return false;
        const int nb = (n / r + lanes - 1) / lanes;
        if (nb > kMaxE / r) return false;
        prod *= r;
    }
    return prod == n;
}

bool radix_plan_relaxed(int n, int lanes, std::vector<int>& out) {
    if (radix_plan(n, lanes, out)) return true;
    // replace 16s by 4x4 (more passes, smaller register blocks)
    std::vector<int> alt;
    for (int r : out) {
        if (r == 16) {
            alt.push_back(4);
            alt.push_back(4);
        } else {
            alt.push_back(r);
        }
    }
    if ((int)alt.size() > kMaxPasses) return false;
    for (int r : alt) {
        const int nb = (n / r + lanes - 1) / lanes;
        if (nb > kMaxE / r) return false;
    }
    out = alt;
    return !out.empty();
}

float2* get_twiddles(asarf_ctx* c, int n, int* status) {
    auto it = c->tw_tables.find(n);
    if (it != c->tw_tables.end()) return it->second;
    std::vector<float2> h(n);
    for (int m = 0; m < n; ++m) {
        const double a = -2.0 * M_PI * (double)m / (double)n;
        h[m] = make_float2((float)std::cos(a), (float)std::sin(a));
    }
    float2* d = nullptr;
    hipError_t e = hipMalloc(&d, sizeof(float2) * (size_t)n);
    if (e != hipSuccess) {
        *status = set_error(ASARF_ENOMEM, "twiddle table %d: %s", n, hipGetErrorString(e));
        return nullptr;
    }
    e = hipMemcpy(d, h.data(), sizeof(float2) * (size_t)n, hipMemcpyHostToDevice);
    if (e != hipSuccess) {
        *status = set_error(ASARF_EHIP, "twiddle upload: %s", hipGetErrorString(e));
        return nullptr;
    }
    c->tw_tables[n] = d;
    return d;
}

const FftPlanHost* get_plan(asarf_ctx* c, int n, int lanes, int* status) {
    if (lanes == kRowThreads) {
        auto t = c->tuned_rows.find(n);
        if (t != c->tuned_rows.end() && t->second.plan) return t->second.plan;
    }
    const auto key = std::make_pair(n, lanes);
    auto it = c->plans.find(key);
    if (it != c->plans.end()) return it->second.get();
    auto p = std::make_unique<FftPlanHost>();
    p->n = n;
    if (!radix_plan_relaxed(n, lanes, p->radix)) {
        // a column length with a compile-time kernel needs only the twiddle table (no runtime radix plan)
        if (!(lanes == kThreads / kColTile && c->col_static && col_static_length(n))) {
            *status = set_error(ASARF_EUNSUPPORTED, "no Stockham plan for length %d with %d lanes per transform", n,
                                lanes);
            return nullptr;
        }
        p->radix.clear();
    }
    p->d_tw = get_twiddles(c, n, status);
    if (!p->d_tw) return nullptr;
    const FftPlanHost* raw = p.get();
    c->plans[key] = std::move(p);
    return raw;
}

int get_col_split(asarf_ctx* c, int n, ColSplit* cs, int n2_max) {
    const int lanes = kThreads / kColTile;
    const size_t lds_cap = c->max_lds / (sizeof(float2) * (kColTile + 1));
    std::vector<int> tmp;
    cs->n = n;
    int st = ASARF_OK;
    const auto tuned_it = c->tuned_col_n1.find(n);
    const int tuned = (n2_max <= 0 && tuned_it != c->tuned_col_n1.end()) ? tuned_it->second : 0;
    if (n2_max <= 0 && (tuned == 0 || tuned == n) && (size_t)n <= lds_cap && radix_plan_relaxed(n, lanes, tmp)) {
        cs->n1 = n;
        cs->n2 = 1;
        cs->p1 = get_plan(c, n, lanes, &st);
        cs->p2 = nullptr;
        cs->full = nullptr;
        return cs->p1 ? ASARF_OK : st;
    }
    // n2_max <= 0: the autotuned split if there is one, else the most balanced split (smallest max(n1, n2));
    // otherwise the largest n2 <= n2_max
    int best1 = 0, best2 = 0;
    if (tuned > 1 && tuned < n && n % tuned == 0) {
        best1 = tuned;
        best2 = n / tuned;
    }
    const bool use_tuned = best1 != 0;
    // a length is usable with a runtime radix plan, or with a compile-time column kernel when those are enabled
    auto usable = [&](int m) { return radix_plan_relaxed(m, lanes, tmp) || (c->col_static && col_static_length(m)); };
    for (int n1 = 2; n1 <= n / 2 && !use_tuned; ++n1) {
        if (n % n1) continue;
        const int n2 = n / n1;
        if ((size_t)n1 > lds_cap || (size_t)n2 > lds_cap) continue;
        if (n2_max > 0 && n2 > n2_max) continue;
        if (!usable(n1) || !usable(n2)) continue;
        bool better;
        if (!best1)
            better = true;
        else if (n2_max > 0)
            better = n2 > best2;
        else
            better = std::max(n1, n2) < std::max(best1, best2) ||
                     (std::max(n1, n2) == std::max(best1, best2) && n1 < best1);
        if (better) {
            best1 = n1;
            best2 = n2;
        }
    }
    if (!best1) return set_error(ASARF_EUNSUPPORTED, "no two-pass split for column transform length %d", n);
    cs->n1 = best1;
    cs->n2 = best2;
    cs->p1 = get_plan(c, best1, lanes, &st);
    if (!cs->p1) return st;
    cs->p2 = get_plan(c, best2, lanes, &st);
    if (!cs->p2) return st;
    // the inter-pass twiddle table of length n, cached with the plans (entries are never replaced, so splits
    // handed out earlier stay valid)
    const auto key = std::make_pair(n, -1);
    auto it = c->plans.find(key);
    if (it == c->plans.end()) {
        auto full = std::make_unique<FftPlanHost>();
        full->n = n;
        full->d_tw = get_twiddles(c, n, &st);
        if (!full->d_tw) return st;
        it = c->plans.emplace(key, std::move(full)).first;
    }
    cs->full = it->second.get();
    return ASARF_OK;
}
}  // namespace asarf

extern "C" int asarf_ctx_destroy(asarf_ctx* c) {
    if (!c) return set_error(ASARF_EINVAL, "asarf_ctx_destroy: null");
    if (c->initialized) {
        (void)hipSetDevice(c->device);
        (void)hipStreamSynchronize(c->stream);
        for (auto& kv : c->tw_tables) (void)hipFree(kv.second);
        if (c->d_scratch) (void)hipFree(c->d_scratch);
        if (c->h_pinned) (void)hipHostFree(c->h_pinned);
        if (c->own_stream && c->stream) (void)hipStreamDestroy(c->stream);
    }
    delete c;
    return clear_error();
}

extern "C" int asarf_sync(asarf_ctx* c) {
    if (!c) return set_error(ASARF_EINVAL, "asarf_sync: null ctx");
    if (!c->initialized) return clear_error();
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return clear_error();
}

extern "C" void* asarf_ctx_stream(asarf_ctx* c) {
    if (!c) return nullptr;
    if (ctx_init(c) != ASARF_OK) return nullptr;
    return c->stream;
}

// ---------------------------------------------------------------------------------------- memory model ----
static size_t elem_bytes(int dtype) { return dtype == ASARF_F32 ? 4 : 8; }

extern "C" size_t asarf_image_bytes(const asarf_image* img) {
    if (!img) return 0;
    return (size_t)img->x_stride * (size_t)img->y_stride * elem_bytes(img->dtype);
}

static int check_geometry(int xs, int ys, int xstr, int ystr, int dtype, const char* who) {
    if (dtype != ASARF_CF32 && dtype != ASARF_F32) return set_error(ASARF_EINVAL, "%s: dtype %d", who, dtype);
    if (xs < 0 || ys < 0 || xstr <= 0 || ystr <= 0) return set_error(ASARF_ESHAPE, "%s: negative size", who);
    if (xstr < xs || ystr < ys) return set_error(ASARF_ESHAPE, "%s: Stride smaller than dimension", who);
    return ASARF_OK;
}

extern "C" int asarf_image_alloc(asarf_ctx* c, int xs, int ys, int xstr, int ystr, int dtype, asarf_image* out) {
    if (!c || !out) return set_error(ASARF_EINVAL, "asarf_image_alloc: null argument");
    ST_CHECK(check_geometry(xs, ys, xstr, ystr, dtype, "asarf_image_alloc"));
    ST_CHECK(ctx_init(c));
    const size_t bytes = (size_t)xstr * ystr * elem_bytes(dtype);
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, bytes));
    HIP_CHECK(hipMemsetAsync(p, 0, bytes, c->stream));
    *out = asarf_image{p, xs, ys, xstr, ystr, dtype};
    return clear_error();
}

extern "C" int asarf_image_free(asarf_ctx* c, asarf_image* img) {
    if (!c || !img) return set_error(ASARF_EINVAL, "asarf_image_free: null argument");
    if (img->data) {
        ST_CHECK(ctx_init(c));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        HIP_CHECK(hipFree(img->data));
    }
    *img = asarf_image{nullptr, 0, 0, 0, 0, img->dtype};
    return clear_error();
}

extern "C" int asarf_image_zero(asarf_ctx* c, asarf_image* img) {
    if (!c || !img || !img->data) return set_error(ASARF_EINVAL, "asarf_image_zero: null argument");
    ST_CHECK(ctx_init(c));
    HIP_CHECK(hipMemsetAsync(img->data, 0, asarf_image_bytes(img), c->stream));
    return clear_error();
}

extern "C" int asarf_workspace_alloc(asarf_ctx* c, size_t bytes, asarf_workspace* out) {
    if (!c || !out) return set_error(ASARF_EINVAL, "asarf_workspace_alloc: null argument");
    if (bytes == 0) return set_error(ASARF_EINVAL, "asarf_workspace_alloc: zero bytes");
    ST_CHECK(ctx_init(c));
    void* p = nullptr;
    HIP_CHECK(hipMalloc(&p, bytes));
    *out = asarf_workspace{p, bytes};
    return clear_error();
}

extern "C" int asarf_workspace_free(asarf_ctx* c, asarf_workspace* ws) {
    if (!c || !ws) return set_error(ASARF_EINVAL, "asarf_workspace_free: null argument");
    if (ws->data) {
        ST_CHECK(ctx_init(c));
        HIP_CHECK(hipStreamSynchronize(c->stream));
        HIP_CHECK(hipFree(ws->data));
    }
    *ws = asarf_workspace{nullptr, 0};
    return clear_error();
}

extern "C" int asarf_image_adopt(asarf_image* img, int xs, int ys, int xstr, int ystr, int dtype, asarf_workspace* ws) {
    if (!img || !ws) return set_error(ASARF_EINVAL, "asarf_image_adopt: null argument");
    if (img->data) return set_error(ASARF_EINVAL, "asarf_image_adopt: DevicePaddedImage already initialized");
    ST_CHECK(check_geometry(xs, ys, xstr, ystr, dtype, "asarf_image_adopt"));
    const size_t need = (size_t)xstr * ystr * elem_bytes(dtype);
    if (!ws->data || ws->bytes < need)
        return set_error(ASARF_EWORKSPACE, "asarf_image_adopt: External workspace does not have enough memory "
                                           "(%zu < %zu)", ws->bytes, need);
    *img = asarf_image{ws->data, xs, ys, xstr, ystr, dtype};
    *ws = asarf_workspace{nullptr, 0};
    return clear_error();
}

extern "C" int asarf_image_release(asarf_image* img, asarf_workspace* ws) {
    if (!img || !ws) return set_error(ASARF_EINVAL, "asarf_image_release: null argument");
    if (ws->data) return set_error(ASARF_EINVAL, "asarf_image_release: workspace not empty");
    *ws = asarf_workspace{img->data, asarf_image_bytes(img)};
    *img = asarf_image{nullptr, 0, 0, 0, 0, img->dtype};
    return clear_error();
}

extern "C" int asarf_image_subset_copy(asarf_ctx* c, const asarf_image* img, int x0, int y0, int w, int h, void* dst,
                                       size_t pitch) {
    if (!c || !img || !img->data || !dst) return set_error(ASARF_EINVAL, "asarf_image_subset_copy: null argument");
    if (x0 < 0 || y0 < 0 || w <= 0 || h <= 0 || x0 + w > img->x_size || y0 + h > img->y_size)
        return set_error(ASARF_ESHAPE, "asarf_image_subset_copy: window (%d,%d,%d,%d) outside %dx%d", x0, y0, w, h,
                         img->x_size, img->y_size);
    const size_t eb = elem_bytes(img->dtype);
    if (pitch < (size_t)w * eb) return set_error(ASARF_ESHAPE, "asarf_image_subset_copy: pitch %zu too small", pitch);
    ST_CHECK(ctx_init(c));
    const char* src = static_cast<const char*>(img->data) + ((size_t)y0 * img->x_stride + x0) * eb;
    HIP_CHECK(hipMemcpy2DAsync(dst, pitch, src, (size_t)img->x_stride * eb, (size_t)w * eb, h, hipMemcpyDefault,
                               c->stream));
    return clear_error();
}

// ---------------------------------------------------------------------------------------------- stages ----
int asarf::check_cf32(const asarf_image* img, const char* who) {
    if (!img || !img->data) return set_error(ASARF_EINVAL, "%s: null image", who);
    if (img->dtype != ASARF_CF32) return set_error(ASARF_EINVAL, "%s: image must be cf32", who);
    ST_CHECK(check_geometry(img->x_size, img->y_size, img->x_stride, img->y_stride, img->dtype, who));
    if (img->x_size <= 0 || img->y_size <= 0) return set_error(ASARF_ESHAPE, "%s: empty image", who);
    return ASARF_OK;
}

// I/Q correction coefficients from logical and padding moments (see k_stats.hip for the derivation).
static IqCoef iq_coefficients(const double* t, int xs, int ys, size_t n_total, int quirks, asarf_iq_results* r) {
    const double nL = t[0], SI = t[1], SQ = t[2], SII = t[3], SQQ = t[4], SIQ = t[5];
    const double nP = t[6], PI_ = t[7], PQ = t[8], PII = t[9], PQQ = t[10], PIQ = t[11];
    const double ntot = n_total ? (double)n_total : (double)xs * (double)ys;
    const double dci = SI / ntot, dcq = SQ / ntot;
    r->dc_i = dci;
    r->dc_q = dcq;
    r->stddev_i = std::sqrt((SII - 2.0 * dci * SI + nL * dci * dci) / ntot);
    r->stddev_q = std::sqrt((SQQ - 2.0 * dcq * SQ + nL * dcq * dcq) / ntot);
    double n = nL, sI = SI, sQ = SQ, sII = SII, sQQ = SQQ, sIQ = SIQ;
    if (quirks & ASARF_Q1_IQ_PADDED_WIDTH) {
        n += nP;
        sI += PI_;
        sQ += PQ;
        sII += PII;
        sQQ += PQQ;
        sIQ += PIQ;
    }
    const double aii = sII - 2.0 * dci * sI + n * dci * dci;
    const double aqq = sQQ - 2.0 * dcq * sQ + n * dcq * dcq;
    const double g = std::sqrt(aii / aqq);
    const double aiq = (sIQ - dcq * sI - dci * sQ + n * dci * dcq) / g;
    const double d = aiq / (aii / (g * g));
    const double cosc = (quirks & ASARF_Q2_COS_NO_SQRT) ? (1.0 - d * d) : std::sqrt(1.0 - d * d);
    r->iq_gain = g;
    r->phase_error = std::asin(d) * 180.0 / M_PI;
    IqCoef c;
    c.dc_i = (float)dci;
    c.dc_q = (float)dcq;
    c.inv_gain = (float)(1.0 / g);
    c.sin_corr = (float)d;
    c.inv_cos = (float)(1.0 / cosc);
    c.identity = 0;
    return c;
}

extern "C" int asarf_iq_correction(asarf_ctx* c, asarf_image* img, size_t n_total, int quirks, asarf_iq_results* res) {
    if (!c || !res) return set_error(ASARF_EINVAL, "asarf_iq_correction: null argument");
    ST_CHECK(check_cf32(img, "asarf_iq_correction"));
    if (quirks & ~ASARF_QUIRKS_REFERENCE) return set_error(ASARF_EINVAL, "asarf_iq_correction: unknown quirk bits");
    ST_CHECK(ctx_init(c));
    const int nwg = std::min(img->y_size, 1024);
    void* scratch = nullptr;
    ST_CHECK(ctx_scratch(c, sizeof(double) * (size_t)(nwg + 1) * 12, &scratch));
    double* partials = static_cast<double*>(scratch);
    double* totals = partials + (size_t)nwg * 12;
    auto* f2 = static_cast<float2*>(img->data);
    HIP_CHECK(launch_moments_cf32(f2, img->x_stride, img->x_size, img->y_size, img->x_stride, partials, nwg, c->stream));
    HIP_CHECK(launch_reduce_f64(partials, nwg, 12, totals, c->stream));
    void* hp = nullptr;
    ST_CHECK(ctx_pinned(c, sizeof(double) * 12, &hp));
    HIP_CHECK(hipMemcpyAsync(hp, totals, sizeof(double) * 12, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    double t[12];
    std::memcpy(t, hp, sizeof(t));
    const IqCoef coef = iq_coefficients(t, img->x_size, img->y_size, n_total, quirks, res);
    const int width = (quirks & ASARF_Q1_IQ_PADDED_WIDTH) ? img->x_stride : img->x_size;
    HIP_CHECK(launch_iq_apply_cf32(f2, img->x_stride, img->x_size, img->y_size, img->x_stride, img->y_stride, width,
                                   coef, c->stream));
    return clear_error();
}

namespace asarf {
int dc_layout(const asarf_image* img, double prf, int quirks, const char* who, DcLayout* L) {
    ST_CHECK(check_cf32(img, who));
    if (!(prf > 0.0)) return set_error(ASARF_EINVAL, "%s: prf must be > 0", who);
    const int xs = img->x_size, ys = img->y_size;
    const int step = xs / 32;
    if (step <= 0) return set_error(ASARF_ESHAPE, "%s: range size %d < 32", who, xs);
    int nb = 0;
    for (int xb = 0; xb < xs; xb += step) {
        if (!(quirks & ASARF_Q4_DC_EXTRA_BLOCK) && xb + step > xs) break;
        ++nb;
    }
    if (nb > kMaxDcBlocks) return set_error(ASARF_EUNSUPPORTED, "%s: %d blocks", who, nb);
    if (ys < 2) return set_error(ASARF_ESHAPE, "%s: need at least 2 lines", who);
    L->step = step;
    L->nb = nb;
    L->nstrips = std::max(1, std::min(ys - 1, kDcMaxPartials / nb));
    L->nvals = (size_t)L->nstrips * nb * 2;
    return ASARF_OK;
}

int dc_run(asarf_ctx* c, const asarf_image* img, double prf, const DcLayout& L, double* d_partials,
           double* h_partials, double* poly3, double* block_freq, double* block_x, int max_blocks, int* n_blocks) {
    const int nb = L.nb, nstrips = L.nstrips;
    HIP_CHECK(launch_dc_cf32(static_cast<const float2*>(img->data), img->x_stride, img->y_size, 0, L.step, nb,
                             img->x_stride, nstrips, d_partials, c->stream));
    HIP_CHECK(hipMemcpyAsync(h_partials, d_partials, sizeof(double) * L.nvals, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    double fx[kMaxDcBlocks], fy[kMaxDcBlocks];
    for (int b = 0; b < nb; ++b) {
        double re = 0.0, im = 0.0;
        for (int s = 0; s < nstrips; ++s) {
            re += h_partials[((size_t)s * nb + b) * 2];
            im += h_partials[((size_t)s * nb + b) * 2 + 1];
        }
        const int xb = b * L.step;
        fx[b] = (double)((xb + xb + L.step) / 2);
        fy[b] = prf * std::atan2(im, re) / (2.0 * M_PI);
        if (block_freq && b < max_blocks) block_freq[b] = fy[b];
        if (block_x && b < max_blocks) block_x[b] = fx[b];
    }
    if (n_blocks) *n_blocks = nb;
    if (polyfit_qr(fx, fy, nb, 2, poly3) != 0) return set_error(ASARF_EINVAL, "Doppler centroid: polynomial fit failed");
    return ASARF_OK;
}
}  // namespace asarf

extern "C" int asarf_doppler_centroid(asarf_ctx* c, const asarf_image* img, double prf, int quirks, double* poly3,
                                      double* block_freq, double* block_x, int max_blocks, int* n_blocks) {
    if (!c || !poly3) return set_error(ASARF_EINVAL, "asarf_doppler_centroid: null argument");
    DcLayout L;
    ST_CHECK(dc_layout(img, prf, quirks, "asarf_doppler_centroid", &L));
    ST_CHECK(ctx_init(c));
    void* scratch = nullptr;
    ST_CHECK(ctx_scratch(c, sizeof(double) * L.nvals, &scratch));
    void* hp = nullptr;
    ST_CHECK(ctx_pinned(c, sizeof(double) * L.nvals, &hp));
    ST_CHECK(dc_run(c, img, prf, L, static_cast<double*>(scratch), static_cast<double*>(hp), poly3, block_freq,
                    block_x, max_blocks, n_blocks));
    return clear_error();
}

// conj(FFT(chirp)) / n into d_h (device, n complex). Uses d_h as the transform buffer.
static int chirp_spectrum(asarf_ctx* c, const float* chirp_iq, int n, float2* d_h) {
    int st = ASARF_OK;
    const FftPlanHost* p = get_plan(c, n, kRowThreads, &st);
    if (!p) return st;
    std::vector<float2> h(n);
    for (int i = 0; i < n; ++i) h[i] = make_float2(chirp_iq[2 * i], -chirp_iq[2 * i + 1]);
    // conj(FFT(x)) = conj(sum x W^nk) = inverse-direction transform of conj(x), i.e. sum conj(x) W^-nk
    HIP_CHECK(hipMemcpyAsync(d_h, h.data(), sizeof(float2) * (size_t)n, hipMemcpyHostToDevice, c->stream));
    RowFftArgs a{};
    a.in = d_h;
    a.in_stride = n;
    a.out = d_h;
    a.out_stride = n;
    a.rows = 1;
    a.scale = 1.0f / (float)n;
    a.plan = p->dev();
    HIP_CHECK(launch_row_fft(a, true, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));  // h (host) must outlive the copy
    return ASARF_OK;
}

extern "C" int asarf_range_compression(asarf_ctx* c, asarf_image* img, const float* chirp_iq, int n_chirp, int nc,
                                       asarf_workspace* ws) {
    if (!c || !chirp_iq) return set_error(ASARF_EINVAL, "asarf_range_compression: null argument");
    ST_CHECK(check_cf32(img, "asarf_range_compression"));
    if (n_chirp != img->x_stride)
        return set_error(ASARF_ESHAPE, "asarf_range_compression: Range padding and chirp size mismatch (%d vs %d)",
                         n_chirp, img->x_stride);
    if (nc <= 0 || nc >= img->x_size)
        return set_error(ASARF_EINVAL, "asarf_range_compression: chirp samples %d vs range size %d", nc, img->x_size);
    ST_CHECK(ctx_init(c));
    int st = ASARF_OK;
    const FftPlanHost* p = get_plan(c, n_chirp, kRowThreads, &st);
    if (!p) return st;
    void* scratch = nullptr;
    ST_CHECK(ctx_scratch(c, sizeof(float2) * (size_t)n_chirp, &scratch));
    float2* d_h = static_cast<float2*>(scratch);
    ST_CHECK(chirp_spectrum(c, chirp_iq, n_chirp, d_h));
    RcArgs a{};
    a.in = img->data;
    a.in_stride = img->x_stride;
    a.nr_in = img->x_stride;
    a.out = static_cast<float2*>(img->data);
    a.out_stride = img->x_stride;
    a.rows = img->y_size;
    a.nrv = img->x_size - nc;
    a.zero_to = img->x_stride;
    a.coef = nullptr;
    a.hconj = d_h;
    a.plan = p->dev();
    HIP_CHECK(launch_range_compress(a, RAW_CF32, c->stream));
    img->x_size = a.nrv;
    return clear_error();
}

namespace asarf {
ColFftArgs col_args(const asarf_ctx* c, const float2* in, long in_stride, float2* out, long out_stride, int ncols,
                    int groups, long in_g, long in_m, int in_rows, long out_g, long out_m, int out_rows,
                    const float2* twf, float scale, const FftPlanHost* sub) {
    ColFftArgs a{};
    a.no_static = c->col_static ? 0 : 1;
    a.col_width = c->col_width;
    a.in = in;
    a.in_stride = in_stride;
    a.out = out;
    a.out_stride = out_stride;
    a.ncols = ncols;
    a.ntiles = (ncols + kColTile - 1) / kColTile;
    a.groups = groups;
    a.in_g = in_g;
    a.in_m = in_m;
    a.in_rows = in_rows;
    a.out_g = out_g;
    a.out_m = out_m;
    a.out_rows = out_rows;
    a.tw_full = twf;
    a.scale = scale;
    a.sub = sub->dev();
    return a;
}
}  // namespace asarf

namespace asarf {
static ColFftArgs with_mfma(int mfma, ColFftArgs a) {
    a.mfma = mfma;
    return a;
}

// Forward column transform into permuted storage: row k1*n2 + k2 holds bin k1 + n1*k2. pass 0 = A (strided
// length-n1 transforms with the inter-pass twiddle), pass 1 = B (contiguous length-n2 transforms). Each pass reads
// `in` (row pitch in_stride; pass 0 reads rows >= in_rows as zero, pass 1 reads all n rows) and writes img; in
// and img may be the same buffer (in place) or two buffers (out of place, which streams slightly faster).
int col_forward_pass(asarf_ctx* c, const ColSplit& cs, int pass, const float2* in, long in_stride, int in_rows,
                     float2* img, long stride, int ncols, int mfma, const int* fused_gate) {
    const int n = cs.n;
    if (cs.n2 == 1) {
        if (pass == 0) {
            const ColFftArgs a = col_args(c, in, in_stride, img, stride, ncols, 1, 0, 1, in_rows, 0, 1, n, nullptr, 1.f,
                                          cs.p1);
            HIP_CHECK(launch_col_fft(with_mfma(mfma, a), false, c->stream));
        }
        return ASARF_OK;
    }
    if (pass == 0) {
        const ColFftArgs a = col_args(c, in, in_stride, img, stride, ncols, cs.n2, 1, cs.n2, in_rows, 1, cs.n2, n,
                                      cs.full->d_tw, 1.f, cs.p1);
        HIP_CHECK(launch_col_fft(with_mfma(mfma, a), false, c->stream));
    } else {
        ColFftArgs a = col_args(c, in, in_stride, img, stride, ncols, cs.n1, cs.n2, 1, n, cs.n2, 1, n, nullptr, 1.f,
                                cs.p2);
        a.fused_gate = fused_gate;  // pass B gated off where the fused azimuth core runs it
        HIP_CHECK(launch_col_fft(with_mfma(mfma, a), false, c->stream));
    }
    return ASARF_OK;
}

// Out-of-place inverse pass of a four-step split (n2 > 1): pass 0 = B' (in -> out, inverse inter-pass twiddle),
// pass 1 = A' (in -> out, scale, rows < out_rows written).
int col_inverse_pass_to(asarf_ctx* c, const ColSplit& cs, int pass, const float2* in, long in_stride, float2* out,
                        long out_stride, int ncols, int out_rows, float scale) {
    const int n = cs.n;
    if (cs.n2 == 1) return set_error(ASARF_EINVAL, "col_inverse_pass_to: no four-step split");
    const ColFftArgs a =
        pass == 0 ? col_args(c, in, in_stride, out, out_stride, ncols, cs.n1, cs.n2, 1, n, cs.n2, 1, n, cs.full->d_tw,
                             1.f, cs.p2)
                  : col_args(c, in, in_stride, out, out_stride, ncols, cs.n2, 1, cs.n2, n, 1, cs.n2, out_rows, nullptr,
                             scale, cs.p1);
    HIP_CHECK(launch_col_fft(a, true, c->stream));
    return ASARF_OK;
}

// Inverse column transform from permuted storage; writes rows < out_rows scaled by `scale`.
// pass 0 = B' (contiguous, inverse inter-pass twiddle), pass 1 = A' (strided, scale and row limit).
int col_inverse_pass(asarf_ctx* c, const ColSplit& cs, int pass, float2* img, long stride, int ncols, int out_rows,
                     float scale, int mfma) {
    const int n = cs.n;
    if (cs.n2 == 1) {
        if (pass == 0) {
            const ColFftArgs a = col_args(c, img, stride, img, stride, ncols, 1, 0, 1, n, 0, 1, out_rows, nullptr, scale,
                                          cs.p1);
            HIP_CHECK(launch_col_fft(with_mfma(mfma, a), true, c->stream));
        }
        return ASARF_OK;
    }
    if (pass == 0) {
        const ColFftArgs a = col_args(c, img, stride, img, stride, ncols, cs.n1, cs.n2, 1, n, cs.n2, 1, n, cs.full->d_tw,
                                      1.f, cs.p2);
        HIP_CHECK(launch_col_fft(with_mfma(mfma, a), true, c->stream));
    } else {
        const ColFftArgs a = col_args(c, img, stride, img, stride, ncols, cs.n2, 1, cs.n2, n, 1, cs.n2, out_rows, nullptr,
                                      scale, cs.p1);
        HIP_CHECK(launch_col_fft(with_mfma(mfma, a), true, c->stream));
    }
    return ASARF_OK;
}
}  // namespace asarf

extern "C" int asarf_range_doppler(asarf_ctx* c, const asarf_rda_params* p, asarf_image* src, asarf_image* out,
                                   asarf_workspace* ws) {
    if (!c || !p || !out) return set_error(ASARF_EINVAL, "asarf_range_doppler: null argument");
    ST_CHECK(check_cf32(src, "asarf_range_doppler"));
    if (!(p->prf > 0) || !(p->wavelength > 0) || !(p->range_spacing > 0) || !(p->slant_range_first > 0) ||
        !(p->az_bandwidth_fraction > 0))
        return set_error(ASARF_EINVAL, "asarf_range_doppler: non-positive radar parameter");
    for (int i = 0; i < 3; ++i)
        if (!std::isfinite(p->dc_poly[i]) || !std::isfinite(p->vr_poly[i]))
            return set_error(ASARF_EINVAL, "asarf_range_doppler: non-finite polynomial");
    const double vr0 = polyval(p->vr_poly, 3, 0.0), vr1 = polyval(p->vr_poly, 3, (double)src->x_size);
    if (!(vr0 > 0) || !(vr1 > 0)) return set_error(ASARF_EINVAL, "asarf_range_doppler: Vr must be > 0");
    const bool adopt = (out->data == nullptr);
    if (adopt) {
        if (!ws) return set_error(ASARF_EINVAL, "asarf_range_doppler: no workspace to adopt");
        const size_t need = asarf_image_bytes(src);
        if (!ws->data || ws->bytes < need)
            return set_error(ASARF_EWORKSPACE, "asarf_range_doppler: workspace %zu < %zu bytes", ws ? ws->bytes : 0,
                             need);
    } else {
        ST_CHECK(check_cf32(out, "asarf_range_doppler(out)"));
        if (out->x_stride < src->x_size || out->y_stride < src->y_stride)
            return set_error(ASARF_ESHAPE, "asarf_range_doppler: output smaller than the padded source");
    }
    if ((size_t)src->x_size * sizeof(float2) > 0 && src->x_size > 8192)
        return set_error(ASARF_EUNSUPPORTED, "asarf_range_doppler: range size %d", src->x_size);
    ST_CHECK(ctx_init(c));
    if ((size_t)src->x_size * sizeof(float2) > c->max_lds)
        return set_error(ASARF_EUNSUPPORTED, "asarf_range_doppler: row of %d samples exceeds LDS", src->x_size);
    ColSplit cs;
    ST_CHECK(get_col_split(c, src->y_stride, &cs));
    if (adopt) ST_CHECK(asarf_image_adopt(out, src->x_size, src->y_size, src->x_stride, src->y_stride, ASARF_CF32, ws));
    const int nrv = src->x_size, na = src->y_size, nap = src->y_stride;
    void* scratch = nullptr;
    ST_CHECK(ctx_scratch(c, sizeof(ColConst) * (size_t)nrv + 64, &scratch));
    auto* col = static_cast<ColConst*>(scratch);
    // device-side bound on the RCMC shift (pixels), cleared and then maximised by the column-constant kernels
    int* d_reach = reinterpret_cast<int*>(static_cast<char*>(scratch) + sizeof(ColConst) * (size_t)nrv);
    ColConstArgs ca{};
    ca.col = col;
    ca.nrv = nrv;
    ca.dc_poly = nullptr;  // by value
    for (int i = 0; i < 3; ++i) ca.dc_val[i] = p->dc_poly[i];
    ca.clear_reach = 1;
    for (int i = 0; i < 3; ++i) ca.vr_poly[i] = p->vr_poly[i];
    ca.wavelength = p->wavelength;
    ca.range_spacing = p->range_spacing;
    ca.r_first = p->slant_range_first;
    ca.prf = p->prf;
    ca.nbins = nap;
    ca.reach = d_reach;
    HIP_CHECK(launch_col_const(ca, c->stream));
    auto* s = static_cast<float2*>(src->data);
    auto* o = static_cast<float2*>(out->data);
    const long so = out->x_stride;
    if (cs.n2 > 1 && src->x_stride == out->x_stride && s != o) {
        // src is consumed (the reference trades it away as the new workspace), so it serves as the second buffer
        // and every pass runs out of place, as in the fused plan: A s -> o, B o -> s, RCMC + matched filter
        // s -> o (column-tiled kernel when the device-side shift bound fits its halo, else the row kernel; both
        // are enqueued and the bound makes exactly one of them work), B' o -> s, A' s -> o
        ST_CHECK(col_forward_pass(c, cs, 0, s, so, na, o, so, nrv));
        ST_CHECK(col_forward_pass(c, cs, 1, o, so, nap, s, so, nrv));
        RcmcArgs rt{};
        rt.img = s;
        rt.out = o;
        rt.stride = so;
        rt.nrv = nrv;
        rt.nbins = nap;
        rt.n1 = cs.n1;
        rt.n2 = cs.n2;
        rt.col = col;
        rt.prf = p->prf;
        rt.beta = p->az_bandwidth_fraction;
        rt.mode = 0;
        rt.tile_gate = d_reach;
        HIP_CHECK(launch_rcmc_tile(rt, c->stream));
        HIP_CHECK(launch_rcmc_mf(rt, c->stream));
        ST_CHECK(col_inverse_pass_to(c, cs, 0, o, so, s, so, nrv, na, 1.f));
        ST_CHECK(col_inverse_pass_to(c, cs, 1, s, so, o, so, nrv, na, 1.0f / (float)nap));
        HIP_CHECK(launch_zero_cols(o, so, out->y_stride, nrv, out->x_stride, c->stream));
        out->x_size = nrv;
        out->y_size = na;
        if (adopt) {
            ws->data = src->data;
            ws->bytes = asarf_image_bytes(src);
            *src = asarf_image{nullptr, 0, 0, 0, 0, ASARF_CF32};
        }
        return clear_error();
    }
    ST_CHECK(col_forward_pass(c, cs, 0, s, src->x_stride, na, o, so, nrv));
    ST_CHECK(col_forward_pass(c, cs, 1, o, so, nap, o, so, nrv));
    RcmcArgs ra{};
    ra.img = o;
    ra.stride = so;
    ra.nrv = nrv;
    ra.nbins = nap;
    ra.n1 = cs.n1;
    ra.n2 = cs.n2;
    ra.col = col;
    ra.prf = p->prf;
    ra.beta = p->az_bandwidth_fraction;
    ra.mode = 0;
    HIP_CHECK(launch_rcmc_mf(ra, c->stream));
    ST_CHECK(col_inverse_pass(c, cs, 0, o, so, nrv, na, 1.0f / (float)nap));
    ST_CHECK(col_inverse_pass(c, cs, 1, o, so, nrv, na, 1.0f / (float)nap));
    HIP_CHECK(launch_zero_cols(o, so, out->y_stride, nrv, out->x_stride, c->stream));
    out->x_size = nrv;
    out->y_size = na;
    if (adopt) {
        // the source buffer becomes the new workspace (range_doppler_algorithm.cu:378)
        ws->data = src->data;
        ws->bytes = asarf_image_bytes(src);
        *src = asarf_image{nullptr, 0, 0, 0, 0, ASARF_CF32};
    }
    return clear_error();
}

// Level-0 decode (k_ingest.hip): validation shared by both formats.
static int check_ingest(asarf_ctx* c, const void* bytes, long line_bytes, int lines, const unsigned short* swst,
                        asarf_image* out, const char* who) {
    if (!c || !bytes || !swst || !out) return set_error(ASARF_EINVAL, "%s: null argument", who);
    ST_CHECK(check_cf32(out, who));
    if (lines <= 0 || line_bytes <= 0) return set_error(ASARF_ESHAPE, "%s: no lines", who);
    if (out->y_size != lines) return set_error(ASARF_ESHAPE, "%s: image has %d lines, input %d", who, out->y_size, lines);
    return ASARF_OK;
}

static IngestArgs ingest_args(const void* bytes, long line_bytes, int lines, const unsigned short* swst, int min_swst,
                              const asarf_image* out) {
    IngestArgs a{};
    a.bytes = static_cast<const uint8_t*>(bytes);
    a.line_bytes = line_bytes;
    a.lines = lines;
    a.swst = swst;
    a.min_swst = min_swst;
    a.out = static_cast<float2*>(out->data);
    a.x_size = out->x_size;
    a.x_stride = out->x_stride;
    a.y_stride = out->y_stride;
    return a;
}

extern "C" int asarf_ingest_ers_im(asarf_ctx* c, const unsigned char* samples, long line_bytes, int samples_per_line,
                                   int lines, const unsigned short* swst_codes, int min_swst, asarf_image* out) {
    ST_CHECK(check_ingest(c, samples, line_bytes, lines, swst_codes, out, "asarf_ingest_ers_im"));
    if (samples_per_line <= 0 || 2L * samples_per_line > line_bytes)
        return set_error(ASARF_ESHAPE, "asarf_ingest_ers_im: %d sample pairs do not fit %ld-byte lines",
                         samples_per_line, line_bytes);
    ST_CHECK(ctx_init(c));
    IngestArgs a = ingest_args(samples, line_bytes, lines, swst_codes, min_swst, out);
    a.samples_per_line = samples_per_line;
    a.swst_mult = 4;  // ERS SWST code step in samples (envisat_format_kernels.cu:126)
    HIP_CHECK(launch_ingest(a, 0, c->stream));
    return clear_error();
}

extern "C" int asarf_ingest_asar_fbaq4(asarf_ctx* c, const unsigned char* records, long record_bytes, int records_n,
                                       const unsigned short* swst_codes, int min_swst, const float* i_lut,
                                       const float* q_lut, int lut_items, asarf_image* out) {
    ST_CHECK(check_ingest(c, records, record_bytes, records_n, swst_codes, out, "asarf_ingest_asar_fbaq4"));
    if (!i_lut || !q_lut || lut_items <= 0) return set_error(ASARF_EINVAL, "asarf_ingest_asar_fbaq4: no tables");
    if (record_bytes < 3) return set_error(ASARF_ESHAPE, "asarf_ingest_asar_fbaq4: records too short");
    ST_CHECK(ctx_init(c));
    IngestArgs a = ingest_args(records, record_bytes, records_n, swst_codes, min_swst, out);
    a.swst_mult = 1;
    a.i_lut = i_lut;
    a.q_lut = q_lut;
    a.lut_items = lut_items;
    if (lut_items >= 4096) {
        void* lt = nullptr;
        ST_CHECK(ctx_scratch(c, sizeof(float2) * 4096, &lt));
        a.lut_t = static_cast<const float2*>(lt);
    }
    HIP_CHECK(launch_ingest(a, 1, c->stream));
    return clear_error();
}

extern "C" int asarf_fbaq4_index(int block, int code) { return fbaq4_index_host(block, code); }

// Output conditioning (k_output.hip).
extern "C" int asarf_condition_slc(asarf_ctx* c, const asarf_image* slc, float calibration, int header_bytes,
                                   void* dest, size_t dest_bytes) {
    if (!c || !dest) return set_error(ASARF_EINVAL, "asarf_condition_slc: null argument");
    ST_CHECK(check_cf32(slc, "asarf_condition_slc"));
    if (header_bytes < 0 || header_bytes > slc->x_size)
        return set_error(ASARF_ESHAPE, "asarf_condition_slc: header of %d bytes", header_bytes);
    const long rec = (long)header_bytes + 4L * slc->x_size;
    if (dest_bytes < (size_t)rec * slc->y_size)
        return set_error(ASARF_EWORKSPACE, "asarf_condition_slc: output %zu < %ld bytes", dest_bytes,
                         rec * slc->y_size);
    ST_CHECK(ctx_init(c));
    OutputArgs a{};
    a.slc = static_cast<const float2*>(slc->data);
    a.x_stride = slc->x_stride;
    a.x_size = slc->x_size;
    a.y_size = slc->y_size;
    a.calibration = calibration;
    a.header = header_bytes;
    a.record_bytes = rec;
    a.dest = static_cast<unsigned char*>(dest);
    HIP_CHECK(launch_condition(a, 0, c->stream));
    return clear_error();
}

extern "C" int asarf_condition_detected(asarf_ctx* c, const float* img, int range_size, int azimuth_size,
                                        float calibration, int header_bytes, void* dest, size_t dest_bytes) {
    if (!c || !img || !dest) return set_error(ASARF_EINVAL, "asarf_condition_detected: null argument");
    if (range_size <= 0 || azimuth_size <= 0) return set_error(ASARF_ESHAPE, "asarf_condition_detected: empty image");
    if (header_bytes < 0 || header_bytes > range_size)
        return set_error(ASARF_ESHAPE, "asarf_condition_detected: header of %d bytes", header_bytes);
    const long rec = (long)header_bytes + 2L * range_size;
    if (dest_bytes < (size_t)rec * azimuth_size)
        return set_error(ASARF_EWORKSPACE, "asarf_condition_detected: output %zu < %ld bytes", dest_bytes,
                         rec * azimuth_size);
    ST_CHECK(ctx_init(c));
    OutputArgs a{};
    a.det = img;
    a.x_stride = range_size;
    a.x_size = range_size;
    a.y_size = azimuth_size;
    a.calibration = calibration;
    a.header = header_bytes;
    a.record_bytes = rec;
    a.dest = static_cast<unsigned char*>(dest);
    HIP_CHECK(launch_condition(a, 1, c->stream));
    return clear_error();
}

extern "C" int asarf_image_statistics(asarf_ctx* c, const asarf_image* img, int imaged, double* out4) {
    if (!c || !out4) return set_error(ASARF_EINVAL, "asarf_image_statistics: null argument");
    ST_CHECK(check_cf32(img, "asarf_image_statistics"));
    ST_CHECK(ctx_init(c));
    const int nwg = std::min(img->y_size, 1024);
    void* scratch = nullptr;
    ST_CHECK(ctx_scratch(c, sizeof(double) * (size_t)(nwg + 1) * 12, &scratch));
    double* partials = static_cast<double*>(scratch);
    double* totals = partials + (size_t)nwg * 12;
    auto* f2 = static_cast<const float2*>(img->data);
    HIP_CHECK(launch_moments_cf32(f2, img->x_stride, img->x_size, img->y_size, img->x_stride, partials, nwg, c->stream));
    HIP_CHECK(launch_reduce_f64(partials, nwg, 12, totals, c->stream));
    void* hp = nullptr;
    ST_CHECK(ctx_pinned(c, sizeof(double) * 12, &hp));
    HIP_CHECK(hipMemcpyAsync(hp, totals, sizeof(double) * 12, hipMemcpyDeviceToHost, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    double t[12];
    std::memcpy(t, hp, sizeof(t));
    // CalcMean / CalcStdDev (device_padded_image.cu:331-414): NaN-skipping sums over the logical region, divided
    // by its full sample count
    const double n = (double)img->x_size * (double)img->y_size;
    const double nL = t[0], mi = t[1] / n, mq = t[2] / n;
    const double si = std::sqrt(std::max(0.0, (t[3] - 2.0 * mi * t[1] + nL * mi * mi) / n));
    const double sq = std::sqrt(std::max(0.0, (t[4] - 2.0 * mq * t[2] + nL * mq * mq) / n));
    if (imaged) {  // CalculateStatisticsImaged (main_flow.cc:672-676)
        out4[0] = (mi + mq) / 2;
        out4[1] = 0.0;
        out4[2] = (si + sq) / 2;
        out4[3] = 0.0;
    } else {  // CalculateStatistics (main_flow.cc:666-670)
        out4[0] = mi;
        out4[1] = mq;
        out4[2] = si;
        out4[3] = sq;
    }
    return clear_error();
}

extern "C" int asarf_az_multilook(asarf_ctx* c, const asarf_image* slc, double az_spacing, asarf_workspace* out,
                                  int* out_rows, double* new_spacing) {
    if (!c || !out || !out_rows) return set_error(ASARF_EINVAL, "asarf_az_multilook: null argument");
    ST_CHECK(check_cf32(slc, "asarf_az_multilook"));
    if (!(az_spacing > 0)) return set_error(ASARF_EINVAL, "asarf_az_multilook: azimuth spacing must be > 0");
    const float ratio = (float)(12.5 / az_spacing);
    const int rows = (int)std::round((float)slc->y_size / ratio);
    const size_t need = (size_t)rows * slc->x_size * sizeof(float);
    if (!out->data || out->bytes < need)
        return set_error(ASARF_EWORKSPACE, "asarf_az_multilook: output %zu < %zu bytes", out->bytes, need);
    ST_CHECK(ctx_init(c));
    MultilookArgs a{};
    a.slc = static_cast<const float2*>(slc->data);
    a.stride = slc->x_stride;
    a.in_rows = slc->y_size;
    a.nr = slc->x_size;
    a.out_rows = rows;
    a.ratio = ratio;
    a.out = static_cast<float*>(out->data);
    HIP_CHECK(launch_multilook(a, c->stream));
    *out_rows = rows;
    if (new_spacing) *new_spacing = az_spacing * (double)ratio;
    return clear_error();
}

extern "C" int asarf_range_spreading_loss(asarf_ctx* c, float* img, int nr, int rows, double r_first, double dr,
                                          double r_ref) {
    if (!c || !img) return set_error(ASARF_EINVAL, "asarf_range_spreading_loss: null argument");
    if (nr <= 0 || rows < 0 || !(r_first > 0) || !(dr > 0) || !(r_ref > 0))
        return set_error(ASARF_EINVAL, "asarf_range_spreading_loss: bad parameters");
    ST_CHECK(ctx_init(c));
    std::vector<float> tab(nr);
    for (int i = 0; i < nr; ++i) {
        const double r = r_first + (double)i * dr;
        tab[i] = (float)(1.0 / std::sqrt(std::pow(r_ref / r, 3.0)));
    }
    void* scratch = nullptr;
    ST_CHECK(ctx_scratch(c, sizeof(float) * (size_t)nr, &scratch));
    HIP_CHECK(hipMemcpyAsync(scratch, tab.data(), sizeof(float) * (size_t)nr, hipMemcpyHostToDevice, c->stream));
    HIP_CHECK(launch_rsl(img, nr, rows, static_cast<float*>(scratch), c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));  // tab (host) must outlive the copy
    return clear_error();
}

extern "C" int asarf_srgr(asarf_ctx* c, const float* in, int in_nr, int rows, const double* grsr5, double dr, double gs,
                          int out_nr, float* out) {
    if (!c || !in || !out || !grsr5) return set_error(ASARF_EINVAL, "asarf_srgr: null argument");
    if (in_nr <= 1 || rows < 0 || out_nr <= 0 || !(dr > 0) || !(gs > 0))
        return set_error(ASARF_EINVAL, "asarf_srgr: bad parameters");
    ST_CHECK(ctx_init(c));
    SrgrArgs a{};
    a.in = in;
    a.in_nr = in_nr;
    a.rows = rows;
    for (int i = 0; i < 5; ++i) a.grsr[i] = grsr5[i];
    a.dr = dr;
    a.gs = gs;
    a.out_nr = out_nr;
    a.out = out;
    HIP_CHECK(launch_srgr(a, c->stream));
    return clear_error();
}

extern "C" int asarf_fft_c2c(asarf_ctx* c, void* data, int n, int batch, long dist, int layout, int direction,
                             float scale) {
    if (!c || !data) return set_error(ASARF_EINVAL, "asarf_fft_c2c: null argument");
    if (n < 2 || batch <= 0) return set_error(ASARF_EINVAL, "asarf_fft_c2c: n=%d batch=%d", n, batch);
    if (direction != ASARF_FORWARD && direction != ASARF_INVERSE)
        return set_error(ASARF_EINVAL, "asarf_fft_c2c: direction %d", direction);
    const bool inv = direction == ASARF_INVERSE;
    auto* d = static_cast<float2*>(data);
    if (layout == ASARF_FFT_ROWS) {
        if (dist < n) return set_error(ASARF_ESHAPE, "asarf_fft_c2c: dist %ld < n %d", dist, n);
        ST_CHECK(ctx_init(c));
        int st = ASARF_OK;
        const FftPlanHost* p = get_plan(c, n, kRowThreads, &st);
        if (!p) return st;
        RowFftArgs a{};
        a.in = d;
        a.in_stride = dist;
        a.out = d;
        a.out_stride = dist;
        a.rows = batch;
        a.scale = scale;
        a.plan = p->dev();
        HIP_CHECK(launch_row_fft(a, inv, c->stream));
        return clear_error();
    }
    if (layout != ASARF_FFT_COLUMNS) return set_error(ASARF_EINVAL, "asarf_fft_c2c: layout %d", layout);
    if (dist < batch) return set_error(ASARF_ESHAPE, "asarf_fft_c2c: row stride %ld < columns %d", dist, batch);
    ST_CHECK(ctx_init(c));
    ColSplit cs;
    ST_CHECK(get_col_split(c, n, &cs));
    if (cs.n2 == 1) {
        HIP_CHECK(launch_col_fft(col_args(c, d, dist, d, dist, batch, 1, 0, 1, n, 0, 1, n, nullptr, scale, cs.p1), inv,
                                 c->stream));
        return clear_error();
    }
    // natural order through a temporary: forward A -> tmp, B -> data (bin k1 + n1 k2 at its own row)
    void* scratch = nullptr;
    ST_CHECK(ctx_scratch(c, sizeof(float2) * (size_t)n * batch, &scratch));
    auto* t = static_cast<float2*>(scratch);
    const long ts = batch;
    if (!inv) {
        HIP_CHECK(launch_col_fft(col_args(c, d, dist, t, ts, batch, cs.n2, 1, cs.n2, n, 1, cs.n2, n, cs.full->d_tw, 1.f,
                                          cs.p1),
                                 false, c->stream));
        HIP_CHECK(launch_col_fft(col_args(c, t, ts, d, dist, batch, cs.n1, cs.n2, 1, n, 1, cs.n1, n, nullptr, scale,
                                          cs.p2),
                                 false, c->stream));
    } else {
        HIP_CHECK(launch_col_fft(col_args(c, d, dist, t, ts, batch, cs.n1, 1, cs.n1, n, cs.n2, 1, n, cs.full->d_tw, 1.f,
                                          cs.p2),
                                 true, c->stream));
        HIP_CHECK(launch_col_fft(col_args(c, t, ts, d, dist, batch, cs.n2, 1, cs.n2, n, 1, cs.n2, n, nullptr, scale,
                                          cs.p1),
                                 true, c->stream));
    }
    HIP_CHECK(hipStreamSynchronize(c->stream));  // the scratch may be reused by the next call
    return clear_error();
}

extern "C" int asarf_fft_c2c_split_bf16(asarf_ctx* c, void* data, int n, int batch, long dist, int layout,
                                        int direction, float scale) {
    if (!c || !data) return set_error(ASARF_EINVAL, "asarf_fft_c2c_split_bf16: null argument");
    if (batch <= 0) return set_error(ASARF_EINVAL, "asarf_fft_c2c_split_bf16: batch %d", batch);
    if (direction != ASARF_FORWARD && direction != ASARF_INVERSE)
        return set_error(ASARF_EINVAL, "asarf_fft_c2c_split_bf16: direction %d", direction);
    const bool inv = direction == ASARF_INVERSE;
    auto* d = static_cast<float2*>(data);
    if (layout == ASARF_FFT_ROWS) {
        if (n != 8192) return set_error(ASARF_EUNSUPPORTED, "asarf_fft_c2c_split_bf16: rows of %d (8192 only)", n);
        if (dist < n) return set_error(ASARF_ESHAPE, "asarf_fft_c2c_split_bf16: dist %ld < n %d", dist, n);
        ST_CHECK(ctx_init(c));
        int st = ASARF_OK;
        const FftPlanHost* p = get_plan(c, n, kRowThreads, &st);
        if (!p) return st;
        RowFftArgs a{};
        a.in = d;
        a.in_stride = dist;
        a.out = d;
        a.out_stride = dist;
        a.rows = batch;
        a.scale = scale;
        a.plan = p->dev();
        a.mfma = 1;
        HIP_CHECK(launch_row_fft(a, inv, c->stream));
        return clear_error();
    }
    if (layout != ASARF_FFT_COLUMNS) return set_error(ASARF_EINVAL, "asarf_fft_c2c_split_bf16: layout %d", layout);
    if (dist < batch) return set_error(ASARF_ESHAPE, "asarf_fft_c2c_split_bf16: row stride %ld < columns %d", dist, batch);
    ST_CHECK(ctx_init(c));
    auto col_ok = [](int m) { return m == 128 || m == 256; };
    if (col_ok(n)) {
        int st = ASARF_OK;
        const FftPlanHost* p = get_plan(c, n, kThreads / kColTile, &st);
        if (!p) return st;
        ColFftArgs a = col_args(c, d, dist, d, dist, batch, 1, 0, 1, n, 0, 1, n, nullptr, scale, p);
        a.mfma = 1;
        HIP_CHECK(launch_col_fft(a, inv, c->stream));
        return clear_error();
    }
    ColSplit cs;
    ST_CHECK(get_col_split(c, n, &cs, 256));
    if (cs.n2 == 1 || !col_ok(cs.n1) || !col_ok(cs.n2))
        return set_error(ASARF_EUNSUPPORTED, "asarf_fft_c2c_split_bf16: columns of %d (128, 256 or their products)",
                         n);
    // natural order through a temporary, as asarf_fft_c2c
    void* scratch = nullptr;
    ST_CHECK(ctx_scratch(c, sizeof(float2) * (size_t)n * batch, &scratch));
    auto* t = static_cast<float2*>(scratch);
    const long ts = batch;
    ColFftArgs a1, a2;
    if (!inv) {
        a1 = col_args(c, d, dist, t, ts, batch, cs.n2, 1, cs.n2, n, 1, cs.n2, n, cs.full->d_tw, 1.f, cs.p1);
        a2 = col_args(c, t, ts, d, dist, batch, cs.n1, cs.n2, 1, n, 1, cs.n1, n, nullptr, scale, cs.p2);
    } else {
        a1 = col_args(c, d, dist, t, ts, batch, cs.n1, 1, cs.n1, n, cs.n2, 1, n, cs.full->d_tw, 1.f, cs.p2);
        a2 = col_args(c, t, ts, d, dist, batch, cs.n2, 1, cs.n2, n, 1, cs.n2, n, nullptr, scale, cs.p1);
    }
    a1.mfma = a2.mfma = 1;
    HIP_CHECK(launch_col_fft(a1, inv, c->stream));
    HIP_CHECK(launch_col_fft(a2, inv, c->stream));
    HIP_CHECK(hipStreamSynchronize(c->stream));
    return clear_error();
}
