// Row transforms, fused range compression and column (azimuth) transform passes for gfx950.
//
// Row kernels: one workgroup (256 lanes, 4 wave64) owns one row; the row lives in LDS between Stockham passes.
// The first pass reads straight from HBM (with the I/Q correction applied on load for range compression) and
// the last pass writes straight to HBM, so a row costs one HBM read and one HBM write.
//
// Column kernels: a workgroup owns a tile of 16 adjacent columns (one 128 B line per row) and one
// sub-transform of length n <= 512 along the column; 16 lanes cooperate per column. Long azimuth transforms run
// as two such passes (four-step decomposition N = N1*N2 with the inter-pass twiddle fused into the first pass's
// store), so there is no transpose.
#include <atomic>

#include "kernels.h"

namespace asarf {

constexpr int kMaxDevices = 64;

// LDS bytes a workgroup may allocate on the current device (queried once per device; 64 KiB if the query fails)
size_t max_lds_bytes() {
    static std::atomic<size_t> cache[kMaxDevices];
    int d = 0;
    if (hipGetDevice(&d) != hipSuccess || d < 0 || d >= kMaxDevices) return 65536;
    size_t v = cache[d].load(std::memory_order_relaxed);
    if (v == 0) {
        int lds = 0;
        v = hipDeviceGetAttribute(&lds, hipDeviceAttributeMaxSharedMemoryPerBlock, d) == hipSuccess && lds > 0
                ? (size_t)lds
                : 65536;
        cache[d].store(v, std::memory_order_relaxed);
    }
    return v;
}

// LDS index with one pad slot every 32 elements (spreads power-of-two strides over banks).
__device__ __forceinline__ int lpad(int i) { return i + (i >> 5); }
__host__ __device__ inline int lds_row_elems(int n) { return n + (n >> 5) + 1; }

// ------------------------------------------------------------------------------------------------ row FFT ----
template <bool INV, int RSET>
__global__ __launch_bounds__(kRowThreads) void k_row_fft(RowFftArgs a) {
    extern __shared__ float2 lds[];
    const int row = blockIdx.x;
    const int lane = threadIdx.x, nlanes = kRowThreads;
    const int fft_n = a.plan.n;
    const TwLds tw_at = tw_fill(lds + lds_row_elems(fft_n), a.plan.tw, fft_n, lane, nlanes);
    const float2* in = a.in + (long)row * a.in_stride;
    float2* out = a.out + (long)row * a.out_stride;
    float2 v[kMaxE];
    const FftPlanDev pl = a.plan;
    auto radix_at = [&](int p) { return plan_radix(pl, INV ? pl.npass - 1 - p : p); };
    auto gl_load = [&](int i) { return in[i]; };
    auto lds_store = [&](int i, float2 x) { lds[lpad(i)] = x; };
    auto lds_load = [&](int i) { return lds[lpad(i)]; };
    const int r0 = radix_at(0);
#define LOAD_GL_(RR) pass_load<RR>(v, lane, nlanes, fft_n, gl_load)
    ASARF_DISPATCH(r0, LOAD_GL_)
#undef LOAD_GL_
    __syncthreads();  // in == out is allowed: every read of the row precedes the first write
    if (INV) {
        ASARF_DISPATCH(r0, ASARF_COMPUTE1_INV)
    } else {
        ASARF_DISPATCH(r0, ASARF_COMPUTE1_FWD)
    }
    ASARF_RUN_PASSES(INV, a.plan.npass, radix_at)
    const float sc = a.scale;
    auto out_write = [&](int i, float2& x, bool ok) {
        if (ok) out[i] = cscale(x, sc);
    };
    const int rl = radix_at(a.plan.npass - 1);
#define VISIT_WR_(RR) pass_visit<RR>(v, lane, nlanes, fft_n, out_write)
    ASARF_DISPATCH(rl, VISIT_WR_)
#undef VISIT_WR_
}

// Radix-set instantiations, most specific first; a plan runs on the first set containing all its radices.
#define ROW_SETS(X) X(72) X(98) X(101) X(kRadixAll)
#define COL_SETS(X) X(50) X(14) X(82) X(22) X(72) X(101) X(kRadixAll)

static int plan_mask(const FftPlanDev& p) {
    int m = 0;
    for (int i = 0; i < p.npass; ++i) m |= radix_bit(plan_radix(p, i));
    return m;
}

hipError_t launch_row_fft(const RowFftArgs& a, bool inverse, hipStream_t s) {
    if (a.mfma) return launch_row_mfma(a, inverse, s);
    const size_t lds = sizeof(float2) * (size_t)(lds_row_elems(a.plan.n) + tw_lds_elems(a.plan.n));
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    const int m = plan_mask(a.plan);
#define X_(SET)                                                                                          \
    if ((m & ~(SET)) == 0) {                                                                             \
        if (inverse)                                                                                     \
            hipLaunchKernelGGL((k_row_fft<true, SET>), dim3(a.rows), dim3(kRowThreads), lds, s, a);      \
        else                                                                                             \
            hipLaunchKernelGGL((k_row_fft<false, SET>), dim3(a.rows), dim3(kRowThreads), lds, s, a);     \
        return hipGetLastError();                                                                        \
    }
    ROW_SETS(X_)
#undef X_
    return hipErrorInvalidValue;
}

// ------------------------------------------------------------------------------------ range compression ----
// Per row: load (raw bytes -> corrected complex | cf32), zero pad to n, forward FFT, multiply by conj(H)/n,
// inverse FFT, store x < nrv and zero [nrv, zero_to). Follows sar/range_compression.cu:57-111.
template <typename T>
struct RawPair;
template <>
struct RawPair<uint8_t> {
    using type = uchar2;
};
template <>
struct RawPair<int8_t> {
    using type = char2;
};

template <int FMT, int RSET>
__global__ __launch_bounds__(kRowThreads) void k_range_compress(RcArgs a) {
    extern __shared__ float2 lds[];
    const int row = blockIdx.x;
    const int lane = threadIdx.x, nlanes = kRowThreads;
    const int fft_n = a.plan.n;
    const TwLds tw_at = tw_fill(lds + lds_row_elems(fft_n), a.plan.tw, fft_n, lane, nlanes);
    float2 v[kMaxE];
    IqCoef c;
    if (a.coef) {
        c = *a.coef;
    } else {
        c.dc_i = c.dc_q = c.sin_corr = 0.f;
        c.inv_gain = c.inv_cos = 1.f;
        c.identity = 1;
    }
    const int nr_in = a.nr_in;
    auto correct = [&](float i, float q) {
        if (c.identity) return make_float2(i, q);
        const float ic = (i - c.dc_i) * c.inv_gain;
        const float qc = (q - c.dc_q - ic * c.sin_corr) * c.inv_cos;
        return make_float2(ic, qc);
    };
    const bool lines_fmt = FMT == RAW_ERS_LINES || FMT == RAW_ASAR_FBAQ4;
    const int shift = lines_fmt ? a.lines.shift(row) : 0;
    const unsigned char* rec = static_cast<const unsigned char*>(a.in) + (lines_fmt ? (long)row * a.lines.line_bytes : 0L);
    const int rec_len = FMT == RAW_ASAR_FBAQ4 ? fbaq4_record_length(rec) : 0;
    auto gl_load = [&](int x) -> float2 {
        if constexpr (FMT == RAW_CF32) {
            const float2* in = static_cast<const float2*>(a.in) + (long)row * a.in_stride;
            const float2 r = in[x];
            return correct(r.x, r.y);
        } else if constexpr (FMT == RAW_ERS_LINES) {
            // NaN where no line sample lands becomes 0 (ZeroNaNs); padding is 0
            const int sx = x - shift;
            if (x >= nr_in || sx < 0 || sx >= a.lines.spl) return make_float2(0.f, 0.f);
            const uchar2 r = reinterpret_cast<const uchar2*>(rec)[sx];
            return correct((float)r.x, (float)r.y);
        } else if constexpr (FMT == RAW_ASAR_FBAQ4) {
            // the record's sample shifted by the SWST; NaN where none lands becomes 0 (ZeroNaNs); padding is 0
            float2 r;
            if (x >= nr_in || !fbaq4_sample(rec, rec_len, a.lines.line_bytes, a.lines.lut, x - shift, &r))
                return make_float2(0.f, 0.f);
            return correct(r.x, r.y);
        } else {
            if (x >= nr_in) return make_float2(0.f, 0.f);
            using P = typename RawPair<typename std::conditional<FMT == RAW_U8, uint8_t, int8_t>::type>::type;
            const P* in = static_cast<const P*>(a.in) + (long)row * a.in_stride;
            const P r = in[x];
            return correct((float)r.x, (float)r.y);
        }
    };
    const int np = a.plan.npass;
    const FftPlanDev pl = a.plan;
    auto fr = [&](int p) { return plan_radix(pl, p); };
    auto ir = [&](int p) { return plan_radix(pl, np - 1 - p); };
    auto lds_store = [&](int i, float2 x) { lds[lpad(i)] = x; };
    auto lds_load = [&](int i) { return lds[lpad(i)]; };

    const int rf0 = fr(0);
#define LOAD_GL_(RR) pass_load<RR>(v, lane, nlanes, fft_n, gl_load)
    ASARF_DISPATCH(rf0, LOAD_GL_)
#undef LOAD_GL_
    ASARF_DISPATCH(rf0, ASARF_COMPUTE1_FWD)
    ASARF_RUN_PASSES(false, np, fr)
    // spectrum multiply at natural positions of the last forward pass
    const float2* h = a.hconj;
    auto spec_mul = [&](int k, float2& x, bool) { x = cmul(x, h[k]); };
    const int rfl = fr(np - 1);
#define VISIT_MUL_(RR) pass_visit<RR>(v, lane, nlanes, fft_n, spec_mul)
    ASARF_DISPATCH(rfl, VISIT_MUL_)
#undef VISIT_MUL_
    // inverse: its first pass has the radix of the last forward pass, so the inputs are already in registers
    __syncthreads();  // all LDS reads of the forward transform are done before the inverse stores
    ASARF_DISPATCH(rfl, ASARF_COMPUTE1_INV)
    ASARF_RUN_PASSES(true, np, ir)
    float2* out = a.out + (long)row * a.out_stride;
    const int nrv = a.nrv, zt = a.zero_to;
    auto out_write = [&](int x, float2& val, bool ok) {
        if (!ok) return;
        if (x < nrv)
            out[x] = val;
        else if (x < zt)
            out[x] = make_float2(0.f, 0.f);
    };
    const int ril = ir(np - 1);
#define VISIT_WR_(RR) pass_visit<RR>(v, lane, nlanes, fft_n, out_write)
    ASARF_DISPATCH(ril, VISIT_WR_)
#undef VISIT_WR_
}

hipError_t launch_range_compress(const RcArgs& a, int fmt, hipStream_t s) {
    if (a.mfma) return launch_rc_mfma(a, fmt, s);
    const size_t lds = sizeof(float2) * (size_t)(lds_row_elems(a.plan.n) + tw_lds_elems(a.plan.n));
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    if (fmt != RAW_CF32 && fmt != RAW_U8 && fmt != RAW_I8 && fmt != RAW_ERS_LINES && fmt != RAW_ASAR_FBAQ4)
        return hipErrorInvalidValue;
    if (fmt == RAW_ASAR_FBAQ4 && !a.lines.lut) return hipErrorInvalidValue;
    hipError_t serr = hipSuccess;
    if (!a.no_static && launch_range_compress_static(a, fmt, s, &serr)) return serr;
    const int m = plan_mask(a.plan);
#define X_(SET)                                                                                                 \
    if ((m & ~(SET)) == 0) {                                                                                    \
        if (fmt == RAW_CF32)                                                                                    \
            hipLaunchKernelGGL((k_range_compress<RAW_CF32, SET>), dim3(a.rows), dim3(kRowThreads), lds, s, a);  \
        else if (fmt == RAW_U8)                                                                                 \
            hipLaunchKernelGGL((k_range_compress<RAW_U8, SET>), dim3(a.rows), dim3(kRowThreads), lds, s, a);    \
        else if (fmt == RAW_ERS_LINES)                                                                          \
            hipLaunchKernelGGL((k_range_compress<RAW_ERS_LINES, SET>), dim3(a.rows), dim3(kRowThreads), lds, s, \
                               a);                                                                              \
        else if (fmt == RAW_ASAR_FBAQ4)                                                                         \
            hipLaunchKernelGGL((k_range_compress<RAW_ASAR_FBAQ4, SET>), dim3(a.rows), dim3(kRowThreads), lds,   \
                               s, a);                                                                           \
        else                                                                                                    \
            hipLaunchKernelGGL((k_range_compress<RAW_I8, SET>), dim3(a.rows), dim3(kRowThreads), lds, s, a);    \
        return hipGetLastError();                                                                               \
    }
    ROW_SETS(X_)
#undef X_
    return hipErrorInvalidValue;
}

// ---------------------------------------------------------------------------------- column transform pass ----
// Workgroup (tile, g): columns [16*tile, 16*tile+16), sub-transform g. Lane t -> column b = t % 16,
// sub-lane s = t / 16 (16 sub-lanes per column). LDS layout element (m, b) at m*17 + b.
template <bool INV, int RSET>
__device__ __forceinline__ void col_fft_item(const ColFftArgs& a, float2* lds, int w) {
    const int tile = w % a.ntiles;
    const int g = w / a.ntiles;
    const int b = threadIdx.x & (kColTile - 1);
    const int lane = threadIdx.x / kColTile;
    constexpr int nlanes = kThreads / kColTile;
    const int col = tile * kColTile + b;
    const bool col_ok = col < a.ncols;
    const int fft_n = a.sub.n;
    const TwLds tw_at = tw_fill(lds + fft_n * (kColTile + 1), a.sub.tw, fft_n, threadIdx.x, kThreads);
    float2 v[kMaxE];
    const float2* in = a.in + (col_ok ? col : 0);
    const long in_base = (long)g * a.in_g;
    const int in_rows = a.in_rows;
    auto gl_load = [&](int m) -> float2 {
        const long r = in_base + (long)m * a.in_m;
        if (!col_ok || r >= in_rows) return make_float2(0.f, 0.f);
        return in[r * a.in_stride];
    };
    const FftPlanDev pl = a.sub;
    auto radix_at = [&](int p) { return plan_radix(pl, INV ? pl.npass - 1 - p : p); };
    auto lds_store = [&](int i, float2 x) { lds[i * (kColTile + 1) + b] = x; };
    auto lds_load = [&](int i) { return lds[i * (kColTile + 1) + b]; };
    const int r0 = radix_at(0);
#define LOAD_GL_(RR) pass_load<RR>(v, lane, nlanes, fft_n, gl_load)
    ASARF_DISPATCH(r0, LOAD_GL_)
#undef LOAD_GL_
    __syncthreads();  // in-place passes: every element of the group is read before any is written
    if (INV) {
        ASARF_DISPATCH(r0, ASARF_COMPUTE1_INV)
    } else {
        ASARF_DISPATCH(r0, ASARF_COMPUTE1_FWD)
    }
    ASARF_RUN_PASSES(INV, a.sub.npass, radix_at)
    float2* out = a.out + (col_ok ? col : 0);
    const long out_base = (long)g * a.out_g;
    const float sc = a.scale;
    const float2* twf = a.tw_full;
    auto out_write = [&](int k, float2& x, bool ok) {
        const long r = out_base + (long)k * a.out_m;
        if (!ok || !col_ok || r >= a.out_rows) return;
        float2 y = x;
        if (twf) {
            const float2 w = twf[(long)g * k];
            y = INV ? cmulc(y, w) : cmul(y, w);
        }
        out[r * a.out_stride] = cscale(y, sc);
    };
    const int rl = radix_at(a.sub.npass - 1);
#define VISIT_WR_(RR) pass_visit<RR>(v, lane, nlanes, fft_n, out_write)
    ASARF_DISPATCH(rl, VISIT_WR_)
#undef VISIT_WR_
}

// One work item per workgroup (LOOP = false), or a gated fallback on a bounded grid whose workgroups take the
// items in turn (LOOP = true; see k_col_static).
template <bool INV, int RSET, bool LOOP>
__global__ __launch_bounds__(kThreads) void k_col_fft(ColFftArgs a) {
    extern __shared__ float2 lds[];
    if (a.fused_gate && *a.fused_gate <= kAzCoreReach) return;
    if constexpr (LOOP) {
        for (int w = blockIdx.x; w < a.ntiles * a.groups; w += gridDim.x) {
            col_fft_item<INV, RSET>(a, lds, w);
            __syncthreads();  // the next item refills LDS
        }
    } else {
        col_fft_item<INV, RSET>(a, lds, blockIdx.x);
    }
}

hipError_t launch_col_fft(const ColFftArgs& a, bool inverse, hipStream_t s) {
    if (a.mfma) return launch_col_mfma(a, inverse, s);
    hipError_t serr = hipSuccess;
    if (!a.no_static && launch_col_static(a, inverse, s, &serr)) return serr;
    if (a.sub.npass == 0) return hipErrorInvalidValue;  // a length planned for its compile-time kernel only
    const size_t lds = sizeof(float2) * (size_t)(a.sub.n * (kColTile + 1) + tw_lds_elems(a.sub.n));
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    const long items = (long)a.ntiles * a.groups;
    const bool loop = a.fused_gate != nullptr;
    const dim3 grid((unsigned)(loop ? std::min(items, 1024L) : items));
    const int m = plan_mask(a.sub);
#define X_(SET)                                                                                         \
    if ((m & ~(SET)) == 0) {                                                                            \
        if (inverse && loop)                                                                            \
            hipLaunchKernelGGL((k_col_fft<true, SET, true>), grid, dim3(kThreads), lds, s, a);          \
        else if (inverse)                                                                               \
            hipLaunchKernelGGL((k_col_fft<true, SET, false>), grid, dim3(kThreads), lds, s, a);         \
        else if (loop)                                                                                  \
            hipLaunchKernelGGL((k_col_fft<false, SET, true>), grid, dim3(kThreads), lds, s, a);         \
        else                                                                                            \
            hipLaunchKernelGGL((k_col_fft<false, SET, false>), grid, dim3(kThreads), lds, s, a);        \
        return hipGetLastError();                                                                       \
    }
    COL_SETS(X_)
#undef X_
    return hipErrorInvalidValue;
}

}  // namespace asarf
