// Fused azimuth core: 16-tap sinc RCMC + azimuth matched filter + inverse pass B' of the azimuth FFT, one
// workgroup per (k1 group, 16-column tile), out of place.
//
// With the four-step split N = N1*N2, forward pass B leaves the bins k1 + N1*k2 (k2 < N2) of one k1 group in
// the contiguous rows k1*N2 + k2. RCMC interpolates along range inside each bin and the matched filter is
// pointwise, so a workgroup that stages the k1 group's rows for its 16 columns plus a range halo in LDS can
// correct them and run the inverse length-N2 transform without another pass over HBM.
//   tile columns:   [x0 - 8, x0 + 32) staged, [x0, x0 + 16) produced
//   halo:           RCMC reads x + p - 8 .. x + p + 7 with p = round(shift) in [0, kAzFusedReach]
// The arithmetic is that of k_rcmc_mf (k_rda.hip) followed by k_col_fft's inverse pass B', operation for
// operation, so both paths give identical results.
#include "interp.h"
#include "kernels.h"

namespace asarf {

constexpr int kFzOut = 16;                                     // produced columns per tile
constexpr int kFzLeft = 8;                                     // first produced column inside the tile
constexpr int kFzCols = kFzLeft + kFzOut + 8 + kAzFusedReach;  // staged columns per tile
constexpr int kFzPitch = kFzCols + 1;
constexpr int kFzBatch = 16;                                   // staging loads in flight per lane
static_assert(kFzCols >= kColTile, "the inverse passes reuse tile columns 0..15 as their exchange buffer");

__device__ __forceinline__ double fz_alias(double f, double dc, double prf) {
    const double half = 0.5 * prf;
    if (dc < 0.0) {
        if (f > half + dc) f -= prf;
    } else if (dc > 0.0) {
        if (f < -half + dc) f += prf;
    }
    return f;
}

template <int RSET>
__global__ __launch_bounds__(kThreads) void k_az_fused(AzFusedArgs a) {
    extern __shared__ float2 lds[];
    if (a.gate && *a.gate > kAzFusedReach) return;
    // XCD-aware work order: consecutive work items (neighbouring tiles of one k1 group, which share halo lines)
    // go to the same XCD, since workgroups are dealt to the 8 XCDs round robin.
    const int per_xcd = gridDim.x / 8;
    const int w = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (w >= a.items) return;
    const int k1 = w / a.ntiles;
    const int tile = w % a.ntiles;
    const int x0 = tile * kFzOut;
    const int n2 = a.sub.n;
    float2* tl = lds;  // [n2][kFzPitch]
    const int fft_n = n2;
    const TwLds tw_at = tw_fill(lds + n2 * kFzPitch, a.sub.tw, n2, threadIdx.x, kThreads);
    const int b = threadIdx.x & (kColTile - 1);
    const int lane = threadIdx.x / kColTile;
    constexpr int nlanes = kThreads / kColTile;
    const FftPlanDev pl = a.sub;
    const long row0 = (long)k1 * n2;

    // ---- stage the spectrum tile with kFzBatch loads in flight per lane. Columns outside [0, nrv) load a
    // clamped neighbour: an output whose taps would reach them is zeroed by the RCMC edge rule (the taps of a
    // kept output lie in [0, nrv - 2]), so those values are never read.
    {
        const int ntile = n2 * kFzCols;
        const float2* base = a.in + row0 * a.stride;
        for (int e0 = 0; e0 < ntile; e0 += kThreads * kFzBatch) {
            float2 t[kFzBatch];
#pragma unroll
            for (int u = 0; u < kFzBatch; ++u) {
                const int e = min(e0 + u * kThreads + (int)threadIdx.x, ntile - 1);
                const int r = e / kFzCols, cidx = e - r * kFzCols;
                const int x = min(max(x0 - kFzLeft + cidx, 0), a.nrv - 1);
                t[u] = base[(long)r * a.stride + x];
            }
#pragma unroll
            for (int u = 0; u < kFzBatch; ++u) {
                const int e = e0 + u * kThreads + (int)threadIdx.x;
                const int r = e / kFzCols, cidx = e - r * kFzCols;
                if (e < ntile) tl[r * kFzPitch + cidx] = t[u];
            }
        }
    }
    __syncthreads();

    // ---- RCMC + matched filter, evaluated as the inverse transform's first pass loads its elements: lane
    // (b, lane) needs bins k1 + n1*m of output column x0 + b, each interpolated from 16 taps of tile row m.
    const int xc = x0 + b;
    const bool xc_ok = xc < a.nrv;
    const ColConst cc = a.col[xc_ok ? xc : 0];
    const double step = a.prf / (double)a.nbins;
    const double half = 0.5 * a.prf;
    float ws[16];  // signed window (-1)^i w[i]
#pragma unroll
    for (int i = 0; i < 16; ++i) ws[i] = (i & 1) ? -a.window[i] : a.window[i];
    const double inv_step = (double)a.nbins / a.prf;
    const double half_bins = 0.5 * a.beta * (double)a.nbins;
    const double m_dc = rint(cc.dc_off * inv_step);
    auto rcmc_load = [&](int m) -> float2 {
        const int k = k1 + a.n1 * m;
        const double fb = (k < a.nbins / 2) ? (double)k * step : (double)(k - a.nbins) * step;
        const double fn = fz_alias(fb, cc.dc, a.prf);
        const double shift = cc.shift_k * fn * fn;
        const double p = round(shift);
        const int pi = (int)fmin(p, 4096.0);  // p >= 0; the cap only guards the integer conversion
        const int cx = xc + pi;
        float2 val = make_float2(0.f, 0.f);
        if (xc_ok && !(cx - 8 < 0 || cx + 8 >= a.nrv)) {
            // the gate guarantees p <= kAzFusedReach; the clamp only keeps the taps inside the tile
            const int lc = min(max(kFzLeft + b + pi - 8, 0), kFzCols - 16);
            const float2* tap = tl + m * kFzPitch + lc;
            const float d = (float)(shift - p);
            val = (d == 0.0f) ? tap[8] : sinc16_interp(tap, ws, d);
        }
        if (outside_band(fn, m_dc, inv_step, half_bins)) return make_float2(0.f, 0.f);
        double fc = fb - cc.dc_off;
        if (fc > half)
            fc -= a.prf;
        else if (fc < -half)
            fc += a.prf;
        const double turns = cc.inv_ka * (0.5 * fc * fc + cc.dc_off * fb);
        const float frc = (float)(turns - rint(turns));
        float sn, cs;
        sincos_turns(frc, &sn, &cs);
        return make_float2(val.x * cs + val.y * sn, val.y * cs - val.x * sn);
    };

    // ---- inverse transform over k2 of the produced columns (tile columns 0..15 serve as its exchange buffer
    // once every tap has been read), inverse inter-pass twiddle, store
    float2 v[kMaxE];
    const int ccol = b;
    auto lds_store = [&](int i, float2 x) { tl[i * kFzPitch + ccol] = x; };
    auto lds_load = [&](int i) { return tl[i * kFzPitch + ccol]; };
    auto ir = [&](int p) { return plan_radix(pl, pl.npass - 1 - p); };
    const int ri0 = ir(0);
#define LOAD_T_(RR) pass_load<RR>(v, lane, nlanes, fft_n, rcmc_load)
    ASARF_DISPATCH(ri0, LOAD_T_)
#undef LOAD_T_
    __syncthreads();  // every tap is read before the passes overwrite the tile
    ASARF_DISPATCH(ri0, ASARF_COMPUTE1_INV)
    ASARF_RUN_PASSES(true, pl.npass, ir)
    float2* dst = a.out + (xc_ok ? xc : 0);
    const float2* twf = a.tw_full;
    auto out_write = [&](int n, float2& val, bool ok) {
        if (!ok || !xc_ok) return;
        const float2 wv = twf[(long)k1 * n];
        dst[(row0 + n) * a.stride] = cmulc(val, wv);
    };
    const int ril = ir(pl.npass - 1);
#define VISIT_WR_(RR) pass_visit<RR>(v, lane, nlanes, fft_n, out_write)
    ASARF_DISPATCH(ril, VISIT_WR_)
#undef VISIT_WR_
}

#define FZ_SETS(X) X(50) X(14) X(82) X(22) X(72) X(101) X(kRadixAll)

size_t az_fused_lds_bytes(int n2) { return sizeof(float2) * (size_t)(n2 * kFzPitch + tw_lds_elems(n2)); }

hipError_t launch_az_fused(const AzFusedArgs& a, hipStream_t s) {
    const size_t lds = az_fused_lds_bytes(a.sub.n);
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    const unsigned grid = (unsigned)((a.items + 7) / 8 * 8);
    int m = 0;
    for (int i = 0; i < a.sub.npass; ++i) m |= radix_bit(plan_radix(a.sub, i));
#define X_(SET)                                                                               \
    if ((m & ~(SET)) == 0) {                                                                  \
        const hipError_t e_ = hipFuncSetAttribute((const void*)k_az_fused<SET>,               \
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, \
                                                  (int)lds);                                  \
        if (e_ != hipSuccess) return e_;                                                      \
        hipLaunchKernelGGL((k_az_fused<SET>), dim3(grid), dim3(kThreads), lds, s, a);         \
        return hipGetLastError();                                                             \
    }
    FZ_SETS(X_)
#undef X_
    return hipErrorInvalidValue;
}

}  // namespace asarf
