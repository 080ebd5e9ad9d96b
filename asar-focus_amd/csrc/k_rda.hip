// Range-Doppler domain kernels: per-column constants and the fused 16-tap sinc RCMC + azimuth matched filter.
//
// RCMC follows sar/range_doppler_algorithm.cu:107-195: shift = lambda^2 f^2 R0 / (8 Vr^2 dr), integer part by
// round-half-away, 16-tap Blackman-windowed sinc normalised by the sum of its weights, zero when a tap leaves
// [0, Nrv). The weights use sinc(d + 8 - i) = (-1)^i sin(pi d) / (pi (d + 8 - i)); the common factor
// sin(pi d)/pi cancels in the normalisation, so no transcendental is evaluated per tap (sinc16_interp).
// The matched filter follows range_doppler_algorithm.cu:197-309 with FP64 geometry and phase (reduced to turns
// before the hardware sine/cosine, whose argument is in turns).
#include "interp.h"
#include "kernels.h"
#include "poly.h"
#include "rcmc_net.h"

namespace asarf {

// numpy.blackman(16) rounded to float (the table at sar/range_doppler_algorithm.cu:49-52 holds the same values).
// The kernels take it by argument (RcmcArgs::window), so no device symbol has to be initialised per device.
void rcmc_window(float w[16]) {
    for (int i = 0; i < 16; ++i) {
        const double a = 2.0 * kPi * (double)i / 15.0;
        w[i] = (float)(0.42 - 0.5 * std::cos(a) + 0.08 * std::cos(2.0 * a));
    }
}

static RcmcArgs with_window(const RcmcArgs& in) {
    RcmcArgs a = in;
    rcmc_window(a.window);
    return a;
}

__global__ void k_col_const(ColConstArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const bool valid = x < a.nrv;
    int r = 0;
    if (valid) {
        const double step = a.prf / (double)a.nbins;
        const double dc = polyval(a.dc_poly ? a.dc_poly : a.dc_val, 3, (double)x);
        const double r0 = a.r_first + (double)x * a.range_spacing;
        const double vr = polyval(a.vr_poly, 3, (double)x);
        ColConst c;
        c.dc = dc;
        c.dc_off = round(dc / step) * step;
        c.shift_k = a.wavelength * a.wavelength * r0 / (8.0 * vr * vr * a.range_spacing);
        c.inv_ka = a.wavelength * r0 / (2.0 * vr * vr);
        a.col[x] = c;
        // The aliased bin frequency stays within prf/2 of the centroid, so |f| <= prf/2 + |dc| bounds the
        // shift of every bin (plus one pixel against rounding).
        const double fmax = (0.5 * a.prf + fabs(dc)) * (1.0 + 1e-12);
        const double smax = c.shift_k * fmax * fmax;
        r = (smax >= 0.0 && smax < 1e6) ? (int)floor(smax + 0.5) + 1 : 1000000;
    }
    if (a.reach) {
        // one atomic per wave; every lane of the wave takes part in the reduction
        for (int o = 32; o > 0; o >>= 1) r = max(r, __shfl_xor(r, o));
        if ((threadIdx.x & 63) == 0 && r > 0) atomicMax(a.reach, r);
    }
}

// Clears the RCMC shift bound before k_col_const takes its maximum (a one-lane vector store; no memset, so the
// sequence can be captured in a graph).
__global__ void k_clear_int(int* p) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *p = 0;
}

hipError_t launch_col_const(const ColConstArgs& a, hipStream_t s) {
    if (a.clear_reach && a.reach) hipLaunchKernelGGL(k_clear_int, dim3(1), dim3(64), 0, s, a.reach);
    hipLaunchKernelGGL(k_col_const, dim3((a.nrv + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// Plain 8-tap Blackman sinc at fractional offset d (d != 0), taps tap[i] at offsets 4 - i (same normalised form
// as the 16-tap interpolator).
__device__ __forceinline__ float2 sinc8_interp(const float2* tap, const float* window8, float d) {
    float2 acc = make_float2(0.f, 0.f);
    float den = 0.f;
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const float wi = ((i & 1) ? -window8[i] : window8[i]) * __builtin_amdgcn_rcpf(d + (float)(4 - i));
        const float2 t = tap[i];
        acc.x += wi * t.x;
        acc.y += wi * t.y;
        den += wi;
    }
    const float inv = __builtin_amdgcn_rcpf(den);
    return (den == 0.f) ? make_float2(0.f, 0.f) : make_float2(acc.x * inv, acc.y * inv);
}

// Learned 8-tap interpolation (complex taps from the network's table, taps tap[i] at offsets 4 - i) where the
// judge accepted the (d, slope) cell; false (nothing written) where it did not.
__device__ __forceinline__ bool learned_interp(const RcmcArgs& a, const float2* tap, float d, double shift_k,
                                               double fn, float2* out) {
    const int cell = net_cell(d, net_slope(shift_k, fn, a.prf), a.lut_nd, a.lut_ns);
    const int jc = (cell / a.lut_nd) * a.judge_nd + (cell % a.lut_nd) * a.judge_nd / a.lut_nd;
    if (!a.lut_ok[jc]) return false;
    const float2* w8 = a.lut + (long)cell * kNetTaps;
    float2 acc = make_float2(0.f, 0.f);
#pragma unroll
    for (int i = 0; i < kNetTaps; ++i) {
        const float2 wi = w8[i], t = tap[i];
        acc.x += wi.x * t.x - wi.y * t.y;
        acc.y += wi.x * t.y + wi.y * t.x;
    }
    *out = acc;
    return true;
}

// One workgroup per frequency row at a time (in place, or out of place into a.out), rows strided over a grid of
// at most 2048 workgroups. The row is staged in LDS so every output can read its 16 taps. 512 lanes per row: the
// LDS row (8 B x nrv) limits the workgroups per CU, so wider workgroups keep more waves in flight. This is the
// kernel for RCMC shifts beyond the column-tiled kernel's halo.
constexpr int kRcmcThreads = 512;
__global__ __launch_bounds__(kRcmcThreads) __attribute__((amdgpu_waves_per_eu(7))) void k_rcmc_mf(RcmcArgs a) {
    extern __shared__ float2 row[];
    if (a.fused_gate && *a.fused_gate <= kAzCoreReach) return;
    if (a.tile_gate && *a.tile_gate <= kRcmcTileReach) return;
    const int nrv = a.nrv;
    const double step = a.prf / (double)a.nbins;
    const double inv_step = (double)a.nbins / a.prf;
    const double half_bins = 0.5 * a.beta * (double)a.nbins;
    const double half = 0.5 * a.prf;
    float w[16];  // signed window (-1)^i w[i]
#pragma unroll
    for (int i = 0; i < 16; ++i) w[i] = (i & 1) ? -a.window[i] : a.window[i];
    // rows strided over a bounded grid, so that the launch costs little when the gate sends the work elsewhere
    for (int q = blockIdx.x; q < a.nbins; q += gridDim.x) {
        const float2* img = a.img + (long)q * a.stride;
        float2* dst = (a.out ? a.out : a.img) + (long)q * a.stride;
        for (int x = threadIdx.x; x < nrv; x += kRcmcThreads) row[x] = img[x];
        __syncthreads();
        const int k = (a.n2 == 1) ? q : (q / a.n2 + a.n1 * (q % a.n2));
        const double fb = (k < a.nbins / 2) ? (double)k * step : (double)(k - a.nbins) * step;
        for (int x = threadIdx.x; x < nrv; x += kRcmcThreads) {
            const ColConst cc = a.col[x];
            const double fn = alias_about_dc(fb, cc.dc, a.prf);
            // outside the azimuth band the matched filter zeroes the pixel: skip its interpolation
            const bool out_band = a.mode != 1 && outside_band(fn, rint(cc.dc_off * inv_step), inv_step, half_bins);
            float2 val = make_float2(0.f, 0.f);
            if (out_band) {
            } else if (a.mode != 2) {
                const double shift = cc.shift_k * fn * fn;
                const double p = floor(shift + 0.5);  // round half away from zero: shift >= 0
                const int cx = x + (int)p;
                if (cx - 8 < 0 || cx + 8 >= nrv) {
                    val = make_float2(0.f, 0.f);
                } else {
                    const float d = (float)(shift - p);
                    bool sinc16 = a.interp == 0;
                    if (a.interp == 1) {
                        // learned taps where the judge accepted the cell, else the 16-tap sinc
                        sinc16 = !learned_interp(a, row + cx - 4, d, cc.shift_k, fn, &val);
                    } else if (a.interp == 2) {
                        val = (d == 0.0f) ? row[cx] : sinc8_interp(row + cx - 4, a.window8, d);
                    }
                    if (sinc16) val = (d == 0.0f) ? row[cx] : sinc16_interp(row + cx - 8, w, d);
                }
            } else {
                val = row[x];
            }
            if (a.mode != 1 && !out_band) {
                double fc = fb - cc.dc_off;
                if (fc > half)
                    fc -= a.prf;
                else if (fc < -half)
                    fc += a.prf;
                // phase / (2 pi) = inv_ka * (fc^2 / 2 + dc_off * f_bin)
                const double turns = cc.inv_ka * (0.5 * fc * fc + cc.dc_off * fb);
                const float fr = (float)(turns - rint(turns));
                float sn, cs;
                sincos_turns(fr, &sn, &cs);
                // H = exp(-i phase)
                val = make_float2(val.x * cs + val.y * sn, val.y * cs - val.x * sn);
            }
            dst[x] = val;
        }
        __syncthreads();  // every lane is done with this row in LDS before the next one is loaded
    }
}

// Column-tiled RCMC + matched filter (16-tap sinc, or the learned / 8-tap interpolator as template INTERP = the
// RcmcArgs::interp code; out of place): a workgroup owns 256 columns for kRtRows
// consecutive storage rows. Each lane loads its column's FP64 constants once and reuses them for every row; the
// rows' range segments (with an 8-column left halo and an 8 + kRcmcTileReach right halo) are double-buffered in
// LDS. Identical arithmetic to k_rcmc_mf, so both give the same result; the plan gates between them on the
// device-side shift bound.
constexpr int kRtCols = 256;
constexpr int kRtRows = 32;
constexpr int kRtSeg = kRtCols + 16 + kRcmcTileReach;

// One tile: kRtCols columns from x0 by kRtRows bins from q0 (item = bin tile * column tiles + column tile).
template <int INTERP>
__device__ __forceinline__ void rcmc_tile_item(const RcmcArgs& a, float2 (*seg)[kRtSeg], int item) {
    const int ntx = (a.nrv + kRtCols - 1) / kRtCols;
    const int x0 = (item % ntx) * kRtCols;
    const int q0 = (item / ntx) * kRtRows;
    const int q1 = min(q0 + kRtRows, a.nbins);
    const int t = threadIdx.x;
    const int nrv = a.nrv;
    const int x = x0 + t;
    const bool xok = x < nrv;
    const ColConst cc = a.col[xok ? x : nrv - 1];
    const double step = a.prf / (double)a.nbins;
    const double inv_step = (double)a.nbins / a.prf;
    const double half_bins = 0.5 * a.beta * (double)a.nbins;
    const double m_dc = rint(cc.dc_off * inv_step);  // the column's quantised centroid in bins
    const double half = 0.5 * a.prf;
    float w[16];  // signed window (-1)^i w[i]
#pragma unroll
    for (int i = 0; i < 16; ++i) w[i] = (i & 1) ? -a.window[i] : a.window[i];
    float2 pre0, pre1;
    auto fetch = [&](int q) {
        const float2* r = a.img + (long)q * a.stride;
        const int xa = x0 - 8 + t, xb = x0 - 8 + t + kRtCols;
        pre0 = (xa >= 0 && xa < nrv) ? r[xa] : make_float2(0.f, 0.f);
        pre1 = (t < kRtSeg - kRtCols && xb < nrv) ? r[xb] : make_float2(0.f, 0.f);
    };
    auto deposit = [&](int buf) {
        seg[buf][t] = pre0;
        if (t < kRtSeg - kRtCols) seg[buf][t + kRtCols] = pre1;
    };
    // one row ahead in registers and one in LDS (a deeper register ring measured slower: fewer waves per SIMD)
    fetch(q0);
    deposit(0);
    if (q0 + 1 < q1) fetch(q0 + 1);
    __syncthreads();
    for (int q = q0; q < q1; ++q) {
        const int buf = (q - q0) & 1;
        if (q + 1 < q1) {
            deposit(buf ^ 1);  // its readers finished before the barrier that ended the previous row
            if (q + 2 < q1) fetch(q + 2);
        }
        const int k = (a.n2 == 1) ? q : (q / a.n2 + a.n1 * (q % a.n2));
        const double fb = (k < a.nbins / 2) ? (double)k * step : (double)(k - a.nbins) * step;
        const double fn = alias_about_dc(fb, cc.dc, a.prf);
        // outside the azimuth band the matched filter zeroes the pixel: skip its interpolation
        const bool out_band = a.mode != 1 && outside_band(fn, m_dc, inv_step, half_bins);
        float2 val = make_float2(0.f, 0.f);
        if (out_band) {
        } else if (a.mode != 2) {
            const double shift = cc.shift_k * fn * fn;
            const double p = floor(shift + 0.5);  // round half away from zero: shift >= 0
            const int pi = (int)fmin(p, 4096.0);
            const int cx = x + pi;
            if (cx - 8 < 0 || cx + 8 >= nrv) {
                val = make_float2(0.f, 0.f);
            } else {
                const float d = (float)(shift - p);
                // the gate guarantees pi <= kRcmcTileReach; the clamp only keeps the taps inside the segment
                const float2* tap = &seg[buf][min(t + pi, kRtSeg - 16)];
                // the same interpolator choices as k_rcmc_mf: learned taps where the judge accepted the cell
                // (at any d), else the sinc (exact copy at d == 0)
                if (INTERP == 1 && learned_interp(a, tap + 4, d, cc.shift_k, fn, &val)) {
                } else if (d == 0.0f) {
                    val = tap[8];
                } else if (INTERP == 2) {
                    val = sinc8_interp(tap + 4, a.window8, d);
                } else {
                    val = sinc16_interp(tap, w, d);
                }
            }
        } else {
            val = seg[buf][t + 8];
        }
        if (a.mode != 1 && !out_band) {
            double fc = fb - cc.dc_off;
            if (fc > half)
                fc -= a.prf;
            else if (fc < -half)
                fc += a.prf;
            const double turns = cc.inv_ka * (0.5 * fc * fc + cc.dc_off * fb);
            const float fr = (float)(turns - rint(turns));
            float sn, cs;
            sincos_turns(fr, &sn, &cs);
            val = make_float2(val.x * cs + val.y * sn, val.y * cs - val.x * sn);
        }
        if (xok) a.out[(long)q * a.stride + x] = val;
        __syncthreads();
    }
}

// One tile per workgroup (LOOP = false), or a gated fallback on a bounded grid whose workgroups take the tiles in
// turn (LOOP = true), as the column kernels do.
template <int INTERP, bool LOOP>
__global__ __launch_bounds__(kRtCols) void k_rcmc_tile(RcmcArgs a) {
    __shared__ float2 seg[2][kRtSeg];
    if (a.tile_gate && *a.tile_gate > kRcmcTileReach) return;
    if (a.fused_gate && *a.fused_gate <= kAzCoreReach) return;
    if constexpr (LOOP) {
        const int nitems = (a.nrv + kRtCols - 1) / kRtCols * ((a.nbins + kRtRows - 1) / kRtRows);
        for (int item = blockIdx.x; item < nitems; item += gridDim.x) rcmc_tile_item<INTERP>(a, seg, item);
    } else {
        rcmc_tile_item<INTERP>(a, seg, blockIdx.y * gridDim.x + blockIdx.x);
    }
}

hipError_t launch_rcmc_tile(const RcmcArgs& args, hipStream_t s) {
    const RcmcArgs a = with_window(args);
    if (!a.out || a.out == a.img) return hipErrorInvalidValue;  // out of place only
    // a gated fallback (it works only when the fused azimuth core does not) runs on a bounded grid
    if (a.fused_gate) {
        const dim3 grid(1024);
        switch (a.interp) {
            case 0: hipLaunchKernelGGL((k_rcmc_tile<0, true>), grid, dim3(kRtCols), 0, s, a); break;
            case 1: hipLaunchKernelGGL((k_rcmc_tile<1, true>), grid, dim3(kRtCols), 0, s, a); break;
            case 2: hipLaunchKernelGGL((k_rcmc_tile<2, true>), grid, dim3(kRtCols), 0, s, a); break;
            default: return hipErrorInvalidValue;
        }
        return hipGetLastError();
    }
    const dim3 grid((a.nrv + kRtCols - 1) / kRtCols, (a.nbins + kRtRows - 1) / kRtRows);
    switch (a.interp) {
        case 0: hipLaunchKernelGGL((k_rcmc_tile<0, false>), grid, dim3(kRtCols), 0, s, a); break;
        case 1: hipLaunchKernelGGL((k_rcmc_tile<1, false>), grid, dim3(kRtCols), 0, s, a); break;
        case 2: hipLaunchKernelGGL((k_rcmc_tile<2, false>), grid, dim3(kRtCols), 0, s, a); break;
        default: return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

hipError_t launch_rcmc_mf(const RcmcArgs& args, hipStream_t s) {
    const RcmcArgs a = with_window(args);
    const size_t lds = sizeof(float2) * (size_t)a.nrv;
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_rcmc_mf, dim3(std::min(a.nbins, 2048)), dim3(kRcmcThreads), lds, s, a);
    return hipGetLastError();
}

}  // namespace asarf
