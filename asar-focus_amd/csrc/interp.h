// Device helpers of the RCMC / matched-filter kernels.
#pragma once

#include <hip/hip_runtime.h>

namespace asarf {

// 16-tap windowed-sinc interpolation at fractional offset d (d != 0) from taps tap[i] at offsets 8 - i, with
// signed window ws[i] = (-1)^i w[i] (sar/range_doppler_algorithm.cu:107-195: sinc(d + 8 - i) w[i], normalised by
// the weight sum). sin(pi (d + 8 - i)) = (-1)^i sin(pi d) is common to every tap and cancels in the
// normalisation, so the weights are ws[i] / (d + 8 - i). d + (8 - i) is exact to one rounding and the 1-ulp
// reciprocal error cancels in the normalisation to first order. (Measured alternatives, both slower: a
// reciprocal-free form, ws[i] times the product of the other fifteen factors, which needs both product chains
// live at once; and the window folded into the reciprocal, 1 / (d / ws[i] + (8 - i) / ws[i]).)
__device__ __forceinline__ float2 sinc16_interp(const float2* tap, const float (&ws)[16], float d) {
    float2 acc = make_float2(0.f, 0.f);
    float den = 0.f;
#pragma unroll
    for (int i = 0; i < 16; ++i) {
        const float wi = ws[i] * __builtin_amdgcn_rcpf(d + (float)(8 - i));
        const float2 t = tap[i];
        acc.x = fmaf(wi, t.x, acc.x);
        acc.y = fmaf(wi, t.y, acc.y);
        den += wi;
    }
    if (den == 0.f) return make_float2(0.f, 0.f);
    const float inv = __builtin_amdgcn_rcpf(den);
    return make_float2(acc.x * inv, acc.y * inv);
}

// sinc16_interp without taps 0 and 15. Their Blackman weights are zero: numpy.blackman(16)[0] = [15] =
// 0.42 - 0.5 + 0.08, -1.4e-17 in floating point, so their products are 1e-17 of the other taps' and vanish in the
// FP32 sums. Two reciprocals and six FMAs fewer per pixel; the caller checks the window's end weights.
__device__ __forceinline__ float2 sinc16_interior(const float2* tap, const float (&ws)[16], float d) {
    float2 acc = make_float2(0.f, 0.f);
    float den = 0.f;
#pragma unroll
    for (int i = 1; i < 15; ++i) {
        const float wi = ws[i] * __builtin_amdgcn_rcpf(d + (float)(8 - i));
        const float2 t = tap[i];
        acc.x = fmaf(wi, t.x, acc.x);
        acc.y = fmaf(wi, t.y, acc.y);
        den += wi;
    }
    if (den == 0.f) return make_float2(0.f, 0.f);
    const float inv = __builtin_amdgcn_rcpf(den);
    return make_float2(acc.x * inv, acc.y * inv);
}

// Bin frequency f aliased into the prf-wide band centred on the Doppler centroid dc
// (sar/range_doppler_algorithm.cu:122-140).
__device__ __forceinline__ double alias_about_dc(double f, double dc, double prf) {
    const double half = 0.5 * prf;
    if (dc < 0.0) {
        if (f > half + dc) f -= prf;
    } else if (dc > 0.0) {
        if (f < -half + dc) f += prf;
    }
    return f;
}

// Azimuth band test of the matched filter (sar/range_doppler_algorithm.cu:238-246): a bin is dropped when its
// aliased frequency fn lies more than beta * prf / 2 from the bin-quantised centroid dc_off. Both are whole
// multiples of the bin spacing, so the test runs on bin offsets, |rint(fn / step) - rint(dc_off / step)| against
// half_bins = beta * nbins / 2: a bin exactly on the edge is kept (the reference's strict inequality in exact
// arithmetic) however the edge would round in Hz, and every kernel decides every bin the same way.
__device__ __forceinline__ bool outside_band(double fn, double m_dc, double inv_step, double half_bins) {
    return fabs(rint(fn * inv_step) - m_dc) > half_bins + 1e-9;
}

// sin(2 pi x) and cos(2 pi x) for |x| <= 0.5 from the hardware sine/cosine, whose argument is in turns
// (measured on gfx950 against FP64 over [-0.5, 0.5]: max abs error 1.3e-7).
__device__ __forceinline__ void sincos_turns(float x, float* s, float* c) {
    *s = __builtin_amdgcn_sinf(x);
    *c = __builtin_amdgcn_cosf(x);
}

}  // namespace asarf
