// Learned RCMC interpolator: network shape, parameter layout, training-pair layout and the device helpers that
// the pair generator, the trainer, the table builder and the RCMC kernel must compute identically.
//
// The network maps (u0, u1) = (2 d, slope) to 8 complex interpolation taps, d = shift - round(shift) in
// [-0.5, 0.5] the fractional RCMC shift and slope = (d shift / d f) * (PRF / 2) / 4 the local migration slope
// along the Doppler axis, clamped to [-1, 1]. Taps apply to the 8 samples at cx - 4 .. cx + 3 (cx = x + round
// (shift)), the centre of the reference's 16-tap window cx - 8 .. cx + 7 (range_doppler_algorithm.cu:171-190).
// MLP 2 -> 64 -> 64 -> 16, tanh hidden units, linear output; weights are used rounded to bfloat16, sums in FP32.
// Interpolation weights = plain 8-tap Blackman sinc weights (real) + network outputs (complex).
#pragma once

#include <hip/hip_runtime.h>

#include <cmath>

#include "kernels.h"

namespace asarf {

constexpr int kNetIn = 2, kNetH = 64, kNetOut = 16, kNetTaps = 8;
constexpr int kNetOW1 = 0;                              // W1[64][2]
constexpr int kNetOB1 = kNetOW1 + kNetH * kNetIn;       // b1[64]
constexpr int kNetOW2 = kNetOB1 + kNetH;                // W2[64][64] (row = output unit)
constexpr int kNetOB2 = kNetOW2 + kNetH * kNetH;        // b2[64]
constexpr int kNetOW3 = kNetOB2 + kNetH;                // W3[16][64]
constexpr int kNetOB3 = kNetOW3 + kNetOut * kNetH;      // b3[16]
constexpr int kNetP = kNetOB3 + kNetOut;                // 5392 parameters
// Training pair: u0, u1, 8 complex samples (re, im), complex target; samples and target divided by the RMS of
// the 8 samples (interpolation is linear, so the pair is scale free).
constexpr int kPairFloats = 20;

// Round-to-nearest-even to bfloat16, returned as float (finite inputs).
__host__ __device__ inline float bf16_round(float x) {
    union {
        float f;
        unsigned u;
    } v;
    v.f = x;
    const unsigned lsb = (v.u >> 16) & 1u;
    v.u = (v.u + 0x7FFFu + lsb) & 0xFFFF0000u;
    return v.f;
}

__host__ __device__ inline float net_slope(double shift_k, double fn, double prf) {
    const double s = shift_k * fn * prf * 0.25;
    return (float)(s > 1.0 ? 1.0 : (s < -1.0 ? -1.0 : s));
}

// Table cell of (d, slope): nd cells over d in [-0.5, 0.5], ns cells over slope in [-1, 1].
__host__ __device__ inline int net_cell(float d, float slope, int nd, int ns) {
    int id = (int)floorf((d + 0.5f) * (float)nd);
    id = id < 0 ? 0 : (id >= nd ? nd - 1 : id);
    int is = (int)floorf((slope + 1.0f) * 0.5f * (float)ns);
    is = is < 0 ? 0 : (is >= ns ? ns - 1 : is);
    return is * nd + id;
}

// Plain 8-tap Blackman-windowed sinc weights (taps cx - 4 .. cx + 3) at fractional shift d, normalised to unit
// sum: sinc(d + 4 - j) = (-1)^j sin(pi d) / (pi (d + 4 - j)) and the common factor cancels. The network
// outputs corrections to these weights (real part), so its outputs stay small and bfloat16 weights keep
// their relative precision where it matters.
__device__ inline void sinc8_weights(float d, const float* w8, float* out) {
    if (d == 0.0f) {
#pragma unroll
        for (int j = 0; j < 8; ++j) out[j] = j == 4 ? 1.0f : 0.0f;
        return;
    }
    float den = 0.f;
#pragma unroll
    for (int j = 0; j < 8; ++j) {
        out[j] = ((j & 1) ? -w8[j] : w8[j]) * __builtin_amdgcn_rcpf(d + (float)(4 - j));
        den += out[j];
    }
    const float inv = __builtin_amdgcn_rcpf(den);
#pragma unroll
    for (int j = 0; j < 8; ++j) out[j] *= inv;
}

// numpy.blackman(M) value n, computed in FP64 and rounded to float (host side).
inline float blackman_value(int n, int m) {
    const double a = 2.0 * 3.14159265358979323846 * (double)n / (double)(m - 1);
    return (float)(0.42 - 0.5 * cos(a) + 0.08 * cos(2.0 * a));
}

struct PairArgs {
    const float2* img;  // range-Doppler image after the azimuth FFT (row q holds bin q/n2 + n1*(q%n2))
    long stride;
    int nrv, nbins, n1, n2;
    const ColConst* col;
    double prf;
    float window[16];   // Blackman-16 of the sinc interpolator
    long n_pairs;
    unsigned long long seed;
    float* pairs;       // n_pairs x kPairFloats
};

struct NetLutArgs {
    const float* wq;    // bf16-rounded parameters
    int nd, ns;
    float2* lut;        // [ns][nd][8] complex taps at the cell centres
    float window8[8];
};

struct NetGradArgs {
    const float* wq;
    const float* pairs;
    long n;
    float* partials;    // [groups][kNetP]
    float* loss_part;   // [groups]
    float window8[8];
};

struct NetJudgeArgs {
    const float2* lut;
    int nd, ns, judge_nd;      // judge cells: judge_nd x ns, each covering nd / judge_nd table cells
    const float* pairs;
    long n;
    float window8[8];          // Blackman-8 of the plain 8-tap sinc
    float* err_sum;            // [ns][judge_nd] squared error of the learned taps
    float* count;              // [ns][judge_nd]
    float* sinc8_err_sum;      // [ns][judge_nd] squared error of the plain 8-tap sinc
};

hipError_t launch_rcmc_pairs(const PairArgs& a, hipStream_t s);
hipError_t launch_net_grad(const NetGradArgs& a, hipStream_t s);
hipError_t launch_net_reduce(const float* partials, int groups, float* out, hipStream_t s);
// loss_slot (may be null) receives the mean loss of the batch whose partials are reduced
hipError_t launch_net_adam(float* master, float* m, float* v, float* wq, const float* partials, int groups,
                           float inv_batch, float lr, int step, const float* loss_part, float* loss_slot,
                           hipStream_t s);
hipError_t launch_net_lut(const NetLutArgs& a, hipStream_t s);
hipError_t launch_net_judge(const NetJudgeArgs& a, hipStream_t s);
constexpr int kNetGroupSamples = 64;  // samples per gradient workgroup

}  // namespace asarf
