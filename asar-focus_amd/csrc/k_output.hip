// Output conditioning (the step after the focusing path): the SLC as big-endian int16 I/Q records and the
// detected image as uint16 dB records, each record led by a zeroed MDS record header
// (envisat_format/src/envisat_format_kernels.cu:28-114).
#include <cmath>

#include "kernels.h"

namespace asarf {

constexpr int kOutThreads = 256;

// float -> integer the way the reference's device casts behave: truncation toward zero, saturating at the
// destination range, NaN -> 0
__device__ __forceinline__ int sat_trunc(float v, float lo, float hi) {
    if (!(v == v)) return 0;
    v = fminf(fmaxf(v, lo), hi);
    return (int)v;  // |v| <= 65535: exact truncation
}

// CalibrateClampToBigEndian (:28-68): record y = header zero bytes, then x_size pixels of (I, Q) as big-endian
// int16 of clamp(value * calibration). One lane per pixel; the record pitch (header + 4 x_size bytes) is not a
// multiple of 4 for the 17-byte ASAR header, so the pixels are stored byte by byte.
__global__ __launch_bounds__(kOutThreads) void k_condition_slc(OutputArgs a) {
    const int x = blockIdx.x * kOutThreads + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.x_size) return;
    unsigned char* rec = a.dest + (long)y * a.record_bytes;
    if (x < a.header) rec[x] = 0;  // header bytes (every header is shorter than a row of pixels)
    const float2 v = a.slc[(long)y * a.x_stride + x];
    const int i = sat_trunc(v.x * a.calibration, -32768.f, 32767.f);
    const int q = sat_trunc(v.y * a.calibration, -32768.f, 32767.f);
    unsigned char* p = rec + a.header + 4 * x;
    p[0] = (unsigned char)((i >> 8) & 0xFF);
    p[1] = (unsigned char)(i & 0xFF);
    p[2] = (unsigned char)((q >> 8) & 0xFF);
    p[3] = (unsigned char)(q & 0xFF);
}

// CalibrateClampToBigEndianDetected (:70-98): record y = header zero bytes, then x_size uint16 of
// 10 log10(value * calibration) in single precision, saturated to [0, 65535] (negative and NaN -> 0). The
// reference stores these low byte first (its detected kernel does not byte-swap, despite its name); so does this.
__global__ __launch_bounds__(kOutThreads) void k_condition_detected(OutputArgs a) {
    const int x = blockIdx.x * kOutThreads + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= a.x_size) return;
    unsigned char* rec = a.dest + (long)y * a.record_bytes;
    if (x < a.header) rec[x] = 0;
    const float v = a.det[(long)y * a.x_stride + x] * a.calibration;
    const int u = sat_trunc(10.0f * log10f(v), 0.f, 65535.f);
    unsigned char* p = rec + a.header + 2 * x;
    p[0] = (unsigned char)(u & 0xFF);
    p[1] = (unsigned char)((u >> 8) & 0xFF);
}

hipError_t launch_condition(const OutputArgs& a, int detected, hipStream_t s) {
    if (a.header > a.x_size) return hipErrorInvalidValue;  // the header is zeroed by the row's first lanes
    const dim3 grid((a.x_size + kOutThreads - 1) / kOutThreads, a.y_size);
    if (detected)
        hipLaunchKernelGGL(k_condition_detected, grid, dim3(kOutThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_condition_slc, grid, dim3(kOutThreads), 0, s, a);
    return hipGetLastError();
}

}  // namespace asarf
