// IMP tail: azimuth multilook (sar/az_multilook.cu:25-80) with the range-spreading-loss factor fused into its
// store (sar/range_spreading_loss.cu:29-40), the standalone RSL pass, and slant-to-ground resampling
// (sar/srgr.cu:49-71, polynomial evaluated in FP64).
#include "kernels.h"
#include "poly.h"

namespace asarf {

// Each lane owns one range column and walks a chunk of output rows; the |s|^2 of every input row is read once
// and kept in a per-lane ring of 8 rows in LDS (a window spans at most 5 input rows).
constexpr int kMlRows = 64;  // output rows per workgroup

__global__ __launch_bounds__(256) void k_multilook(MultilookArgs a) {
    __shared__ float ring[8][256];
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int out_rows = min(a.out_rows_dev ? *a.out_rows_dev : a.out_rows, a.out_rows);
    const int y0 = blockIdx.y * kMlRows;
    const int y1 = min(y0 + kMlRows, out_rows);
    if (x >= a.nr || y0 >= y1) return;
    const int row0 = a.row0_dev ? *a.row0_dev : a.row0;
    const int in_rows = a.in_rows_dev ? *a.in_rows_dev : a.in_rows;
    const float2* col = a.slc + (long)row0 * a.stride + a.col0 + x;
    float* rg = &ring[0][threadIdx.x];
    int next = -1;
    for (int y = y0; y < y1; ++y) {
        // resampling position in float, exactly as the reference computes it
        const float in_idx = (float)y * a.ratio;
        const float mu = in_idx - (float)(int)in_idx;
        int s0 = (int)(in_idx - 2.0f), e0 = (int)(in_idx + 2.0f);
        int s1 = (int)(in_idx - 1.0f), e1 = (int)(in_idx + 3.0f);
        s0 = max(s0, 0);
        e0 = min(e0, in_rows);
        s1 = max(s1, 0);
        e1 = min(e1, in_rows);
        if (next < s0) next = s0;
        const int hi = max(e0, e1);
        for (; next < hi; ++next) {
            const float2 v = col[(long)next * a.stride];
            rg[(next & 7) * 256] = v.x * v.x + v.y * v.y;
        }
        float sum0 = 0.f, sum1 = 0.f;
        for (int i = s0; i < e0; ++i) sum0 += rg[(i & 7) * 256];
        for (int i = s1; i < e1; ++i) sum1 += rg[(i & 7) * 256];
        const int c0 = max(e0 - s0, 0), c1 = max(e1 - s1, 0);
        if (c0) sum0 /= (float)c0;
        if (c1) sum1 /= (float)c1;
        float val = (c0 && c1) ? sum0 + (sum1 - sum0) * mu : 0.f;
        if (a.rsl) val *= a.rsl[x];
        a.out[(long)y * a.nr + x] = val;
    }
}

hipError_t launch_multilook(const MultilookArgs& a, hipStream_t s) {
    if (a.out_rows <= 0 || a.nr <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_multilook, dim3((a.nr + 255) / 256, (a.out_rows + kMlRows - 1) / kMlRows), dim3(256), 0, s,
                       a);
    return hipGetLastError();
}

__global__ void k_rsl(float* img, int nr, int rows, const float* tab) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= nr || y >= rows) return;
    img[(long)y * nr + x] *= tab[x];
}

hipError_t launch_rsl(float* img, int nr, int rows, const float* tab, hipStream_t s) {
    if (rows <= 0 || nr <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rsl, dim3((nr + 255) / 256, rows), dim3(256), 0, s, img, nr, rows, tab);
    return hipGetLastError();
}

__global__ void k_srgr(SrgrArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    const int rows = a.rows_dev ? *a.rows_dev : a.rows;
    if (x >= a.out_nr || y >= rows || y >= a.rows) return;
    const double rd = polyval(a.grsr, 5, (double)x * a.gs);
    const double ri = rd / a.dr;
    int r0 = (int)ri;
    float fx = (float)(ri - (double)r0);
    if (r0 < 0) {
        r0 = 0;
        fx = 0.f;
    }
    if (r0 > a.in_nr - 1) {
        r0 = a.in_nr - 1;
        fx = 0.f;
    }
    const int r1 = min(r0 + 1, a.in_nr - 1);
    const float* row = a.in + (long)y * a.in_nr;
    const float v0 = row[r0], v1 = row[r1];
    a.out[(long)y * a.out_nr + x] = v0 + fx * (v1 - v0);
}

hipError_t launch_srgr(const SrgrArgs& a, hipStream_t s) {
    if (a.rows <= 0 || a.out_nr <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_srgr, dim3((a.out_nr + 255) / 256, a.rows), dim3(256), 0, s, a);
    return hipGetLastError();
}

__global__ void k_zero_cols(float2* img, long stride, int rows, int c0, int c1) {
    const int x = c0 + blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= c1 || y >= rows) return;
    img[(long)y * stride + x] = make_float2(0.f, 0.f);
}

hipError_t launch_zero_cols(float2* img, long stride, int rows, int c0, int c1, hipStream_t s) {
    if (c1 <= c0 || rows <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_zero_cols, dim3((c1 - c0 + 255) / 256, rows), dim3(256), 0, s, img, stride, rows, c0, c1);
    return hipGetLastError();
}

}  // namespace asarf
