// IMP tail: azimuth multilook (sar/az_multilook.cu:25-80) with the range-spreading-loss factor fused into its
// store (sar/range_spreading_loss.cu:29-40), the standalone RSL pass, and slant-to-ground resampling
// (sar/srgr.cu:49-71, polynomial evaluated in FP64).
#include "kernels.h"
#include "poly.h"

namespace asarf {

// Each lane owns one range column of a tile of `rows` output rows. The |s|^2 of every input row the tile needs
// (at most kMlSpan) is loaded first, all loads independent and in flight together, into the lane's own LDS
// column (no barrier: a lane only reads what it wrote); the boxcars and the interpolation then run from LDS.
// Neighbouring tiles share the rows of one output's boxcars (about 6 rows): with a 40-row span a tile produces 12
// outputs of the ASAR scene (ratio 2.9) from 38 rows, about 10 % more reads than the outputs consume. The span
// also sets the LDS of the fused multilook + SRGR kernel ((span + rows) x 1 KB): 40 rows leave room for three of
// its workgroups per CU, 48 rows for two (0.376 -> 0.332 ms on the ASAR scene; 32 rows: 0.334 ms). (A ring that
// walks each column once, 24 rows a batch, measured slower: each batch waits for its own loads.)
constexpr int kMlSpan = 40;     // input rows per tile
constexpr int kMlLanes = 256;

// Output rows per tile: the span (rows - 1) * ratio + 6 must fit kMlSpan (with one row of margin for rounding).
static int ml_rows_per_tile(float ratio) {
    const int rows = (int)((float)(kMlSpan - 7) / ratio) + 1;
    return max(1, min(rows, 24));
}

__global__ __launch_bounds__(kMlLanes) void k_multilook(MultilookArgs a, int rows) {
    __shared__ float pw[kMlSpan][kMlLanes];
    const int x = blockIdx.x * kMlLanes + threadIdx.x;
    const int out_rows = min(a.out_rows_dev ? *a.out_rows_dev : a.out_rows, a.out_rows);
    const int y0 = blockIdx.y * rows;
    const int y1 = min(y0 + rows, out_rows);
    if (y0 >= y1) return;
    const int row0 = a.row0_dev ? *a.row0_dev : a.row0;
    const int in_rows = a.in_rows_dev ? *a.in_rows_dev : a.in_rows;
    // input rows of the tile: resampling positions in float, exactly as the reference computes them
    const float first = (float)y0 * a.ratio, last = (float)(y1 - 1) * a.ratio;
    const int lo = max(min((int)(first - 2.0f), (int)(first - 1.0f)), 0);
    const int hi = min(max((int)(last + 2.0f), (int)(last + 3.0f)), in_rows);
    const int span = min(hi - lo, kMlSpan);  // the host sizes tiles so that hi - lo <= kMlSpan
    const bool xok = x < a.nr;
    const float2* col = a.slc + (long)(row0 + lo) * a.stride + a.col0 + (xok ? x : 0);
    // rows past the span load a clamped copy that no output reads (unconditional loads stay in flight together)
    float2 v[kMlSpan];
#pragma unroll
    for (int i = 0; i < kMlSpan; ++i) v[i] = col[(long)min(i, max(span - 1, 0)) * a.stride];
#pragma unroll
    for (int i = 0; i < kMlSpan; ++i) pw[i][threadIdx.x] = v[i].x * v[i].x + v[i].y * v[i].y;
    if (!xok) return;
    for (int y = y0; y < y1; ++y) {
        const float in_idx = (float)y * a.ratio;
        const float mu = in_idx - (float)(int)in_idx;
        int s0 = (int)(in_idx - 2.0f), e0 = (int)(in_idx + 2.0f);
        int s1 = (int)(in_idx - 1.0f), e1 = (int)(in_idx + 3.0f);
        s0 = max(s0, 0);
        e0 = min(e0, in_rows);
        s1 = max(s1, 0);
        e1 = min(e1, in_rows);
        float sum0 = 0.f, sum1 = 0.f;
        for (int i = s0; i < e0; ++i) sum0 += pw[i - lo][threadIdx.x];
        for (int i = s1; i < e1; ++i) sum1 += pw[i - lo][threadIdx.x];
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
    if (!(a.ratio > 0.f) || a.ratio > (float)(kMlSpan - 8)) return hipErrorInvalidValue;
    const int rows = ml_rows_per_tile(a.ratio);
    hipLaunchKernelGGL(k_multilook, dim3((a.nr + kMlLanes - 1) / kMlLanes, (a.out_rows + rows - 1) / rows),
                       dim3(kMlLanes), 0, s, a, rows);
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

// One lane per output column and a tile of kSrgrRows rows: the FP64 polynomial and the source index are
// computed once per column and reused for every row of the tile. 1024 columns per workgroup: the input span of
// a workgroup row (about 1.6 input columns per output column) starts and ends in partly used cache lines, which
// cost 16 % extra reads with 256-column workgroups and 4 % with 1024.
constexpr int kSrgrRows = 16;
constexpr int kSrgrLanes = 1024;

__global__ __launch_bounds__(kSrgrLanes) void k_srgr(SrgrArgs a) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int rows = min(a.rows_dev ? *a.rows_dev : a.rows, a.rows);
    const int y0 = blockIdx.y * kSrgrRows;
    if (x >= a.out_nr || y0 >= rows) return;
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
    const int y1 = min(y0 + kSrgrRows, rows);
#pragma unroll 4
    for (int y = y0; y < y1; ++y) {
        const float* row = a.in + (long)y * a.in_nr;
        const float v0 = row[r0], v1 = row[r1];
        a.out[(long)y * a.out_nr + x] = v0 + fx * (v1 - v0);
    }
}

hipError_t launch_srgr(const SrgrArgs& a, hipStream_t s) {
    if (a.rows <= 0 || a.out_nr <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_srgr, dim3((a.out_nr + kSrgrLanes - 1) / kSrgrLanes, (a.rows + kSrgrRows - 1) / kSrgrRows),
                       dim3(kSrgrLanes), 0, s, a);
    return hipGetLastError();
}

// Slant-to-ground index of each ground column: k_srgr's source column r0 and fraction fx (same FP64 code).
__global__ void k_srgr_index(SrgrArgs a, int* r0_out, float* fx_out) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    if (x >= a.out_nr) return;
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
    r0_out[x] = r0;
    fx_out[x] = fx;
}

hipError_t launch_srgr_index(const SrgrArgs& a, int* r0, float* fx, hipStream_t s) {
    if (a.out_nr <= 0 || a.in_nr <= 0) return hipErrorInvalidValue;
    hipLaunchKernelGGL(k_srgr_index, dim3((a.out_nr + 255) / 256), dim3(256), 0, s, a, r0, fx);
    return hipGetLastError();
}

// Multilook x RSL fused with SRGR. The multilook half is k_multilook's (same loads, boxcars and interpolation, in
// the same order), with the values of the tile's output rows kept in LDS instead of written out; the SRGR half
// is k_srgr's interpolation between the two source columns, read from LDS. The workgroups' slant columns overlap
// by one (each covers 256 columns but owns the ground columns whose r0 is among its first 255), so both source
// columns of every owned ground column are in its tile. The detected slant-range image never goes to HBM: that
// saves its write and its read back (4 bytes per multilooked pixel each way). Requires r0 non-decreasing in the
// ground column (a monotone slant-to-ground polynomial), which the plan checks before it uses this kernel.
constexpr int kMlsOwn = kMlLanes - 1;  // ground-column ownership stride in slant columns

__global__ __launch_bounds__(kMlLanes) void k_ml_srgr(MlSrgrArgs a, int rows) {
    extern __shared__ float sm[];
    float (*pw)[kMlLanes] = reinterpret_cast<float (*)[kMlLanes]>(sm);                      // [kMlSpan]
    float (*mv)[kMlLanes] = reinterpret_cast<float (*)[kMlLanes]>(sm + kMlSpan * kMlLanes);  // [rows]
    const MultilookArgs& m = a.ml;
    const int x0 = blockIdx.x * kMlsOwn;
    const int x = x0 + threadIdx.x;
    const int out_rows = min(m.out_rows_dev ? *m.out_rows_dev : m.out_rows, m.out_rows);
    const int y0 = blockIdx.y * rows;
    const int y1 = min(y0 + rows, out_rows);
    if (y0 >= y1) return;  // uniform over the workgroup
    const int row0 = m.row0_dev ? *m.row0_dev : m.row0;
    const int in_rows = m.in_rows_dev ? *m.in_rows_dev : m.in_rows;
    const float first = (float)y0 * m.ratio, last = (float)(y1 - 1) * m.ratio;
    const int lo = max(min((int)(first - 2.0f), (int)(first - 1.0f)), 0);
    const int hi = min(max((int)(last + 2.0f), (int)(last + 3.0f)), in_rows);
    const int span = min(hi - lo, kMlSpan);
    const bool xok = x < m.nr;
    const float2* col = m.slc + (long)(row0 + lo) * m.stride + m.col0 + (xok ? x : 0);
    float2 v[kMlSpan];
#pragma unroll
    for (int i = 0; i < kMlSpan; ++i) v[i] = col[(long)min(i, max(span - 1, 0)) * m.stride];
#pragma unroll
    for (int i = 0; i < kMlSpan; ++i) pw[i][threadIdx.x] = v[i].x * v[i].x + v[i].y * v[i].y;
    if (xok) {
        const float rsl = m.rsl ? m.rsl[x] : 1.f;
        for (int y = y0; y < y1; ++y) {
            const float in_idx = (float)y * m.ratio;
            const float mu = in_idx - (float)(int)in_idx;
            int s0 = (int)(in_idx - 2.0f), e0 = (int)(in_idx + 2.0f);
            int s1 = (int)(in_idx - 1.0f), e1 = (int)(in_idx + 3.0f);
            s0 = max(s0, 0);
            e0 = min(e0, in_rows);
            s1 = max(s1, 0);
            e1 = min(e1, in_rows);
            float sum0 = 0.f, sum1 = 0.f;
            for (int i = s0; i < e0; ++i) sum0 += pw[i - lo][threadIdx.x];
            for (int i = s1; i < e1; ++i) sum1 += pw[i - lo][threadIdx.x];
            const int c0 = max(e0 - s0, 0), c1 = max(e1 - s1, 0);
            if (c0) sum0 /= (float)c0;
            if (c1) sum1 /= (float)c1;
            float val = (c0 && c1) ? sum0 + (sum1 - sum0) * mu : 0.f;
            if (m.rsl) val *= rsl;
            mv[y - y0][threadIdx.x] = val;
        }
    }
    __syncthreads();
    const int g1 = a.g_start[blockIdx.x + 1];
    for (int g = a.g_start[blockIdx.x] + (int)threadIdx.x; g < g1; g += kMlLanes) {
        const int r0 = a.r0[g];
        const int r1 = min(r0 + 1, m.nr - 1);
        const float fx = a.fx[g];
        const int c0 = r0 - x0, c1 = r1 - x0;  // 0 <= c0 <= c1 <= 255 for an owned column
        for (int y = y0; y < y1; ++y) {
            const float q0 = mv[y - y0][c0], q1 = mv[y - y0][c1];
            a.out[(long)y * a.out_nr + g] = q0 + fx * (q1 - q0);
        }
    }
}

int ml_srgr_tiles(int nr) { return (nr + kMlsOwn - 1) / kMlsOwn; }

int ml_srgr_owned_columns() { return kMlsOwn; }

hipError_t ml_srgr_prepare(float ratio) {
    if (!(ratio > 0.f) || ratio > (float)(kMlSpan - 8)) return hipErrorInvalidValue;
    const size_t lds = sizeof(float) * (size_t)(kMlSpan + ml_rows_per_tile(ratio)) * kMlLanes;
    return hipFuncSetAttribute((const void*)k_ml_srgr, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
}

hipError_t launch_ml_srgr(const MlSrgrArgs& a, hipStream_t s) {
    const MultilookArgs& m = a.ml;
    if (m.out_rows <= 0 || m.nr <= 0 || a.out_nr <= 0) return hipSuccess;
    if (!(m.ratio > 0.f) || m.ratio > (float)(kMlSpan - 8) || !a.r0 || !a.fx || !a.g_start) return hipErrorInvalidValue;
    const int rows = ml_rows_per_tile(m.ratio);
    const size_t lds = sizeof(float) * (size_t)(kMlSpan + rows) * kMlLanes;
    hipLaunchKernelGGL(k_ml_srgr, dim3(ml_srgr_tiles(m.nr), (m.out_rows + rows - 1) / rows), dim3(kMlLanes), lds, s,
                       a, rows);
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
