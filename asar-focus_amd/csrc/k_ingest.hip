// Level-0 sample decoders (the step before the focusing path): ERS IM 8-bit I/Q samples and ASAR IM FBAQ-4 blocks
// to the cf32 focusing image, with the per-line sampling-window-start (SWST) shift.
//
// The reference decodes into a NaN-filled unpadded buffer by scattering each input sample to its column
// (envisat_format_kernels.cu:116-247), then copies that buffer into a zeroed padded image (main.cc:233-239).
// Here one gather pass writes every element of the padded image once: each lane takes one output column of one
// line, finds which input sample (if any) the shift puts there, and writes the decoded sample, NaN where no sample
// lands inside the logical region, and 0 in the padding. Samples the shift would push past the logical row are
// dropped (the reference's scatter has no bound there).
#include <cmath>

#include "kernels.h"

namespace asarf {

// Fbaq4IndexCalc (envisat_parse_util.cuh:75-84): codes 8..15 map to 7..0 (sign-magnitude negative half), codes
// 0..7 to 8..15; the LUT row is 256 * that + the block's adaptive-quantiser id.
__host__ __device__ inline int fbaq4_index(int block, int code) {
    const int conv = code > 7 ? 15 - code : 8 + code;
    return 256 * conv + block;
}

constexpr int kIngestThreads = 256;
constexpr int kIngestRows = 8;  // lines per workgroup; each lane writes 2 adjacent columns of each

// One kernel for both formats (FMT 0: ERS IM u8 pairs, envisat_format_kernels.cu:116-133; FMT 1: ASAR FBAQ-4,
// :160-210, FMT 2 the same from the block-major table copy). ERS line y holds samples_per_line (I, Q) byte pairs; sample s lands at column s + 4 (swst[y] -
// min_swst). ASAR record y = 2-byte little-endian length L, then L bytes in blocks of 64: one block-id byte and
// 63 code bytes (I code in the high nibble, Q in the low); sample s is code byte 64 (s / 63) + 1 + s % 63 (if
// that is < L) and lands at column s + swst[y] - min_swst. The lines of a workgroup are unrolled so that their
// (dependent: length, code, table) loads are in flight together.
template <int FMT>
__global__ __launch_bounds__(kIngestThreads) void k_ingest(IngestArgs a) {
    const uint8_t* __restrict__ bytes = a.bytes;
    const float* __restrict__ i_lut = a.i_lut;
    const float* __restrict__ q_lut = a.q_lut;
    float2* __restrict__ out = a.out;
    const int x = (blockIdx.x * kIngestThreads + threadIdx.x) * 2;
    if (x >= a.x_stride) return;
    const int y0 = blockIdx.y * kIngestRows;
    const bool pair = x + 1 < a.x_stride;
    // decoded value of logical element (xc, y): NaN when no sample lands there
    auto sample = [&](int xc, int y, int shift, int len) -> float2 {
        const int s = xc - shift;
        if (FMT == 0) {
            if (s < 0 || s >= a.samples_per_line) return make_float2(NAN, NAN);
            const uint8_t* p = bytes + (long)y * a.line_bytes + 2 * s;
            return make_float2((float)p[0], (float)p[1]);
        } else {
            if (s < 0) return make_float2(NAN, NAN);
            const int blk = s / 63;
            const int k = 64 * blk + 1 + (s - 63 * blk);  // byte index within the block data
            if (k >= len || 2 + k >= a.line_bytes) return make_float2(NAN, NAN);
            const uint8_t* rec = bytes + (long)y * a.line_bytes;
            const int block_id = rec[2 + 64 * blk];
            const int code = rec[2 + k];
            if (FMT == 2) {
                // block-major copy of the tables: the 16 entries of one quantiser block share a 128 B line
                const float2* __restrict__ lt = a.lut_t + 16 * block_id;
                const int ci = code >> 4, cq = code & 15;
                return make_float2(lt[ci > 7 ? 15 - ci : 8 + ci].x, lt[cq > 7 ? 15 - cq : 8 + cq].y);
            }
            const int ii = fbaq4_index(block_id, code >> 4), qi = fbaq4_index(block_id, code & 15);
            if (ii >= a.lut_items || qi >= a.lut_items) return make_float2(NAN, NAN);
            return make_float2(i_lut[ii], q_lut[qi]);
        }
    };
#pragma unroll
    for (int u = 0; u < kIngestRows; ++u) {
        const int y = y0 + u;
        if (y >= a.y_stride) break;
        float2 v0 = make_float2(0.f, 0.f), v1 = make_float2(0.f, 0.f);
        if (y < a.lines) {
            const int shift = a.swst_mult * ((int)a.swst[y] - a.min_swst);
            int len = 0;
            if (FMT != 0) {
                const uint8_t* rec = bytes + (long)y * a.line_bytes;
                len = (int)rec[0] | ((int)rec[1] << 8);
            }
            const int s0 = x - shift;
            if (FMT == 0 && a.w32 && x + 1 < a.x_size && s0 >= 0 && s0 + 1 < a.samples_per_line) {
                // both samples inside the line: one aligned 4-byte load (I0, Q0, I1, Q1). x is even and the ERS
                // shift a multiple of 4, so 2 s0 is a multiple of 4; w32 says line pitch and base are too
                const uint32_t w = *reinterpret_cast<const uint32_t*>(bytes + (long)y * a.line_bytes + 2 * s0);
                v0 = make_float2((float)(w & 0xFF), (float)((w >> 8) & 0xFF));
                v1 = make_float2((float)((w >> 16) & 0xFF), (float)(w >> 24));
            } else {
                if (x < a.x_size) v0 = sample(x, y, shift, len);
                if (x + 1 < a.x_size) v1 = sample(x + 1, y, shift, len);
            }
        }
        float2* row = out + (long)y * a.x_stride;
        if (pair && a.vec16) {
            *reinterpret_cast<float4*>(row + x) = make_float4(v0.x, v0.y, v1.x, v1.y);
        } else {
            row[x] = v0;
            if (pair) row[x + 1] = v1;
        }
    }
}

// Block-major copy of the FBAQ-4 tables: lut_t[16 b + c] = (i_lut[256 c + b], q_lut[256 c + b]) for the 256
// quantiser blocks b and the 16 converted codes c (the full 4096-entry tables only).
__global__ void k_fbaq4_lut_t(const float* i_lut, const float* q_lut, float2* lut_t) {
    const int e = blockIdx.x * blockDim.x + threadIdx.x;
    if (e >= 4096) return;
    const int b = e >> 4, c = e & 15;
    lut_t[e] = make_float2(i_lut[256 * c + b], q_lut[256 * c + b]);
}

hipError_t launch_ingest(const IngestArgs& args, int kind, hipStream_t s) {
    IngestArgs a = args;
    // 16-byte stores of column pairs when every pair is 16-byte aligned
    a.vec16 = (a.x_stride % 2 == 0) && (reinterpret_cast<uintptr_t>(a.out) & 15) == 0;
    // 4-byte ERS sample-pair loads when every lane's pair is 4-byte aligned: the shift step is even (4 for ERS),
    // the line pitch and the base are multiples of 4
    a.w32 = (a.swst_mult % 2 == 0) && (a.line_bytes % 4 == 0) && (reinterpret_cast<uintptr_t>(a.bytes) & 3) == 0;
    const dim3 grid((a.x_stride + 2 * kIngestThreads - 1) / (2 * kIngestThreads),
                    (a.y_stride + kIngestRows - 1) / kIngestRows);
    if (kind == 0) {
        hipLaunchKernelGGL(k_ingest<0>, grid, dim3(kIngestThreads), 0, s, a);
    } else if (a.lut_t && a.lut_items >= 4096) {
        hipLaunchKernelGGL(k_fbaq4_lut_t, dim3(16), dim3(256), 0, s, a.i_lut, a.q_lut, const_cast<float2*>(a.lut_t));
        hipLaunchKernelGGL(k_ingest<2>, grid, dim3(kIngestThreads), 0, s, a);
    } else {
        hipLaunchKernelGGL(k_ingest<1>, grid, dim3(kIngestThreads), 0, s, a);
    }
    return hipGetLastError();
}

int fbaq4_index_host(int block, int code) { return fbaq4_index(block, code); }

}  // namespace asarf
