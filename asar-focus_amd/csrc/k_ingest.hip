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
// that is < L) and lands at column s + swst[y] - min_swst.
template <int FMT>
__global__ __launch_bounds__(kIngestThreads) void k_ingest(IngestArgs a) {
    const uint8_t* __restrict__ bytes = a.bytes;
    const float* __restrict__ i_lut = a.i_lut;
    const float* __restrict__ q_lut = a.q_lut;
    float2* __restrict__ out = a.out;
    constexpr int R = kIngestRows;
    const int x = (blockIdx.x * kIngestThreads + threadIdx.x) * 2;
    if (x >= a.x_stride) return;
    const int y0 = blockIdx.y * R;
    const bool pair = x + 1 < a.x_stride;
    const float2 nan2 = make_float2(NAN, NAN);
    // The R lines are processed in phases, each issuing the loads of every line before any result is used, so
    // that the dependent loads (SWST code and record length, then the sample or code bytes, then the table
    // entries) of all R lines are in flight together.
    // phase 1: each line's shift and (ASAR) record length
    int shift[R], len[R];
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int y = y0 + u;
        shift[u] = 0;
        len[u] = 0;
        if (y < a.lines) {
            shift[u] = a.swst_mult * ((int)a.swst[y] - a.min_swst);
            if (FMT != 0) {
                const uint8_t* rec = bytes + (long)y * a.line_bytes;
                len[u] = (int)rec[0] | ((int)rec[1] << 8);
            }
        }
    }
    // v[u][j]: element (x + j, y0 + u); 0 in the padding, NaN where no sample lands in the logical region
    float2 v[R][2];
    if constexpr (FMT == 0) {
        // phase 2: the samples (one aligned 4-byte load per pair where the pitch allows)
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int y = y0 + u;
            v[u][0] = v[u][1] = make_float2(0.f, 0.f);
            if (y >= a.lines) continue;
            const uint8_t* line = bytes + (long)y * a.line_bytes;
            const int s0 = x - shift[u];
            if (a.w32 && x + 1 < a.x_size && s0 >= 0 && s0 + 1 < a.samples_per_line) {
                // x is even and the ERS shift a multiple of 4, so 2 s0 is a multiple of 4; w32 says line pitch
                // and base are too
                const uint32_t w = *reinterpret_cast<const uint32_t*>(line + 2 * s0);
                v[u][0] = make_float2((float)(w & 0xFF), (float)((w >> 8) & 0xFF));
                v[u][1] = make_float2((float)((w >> 16) & 0xFF), (float)(w >> 24));
            } else {
#pragma unroll
                for (int j = 0; j < 2; ++j) {
                    const int s = s0 + j;
                    if (x + j >= a.x_size) continue;
                    v[u][j] = (s < 0 || s >= a.samples_per_line)
                                  ? nan2
                                  : make_float2((float)line[2 * s], (float)line[2 * s + 1]);
                }
            }
        }
    } else {
        // phase 2: block-id and code bytes of both samples of every line; sample s is code byte
        // 64 (s / 63) + 1 + s % 63 of the record's data, valid below the record length
        int bid[R][2], code[R][2];
        bool ok[R][2];
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int y = y0 + u;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                const int s = x + j - shift[u];
                const int blk = s / 63;
                const int k = 64 * blk + 1 + (s - 63 * blk);
                ok[u][j] = y < a.lines && x + j < a.x_size && s >= 0 && k < len[u] && 2 + k < a.line_bytes;
                bid[u][j] = code[u][j] = 0;
                if (ok[u][j]) {
                    const uint8_t* rec = bytes + (long)y * a.line_bytes;
                    bid[u][j] = rec[2 + 64 * blk];
                    code[u][j] = rec[2 + k];
                }
            }
        }
        // phase 3: table entries
#pragma unroll
        for (int u = 0; u < R; ++u) {
            const int y = y0 + u;
#pragma unroll
            for (int j = 0; j < 2; ++j) {
                if (y >= a.lines || x + j >= a.x_size) {
                    v[u][j] = make_float2(0.f, 0.f);
                } else if (!ok[u][j]) {
                    v[u][j] = nan2;
                } else if (FMT == 2) {
                    // block-major copy of the tables: the 16 entries of one quantiser block share a 128 B line
                    const float2* __restrict__ lt = a.lut_t + 16 * bid[u][j];
                    const int ci = code[u][j] >> 4, cq = code[u][j] & 15;
                    v[u][j] = make_float2(lt[ci > 7 ? 15 - ci : 8 + ci].x, lt[cq > 7 ? 15 - cq : 8 + cq].y);
                } else {
                    const int ii = fbaq4_index(bid[u][j], code[u][j] >> 4);
                    const int qi = fbaq4_index(bid[u][j], code[u][j] & 15);
                    v[u][j] = (ii >= a.lut_items || qi >= a.lut_items) ? nan2 : make_float2(i_lut[ii], q_lut[qi]);
                }
            }
        }
    }
    // phase 4: stores
#pragma unroll
    for (int u = 0; u < R; ++u) {
        const int y = y0 + u;
        if (y >= a.y_stride) continue;
        float2* row = out + (long)y * a.x_stride;
        if (pair && a.vec16) {
            *reinterpret_cast<float4*>(row + x) = make_float4(v[u][0].x, v[u][0].y, v[u][1].x, v[u][1].y);
        } else {
            row[x] = v[u][0];
            if (pair) row[x + 1] = v[u][1];
        }
    }
}

// Block-major copy of the FBAQ-4 tables: lut_t[16 b + c] = (i_lut[256 c + b], q_lut[256 c + b]) for the 256
// quantiser blocks b and the 16 converted codes c (the full 4096-entry tables only). (A product table keyed by
// block and code byte, one 8-byte read per sample, measured slower: a wave's reads then spread over 16 cache
// lines per block instead of one.)
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

hipError_t launch_fbaq4_lut_t(const float* i_lut, const float* q_lut, float2* lut_t, hipStream_t s) {
    hipLaunchKernelGGL(k_fbaq4_lut_t, dim3(16), dim3(256), 0, s, i_lut, q_lut, lut_t);
    return hipGetLastError();
}

}  // namespace asarf
