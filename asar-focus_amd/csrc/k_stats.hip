// I/Q statistics, I/Q correction and Doppler-centroid (ACCC) kernels.
//
// Two families:
//  * cf32 image kernels (the per-stage entry points): NaN-aware FP64 moments of the logical region and of the
//    padding columns, the affine correction applied over the padded width, and direct lag-1 products per range
//    block for the Doppler estimate. They follow sar/iq_correction.cu:140-273, cuda_util/device_padded_image.cu
//    :156-257,322-414 and sar/fractional_doppler_centroid.cu:33-96.
//  * raw-byte kernels (the fused plan): one pass over the uint8/int8 samples collects, per Doppler block, exact
//    integer sums of I, Q, I^2, Q^2, IQ and of the lag-1 cross products between consecutive lines. The I/Q
//    correction is affine, so the corrected lag-1 sum of every block is a bilinear form in these sums; one FP64
//    finalisation wave turns them into the DC bias, gain, phase correction, block Doppler estimates and the
//    degree-2 Doppler polynomial without a second read of the raw data.
#include "kernels.h"
#include "poly.h"

namespace asarf {

// ---------------------------------------------------------------------------------- cf32 moments ----
// partials[wg][12]: logical {n, SI, SQ, SII, SQQ, SIQ}, padding {n, SI, SQ, SII, SQQ, SIQ}. A sample with a NaN in
// either component is skipped whole. The logical columns [0, xs) and the padding columns [xs, xstr) run as two
// loops with their own accumulators (a runtime index into one accumulator array would put it in scratch memory),
// each with four loads in flight per lane.
struct Mom6 {
    double n = 0.0, si = 0.0, sq = 0.0, sii = 0.0, sqq = 0.0, siq = 0.0;
    __device__ __forceinline__ void add(float2 v) {
        const bool ok = !(isnan(v.x) || isnan(v.y));
        const double i_ = ok ? (double)v.x : 0.0, q_ = ok ? (double)v.y : 0.0;
        n += ok ? 1.0 : 0.0;
        si += i_;
        sq += q_;
        sii = fma(i_, i_, sii);
        sqq = fma(q_, q_, sqq);
        siq = fma(i_, q_, siq);
    }
    __device__ __forceinline__ void span(const float2* row, int x0, int x1) {
        int x = x0 + (int)threadIdx.x;
        for (; x + 3 * kThreads < x1; x += 4 * kThreads) {
            const float2 v0 = row[x], v1 = row[x + kThreads], v2 = row[x + 2 * kThreads], v3 = row[x + 3 * kThreads];
            add(v0);
            add(v1);
            add(v2);
            add(v3);
        }
        for (; x < x1; x += kThreads) add(row[x]);
    }
};

__global__ __launch_bounds__(kThreads) void k_moments_cf32(const float2* img, long stride, int xs, int ys,
                                                          int xstr, double* partials) {
    Mom6 lg, pd;
    for (int y = blockIdx.x; y < ys; y += gridDim.x) {
        const float2* row = img + (long)y * stride;
        lg.span(row, 0, xs);
        pd.span(row, xs, xstr);
    }
    __shared__ double sh[12][kThreads];
    const double acc[12] = {lg.n, lg.si, lg.sq, lg.sii, lg.sqq, lg.siq, pd.n, pd.si, pd.sq, pd.sii, pd.sqq, pd.siq};
#pragma unroll
    for (int i = 0; i < 12; ++i) sh[i][threadIdx.x] = acc[i];
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
#pragma unroll
            for (int i = 0; i < 12; ++i) sh[i][threadIdx.x] += sh[i][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x < 12) partials[(long)blockIdx.x * 12 + threadIdx.x] = sh[threadIdx.x][0];
}

hipError_t launch_moments_cf32(const float2* img, long stride, int xs, int ys, int xstr, double* partials, int nwg,
                               hipStream_t s) {
    hipLaunchKernelGGL(k_moments_cf32, dim3(nwg), dim3(kThreads), 0, s, img, stride, xs, ys, xstr, partials);
    return hipGetLastError();
}

// Apply the correction over x < width for rows < ys; NaN -> 0 component-wise over the whole padded image.
__global__ void k_iq_apply_cf32(float2* img, long stride, int xs, int ys, int xstr, int ystr, int width, IqCoef c) {
    const int x = blockIdx.x * blockDim.x + threadIdx.x;
    const int y = blockIdx.y;
    if (x >= xstr || y >= ystr) return;
    float2 v = img[(long)y * stride + x];
    if (y < ys && x < width) {
        const float ic = (v.x - c.dc_i) * c.inv_gain;
        const float qc = (v.y - c.dc_q - ic * c.sin_corr) * c.inv_cos;
        v = make_float2(ic, qc);
    }
    if (isnan(v.x)) v.x = 0.f;
    if (isnan(v.y)) v.y = 0.f;
    img[(long)y * stride + x] = v;
}

hipError_t launch_iq_apply_cf32(float2* img, long stride, int xs, int ys, int xstr, int ystr, int width, IqCoef c,
                                hipStream_t s) {
    hipLaunchKernelGGL(k_iq_apply_cf32, dim3((xstr + 255) / 256, ystr), dim3(256), 0, s, img, stride, xs, ys, xstr,
                       ystr, width, c);
    return hipGetLastError();
}

// Lag-1 products per Doppler block: grid (nblocks, nstrips); partials[(strip*nblocks + b)*2 + {re, im}].
__global__ __launch_bounds__(kThreads) void k_dc_cf32(const float2* img, long stride, int ys, int step, int nblocks,
                                                     int xstr, int nstrips, double* partials) {
    const int b = blockIdx.x;
    const int strip = blockIdx.y;
    const int rows_per = (ys + nstrips - 1) / nstrips;
    const int y0 = max(1, strip * rows_per), y1 = min(ys, (strip + 1) * rows_per);
    const int xb = b * step;
    const int xe = min(xb + step, xstr);
    double re = 0.0, im = 0.0;
    auto lag = [&](float2 p, float2 c) {
        re += (double)p.x * c.x + (double)p.y * c.y;
        im += (double)p.x * c.y - (double)p.y * c.x;
    };
    for (int x = xb + threadIdx.x; x < xe; x += kThreads) {
        if (y0 >= y1) break;
        const float2* col = img + x;
        float2 prev = col[(long)(y0 - 1) * stride];
        int y = y0;
        // four rows loaded before they are used (the sums keep their row order)
        for (; y + 3 < y1; y += 4) {
            const float2 c0 = col[(long)y * stride], c1 = col[(long)(y + 1) * stride];
            const float2 c2 = col[(long)(y + 2) * stride], c3 = col[(long)(y + 3) * stride];
            lag(prev, c0);
            lag(c0, c1);
            lag(c1, c2);
            lag(c2, c3);
            prev = c3;
        }
        for (; y < y1; ++y) {
            const float2 cur = col[(long)y * stride];
            lag(prev, cur);
            prev = cur;
        }
    }
    __shared__ double sh[2][kThreads];
    sh[0][threadIdx.x] = re;
    sh[1][threadIdx.x] = im;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) {
            sh[0][threadIdx.x] += sh[0][threadIdx.x + w];
            sh[1][threadIdx.x] += sh[1][threadIdx.x + w];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        partials[((long)strip * nblocks + b) * 2 + 0] = sh[0][0];
        partials[((long)strip * nblocks + b) * 2 + 1] = sh[1][0];
    }
}

hipError_t launch_dc_cf32(const float2* img, long stride, int ys, int x_begin0, int step, int nblocks, int xstr,
                          int nstrips, double* partials, hipStream_t s) {
    hipLaunchKernelGGL(k_dc_cf32, dim3(nblocks, nstrips), dim3(kThreads), 0, s, img, stride, ys, step, nblocks, xstr,
                       nstrips, partials);
    return hipGetLastError();
}

// totals[v] = sum_p partials[p * nvals + v] (fixed order, deterministic).
__global__ __launch_bounds__(kThreads) void k_reduce_f64(const double* partials, int nparts, int nvals,
                                                        double* totals) {
    const int v = blockIdx.x;
    double acc = 0.0;
    for (int p = threadIdx.x; p < nparts; p += kThreads) acc += partials[(long)p * nvals + v];
    __shared__ double sh[kThreads];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) totals[v] = sh[0];
}

hipError_t launch_reduce_f64(const double* partials, int nparts, int nvals, double* totals, hipStream_t s) {
    hipLaunchKernelGGL(k_reduce_f64, dim3(nvals), dim3(kThreads), 0, s, partials, nparts, nvals, totals);
    return hipGetLastError();
}

// ---------------------------------------------------------------------------------- raw moments ----
template <typename T>
__device__ __forceinline__ T wave_sum(T v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) v += __shfl_xor(v, m, 64);
    return v;
}

// Workgroup (tile, strip): lane t owns columns [V*(256*tile + t), +V) over rows [128*strip, +128) and
// accumulates exact integer sums; lag-1 products pair each row with the one above (the strip re-reads its
// preceding row). With V = 4 a row of a lane is one 8-byte load (4 I/Q byte pairs). Statistics are reduced over
// the workgroup; lag sums per Doppler block (slot = block - first block of the tile) through per-column LDS rows.
// Integer ranges: a column's sums over 128 rows stay below 128 * 255^2 < 2^31.
// LINES: the rows are ERS level-0 lines (RawMomentsArgs::lines): row y of range column x holds line sample
// x - shift(y) when that is one of the line's samples, else NaN, which the statistics skip and the lag-1 sums treat
// as the zero ZeroNaNs makes it (a pair counts only when both of its samples are valid). V = 4 needs the shifts
// and the samples per line to be multiples of 4, so a lane's four samples are valid or NaN together.
template <typename P, int V, bool LINES = false>
__global__ __launch_bounds__(kThreads) void k_raw_moments(RawMomentsArgs a) {
    constexpr bool SIGNED = std::is_same<P, char2>::value;
    constexpr int TW = kThreads * V;
    const int tile = blockIdx.x;
    const int strip = blockIdx.y;
    const int x0 = (tile * kThreads + threadIdx.x) * V;
    const int y0 = strip * kStripRows;
    const int y1 = min(y0 + kStripRows, a.na);
    const unsigned char* raw = static_cast<const unsigned char*>(a.raw);
    const int nv = min(V, max(0, a.nr - x0));  // valid columns of this lane: V or 0 (nr % V == 0)
    int st[5] = {0, 0, 0, 0, 0};
    int lg[V][8];
#pragma unroll
    for (int j = 0; j < V; ++j)
#pragma unroll
        for (int i = 0; i < 8; ++i) lg[j][i] = 0;
    int rows = 0, lag_rows = 0;
    // one row of the lane's V samples as raw bytes (V = 4: one 8-byte word; V = 1: one I/Q pair), unpacked later
    using W = typename std::conditional<V == 4, uint2, unsigned short>::type;
    // LINES: whether row y has samples at the lane's columns (all V of them or none) and where they are
    auto line_sample = [&](int y) { return x0 - a.lines.shift(y); };
    auto row_valid = [&](int y) {
        if constexpr (LINES) {
            const int s0 = line_sample(y);
            return s0 >= 0 && s0 + V <= a.lines.spl;
        } else {
            return true;
        }
    };
    auto load_row = [&](int y) -> W {
        if constexpr (LINES) {
            if (!row_valid(y)) return W{};
            return *reinterpret_cast<const W*>(raw + (long)y * a.lines.line_bytes + (long)line_sample(y) * 2);
        } else {
            return *reinterpret_cast<const W*>(raw + ((long)y * a.nr + x0) * 2);
        }
    };
    auto sample = [&](const W& w, int j, int& si, int& sq) {
        unsigned int word;
        if constexpr (V == 4)
            word = (j < 2 ? w.x : w.y) >> ((j & 1) * 16);
        else
            word = w;
        const unsigned int bi = word & 0xff, bq = (word >> 8) & 0xff;
        si = SIGNED ? (int)(signed char)bi : (int)bi;
        sq = SIGNED ? (int)(signed char)bq : (int)bq;
    };
    if (nv == V) {
        constexpr int U = 8;  // rows loaded per iteration (independent loads in flight)
        int pi[V], pq[V];
        bool have_prev = false;
        if (y0 > 0) {
            const W w = load_row(y0 - 1);
#pragma unroll
            for (int j = 0; j < V; ++j) sample(w, j, pi[j], pq[j]);
            have_prev = row_valid(y0 - 1);
        }
        for (int y = y0; y < y1; y += U) {
            W w[U];
#pragma unroll
            for (int u = 0; u < U; ++u) w[u] = (y + u < y1) ? load_row(y + u) : W{};
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (y + u >= y1) break;
                const bool ok = row_valid(y + u);  // always true without LINES
#pragma unroll
                for (int j = 0; j < V; ++j) {
                    int c_i, c_q;
                    sample(w[u], j, c_i, c_q);
                    if (ok) {
                        st[0] += c_i;
                        st[1] += c_q;
                        st[2] += c_i * c_i;
                        st[3] += c_q * c_q;
                        st[4] += c_i * c_q;
                    }
                    if (have_prev && ok) {
                        lg[j][0] += pi[j] * c_i;
                        lg[j][1] += pq[j] * c_q;
                        lg[j][2] += pi[j] * c_q;
                        lg[j][3] += pq[j] * c_i;
                        lg[j][4] += pi[j];
                        lg[j][5] += pq[j];
                        lg[j][6] += c_i;
                        lg[j][7] += c_q;
                    }
                    pi[j] = c_i;
                    pq[j] = c_q;
                }
                if (ok) ++rows;
                if (have_prev && ok) ++lag_rows;
                have_prev = ok;
            }
        }
    }
    const int nval = kRawStats + a.nslots * kRawLag;
    long long* out = a.partials + ((long)strip * a.ntiles + tile) * nval;
    // statistics: wave sums, then across the 4 waves
    __shared__ long long shs[kThreads / 64][kRawStats];
    __shared__ int shl[kRawLag][TW];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < kRawStats; ++i) {
        const long long v = wave_sum<long long>(i < 5 ? (long long)st[i] : (long long)rows * nv);
        if (lane == 0) shs[wave][i] = v;
    }
    // column-major per lane: column c of the tile is lane c / V, slot c % V
#pragma unroll
    for (int j = 0; j < V; ++j) {
        const int c = threadIdx.x * V + j;
#pragma unroll
        for (int i = 0; i < 8; ++i) shl[i][c] = lg[j][i];
        shl[8][c] = nv ? lag_rows : 0;
    }
    __syncthreads();
    if (threadIdx.x < kRawStats) {
        long long t = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) t += shs[w][threadIdx.x];
        out[threadIdx.x] = t;
    }
    // lag sums per block slot: columns of one block form a contiguous range
    const int c0 = tile * TW;
    const int b0 = c0 / a.step;
    for (int k = threadIdx.x; k < a.nslots * kRawLag; k += kThreads) {
        const int slot = k / kRawLag, v = k % kRawLag;
        const int b = b0 + slot;
        long long t = 0;
        if (b < a.nblocks) {
            const int l0 = max(0, b * a.step - c0);
            const int l1 = min(TW, (b + 1) * a.step - c0);
            for (int l = l0; l < l1; ++l) t += shl[v][l];
        }
        out[kRawStats + k] = t;
    }
}

// Eight columns per lane (one 16-byte load per row: four 32-bit words of two I/Q byte pairs), with the sums of
// squares, cross products and lag-1 products of both pairs of a word in one packed 8-bit dot product each:
// with mi = w & 0x00FF00FF (the two I bytes) and mq = (w >> 8) & 0x00FF00FF (the two Q bytes),
// dot4(mi, mi) = I0^2 + I1^2, dot4(mi, mq) = I0 Q0 + I1 Q1, dot4(mi_prev, mq) = I0' Q0 + I1' Q1, ...
// Lag sums are kept per word and filed under the Doppler block of the word's first column; where a block starts
// at an odd column the word straddles two blocks, and the lane holding that column also sums its lag products
// alone (corr[strip][block] of the arguments, plain integer products in a branch only such lanes take), which the reduction
// moves to its own block. The lag-row sums of I and Q follow from the per-word column sums and the rows at the
// strip's edges, so the loop carries nine dot products per word and row.
template <bool SIGNED>
__device__ __forceinline__ int dot4(unsigned a, unsigned b, int c) {
    if constexpr (SIGNED)
        return __builtin_amdgcn_sdot4((int)a, (int)b, c, false);
    else
        return (int)__builtin_amdgcn_udot4(a, b, (unsigned)c, false);
}

template <bool SIGNED>
__global__ __launch_bounds__(kThreads) void k_raw_moments8(RawMomentsArgs a) {
    constexpr int V = 8, NW = V / 2, TW = kThreads * V;
    constexpr unsigned MI = 0x00FF00FFu, ONES = 0x01010101u;
    const int tile = blockIdx.x;
    const int strip = blockIdx.y;
    const int x0 = (tile * kThreads + threadIdx.x) * V;
    const int y0 = strip * kStripRows;
    const int y1 = min(y0 + kStripRows, a.na);
    const bool valid = x0 < a.nr;  // nr % 8 == 0: all eight columns of the lane, or none
    const unsigned char* raw = static_cast<const unsigned char*>(a.raw);
    auto load_row = [&](int y) -> uint4 { return *reinterpret_cast<const uint4*>(raw + ((long)y * a.nr + x0) * 2); };
    int sii = 0, sqq = 0, siq = 0;
    int wi[NW], wq[NW], lii[NW], lqq[NW], liq[NW], lqi[NW];
    unsigned pmi[NW], pmq[NW];
#pragma unroll
    for (int k = 0; k < NW; ++k) wi[k] = wq[k] = lii[k] = lqq[k] = liq[k] = lqi[k] = 0, pmi[k] = pmq[k] = 0;
    int edge_i[NW], edge_q[NW];  // row y0 - 1 (strip > 0) or row 0 (strip 0), per word
    bool have_prev = false;
    // the lane's block start at an odd column c = sb * step (step >= 8: at most one in the lane's eight columns),
    // the second column of word sk; its lag sums over the strip's lag rows
    int sb = -1, sk = 0;
    if (valid) {
        const int bb = (x0 + a.step - 1) / a.step;
        const int cs = bb * a.step;
        if (bb >= 1 && bb < a.nblocks && cs < x0 + V && (cs & 1)) {
            sb = bb;
            sk = (cs - x0) >> 1;
        }
    }
    int s_ii = 0, s_qq = 0, s_iq = 0, s_qi = 0, s_pi = 0, s_pq = 0, s_ci = 0, s_cq = 0, s_n = 0;
    int s_prev_i = 0, s_prev_q = 0;
    auto byte_at = [](unsigned word, int shift) -> int {
        return SIGNED ? (int)(signed char)(word >> shift) : (int)((word >> shift) & 0xFFu);
    };
    auto pick = [&](const unsigned (&ws)[NW]) -> unsigned {
        return sk == 0 ? ws[0] : sk == 1 ? ws[1] : sk == 2 ? ws[2] : ws[3];
    };
    if (valid) {
        const uint4 w = load_row(y0 > 0 ? y0 - 1 : 0);
        const unsigned ws[NW] = {w.x, w.y, w.z, w.w};
        if (sb >= 0) {
            const unsigned e = pick(ws);
            s_prev_i = byte_at(e, 16);
            s_prev_q = byte_at(e, 24);
        }
#pragma unroll
        for (int k = 0; k < NW; ++k) {
            const unsigned mi = ws[k] & MI, mq = (ws[k] >> 8) & MI;
            edge_i[k] = dot4<SIGNED>(mi, ONES, 0);
            edge_q[k] = dot4<SIGNED>(mq, ONES, 0);
            pmi[k] = mi;
            pmq[k] = mq;
        }
        have_prev = y0 > 0;
        constexpr int U = 8;  // rows loaded per iteration (independent loads in flight)
        for (int y = y0; y < y1; y += U) {
            uint4 w8[U];
#pragma unroll
            for (int u = 0; u < U; ++u) w8[u] = (y + u < y1) ? load_row(y + u) : uint4{};
#pragma unroll
            for (int u = 0; u < U; ++u) {
                if (y + u >= y1) break;
                const unsigned ws2[NW] = {w8[u].x, w8[u].y, w8[u].z, w8[u].w};
#pragma unroll
                for (int k = 0; k < NW; ++k) {
                    const unsigned mi = ws2[k] & MI, mq = (ws2[k] >> 8) & MI;
                    sii = dot4<SIGNED>(mi, mi, sii);
                    sqq = dot4<SIGNED>(mq, mq, sqq);
                    siq = dot4<SIGNED>(mi, mq, siq);
                    wi[k] = dot4<SIGNED>(mi, ONES, wi[k]);
                    wq[k] = dot4<SIGNED>(mq, ONES, wq[k]);
                    if (have_prev) {
                        lii[k] = dot4<SIGNED>(pmi[k], mi, lii[k]);
                        lqq[k] = dot4<SIGNED>(pmq[k], mq, lqq[k]);
                        liq[k] = dot4<SIGNED>(pmi[k], mq, liq[k]);
                        lqi[k] = dot4<SIGNED>(pmq[k], mi, lqi[k]);
                    }
                    pmi[k] = mi;
                    pmq[k] = mq;
                }
                if (sb >= 0) {
                    const unsigned e = pick(ws2);
                    const int ci = byte_at(e, 16), cq = byte_at(e, 24);
                    if (have_prev) {
                        s_ii += s_prev_i * ci;
                        s_qq += s_prev_q * cq;
                        s_iq += s_prev_i * cq;
                        s_qi += s_prev_q * ci;
                        s_pi += s_prev_i;
                        s_pq += s_prev_q;
                        s_ci += ci;
                        s_cq += cq;
                        ++s_n;
                    }
                    s_prev_i = ci;
                    s_prev_q = cq;
                }
                have_prev = true;
            }
        }
    }
    if (sb >= 0) {
        const int sv[kRawLag] = {s_ii, s_qq, s_iq, s_qi, s_pi, s_pq, s_ci, s_cq, s_n};
        long long* cp = a.corr + ((long)strip * a.nblocks + sb) * kRawLag;
#pragma unroll
        for (int i = 0; i < kRawLag; ++i) cp[i] = sv[i];
    }
    const int rows = valid ? y1 - y0 : 0;
    const int lag_rows = valid ? (y1 - y0) - (y0 == 0 ? 1 : 0) : 0;
    const int nval = kRawStats + a.nslots * kRawLag;
    long long* out = a.partials + ((long)strip * a.ntiles + tile) * nval;
    __shared__ long long shs[kThreads / 64][kRawStats];
    __shared__ int shl[kRawLag][TW / 2];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
    int si = 0, sq = 0;
#pragma unroll
    for (int k = 0; k < NW; ++k) si += wi[k], sq += wq[k];
    const long long stv[kRawStats] = {si, sq, sii, sqq, siq, (long long)rows * V};
#pragma unroll
    for (int i = 0; i < kRawStats; ++i) {
        const long long v = wave_sum<long long>(stv[i]);
        if (lane == 0) shs[wave][i] = v;
    }
    // lag sums of each word: cur over the lag rows = strip sum less row 0 (strip 0 only); prev = strip sum less
    // the last row plus the row above the strip (strips after the first)
#pragma unroll
    for (int k = 0; k < NW; ++k) {
        const int last_i = dot4<SIGNED>(pmi[k], ONES, 0), last_q = dot4<SIGNED>(pmq[k], ONES, 0);
        const int li2 = y0 == 0 ? wi[k] - edge_i[k] : wi[k];
        const int lq2 = y0 == 0 ? wq[k] - edge_q[k] : wq[k];
        const int li1 = wi[k] - last_i + (y0 > 0 ? edge_i[k] : 0);
        const int lq1 = wq[k] - last_q + (y0 > 0 ? edge_q[k] : 0);
        const int c = threadIdx.x * NW + k;
        const int vals[kRawLag] = {lii[k], lqq[k], liq[k], lqi[k], li1, lq1, li2, lq2, 2 * lag_rows};
#pragma unroll
        for (int i = 0; i < kRawLag; ++i) shl[i][c] = valid ? vals[i] : 0;
    }
    __syncthreads();
    if (threadIdx.x < kRawStats) {
        long long t = 0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) t += shs[w][threadIdx.x];
        out[threadIdx.x] = t;
    }
    // lag sums per block slot: the words whose first column lies in the block
    const int c0 = tile * TW;
    const int b0 = c0 / a.step;
    for (int k = threadIdx.x; k < a.nslots * kRawLag; k += kThreads) {
        const int slot = k / kRawLag, v = k % kRawLag;
        const int b = b0 + slot;
        long long t = 0;
        if (b < a.nblocks) {
            // words [l0, l1): first columns c0 + 2l in [b step, (b + 1) step)
            const int l0 = max(0, (b * a.step - c0 + 1) / 2);
            const int l1 = min(TW / 2, ((b + 1) * a.step - c0 + 1) / 2);
            for (int l = l0; l < l1; ++l) t += shl[v][l];
        }
        out[kRawStats + k] = t;
    }
}

// ASAR FBAQ-4 level-0 records (RAW_ASAR_FBAQ4, RawMomentsArgs::lines): the samples are reconstruction-table
// values, not integers, so the moments are FP64 sums, formed in a fixed order (each lane over its strip's rows,
// the wave butterfly, the workgroup's waves, then k_reduce_raw over the strips and tiles) and stored as FP64 bit
// patterns in the int64 partial slots. One range column per lane. Range samples no record sample lands on are the
// reference decoder's NaN: skipped by the sums, and a lag-1 pair counts only when both of its samples exist. Each
// group of 8 rows is decoded in phases (record lengths and shifts, then block ids and code bytes, then the table
// entries) so that the dependent loads of the 8 rows are in flight together. (Staging the 32 KB table in LDS per
// workgroup measured slower: 0.46 -> 0.67 ms on the ASAR scene; the 16 entries of a block share one cache line.)
__global__ __launch_bounds__(kThreads) void k_raw_moments_fbaq(RawMomentsArgs a) {
    const float2* lut = a.lines.lut;
    const int tile = blockIdx.x;
    const int strip = blockIdx.y;
    const int x = tile * kThreads + threadIdx.x;
    const int y0 = strip * kStripRows;
    const int y1 = min(y0 + kStripRows, a.na);
    const bool col_ok = x < a.nr;
    const unsigned char* base = static_cast<const unsigned char*>(a.raw);
    const LinesDesc L = a.lines;
    double st[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    double lg[8] = {0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    double nvalid = 0.0, npairs = 0.0;
    float2 prev = make_float2(0.f, 0.f);
    bool have_prev = false;
    if (y0 > 0 && col_ok) {
        const unsigned char* rec = base + (long)(y0 - 1) * L.line_bytes;
        have_prev = fbaq4_sample(rec, fbaq4_record_length(rec), L.line_bytes, lut, x - L.shift(y0 - 1), &prev);
    }
    constexpr int U = 8;
    for (int y = y0; y < y1; y += U) {
        int len[U], sh[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            len[u] = 0;
            sh[u] = 0;
            if (y + u < y1) {
                len[u] = fbaq4_record_length(base + (long)(y + u) * L.line_bytes);
                sh[u] = L.shift(y + u);
            }
        }
        int bid[U], code[U];
        bool ok[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            const int smp = x - sh[u];
            const int blk = smp >= 0 ? smp / 63 : 0;
            const int k = 64 * blk + 1 + (smp - 63 * blk);
            ok[u] = col_ok && y + u < y1 && smp >= 0 && k < len[u] && 2 + (long)k < L.line_bytes;
            bid[u] = code[u] = 0;
            if (ok[u]) {
                const unsigned char* rec = base + (long)(y + u) * L.line_bytes;
                bid[u] = rec[2 + 64 * blk];
                code[u] = rec[2 + k];
            }
        }
        float2 v[U];
#pragma unroll
        for (int u = 0; u < U; ++u) {
            v[u] = make_float2(0.f, 0.f);
            if (ok[u]) {
                const int ci = code[u] >> 4, cq = code[u] & 15;
                const float2* lt = lut + 16 * bid[u];
                v[u] = make_float2(lt[ci > 7 ? 15 - ci : 8 + ci].x, lt[cq > 7 ? 15 - cq : 8 + cq].y);
            }
        }
#pragma unroll
        for (int u = 0; u < U; ++u) {
            if (y + u >= y1) break;
            if (ok[u]) {
                const double ci = v[u].x, cq = v[u].y;
                st[0] += ci;
                st[1] += cq;
                st[2] += ci * ci;
                st[3] += cq * cq;
                st[4] += ci * cq;
                nvalid += 1.0;
                if (have_prev) {
                    const double pi = prev.x, pq = prev.y;
                    lg[0] += pi * ci;
                    lg[1] += pq * cq;
                    lg[2] += pi * cq;
                    lg[3] += pq * ci;
                    lg[4] += pi;
                    lg[5] += pq;
                    lg[6] += ci;
                    lg[7] += cq;
                    npairs += 1.0;
                }
            }
            prev = v[u];
            have_prev = ok[u];
        }
    }
    const int nval = kRawStats + a.nslots * kRawLag;
    long long* out = a.partials + ((long)strip * a.ntiles + tile) * nval;
    __shared__ double shs[kThreads / 64][kRawStats];
    __shared__ double shl[kRawLag][kThreads];
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63;
#pragma unroll
    for (int i = 0; i < kRawStats; ++i) {
        const double v = wave_sum<double>(i < 5 ? st[i] : nvalid);
        if (lane == 0) shs[wave][i] = v;
    }
#pragma unroll
    for (int i = 0; i < 8; ++i) shl[i][threadIdx.x] = lg[i];
    shl[8][threadIdx.x] = npairs;
    __syncthreads();
    if (threadIdx.x < kRawStats) {
        double t = 0.0;
#pragma unroll
        for (int w = 0; w < kThreads / 64; ++w) t += shs[w][threadIdx.x];
        out[threadIdx.x] = __double_as_longlong(t);
    }
    // lag sums per block slot: columns of one block form a contiguous range
    const int c0 = tile * kThreads;
    const int b0 = c0 / a.step;
    for (int k = threadIdx.x; k < a.nslots * kRawLag; k += kThreads) {
        const int slot = k / kRawLag, v = k % kRawLag;
        const int b = b0 + slot;
        double t = 0.0;
        if (b < a.nblocks) {
            const int l0 = max(0, b * a.step - c0);
            const int l1 = min(kThreads, (b + 1) * a.step - c0);
            for (int l = l0; l < l1; ++l) t += shl[v][l];
        }
        out[kRawStats + k] = __double_as_longlong(t);
    }
}

hipError_t launch_raw_moments(const RawMomentsArgs& a, int fmt, hipStream_t s) {
    const dim3 grid(a.ntiles, a.nstrips);
    if (a.vec == 8) {
        // 16-byte loads of whole rows of eight pairs, at most one block boundary per lane, the straddle columns
        // into a.corr
        if (a.nr % 8 != 0 || (reinterpret_cast<uintptr_t>(a.raw) & 15) != 0 || a.step < 8 || !a.corr ||
            (fmt != RAW_U8 && fmt != RAW_I8))
            return hipErrorInvalidValue;
        if (fmt == RAW_U8)
            hipLaunchKernelGGL(k_raw_moments8<false>, grid, dim3(kThreads), 0, s, a);
        else
            hipLaunchKernelGGL(k_raw_moments8<true>, grid, dim3(kThreads), 0, s, a);
        return hipGetLastError();
    }
    if (fmt == RAW_ASAR_FBAQ4) {
        if (a.vec != 1 || !a.lines.lut) return hipErrorInvalidValue;
        hipLaunchKernelGGL(k_raw_moments_fbaq, grid, dim3(kThreads), 0, s, a);
        return hipGetLastError();
    }
    if (a.vec != 1 && a.vec != 4) return hipErrorInvalidValue;
    if (fmt == RAW_ERS_LINES) {
        // V = 4: 8-byte aligned rows of whole lane groups, shifted by multiples of 4 samples
        if (a.vec == 4 && (a.nr % 4 != 0 || a.lines.spl % 4 != 0 || a.lines.mult % 4 != 0 ||
                           a.lines.line_bytes % 8 != 0 || (reinterpret_cast<uintptr_t>(a.raw) & 7) != 0))
            return hipErrorInvalidValue;
        if (a.vec == 4)
            hipLaunchKernelGGL((k_raw_moments<uchar2, 4, true>), grid, dim3(kThreads), 0, s, a);
        else
            hipLaunchKernelGGL((k_raw_moments<uchar2, 1, true>), grid, dim3(kThreads), 0, s, a);
        return hipGetLastError();
    }
    // 8-byte loads: rows and the base must be 8-byte aligned
    if (a.vec == 4 && (a.nr % 4 != 0 || (reinterpret_cast<uintptr_t>(a.raw) & 7) != 0)) return hipErrorInvalidValue;
    if (fmt == RAW_U8) {
        if (a.vec == 4)
            hipLaunchKernelGGL((k_raw_moments<uchar2, 4>), grid, dim3(kThreads), 0, s, a);
        else
            hipLaunchKernelGGL((k_raw_moments<uchar2, 1>), grid, dim3(kThreads), 0, s, a);
    } else if (fmt == RAW_I8) {
        if (a.vec == 4)
            hipLaunchKernelGGL((k_raw_moments<char2, 4>), grid, dim3(kThreads), 0, s, a);
        else
            hipLaunchKernelGGL((k_raw_moments<char2, 1>), grid, dim3(kThreads), 0, s, a);
    } else {
        return hipErrorInvalidValue;
    }
    return hipGetLastError();
}

// totals[v]: statistics summed over every (strip, tile); lag value k of block b summed over every strip and the
// tiles whose slot range covers b. One workgroup per output value, fixed summation order.
// F64: the partials are FP64 bit patterns (k_raw_moments_fbaq), summed in FP64 in the same fixed order.
template <bool F64>
__global__ __launch_bounds__(kThreads) void k_reduce_raw(RawReduceArgs a) {
    using T = typename std::conditional<F64, double, long long>::type;
    auto val = [](long long bits) -> T {
        if constexpr (F64)
            return __longlong_as_double(bits);
        else
            return bits;
    };
    const int v = blockIdx.x;
    const int nval = kRawStats + a.nslots * kRawLag;
    T acc = 0;
    if (v < kRawStats) {
        const long parts = (long)a.nstrips * a.ntiles;
        for (long p = threadIdx.x; p < parts; p += kThreads) acc += val(a.partials[p * nval + v]);
    } else {
        const int b = (v - kRawStats) / kRawLag, k = (v - kRawStats) % kRawLag;
        // tiles t with b0(t) <= b < b0(t) + nslots
        const int t_lo = max(0, (b - a.nslots + 1) * a.step / a.tile_cols - 1);
        const int t_hi = min(a.ntiles - 1, ((b + 1) * a.step) / a.tile_cols + 1);
        const int nt = t_hi - t_lo + 1;
        const long parts = (long)a.nstrips * nt;
        for (long p = threadIdx.x; p < parts; p += kThreads) {
            const int strip = (int)(p / nt), t = t_lo + (int)(p % nt);
            const int slot = b - (t * a.tile_cols) / a.step;
            if (slot >= 0 && slot < a.nslots)
                acc += val(a.partials[((long)strip * a.ntiles + t) * nval + kRawStats + slot * kRawLag + k]);
        }
        // eight-column tiling: a column that starts block b at an odd index was filed under block b - 1
        if (!F64 && a.corr) {
            const bool own = b > 0 && ((b * a.step) & 1);
            const bool next = b + 1 < a.nblocks && (((b + 1) * a.step) & 1);
            for (int p = threadIdx.x; p < a.nstrips && (own || next); p += kThreads) {
                const long base = (long)p * a.nblocks;
                if (own) acc += val(a.corr[(base + b) * kRawLag + k]);
                if (next) acc -= val(a.corr[(base + b + 1) * kRawLag + k]);
            }
        }
    }
    __shared__ T sh[kThreads];
    sh[threadIdx.x] = acc;
    __syncthreads();
    for (int w = kThreads / 2; w > 0; w >>= 1) {
        if ((int)threadIdx.x < w) sh[threadIdx.x] += sh[threadIdx.x + w];
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        if constexpr (F64)
            a.totals[v] = __double_as_longlong(sh[0]);
        else
            a.totals[v] = sh[0];
    }
}

hipError_t launch_reduce_raw(const RawReduceArgs& a, hipStream_t s) {
    if (a.f64)
        hipLaunchKernelGGL(k_reduce_raw<true>, dim3(kRawStats + a.nblocks * kRawLag), dim3(kThreads), 0, s, a);
    else
        hipLaunchKernelGGL(k_reduce_raw<false>, dim3(kRawStats + a.nblocks * kRawLag), dim3(kThreads), 0, s, a);
    return hipGetLastError();
}

// Degree-M-1 least squares on one wave (lane i owns point i, n <= 64): Householder QR with wave reductions,
// the same steps as polyfit_qr (util/src/math_utils.cc:23-44), result highest power first.
template <int M>
__device__ bool polyfit_wave(double x, double y, int n, double* out) {
    const int lane = threadIdx.x & 63;
    const bool in = lane < n;
    double a[M];
    double v = 1.0;
#pragma unroll
    for (int c = 0; c < M; ++c) {
        a[c] = in ? v : 0.0;
        v *= x;
    }
    double b = in ? y : 0.0;
    double rdiag[M], rrow[M][M];
#pragma unroll
    for (int k = 0; k < M; ++k) {
        const double ak = lane >= k ? a[k] : 0.0;
        const double nrm = sqrt(wave_sum(ak * ak));
        if (nrm == 0.0) return false;
        const double akk = __shfl(a[k], k, 64);
        const double alpha = akk > 0.0 ? -nrm : nrm;
        const double hv = (lane == k) ? ak - alpha : ak;  // Householder vector (zero above row k)
        const double vv = wave_sum(hv * hv);
#pragma unroll
        for (int c = k + 1; c < M; ++c) {
            const double d = wave_sum(hv * a[c]);
            a[c] -= 2.0 * d / vv * hv;
        }
        const double d = wave_sum(hv * b);
        b -= 2.0 * d / vv * hv;
        rdiag[k] = alpha;
#pragma unroll
        for (int c = k + 1; c < M; ++c) rrow[k][c] = __shfl(a[c], k, 64);
    }
    double qb[M];
#pragma unroll
    for (int k = 0; k < M; ++k) qb[k] = __shfl(b, k, 64);
    double coef[M];
#pragma unroll
    for (int k = M - 1; k >= 0; --k) {
        double t = qb[k];
#pragma unroll
        for (int c = k + 1; c < M; ++c) t -= rrow[k][c] * coef[c];
        coef[k] = t / rdiag[k];
    }
#pragma unroll
    for (int k = 0; k < M; ++k) out[k] = coef[M - 1 - k];
    return true;
}

// One-wave FP64 finalisation of the fused statistics. With the corrected sample u = a I + b, v = c I + d Q + e,
// the lag-1 sum of a Doppler block is
//   sum conj(s_prev) s_cur = sum(u1 u2 + v1 v2) + i sum(u1 v2 - v1 u2),
// each term expanded over the raw lag moments. Padding columns of the last block (Q4) hold the corrected image
// of raw 0 when the correction runs over the padded width (Q1), i.e. the constant (b, e). Lane b handles block b.
__global__ __launch_bounds__(64) void k_finalize(FinalizeArgs f) {
    const int lane = threadIdx.x;
    const long long* T = f.totals;
    // exact integer sums, or (FBAQ-4 records) FP64 sums stored as bit patterns
    auto tot = [&](int i) -> double { return f.f64 ? __longlong_as_double(T[i]) : (double)T[i]; };
    double st[6];
#pragma unroll
    for (int i = 0; i < 6; ++i) st[i] = tot(i);
    const double SI = st[0], SQ = st[1], SII = st[2], SQQ = st[3], SIQ = st[4], nvalid = st[5];
    const double ntot = f.n_total;
    const double dci = SI / ntot, dcq = SQ / ntot;
    const bool q1 = (f.quirks & 1) != 0;
    const double P = q1 ? (double)f.na * (double)(f.x_stride - f.nr) : 0.0;
    const double sii = SII - 2.0 * dci * SI + nvalid * dci * dci + P * dci * dci;
    const double sqq = SQQ - 2.0 * dcq * SQ + nvalid * dcq * dcq + P * dcq * dcq;
    const double g = sqrt(sii / sqq);
    const double siq = (SIQ - dcq * SI - dci * SQ + nvalid * dci * dcq + P * dci * dcq) / g;
    const double delta = siq / (sii / (g * g));
    const double cosc = (f.quirks & 2) ? (1.0 - delta * delta) : sqrt(1.0 - delta * delta);
    const double ca = 1.0 / g, cb = -dci / g;
    const double cc = -delta / (g * cosc), cd = 1.0 / cosc, ce = (-dcq + delta * dci / g) / cosc;
    DevResults* r = f.res;
    double bx = 0.0, by = 0.0;
    if (lane < f.nblocks) {
        const int t = kRawStats + lane * kRawLag;
        const double LII = tot(t), LQQ = tot(t + 1), LIQ = tot(t + 2), LQI = tot(t + 3);
        const double I1 = tot(t + 4), Q1 = tot(t + 5), I2 = tot(t + 6), Q2 = tot(t + 7);
        const double M = tot(t + 8);
        const double uu = ca * ca * LII + ca * cb * (I1 + I2) + cb * cb * M;
        const double vv = cc * cc * LII + cc * cd * (LIQ + LQI) + cd * cd * LQQ + cc * ce * (I1 + I2) +
                          cd * ce * (Q1 + Q2) + ce * ce * M;
        const double uv = ca * cc * LII + ca * cd * LIQ + ca * ce * I1 + cb * cc * I2 + cb * cd * Q2 + cb * ce * M;
        const double vu = cc * ca * LII + cc * cb * I1 + cd * ca * LQI + cd * cb * Q1 + ce * ca * I2 + ce * cb * M;
        double re = uu + vv;
        const double im = uv - vu;
        const int xb = lane * f.step, xe = xb + f.step;
        const int pad0 = max(xb, f.nr), pad1 = min(xe, f.x_stride);
        if (q1 && pad1 > pad0 && f.na > 1) re += (double)(f.na - 1) * (double)(pad1 - pad0) * (cb * cb + ce * ce);
        bx = (double)((xb + xe) / 2);
        by = f.prf * atan2(im, re) / (2.0 * kPi);
        r->block_freq[lane] = by;
        r->block_x[lane] = bx;
    }
    double poly[3];
    if (!polyfit_wave<3>(bx, by, f.nblocks, poly)) poly[0] = poly[1] = poly[2] = 0.0;
    if (lane == 0) {
        r->dc_i = dci;
        r->dc_q = dcq;
        r->stddev_i = sqrt((SII - 2.0 * dci * SI + nvalid * dci * dci) / ntot);
        r->stddev_q = sqrt((SQQ - 2.0 * dcq * SQ + nvalid * dcq * dcq) / ntot);
        r->iq_gain = g;
        r->phase_error = asin(delta) * 180.0 / kPi;
        IqCoef c;
        c.dc_i = (float)dci;
        c.dc_q = (float)dcq;
        c.inv_gain = (float)(1.0 / g);
        c.sin_corr = (float)delta;
        c.inv_cos = (float)(1.0 / cosc);
        c.identity = 0;
        r->coef = c;
        r->nblocks = f.nblocks;
        r->dc_poly[0] = poly[0];
        r->dc_poly[1] = poly[1];
        r->dc_poly[2] = poly[2];
        // results window: lines cut at the start when the constant Doppler term is positive, else at the end
        const bool pos = poly[2] > 0.0;
        r->crop_front = pos ? f.crop_dc_pos : f.crop_dc_neg;
        const int in_rows = pos ? f.imp_in_rows_pos : f.imp_in_rows_neg;
        r->imp_rows = (f.ml_ratio > 0.f) ? (int)roundf((float)in_rows / f.ml_ratio) : 0;
        r->pad = in_rows;
        r->rcmc_reach = 0;
    }
}

hipError_t launch_finalize(const FinalizeArgs& a, hipStream_t s) {
    hipLaunchKernelGGL(k_finalize, dim3(1), dim3(64), 0, s, a);
    return hipGetLastError();
}

}  // namespace asarf
