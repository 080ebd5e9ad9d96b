// Split-precision FFT passes on the matrix cores: every radix-16 Stockham pass is a batched 16-point DFT done as
// a real GEMM on v_mfma_f32_16x16x32_bf16, with both operands split into bfloat16 hi/lo pairs and FP32
// accumulation (x ~ hi + lo keeps 16 significant bits; the product drops only lo*lo).
//
//   [Yr]   [Fr  -Fi] [Xr]        A1 = [Fr | -Fi], A2 = [Fi | Fr]   (16 x 32, constants per lane)
//   [Yi] = [Fi   Fr] [Xi]        B  = [Xr ; Xi] of 16 butterflies  (32 x 16, one butterfly per column)
//   Y ~ Ah Bh + Ah Bl + Al Bh    (three MFMAs per output half, chained in the accumulator)
//
// Lane layout of the 16x16x32 instruction (checked on the device): lane l holds A[l % 16][8 (l / 16) + i] and
// B[8 (l / 16) + i][l % 16] for i < 8, and D[4 (l / 16) + j][l % 16] for j < 4. So lane l loads inputs
// r = 8 ((l / 16) & 1) .. + 7 of butterfly l % 16 (real parts for l < 32, imaginary parts above) and receives
// outputs s = 4 (l / 16) .. + 3 of that butterfly.
//
// Twiddles (W_n^e from the LDS two-level table), the pass structure and the output order are those of the FP32
// Stockham engine (fft_core.h), so the split-precision transforms are drop-in replacements:
//   range compression rows of 8192 = 16 * 16 * 16 * 2 (three MFMA passes + one FP32 radix-2 pass),
//   column passes of 256 = 16 * 16 and 128 = 16 * 8 (the four-step halves of a 32768-point azimuth FFT).
#include "kernels.h"

namespace asarf {

typedef __bf16 bf16x8 __attribute__((ext_vector_type(8)));
typedef float f32x4 __attribute__((ext_vector_type(4)));

__constant__ float c_cos16[16] = {
    (float)cx_cos(0 * kPi / 8),  (float)cx_cos(1 * kPi / 8),  (float)cx_cos(2 * kPi / 8),  (float)cx_cos(3 * kPi / 8),
    (float)cx_cos(4 * kPi / 8),  (float)cx_cos(5 * kPi / 8),  (float)cx_cos(6 * kPi / 8),  (float)cx_cos(7 * kPi / 8),
    (float)cx_cos(8 * kPi / 8),  (float)cx_cos(9 * kPi / 8),  (float)cx_cos(10 * kPi / 8), (float)cx_cos(11 * kPi / 8),
    (float)cx_cos(12 * kPi / 8), (float)cx_cos(13 * kPi / 8), (float)cx_cos(14 * kPi / 8), (float)cx_cos(15 * kPi / 8)};
__constant__ float c_sin16[16] = {
    (float)cx_sin(0 * kPi / 8),  (float)cx_sin(1 * kPi / 8),  (float)cx_sin(2 * kPi / 8),  (float)cx_sin(3 * kPi / 8),
    (float)cx_sin(4 * kPi / 8),  (float)cx_sin(5 * kPi / 8),  (float)cx_sin(6 * kPi / 8),  (float)cx_sin(7 * kPi / 8),
    (float)cx_sin(8 * kPi / 8),  (float)cx_sin(9 * kPi / 8),  (float)cx_sin(10 * kPi / 8), (float)cx_sin(11 * kPi / 8),
    (float)cx_sin(12 * kPi / 8), (float)cx_sin(13 * kPi / 8), (float)cx_sin(14 * kPi / 8), (float)cx_sin(15 * kPi / 8)};

struct Dft16Ops {
    bf16x8 a1h, a1l, a2h, a2l;
};

__device__ __forceinline__ void split_bf16(const float (&v)[8], bf16x8& h, bf16x8& l) {
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const __bf16 hb = (__bf16)v[i];
        h[i] = hb;
        l[i] = (__bf16)(v[i] - (float)hb);
    }
}

template <bool INV>
__device__ __forceinline__ Dft16Ops dft16_ops(int lane) {
    const int m = lane & 15, kb = 8 * (lane >> 4);
    float a1[8], a2[8];
#pragma unroll
    for (int i = 0; i < 8; ++i) {
        const int k = kb + i;
        const int idx = (m * (k & 15)) & 15;
        const float c = c_cos16[idx], s = c_sin16[idx];
        const float fi = INV ? s : -s;  // imaginary part of exp(-+2 pi i m k / 16)
        a1[i] = k < 16 ? c : -fi;
        a2[i] = k < 16 ? fi : c;
    }
    Dft16Ops o;
    split_bf16(a1, o.a1h, o.a1l);
    split_bf16(a2, o.a2h, o.a2l);
    return o;
}

__device__ __forceinline__ f32x4 mfma3(const bf16x8& ah, const bf16x8& al, const bf16x8& bh, const bf16x8& bl) {
    f32x4 d = {0.f, 0.f, 0.f, 0.f};
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bh, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(ah, bl, d, 0, 0, 0);
    d = __builtin_amdgcn_mfma_f32_16x16x32_bf16(al, bh, d, 0, 0, 0);
    return d;
}

// One in-place radix-16 Stockham pass (prefix product ns) over the butterflies of the buffer. Butterfly
// (grp, l % 16) -> (transform t, index j) by BF; element e of transform t lives at ADDR(t, e). A wave handles
// groups wave, wave + nwaves, ...; GPW bounds the groups per wave. Ends with a barrier.
template <int GPW, bool INV, class BF, class ADDR>
__device__ __forceinline__ void mfma_pass16(float2* buf, int n, int ns, int ngroups, const TwLds& tw,
                                            const Dft16Ops& ops, int wave, int nwaves, int lane, BF&& bf,
                                            ADDR&& addr) {
    const int nbf = n / 16;
    const int q = (lane >> 4) & 1, comp = lane >> 5, nl = lane & 15;
    const int tstride = n / (16 * ns);
    f32x4 d1[GPW], d2[GPW];
    static_for<0, GPW>([&](auto gc) __attribute__((always_inline)) {
        constexpr int g = decltype(gc)::value;
        const int grp = wave + g * nwaves;
        if (grp < ngroups) {  // uniform over the wave
            int t, j;
            bf(grp, nl, t, j);
            const int k = j % ns;
            float v[8];
#pragma unroll
            for (int i = 0; i < 8; ++i) {
                const int r = 8 * q + i;
                float2 x = buf[addr(t, j + r * nbf)];
                if (ns > 1) {
                    const float2 w = tw(r * k * tstride);
                    x = INV ? cmulc(x, w) : cmul(x, w);
                }
                v[i] = comp ? x.y : x.x;
            }
            bf16x8 bh, bl;
            split_bf16(v, bh, bl);
            d1[g] = mfma3(ops.a1h, ops.a1l, bh, bl);
            d2[g] = mfma3(ops.a2h, ops.a2l, bh, bl);
        }
    });
    __syncthreads();
    static_for<0, GPW>([&](auto gc) __attribute__((always_inline)) {
        constexpr int g = decltype(gc)::value;
        const int grp = wave + g * nwaves;
        if (grp < ngroups) {
            int t, j;
            bf(grp, nl, t, j);
            const int k = j % ns;
            const int base = (j - k) * 16 + k;
#pragma unroll
            for (int jj = 0; jj < 4; ++jj) {
                const int s = 4 * (lane >> 4) + jj;
                buf[addr(t, base + s * ns)] = make_float2(d1[g][jj], d2[g][jj]);
            }
        }
    });
    __syncthreads();
}

// ------------------------------------------------------------------------------------ rows of 8192 ---------
constexpr int kMfRowN = 8192;
constexpr int kMfRowT = 512;  // 8 waves: 4 radix-16 groups per wave and pass
__host__ __device__ constexpr int mf_pad(int i) { return i + (i >> 4); }
constexpr int kMfRowLds = mf_pad(kMfRowN - 1) + 1;

template <bool INV>
__device__ __forceinline__ void row8192_mfma_passes(float2* buf, const TwLds& tw, const Dft16Ops& ops, int wave,
                                                    int lane) {
    auto bf = [](int grp, int nl, int& t, int& j) {
        t = 0;
        j = grp * 16 + nl;
    };
    auto addr = [](int, int e) { return mf_pad(e); };
    constexpr int ngroups = kMfRowN / 16 / 16;  // 32
    constexpr int nwaves = kMfRowT / 64;
    mfma_pass16<ngroups / nwaves, INV>(buf, kMfRowN, 1, ngroups, tw, ops, wave, nwaves, lane, bf, addr);
    mfma_pass16<ngroups / nwaves, INV>(buf, kMfRowN, 16, ngroups, tw, ops, wave, nwaves, lane, bf, addr);
    mfma_pass16<ngroups / nwaves, INV>(buf, kMfRowN, 256, ngroups, tw, ops, wave, nwaves, lane, bf, addr);
}

// Last (radix-2, ns = 4096) pass: butterfly j pairs positions j and j + 4096, which are also its outputs.
template <bool INV, class F>
__device__ __forceinline__ void row8192_radix2(float2* buf, const TwLds& tw, F&& sink) {
    constexpr int h = kMfRowN / 2;
#pragma unroll 4
    for (int j = threadIdx.x; j < h; j += kMfRowT) {
        const float2 x0 = buf[mf_pad(j)];
        float2 x1 = buf[mf_pad(j + h)];
        const float2 w = tw(j);
        x1 = INV ? cmulc(x1, w) : cmul(x1, w);
        sink(j, cadd(x0, x1));
        sink(j + h, csub(x0, x1));
    }
}

template <typename T>
struct RawPairM;
template <>
struct RawPairM<uint8_t> {
    using type = uchar2;
};
template <>
struct RawPairM<int8_t> {
    using type = char2;
};

// Range compression of one line with split-precision transforms (same steps as k_rc_static).
template <int FMT>
__global__ __launch_bounds__(kMfRowT) void k_rc_mfma(RcArgs a) {
    extern __shared__ float2 lds[];
    float2* buf = lds;
    const int row = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const TwLds tw = tw_fill(lds + kMfRowLds, a.plan.tw, kMfRowN, threadIdx.x, kMfRowT);
    IqCoef c;
    if (a.coef) {
        c = *a.coef;
    } else {
        c.dc_i = c.dc_q = c.sin_corr = 0.f;
        c.inv_gain = c.inv_cos = 1.f;
        c.identity = 1;
    }
    auto correct = [&](float i, float q) {
        if (c.identity) return make_float2(i, q);
        const float ic = (i - c.dc_i) * c.inv_gain;
        const float qc = (q - c.dc_q - ic * c.sin_corr) * c.inv_cos;
        return make_float2(ic, qc);
    };
    for (int x = threadIdx.x; x < kMfRowN; x += kMfRowT) {
        float2 v = make_float2(0.f, 0.f);
        if constexpr (FMT == RAW_CF32) {
            const float2* in = static_cast<const float2*>(a.in) + (long)row * a.in_stride;
            if (x < a.nr_in) v = correct(in[x].x, in[x].y);
        } else {
            using P = typename RawPairM<typename std::conditional<FMT == RAW_U8, uint8_t, int8_t>::type>::type;
            const P* in = static_cast<const P*>(a.in) + (long)row * a.in_stride;
            if (x < a.nr_in) {
                const P r = in[x];
                v = correct((float)r.x, (float)r.y);
            }
        }
        buf[mf_pad(x)] = v;
    }
    __syncthreads();
    const Dft16Ops fwd = dft16_ops<false>(lane);
    row8192_mfma_passes<false>(buf, tw, fwd, wave, lane);
    const float2* h = a.hconj;
    row8192_radix2<false>(buf, tw, [&](int k, float2 y) { buf[mf_pad(k)] = cmul(y, h[k]); });
    __syncthreads();
    const Dft16Ops inv = dft16_ops<true>(lane);
    row8192_mfma_passes<true>(buf, tw, inv, wave, lane);
    float2* out = a.out + (long)row * a.out_stride;
    const int nrv = a.nrv, zt = a.zero_to;
    row8192_radix2<true>(buf, tw, [&](int x, float2 y) {
        if (x < nrv)
            out[x] = y;
        else if (x < zt)
            out[x] = make_float2(0.f, 0.f);
    });
}

// Plain row transform of 8192 points (test hook and building block): in place or out of place, scaled.
template <bool INV>
__global__ __launch_bounds__(kMfRowT) void k_row_mfma(RowFftArgs a) {
    extern __shared__ float2 lds[];
    float2* buf = lds;
    const int row = blockIdx.x;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const TwLds tw = tw_fill(lds + kMfRowLds, a.plan.tw, kMfRowN, threadIdx.x, kMfRowT);
    const float2* in = a.in + (long)row * a.in_stride;
    for (int x = threadIdx.x; x < kMfRowN; x += kMfRowT) buf[mf_pad(x)] = in[x];
    __syncthreads();
    const Dft16Ops ops = dft16_ops<INV>(lane);
    row8192_mfma_passes<INV>(buf, tw, ops, wave, lane);
    float2* out = a.out + (long)row * a.out_stride;
    const float sc = a.scale;
    row8192_radix2<INV>(buf, tw, [&](int x, float2 y) { out[x] = cscale(y, sc); });
}

// ------------------------------------------------------------------------------ column passes 128, 256 -----
// Workgroup (tile, g): 16 columns x one sub-transform of length N staged in LDS as [N][17]; lanes of a 16-lane
// group map to the 16 columns, so every LDS access of a pass touches 16 consecutive banks.
template <int N, bool INV>
__global__ __launch_bounds__(kThreads) void k_col_mfma(ColFftArgs a) {
    extern __shared__ float2 lds[];
    if (a.fused_gate && *a.fused_gate <= kAzCoreReach) return;
    constexpr int P = kColTile + 1;
    float2* buf = lds;
    const int tile = blockIdx.x % a.ntiles;
    const int g = blockIdx.x / a.ntiles;
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const TwLds tw = tw_fill(lds + N * P, a.sub.tw, N, threadIdx.x, kThreads);
    const int b = threadIdx.x & (kColTile - 1);
    const int col = tile * kColTile + b;
    const bool col_ok = col < a.ncols;
    const float2* in = a.in + (col_ok ? col : 0);
    const long in_base = (long)g * a.in_g;
#pragma unroll 4
    for (int m = threadIdx.x / kColTile; m < N; m += kThreads / kColTile) {
        const long r = in_base + (long)m * a.in_m;
        buf[m * P + b] = (col_ok && r < a.in_rows) ? in[r * a.in_stride] : make_float2(0.f, 0.f);
    }
    __syncthreads();
    auto bf = [](int grp, int nl, int& t, int& j) {
        t = nl;
        j = grp;
    };
    auto addr = [](int t, int e) { return e * P + t; };
    const Dft16Ops ops = dft16_ops<INV>(lane);
    constexpr int ngroups = N / 16;
    constexpr int gpw = (ngroups + 3) / 4;
    mfma_pass16<gpw, INV>(buf, N, 1, ngroups, tw, ops, wave, 4, lane, bf, addr);
    if constexpr (N == 256) {
        mfma_pass16<gpw, INV>(buf, N, 16, ngroups, tw, ops, wave, 4, lane, bf, addr);
    } else {
        static_assert(N == 128, "column lengths 128 and 256");
        // radix-8 pass, ns = 16: thread -> (column, butterfly j = 0..15); outputs land on the input positions
        const int t = threadIdx.x & 15, j = threadIdx.x >> 4;
        float2 v[8];
#pragma unroll
        for (int r = 0; r < 8; ++r) {
            v[r] = buf[(j + r * 16) * P + t];
            if (r > 0) {
                const float2 w = tw(r * j);
                v[r] = INV ? cmulc(v[r], w) : cmul(v[r], w);
            }
        }
        Dft<8, INV>::run(v);
        __syncthreads();
#pragma unroll
        for (int r = 0; r < 8; ++r) buf[(j + r * 16) * P + t] = v[r];
        __syncthreads();
    }
    float2* out = a.out + (col_ok ? col : 0);
    const long out_base = (long)g * a.out_g;
    const float sc = a.scale;
    const float2* twf = a.tw_full;
#pragma unroll 4
    for (int k = threadIdx.x / kColTile; k < N; k += kThreads / kColTile) {
        const long r = out_base + (long)k * a.out_m;
        if (!col_ok || r >= a.out_rows) continue;
        float2 y = buf[k * P + b];
        if (twf) {
            const float2 w = twf[(long)g * k];
            y = INV ? cmulc(y, w) : cmul(y, w);
        }
        out[r * a.out_stride] = cscale(y, sc);
    }
}

// --------------------------------------------------------------------------------------------- launchers --
hipError_t launch_rc_mfma(const RcArgs& a, int fmt, hipStream_t s) {
    if (a.plan.n != kMfRowN) return hipErrorInvalidValue;
    if (fmt != RAW_CF32 && fmt != RAW_U8 && fmt != RAW_I8) return hipErrorInvalidValue;
    const size_t lds = sizeof(float2) * (size_t)(kMfRowLds + tw_lds_elems(kMfRowN));
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
#define RC_(F)                                                                                                   \
    {                                                                                                            \
        const hipError_t e_ = hipFuncSetAttribute((const void*)k_rc_mfma<F>,                                     \
                                                  hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);         \
        if (e_ != hipSuccess) return e_;                                                                         \
        hipLaunchKernelGGL((k_rc_mfma<F>), dim3(a.rows), dim3(kMfRowT), lds, s, a);                             \
    }
    if (fmt == RAW_CF32)
        RC_(RAW_CF32)
    else if (fmt == RAW_U8)
        RC_(RAW_U8)
    else
        RC_(RAW_I8)
#undef RC_
    return hipGetLastError();
}

hipError_t launch_row_mfma(const RowFftArgs& a, bool inverse, hipStream_t s) {
    if (a.plan.n != kMfRowN) return hipErrorInvalidValue;
    const size_t lds = sizeof(float2) * (size_t)(kMfRowLds + tw_lds_elems(kMfRowN));
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    const void* kf = inverse ? (const void*)k_row_mfma<true> : (const void*)k_row_mfma<false>;
    const hipError_t e = hipFuncSetAttribute(kf, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
    if (e != hipSuccess) return e;
    if (inverse)
        hipLaunchKernelGGL((k_row_mfma<true>), dim3(a.rows), dim3(kMfRowT), lds, s, a);
    else
        hipLaunchKernelGGL((k_row_mfma<false>), dim3(a.rows), dim3(kMfRowT), lds, s, a);
    return hipGetLastError();
}

hipError_t launch_col_mfma(const ColFftArgs& a, bool inverse, hipStream_t s) {
    const int n = a.sub.n;
    if (n != 128 && n != 256) return hipErrorInvalidValue;
    const size_t lds = sizeof(float2) * (size_t)(n * (kColTile + 1) + tw_lds_elems(n));
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    const dim3 grid((unsigned)((long)a.ntiles * a.groups));
    if (n == 128) {
        if (inverse)
            hipLaunchKernelGGL((k_col_mfma<128, true>), grid, dim3(kThreads), lds, s, a);
        else
            hipLaunchKernelGGL((k_col_mfma<128, false>), grid, dim3(kThreads), lds, s, a);
    } else {
        if (inverse)
            hipLaunchKernelGGL((k_col_mfma<256, true>), grid, dim3(kThreads), lds, s, a);
        else
            hipLaunchKernelGGL((k_col_mfma<256, false>), grid, dim3(kThreads), lds, s, a);
    }
    return hipGetLastError();
}

}  // namespace asarf
