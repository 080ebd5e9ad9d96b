// Range compression with exchange addresses that are compile-time offsets.
//
// Same stage as k_rc_static (correct on load, forward FFT, x conj(H)/n, inverse FFT, keep nrv columns; the
// reference's RangeCompression, sar/range_compression.cu:57-111), with two changes to the transform engine that
// remove most of its integer work:
//
// * LDS layout i -> i + (i >> PS): one padding element every 2^PS. A Stockham pass gathers x[j + r*N/R] and
//   scatters to (j / NS)*NS*R + j % NS + r*NS; for the production radix sequences both position sets, mapped
//   through the layout, are a per-butterfly base plus an offset that depends on r alone (checked at compile
//   time per pass), so every LDS access is one base address and an immediate offset.
// * The inverse transform runs as the transpose of the forward decimation-in-time passes (decimation in
//   frequency, last forward pass first): it loads where the forward pass stored and stores where it loaded, so
//   it inherits the same compile-time offsets, and its first pass consumes the forward transform's registers
//   with the twiddles of the forward's last pass, conjugated.
//
// With F = P_{m-1} ... P_1 P_0 (P_p = scatter * DFT_R * twiddle * gather), the unnormalised inverse transform is
// conj(F) = conj(F)^T = conj(P_0)^T ... conj(P_{m-1})^T, and conj(P_p)^T = gather^T * conj(twiddle) *
// inverse DFT_R * scatter^T: load from the scatter positions, inverse DFT, multiply by the conjugate twiddles,
// store to the gather positions.
#include "kernels.h"

namespace asarf {
namespace rcpad {

template <int PS>
__host__ __device__ constexpr int lpos(int i) {
    return i + (i >> PS);
}

template <int R, int NS>
__host__ __device__ constexpr int scat_base(int j) {
    return (j / NS) * NS * R + j % NS;
}

// True when lpos(j + r*(N/R)) - lpos(j) depends on r alone.
template <int N, int R, int PS>
constexpr bool gather_const() {
    constexpr int M = N / R;
    for (int j = 0; j < M; ++j)
        for (int r = 1; r < R; ++r)
            if (lpos<PS>(j + r * M) - lpos<PS>(j) != lpos<PS>(r * M)) return false;
    return true;
}

// True when lpos(scat_base(j) + r*NS) - lpos(scat_base(j)) depends on r alone.
template <int N, int R, int NS, int PS>
constexpr bool scatter_const() {
    for (int j = 0; j < N / R; ++j) {
        const int b = scat_base<R, NS>(j);
        for (int r = 1; r < R; ++r)
            if (lpos<PS>(b + r * NS) - lpos<PS>(b) != lpos<PS>(r * NS)) return false;
    }
    return true;
}

// Butterflies j = lane + i*T, i < NB. Lanes past the last butterfly work on a clamped copy (registers written
// unconditionally keep the array in registers); a butterfly past the end for the whole wave is skipped.
template <int N, int R, int T>
struct Pass {
    static constexpr int M = N / R;
    static constexpr int NB = (M + T - 1) / T;
    static constexpr bool partial(int i) { return (i + 1) * T > M; }
};

template <int N, int R, int T, int PS, int E>
__device__ __forceinline__ void ld_gather(float2 (&v)[E], int lane, const float2* lds) {
    using P = Pass<N, R, T>;
    static_assert(P::NB * R <= E, "register array too small");
    constexpr bool C = gather_const<N, R, PS>();
#pragma unroll
    for (int i = 0; i < P::NB; ++i) {
        if (P::partial(i) && !wave_any(lane + i * T < P::M)) break;
        const int j = min(lane + i * T, P::M - 1);
        if constexpr (C) {
            const float2* p = lds + lpos<PS>(j);
#pragma unroll
            for (int r = 0; r < R; ++r) v[i * R + r] = p[lpos<PS>(r * P::M)];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i * R + r] = lds[lpos<PS>(j + r * P::M)];
        }
    }
}

template <int N, int R, int T, int PS, int E>
__device__ __forceinline__ void st_gather(const float2 (&v)[E], int lane, float2* lds) {
    using P = Pass<N, R, T>;
    constexpr bool C = gather_const<N, R, PS>();
#pragma unroll
    for (int i = 0; i < P::NB; ++i) {
        const int j = lane + i * T;
        if (P::partial(i) && j >= P::M) continue;
        if constexpr (C) {
            float2* p = lds + lpos<PS>(j);
#pragma unroll
            for (int r = 0; r < R; ++r) p[lpos<PS>(r * P::M)] = v[i * R + r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) lds[lpos<PS>(j + r * P::M)] = v[i * R + r];
        }
    }
}

template <int N, int R, int NS, int T, int PS, int E>
__device__ __forceinline__ void ld_scatter(float2 (&v)[E], int lane, const float2* lds) {
    using P = Pass<N, R, T>;
    static_assert(P::NB * R <= E, "register array too small");
    constexpr bool C = scatter_const<N, R, NS, PS>();
#pragma unroll
    for (int i = 0; i < P::NB; ++i) {
        if (P::partial(i) && !wave_any(lane + i * T < P::M)) break;
        const int j = min(lane + i * T, P::M - 1);
        const int b = scat_base<R, NS>(j);
        if constexpr (C) {
            const float2* p = lds + lpos<PS>(b);
#pragma unroll
            for (int r = 0; r < R; ++r) v[i * R + r] = p[lpos<PS>(r * NS)];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) v[i * R + r] = lds[lpos<PS>(b + r * NS)];
        }
    }
}

template <int N, int R, int NS, int T, int PS, int E>
__device__ __forceinline__ void st_scatter(const float2 (&v)[E], int lane, float2* lds) {
    using P = Pass<N, R, T>;
    constexpr bool C = scatter_const<N, R, NS, PS>();
#pragma unroll
    for (int i = 0; i < P::NB; ++i) {
        const int j = lane + i * T;
        if (P::partial(i) && j >= P::M) continue;
        const int b = scat_base<R, NS>(j);
        if constexpr (C) {
            float2* p = lds + lpos<PS>(b);
#pragma unroll
            for (int r = 0; r < R; ++r) p[lpos<PS>(r * NS)] = v[i * R + r];
        } else {
#pragma unroll
            for (int r = 0; r < R; ++r) lds[lpos<PS>(b + r * NS)] = v[i * R + r];
        }
    }
}

// Twiddles w^r = W_N^(r * (j % NS) * N / (NS * R)), r = 1..R-1, of butterfly j: one table lookup for w and the
// other powers by products (w^(2a) = (w^a)^2, w^r = w^hb w^(r - hb)), as k_rc_static's TW1 variant.
template <int N, int R, int NS>
__device__ __forceinline__ void twiddles(int j, const TwLds& tw, float2 (&pw)[R]) {
    const int k = j % NS;
    pw[1] = tw(k * (N / (NS * R)));
#pragma unroll
    for (int r = 2; r < R; ++r) {
        if ((r & (r - 1)) == 0) {
            pw[r] = cmul(pw[r / 2], pw[r / 2]);
        } else {
            int hb = 1;
            while (hb * 2 <= r) hb *= 2;
            pw[r] = cmul(pw[hb], pw[r - hb]);
        }
    }
}

// Forward decimation-in-time pass: twiddle (NS > 1), then DFT_R.
template <int N, int R, int NS, int T, int E>
__device__ __forceinline__ void dit_pass(float2 (&v)[E], int lane, const TwLds& tw) {
    using P = Pass<N, R, T>;
#pragma unroll
    for (int i = 0; i < P::NB; ++i) {
        if (P::partial(i) && !wave_any(lane + i * T < P::M)) break;
        if constexpr (NS > 1) {
            const int j = min(lane + i * T, P::M - 1);
            float2 pw[R];
            twiddles<N, R, NS>(j, tw, pw);
#pragma unroll
            for (int r = 1; r < R; ++r) v[i * R + r] = cmul(v[i * R + r], pw[r]);
        }
        Dft<R, false>::run(&v[i * R]);
        __builtin_amdgcn_sched_barrier(0);
    }
}

// Inverse decimation-in-frequency pass (transpose of the forward pass (R, NS)): inverse DFT_R, then the
// conjugate twiddles.
template <int N, int R, int NS, int T, int E>
__device__ __forceinline__ void dif_pass(float2 (&v)[E], int lane, const TwLds& tw) {
    using P = Pass<N, R, T>;
#pragma unroll
    for (int i = 0; i < P::NB; ++i) {
        if (P::partial(i) && !wave_any(lane + i * T < P::M)) break;
        Dft<R, true>::run(&v[i * R]);
        if constexpr (NS > 1) {
            const int j = min(lane + i * T, P::M - 1);
            float2 pw[R];
            twiddles<N, R, NS>(j, tw, pw);
#pragma unroll
            for (int r = 1; r < R; ++r) v[i * R + r] = cmulc(v[i * R + r], pw[r]);
        }
        __builtin_amdgcn_sched_barrier(0);
    }
}

template <typename T>
struct RawPair;
template <>
struct RawPair<uint8_t> {
    using type = uchar2;
};
template <>
struct RawPair<int8_t> {
    using type = char2;
};

template <int N, int PS>
__host__ __device__ constexpr int data_elems() {
    return lpos<PS>(N - 1) + 1;
}

template <int FMT, int N, int T, int PS, int WAVES, class Seq>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_rc_pad(RcArgs a) {
    static_assert(Seq::product() == N, "radix sequence must factor N");
    constexpr int E = seq_max_elems<N, Seq, T>();
    constexpr int NP = Seq::count;
    extern __shared__ float2 lds[];
    const int row = blockIdx.x;
    const int lane = threadIdx.x;
    const TwLds tw = tw_fill(lds + data_elems<N, PS>(), a.plan.tw, N, lane, T);
    IqCoef c;
    if (a.coef) {
        c = *a.coef;
    } else {
        c.dc_i = c.dc_q = c.sin_corr = 0.f;
        c.inv_gain = c.inv_cos = 1.f;
        c.identity = 1;
    }
    const int nr_in = a.nr_in;
    auto correct = [&](float i, float q) {
        if (c.identity) return make_float2(i, q);
        const float ic = (i - c.dc_i) * c.inv_gain;
        const float qc = (q - c.dc_q - ic * c.sin_corr) * c.inv_cos;
        return make_float2(ic, qc);
    };
    // ---- the row into LDS in natural order, corrected on load
    if constexpr (FMT == RAW_CF32) {
        const float2* in = static_cast<const float2*>(a.in) + (long)row * a.in_stride;
        for (int x = lane; x < N; x += T) {
            const float2 r = in[x];
            lds[lpos<PS>(x)] = correct(r.x, r.y);
        }
    } else if constexpr (FMT == RAW_ASAR_FBAQ4) {
        // the row's record decoded and shifted by its SWST: NaN where no sample lands becomes 0, padding is 0
        const unsigned char* rec = static_cast<const unsigned char*>(a.in) + (long)row * a.lines.line_bytes;
        fbaq4_row_decode<N, T>(rec, a.lines.line_bytes, a.lines.lut, a.lines.shift(row), nr_in, lane, correct,
                               [&](int x, float2 v) { lds[lpos<PS>(x)] = v; });
    } else if constexpr (FMT == RAW_ERS_LINES) {
        // the row's line shifted by its SWST: NaN where no line sample lands becomes 0 (ZeroNaNs), padding is 0
        const uchar2* line =
            reinterpret_cast<const uchar2*>(static_cast<const uint8_t*>(a.in) + (long)row * a.lines.line_bytes);
        const int shift = a.lines.shift(row);
        for (int x = lane; x < N; x += T) {
            const int sx = x - shift;
            float2 v = make_float2(0.f, 0.f);
            if (x < nr_in && sx >= 0 && sx < a.lines.spl) v = correct((float)line[sx].x, (float)line[sx].y);
            lds[lpos<PS>(x)] = v;
        }
    } else {
        using Pr = typename RawPair<typename std::conditional<FMT == RAW_U8, uint8_t, int8_t>::type>::type;
        const uint8_t* rb = static_cast<const uint8_t*>(a.in) + (long)row * a.in_stride * 2;
        const int nbytes = nr_in * 2;
        static_assert(PS >= 3, "eight consecutive elements stay contiguous in the layout");
        if ((reinterpret_cast<uintptr_t>(rb) & 15) == 0) {
            // 16-byte loads: 8 I/Q pairs each, stored at consecutive LDS positions (x is a multiple of 8)
            for (int off = lane * 16; off < nbytes; off += T * 16) {
                if (off + 16 <= nbytes) {
                    const uint4 w = *reinterpret_cast<const uint4*>(rb + off);
                    const unsigned ws[4] = {w.x, w.y, w.z, w.w};
                    float2* p = lds + lpos<PS>(off / 2);
#pragma unroll
                    for (int k = 0; k < 8; ++k) {
                        const unsigned word = ws[k >> 1] >> (16 * (k & 1));
                        float fi, fq;
                        if constexpr (FMT == RAW_U8) {
                            fi = (float)(word & 0xFFu);
                            fq = (float)((word >> 8) & 0xFFu);
                        } else {
                            fi = (float)(int8_t)(word & 0xFFu);
                            fq = (float)(int8_t)((word >> 8) & 0xFFu);
                        }
                        p[k] = correct(fi, fq);
                    }
                } else {
                    const Pr* in = reinterpret_cast<const Pr*>(rb);
                    for (int x = off / 2; x < nr_in; ++x) lds[lpos<PS>(x)] = correct((float)in[x].x, (float)in[x].y);
                }
            }
        } else {
            const Pr* in = reinterpret_cast<const Pr*>(rb);
            for (int x = lane; x < nr_in; x += T) lds[lpos<PS>(x)] = correct((float)in[x].x, (float)in[x].y);
        }
        for (int x = nr_in + lane; x < N; x += T) lds[lpos<PS>(x)] = make_float2(0.f, 0.f);
    }
    __syncthreads();

    float2 v[E];
    // ---- forward transform, decimation in time: pass p gathers, twiddles (p > 0), DFT, scatters
    static_for<0, NP>([&](auto pc) __attribute__((always_inline)) {
        constexpr int p = decltype(pc)::value;
        constexpr int R = Seq::at(p), NS = Seq::prefix(p);
        ld_gather<N, R, T, PS>(v, lane, lds);
        __syncthreads();  // every element is read before the scatter of this pass overwrites LDS
        dit_pass<N, R, NS, T>(v, lane, tw);
        if constexpr (p + 1 < NP) {
            st_scatter<N, R, NS, T, PS>(v, lane, lds);
            __syncthreads();
        }
    });
    // ---- matched filter: the last forward pass left X[k] at its scatter positions k = scat_base(j) + r*NS
    {
        constexpr int R = Seq::at(NP - 1), NS = Seq::prefix(NP - 1);
        using P = Pass<N, R, T>;
        const float2* h = a.hconj;
#pragma unroll
        for (int i = 0; i < P::NB; ++i) {
            const int j = min(lane + i * T, P::M - 1);
            const int b = scat_base<R, NS>(j);
#pragma unroll
            for (int r = 0; r < R; ++r) v[i * R + r] = cmul(v[i * R + r], h[b + r * NS]);
        }
    }
    // ---- inverse transform, decimation in frequency: the transposes of the forward passes, last one first.
    // Pass q inverts forward pass p = NP-1-q: its input sits at p's scatter positions (in registers for q = 0,
    // else in LDS), its output goes to p's gather positions (LDS), except for the last, which stays in registers.
    static_for<0, NP>([&](auto qc) __attribute__((always_inline)) {
        constexpr int q = decltype(qc)::value;
        constexpr int p = NP - 1 - q;
        constexpr int R = Seq::at(p), NS = Seq::prefix(p);
        if constexpr (q > 0) {
            ld_scatter<N, R, NS, T, PS>(v, lane, lds);
            __syncthreads();  // every element is read before the gather stores of this pass overwrite LDS
        }
        dif_pass<N, R, NS, T>(v, lane, tw);
        if constexpr (q + 1 < NP) {
            st_gather<N, R, T, PS>(v, lane, lds);
            __syncthreads();
        }
    });
    // ---- output: the last inverse pass (transpose of forward pass 0, NS = 1) holds x[j + r*N/R]
    {
        constexpr int R = Seq::at(0);
        using P = Pass<N, R, T>;
        float2* out = a.out + (long)row * a.out_stride;
        const int nrv = a.nrv, zt = a.zero_to;
#pragma unroll
        for (int i = 0; i < P::NB; ++i) {
            const int j = lane + i * T;
            if (P::partial(i) && j >= P::M) continue;
#pragma unroll
            for (int r = 0; r < R; ++r) {
                const int x = j + r * P::M;
                if (x < nrv)
                    out[x] = v[i * R + r];
                else if (x < zt)
                    out[x] = make_float2(0.f, 0.f);
            }
        }
    }
}

template <int N, int T, int PS, int WAVES, class Seq>
static hipError_t launch(const RcArgs& a, int fmt, hipStream_t s) {
    const size_t lds = sizeof(float2) * (size_t)(data_elems<N, PS>() + tw_lds_elems(N));
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    if (fmt == RAW_CF32)
        hipLaunchKernelGGL((k_rc_pad<RAW_CF32, N, T, PS, WAVES, Seq>), dim3(a.rows), dim3(T), lds, s, a);
    else if (fmt == RAW_U8)
        hipLaunchKernelGGL((k_rc_pad<RAW_U8, N, T, PS, WAVES, Seq>), dim3(a.rows), dim3(T), lds, s, a);
    else if (fmt == RAW_ERS_LINES)
        hipLaunchKernelGGL((k_rc_pad<RAW_ERS_LINES, N, T, PS, WAVES, Seq>), dim3(a.rows), dim3(T), lds, s, a);
    else if (fmt == RAW_ASAR_FBAQ4)
        hipLaunchKernelGGL((k_rc_pad<RAW_ASAR_FBAQ4, N, T, PS, WAVES, Seq>), dim3(a.rows), dim3(T), lds, s, a);
    else
        hipLaunchKernelGGL((k_rc_pad<RAW_I8, N, T, PS, WAVES, Seq>), dim3(a.rows), dim3(T), lds, s, a);
    return hipGetLastError();
}

}  // namespace rcpad

// Padded-layout variants per production length (see range_compress_pad_variants): 0 = the default.
int range_compress_pad_variants(int n) {
    switch (n) {
        case 6144: return 2;
        case 8192: return 1;
        default: return 0;
    }
}

bool launch_range_compress_pad(const RcArgs& a, int fmt, int variant, hipStream_t s, hipError_t* err) {
    using namespace rcpad;
    switch (a.plan.n) {
        case 6144:
            switch (variant) {
                case 1: *err = launch<6144, 512, 5, 6, RadixSeq<8, 8, 8, 12>>(a, fmt, s); return true;
                default: *err = launch<6144, 512, 4, 6, RadixSeq<8, 8, 8, 12>>(a, fmt, s); return true;
            }
        case 8192:
            *err = launch<8192, 512, 5, 4, RadixSeq<16, 8, 8, 8>>(a, fmt, s);
            return true;
        default:
            return false;
    }
}

}  // namespace asarf
