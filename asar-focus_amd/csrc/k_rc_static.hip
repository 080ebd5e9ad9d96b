// Range compression with compile-time transform plans for the production row lengths.
//
// Same algorithm as k_range_compress (correct on load, forward Stockham FFT, x conj(H)/n, inverse FFT, store the
// kept columns), but the radix sequence, the lane count, the LDS padding and the occupancy target are template
// parameters: no runtime radix dispatch and compile-time index arithmetic. A lane owns R * ceil((N/R)/T) elements
// in each pass (lanes beyond a pass's butterfly count compute clamped copies). Lengths without a specialisation
// use the generic kernel.
//   6144 (ASAR IM, 5616 + 521 samples), 6400 (ERS IM, 5616 + 704 samples), 8192 (split-precision ERS form);
//   several radix sequences and lane counts per length (the autotuner picks among them; see the table below).
#include "kernels.h"

namespace asarf {

template <typename T>
struct RawPairS;
template <>
struct RawPairS<uint8_t> {
    using type = uchar2;
};
template <>
struct RawPairS<int8_t> {
    using type = char2;
};

// LDS position of transform element i. PADSHIFT > 0: one padding element every 2^PADSHIFT. PADSHIFT == 0: no
// padding, element i sits at i ^ ((i >> 3) & 31), a permutation of every aligned 256-element block that spreads
// the strided Stockham stores (stride R at the first pass, runs of ns at the next) over the banks. Offsets that
// are multiples of 256 (n / R for the production radices) pass through unchanged, so loads at j + r n / R keep
// constant offsets.
template <int PADSHIFT>
__host__ __device__ constexpr int lds_pos(int i) {
    return PADSHIFT > 0 ? i + (i >> PADSHIFT) : i ^ ((i >> 3) & 31);
}

template <int N, int PADSHIFT>
__host__ __device__ constexpr int static_lds_elems() {
    static_assert(PADSHIFT > 0 || N % 256 == 0, "the swizzled layout permutes whole 256-element blocks");
    return PADSHIFT > 0 ? N + (N >> PADSHIFT) + 1 : N;
}

template <int FMT, int N, int T, int PADSHIFT, int WAVES, class Seq, bool STAGE = false, bool TW1 = false>
__global__ __launch_bounds__(T) __attribute__((amdgpu_waves_per_eu(WAVES))) void k_rc_static(RcArgs a) {
    static_assert(Seq::product() == N, "radix sequence must factor N");
    // elements per lane: N / T when every radix divides it, else the largest R * ceil((N / R) / T) (lanes
    // beyond a pass's butterfly count compute clamped copies and store nothing)
    constexpr int E = seq_max_elems<N, Seq, T>();
    // vote on a lane's first butterfly only where some pass has fewer butterflies than lanes
    constexpr bool V0 = [] {
        for (int p = 0; p < Seq::count; ++p)
            if (N / Seq::at(p) < T) return true;
        return false;
    }();
    extern __shared__ float2 lds[];
    const int row = blockIdx.x;
    const int lane = threadIdx.x;
    const TwLds tw_at = tw_fill(lds + static_lds_elems<N, PADSHIFT>(), a.plan.tw, N, lane, T);
    float2 v[E];
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
    constexpr bool LINES_FMT = FMT == RAW_ERS_LINES || FMT == RAW_ASAR_FBAQ4;
    const int shift = LINES_FMT ? a.lines.shift(row) : 0;
    const unsigned char* rec = static_cast<const unsigned char*>(a.in) + (LINES_FMT ? (long)row * a.lines.line_bytes : 0L);
    const int rec_len = FMT == RAW_ASAR_FBAQ4 ? fbaq4_record_length(rec) : 0;
    auto gl_load = [&](int x) -> float2 {
        if constexpr (FMT == RAW_CF32) {
            const float2* in = static_cast<const float2*>(a.in) + (long)row * a.in_stride;
            const float2 r = in[x];
            return correct(r.x, r.y);
        } else if constexpr (FMT == RAW_ERS_LINES) {
            // the line's sample shifted by the SWST; NaN where none lands becomes 0 (ZeroNaNs); padding is 0
            const int sx = x - shift;
            if (x >= nr_in || sx < 0 || sx >= a.lines.spl) return make_float2(0.f, 0.f);
            const uchar2 r = reinterpret_cast<const uchar2*>(rec)[sx];
            return correct((float)r.x, (float)r.y);
        } else if constexpr (FMT == RAW_ASAR_FBAQ4) {
            float2 r;
            if (x >= nr_in || !fbaq4_sample(rec, rec_len, a.lines.line_bytes, a.lines.lut, x - shift, &r))
                return make_float2(0.f, 0.f);
            return correct(r.x, r.y);
        } else {
            if (x >= nr_in) return make_float2(0.f, 0.f);
            using P = typename RawPairS<typename std::conditional<FMT == RAW_U8, uint8_t, int8_t>::type>::type;
            const P* in = static_cast<const P*>(a.in) + (long)row * a.in_stride;
            const P r = in[x];
            return correct((float)r.x, (float)r.y);
        }
    };
    auto lds_store = [&](int i, float2 x) { lds[lds_pos<PADSHIFT>(i)] = x; };
    auto lds_load = [&](int i) { return lds[lds_pos<PADSHIFT>(i)]; };

    // forward transform
    constexpr int R0 = Seq::at(0);
    if constexpr (FMT == RAW_ASAR_FBAQ4) {
        // the record decoded into LDS in natural order (phased loads), then the first pass reads LDS
        fbaq4_row_decode<N, T>(rec, a.lines.line_bytes, a.lines.lut, shift, nr_in, lane, correct, lds_store);
        __syncthreads();
        pass_load<R0, E, V0>(v, lane, T, N, lds_load);
        __syncthreads();  // every staged element is read before the passes overwrite LDS
    } else if constexpr (STAGE && (FMT == RAW_U8 || FMT == RAW_I8)) {
        // raw bytes with 16-byte loads (8 I/Q pairs per load) into LDS in natural order, then the first pass
        // reads LDS: one wide load instead of eight 2-byte loads per lane
        const uint8_t* rb = static_cast<const uint8_t*>(a.in) + (long)row * a.in_stride * 2;
        const int nbytes = nr_in * 2;
        if ((reinterpret_cast<uintptr_t>(rb) & 15) == 0) {
            for (int off = lane * 16; off < nbytes; off += T * 16) {
                if (off + 16 <= nbytes) {
                    const uint4 w = *reinterpret_cast<const uint4*>(rb + off);
                    const unsigned ws[4] = {w.x, w.y, w.z, w.w};
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
                        lds_store(off / 2 + k, correct(fi, fq));
                    }
                } else {
                    for (int x = off / 2; x < nr_in; ++x) lds_store(x, gl_load(x));
                }
            }
            for (int x = nr_in + lane; x < N; x += T) lds_store(x, make_float2(0.f, 0.f));
        } else {
            for (int x = lane; x < N; x += T) lds_store(x, gl_load(x));
        }
        __syncthreads();
        pass_load<R0, E, V0>(v, lane, T, N, lds_load);
        __syncthreads();  // every staged element is read before the passes overwrite LDS
    } else {
        pass_load<R0, E, V0>(v, lane, T, N, gl_load);
    }
    pass_compute<R0, E, false, V0, TW1>(v, lane, T, N, 1, tw_at);
    static_for<1, Seq::count>([&](auto pc) __attribute__((always_inline)) {
        constexpr int p = decltype(pc)::value;
        constexpr int Rp = Seq::at(p - 1), R = Seq::at(p);
        pass_store<Rp>(v, lane, T, N, Seq::prefix(p - 1), lds_store);
        __syncthreads();
        pass_load<R, E, V0>(v, lane, T, N, lds_load);
        __syncthreads();
        pass_compute<R, E, false, V0, TW1>(v, lane, T, N, Seq::prefix(p), tw_at);
    });
    // matched filter in the range frequency domain
    constexpr int RL = Seq::at(Seq::count - 1);
    const float2* h = a.hconj;
    auto spec_mul = [&](int k, float2& x, bool) { x = cmul(x, h[k]); };
    pass_visit<RL>(v, lane, T, N, spec_mul);
    __syncthreads();  // all LDS reads of the forward transform precede the inverse stores
    // inverse transform (radices in reverse order; its first pass consumes the registers as they are)
    pass_compute<RL, E, true, V0, TW1>(v, lane, T, N, 1, tw_at);
    static_for<1, Seq::count>([&](auto pc) __attribute__((always_inline)) {
        constexpr int p = decltype(pc)::value;
        constexpr int Rp = Seq::rat(p - 1), R = Seq::rat(p);
        pass_store<Rp>(v, lane, T, N, Seq::rprefix(p - 1), lds_store);
        __syncthreads();
        pass_load<R, E, V0>(v, lane, T, N, lds_load);
        __syncthreads();
        pass_compute<R, E, true, V0, TW1>(v, lane, T, N, Seq::rprefix(p), tw_at);
    });
    float2* out = a.out + (long)row * a.out_stride;
    const int nrv = a.nrv, zt = a.zero_to;
    auto out_write = [&](int x, float2& val, bool ok) {
        if (!ok) return;
        if (x < nrv)
            out[x] = val;
        else if (x < zt)
            out[x] = make_float2(0.f, 0.f);
    };
    constexpr int RIL = Seq::rat(Seq::count - 1);
    pass_visit<RIL>(v, lane, T, N, out_write);
}

template <int N, int T, int PADSHIFT, int WAVES, class Seq, bool STAGE = false, bool TW1 = false>
static hipError_t launch_static(const RcArgs& a, int fmt, hipStream_t s) {
    const size_t lds = sizeof(float2) * (size_t)(static_lds_elems<N, PADSHIFT>() + tw_lds_elems(N));
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    if (fmt == RAW_CF32)
        hipLaunchKernelGGL((k_rc_static<RAW_CF32, N, T, PADSHIFT, WAVES, Seq, STAGE, TW1>), dim3(a.rows), dim3(T), lds,
                           s, a);
    else if (fmt == RAW_U8)
        hipLaunchKernelGGL((k_rc_static<RAW_U8, N, T, PADSHIFT, WAVES, Seq, STAGE, TW1>), dim3(a.rows), dim3(T), lds,
                           s, a);
    else if (fmt == RAW_ERS_LINES)
        hipLaunchKernelGGL((k_rc_static<RAW_ERS_LINES, N, T, PADSHIFT, WAVES, Seq, STAGE, TW1>), dim3(a.rows), dim3(T),
                           lds, s, a);
    else if (fmt == RAW_ASAR_FBAQ4)
        hipLaunchKernelGGL((k_rc_static<RAW_ASAR_FBAQ4, N, T, PADSHIFT, WAVES, Seq, STAGE, TW1>), dim3(a.rows),
                           dim3(T), lds, s, a);
    else
        hipLaunchKernelGGL((k_rc_static<RAW_I8, N, T, PADSHIFT, WAVES, Seq, STAGE, TW1>), dim3(a.rows), dim3(T), lds,
                           s, a);
    return hipGetLastError();
}

// Compile-time variants per production length (radix sequence, lanes per line); the autotuner times them and
// records the fastest (RcArgs::static_variant). Variant 0 is the default.
static int stockham_variants(int n) {
    switch (n) {
        case 6144: return 8;
        case 6400: return 6;
        case 8192: return 6;
        default: return 0;
    }
}

// Variants past the Stockham ones are the padded-layout kernels of k_rc_pad.hip.
int range_compress_static_variants(int n) { return stockham_variants(n) + range_compress_pad_variants(n); }

// Returns true (and the launch status in *err) when a specialisation exists for the plan's length. Variant 0 is
// the fastest measured on the ASAR/ERS scenes (512 lanes with a 4-wave occupancy target keeps 16 waves per CU;
// for 6144 the raw bytes are also staged through LDS with 16-byte loads).
bool launch_range_compress_static(const RcArgs& a, int fmt, hipStream_t s, hipError_t* err) {
    int v = a.static_variant;
    // 6144: variant 0 is the padded-layout kernel (measured fastest on the ASAR scene); the Stockham kernel that was
    // the default before it takes index 8
    if (a.plan.n == 6144) v = (v == 0) ? stockham_variants(6144) : (v == stockham_variants(6144)) ? 0 : v;
    if (v >= stockham_variants(a.plan.n) && v > 0)
        return launch_range_compress_pad(a, fmt, v - stockham_variants(a.plan.n), s, err);
    switch (a.plan.n) {
        case 6144:
            switch (v) {
                case 1: *err = launch_static<6144, 256, 5, 2, RadixSeq<8, 8, 8, 12>>(a, fmt, s); return true;
                case 2: *err = launch_static<6144, 384, 5, 2, RadixSeq<24, 16, 16>>(a, fmt, s); return true;
                case 3: *err = launch_static<6144, 512, 5, 4, RadixSeq<12, 8, 8, 8>>(a, fmt, s); return true;
                case 4: *err = launch_static<6144, 512, 5, 6, RadixSeq<8, 8, 8, 12>>(a, fmt, s); return true;
                case 5: *err = launch_static<6144, 512, 5, 4, RadixSeq<8, 8, 8, 12>, true>(a, fmt, s); return true;
                case 6: *err = launch_static<6144, 512, 5, 6, RadixSeq<8, 8, 8, 12>, true>(a, fmt, s); return true;
                case 7: *err = launch_static<6144, 512, 0, 6, RadixSeq<8, 8, 8, 12>, true>(a, fmt, s); return true;
                default: *err = launch_static<6144, 512, 0, 6, RadixSeq<8, 8, 8, 12>, true, true>(a, fmt, s); return true;
            }
        case 6400:
            switch (v) {
                case 1: *err = launch_static<6400, 512, 5, 4, RadixSeq<20, 20, 16>>(a, fmt, s); return true;
                case 2: *err = launch_static<6400, 400, 5, 2, RadixSeq<25, 16, 16>>(a, fmt, s); return true;
                case 3: *err = launch_static<6400, 320, 7, 3, RadixSeq<20, 20, 4, 4>>(a, fmt, s); return true;
                case 4: *err = launch_static<6400, 512, 5, 4, RadixSeq<4, 4, 20, 20>>(a, fmt, s); return true;
                case 5: *err = launch_static<6400, 512, 5, 4, RadixSeq<16, 20, 20>, true>(a, fmt, s); return true;
                default: *err = launch_static<6400, 512, 5, 4, RadixSeq<16, 20, 20>>(a, fmt, s); return true;
            }
        case 8192:
            switch (v) {
                case 1: *err = launch_static<8192, 256, 5, 2, RadixSeq<8, 8, 8, 16>>(a, fmt, s); return true;
                case 2: *err = launch_static<8192, 512, 5, 2, RadixSeq<16, 16, 16, 2>>(a, fmt, s); return true;
                case 3: *err = launch_static<8192, 512, 5, 2, RadixSeq<8, 8, 8, 16>>(a, fmt, s); return true;
                case 4: *err = launch_static<8192, 512, 5, 4, RadixSeq<8, 8, 8, 16>>(a, fmt, s); return true;
                case 5: *err = launch_static<8192, 512, 5, 4, RadixSeq<16, 8, 8, 8>, true>(a, fmt, s); return true;
                default: *err = launch_static<8192, 512, 5, 4, RadixSeq<16, 8, 8, 8>>(a, fmt, s); return true;
            }
        default:
            return false;
    }
}

}  // namespace asarf
