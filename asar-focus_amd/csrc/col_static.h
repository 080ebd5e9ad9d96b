// Column transforms with compile-time radix sequences (the kernel template and its launcher). The sub-transform
// lengths of the four-step azimuth FFT that have such a kernel are every 2^a 3^b 5^c 7^d from 16 to 400, listed
// below with their radix sequences and split over six compilation units (k_col_static*.hip) so they build in
// parallel. Same workgroup layout and global access as k_col_fft (16 lanes per column, 16, 32 or 48 columns, LDS
// rows of COLS + 1 elements), but two large radices (composite DFTs in registers) replace three or four
// runtime-dispatched passes: one LDS exchange per transform instead of two or three, and compile-time index
// arithmetic.
//
// Radix sequences: the fewest passes (two radices of at most 32; three for 245 and 343), then the fewest elements
// per lane (max over passes of R * ceil(N / R / 16)), then the most balanced pair. The lengths of the production
// scenes keep the sequences measured on them: 168 = 12 x 14, 180 = 12 x 15 (ASAR scene, 30240 = 168 x 180),
// 84 = 12 x 7, 49 = 7 x 7 (ERS block, 4116), 112 = 16 x 7, 147 = 21 x 7 (high-migration scene, 16464), 128 = 16 x 8,
// 256 = 16 x 16 (32768), 360 = 15 x 24, 196 = 14 x 14, 336 = 12 x 28 (first passes of the fused azimuth core's
// splits).
#pragma once
#include "kernels.h"

#define ASARF_COL_LENGTHS_0(X) \
    X(18, 6, 3) X(20, 5, 4) X(35, 7, 5) X(50, 10, 5) X(81, 9, 9) X(90, 10, 9) X(147, 21, 7) X(168, 12, 14) \
    X(180, 12, 15) X(189, 21, 9) X(240, 16, 15) X(320, 20, 16) X(324, 27, 12) X(343, 7, 7, 7)
#define ASARF_COL_LENGTHS_1(X) \
    X(24, 6, 4) X(25, 5, 5) X(32, 8, 4) X(45, 9, 5) X(48, 8, 6) X(125, 25, 5) X(140, 14, 10) X(144, 12, 12) \
    X(150, 15, 10) X(162, 18, 9) X(224, 16, 14) X(250, 25, 10) X(350, 25, 14) X(384, 24, 16)
#define ASARF_COL_LENGTHS_2(X) \
    X(16, 4, 4) X(30, 6, 5) X(54, 9, 6) X(56, 8, 7) X(70, 10, 7) X(96, 12, 8) X(98, 14, 7) X(105, 15, 7) \
    X(112, 16, 7) X(160, 16, 10) X(175, 25, 7) X(300, 25, 12) X(378, 27, 14) X(400, 25, 16)
#define ASARF_COL_LENGTHS_3(X) \
    X(36, 6, 6) X(63, 9, 7) X(64, 8, 8) X(80, 10, 8) X(108, 12, 9) X(126, 14, 9) X(128, 16, 8) X(135, 15, 9) \
    X(192, 16, 12) X(280, 20, 14) X(288, 24, 12) X(315, 21, 15) X(392, 28, 14)
#define ASARF_COL_LENGTHS_4(X) \
    X(21, 7, 3) X(40, 8, 5) X(42, 7, 6) X(60, 10, 6) X(75, 15, 5) X(100, 10, 10) X(196, 14, 14) X(210, 15, 14) \
    X(216, 18, 12) X(243, 27, 9) X(245, 7, 5, 7) X(294, 21, 14) X(360, 15, 24)
#define ASARF_COL_LENGTHS_5(X) \
    X(27, 9, 3) X(28, 7, 4) X(49, 7, 7) X(72, 9, 8) X(84, 12, 7) X(120, 12, 10) X(200, 20, 10) X(225, 15, 15) \
    X(252, 21, 12) X(256, 16, 16) X(270, 27, 10) X(336, 12, 28) X(375, 25, 15)

// case label of a listed length (for switch statements over a list)
#define ASARF_COL_LABEL(N, ...) case N:
// launch of a listed length inside a dispatch switch (a, inverse, s, err in scope)
#define ASARF_COL_LAUNCH(N, ...) \
    case N: *err = launch_cs<N, RadixSeq<__VA_ARGS__>>(a, inverse, s); return true;

namespace asarf {


// Lanes per column. (32 lanes for the long first-pass lengths 336 and 360 halve the register block, but the
// compiler then schedules those kernels into 188-239 VGPRs, or spills under a 4-wave bound: 16 stays.)
template <int N>
constexpr int col_lanes() {
    return 16;
}

// One work item: the COLS-column tile `tile` of transform group `g` (w = g * ntiles + tile).
template <bool INV, int N, class Seq, int COLS>
__device__ __forceinline__ void col_static_item(const ColFftArgs& a, float2* lds, int w) {
    static_assert(Seq::product() == N, "radix sequence must factor N");
    constexpr int L = col_lanes<N>();  // lanes per column
    constexpr int T = L * COLS;   // threads per workgroup
    constexpr int E = seq_max_elems<N, Seq, L>();
    constexpr int P = COLS + 1;
    const int ntiles = (a.ncols + COLS - 1) / COLS;
    const int tile = w % ntiles;
    const int g = w / ntiles;
    const int b = threadIdx.x % COLS;
    const int lane = threadIdx.x / COLS;
    const int col = tile * COLS + b;
    const bool col_ok = col < a.ncols;
    const TwLds tw_at = tw_fill(lds + N * P, a.sub.tw, N, threadIdx.x, T);
    // inter-pass twiddles W_n^(g k) of the full length n = groups * N from a second two-level table in LDS
    const int n_full = a.groups * N;
    TwLds tw_full{};
    if (a.tw_full) tw_full = tw_fill(lds + N * P + tw_lds_elems(N), a.tw_full, n_full, threadIdx.x, T);
    float2 v[E];
    const float2* in = a.in + (col_ok ? col : 0);
    const long in_base = (long)g * a.in_g;
    const int in_rows = a.in_rows;
    auto gl_load = [&](int m) -> float2 {
        const long r = in_base + (long)m * a.in_m;
        if (!col_ok || r >= in_rows) return make_float2(0.f, 0.f);
        return in[r * a.in_stride];
    };
    auto lds_store = [&](int i, float2 x) { lds[i * P + b] = x; };
    auto lds_load = [&](int i) { return lds[i * P + b]; };
    constexpr int R0 = INV ? Seq::rat(0) : Seq::at(0);
    pass_load<R0>(v, lane, L, N, gl_load);
    __syncthreads();  // in-place passes: every element of the group is read before any is written
    pass_compute<R0, E, INV>(v, lane, L, N, 1, tw_at);
    static_for<1, Seq::count>([&](auto pc) __attribute__((always_inline)) {
        constexpr int p = decltype(pc)::value;
        constexpr int Rp = INV ? Seq::rat(p - 1) : Seq::at(p - 1);
        constexpr int R = INV ? Seq::rat(p) : Seq::at(p);
        constexpr int ns_prev = INV ? Seq::rprefix(p - 1) : Seq::prefix(p - 1);
        constexpr int ns = INV ? Seq::rprefix(p) : Seq::prefix(p);
        pass_store<Rp>(v, lane, L, N, ns_prev, lds_store);
        __syncthreads();
        pass_load<R>(v, lane, L, N, lds_load);
        __syncthreads();
        pass_compute<R, E, INV>(v, lane, L, N, ns, tw_at);
    });
    float2* out = a.out + (col_ok ? col : 0);
    const long out_base = (long)g * a.out_g;
    const float sc = a.scale;
    const bool twf = a.tw_full != nullptr;
    auto out_write = [&](int k, float2& x, bool ok) {
        const long r = out_base + (long)k * a.out_m;
        if (!ok || !col_ok || r >= a.out_rows) return;
        float2 y = x;
        if (twf) {
            const float2 w = tw_full(g * k);  // g * k < n_full
            y = INV ? cmulc(y, w) : cmul(y, w);
        }
        out[r * a.out_stride] = cscale(y, sc);
    };
    constexpr int RL = INV ? Seq::rat(Seq::count - 1) : Seq::at(Seq::count - 1);
    pass_visit<RL>(v, lane, L, N, out_write);
}

// LOOP = false: one work item per workgroup (the grid covers every item). LOOP = true: a gated fallback (it works
// only when the fused azimuth core does not), launched on a bounded grid whose workgroups take the items in turn,
// so that skipping it costs few workgroup launches. Kept as two kernels: the loop costs the ungated passes 20 % (the
// 360-point pass A: 0.57 -> 0.71 ms).
template <bool INV, int N, class Seq, int COLS, bool LOOP>
__global__ __launch_bounds__(col_lanes<N>() * COLS) void k_col_static(ColFftArgs a) {
    extern __shared__ float2 lds[];
    if (a.fused_gate && *a.fused_gate <= kAzCoreReach) return;
    if constexpr (LOOP) {
        const int items = (a.ncols + COLS - 1) / COLS * a.groups;
        for (int w = blockIdx.x; w < items; w += gridDim.x) {
            col_static_item<INV, N, Seq, COLS>(a, lds, w);
            __syncthreads();  // the next item refills LDS
        }
    } else {
        col_static_item<INV, N, Seq, COLS>(a, lds, blockIdx.x);
    }
}


template <int N, class Seq, int COLS>
inline hipError_t launch_cs_w(const ColFftArgs& a, bool inverse, hipStream_t s) {
    const size_t lds =
        sizeof(float2) * (size_t)(N * (COLS + 1) + tw_lds_elems(N) + (a.tw_full ? tw_lds_elems(a.groups * N) : 0));
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    const long ntiles = (a.ncols + COLS - 1) / COLS;
    const long items = ntiles * a.groups;
    const dim3 block(col_lanes<N>() * COLS);
    // gated fallbacks of the contiguous passes (16 columns wide) take the bounded-grid loop; any other launch
    // covers every item (a gated one still exits at once, on a full grid)
    if constexpr (COLS == 16) {
        if (a.fused_gate) {
            const dim3 grid((unsigned)std::min(items, 1024L));
            if (inverse)
                hipLaunchKernelGGL((k_col_static<true, N, Seq, COLS, true>), grid, block, lds, s, a);
            else
                hipLaunchKernelGGL((k_col_static<false, N, Seq, COLS, true>), grid, block, lds, s, a);
            return hipGetLastError();
        }
    }
    {
        const dim3 grid((unsigned)items);
        if (inverse)
            hipLaunchKernelGGL((k_col_static<true, N, Seq, COLS, false>), grid, block, lds, s, a);
        else
            hipLaunchKernelGGL((k_col_static<false, N, Seq, COLS, false>), grid, block, lds, s, a);
    }
    return hipGetLastError();
}

// Columns per workgroup: ColFftArgs::col_width, or by access pattern when 0 (32 where the transform's elements are
// strided rows, as in passes A and A' of the four-step split, so each row access is 256 contiguous bytes; 16 for
// contiguous sub-transforms, whose 16-column tiles measured faster). The strided passes stay at least 32 wide even
// where the tile leaves room for one workgroup per CU (N = 336, 360: 0.57 ms against 1.2 ms for 16-column tiles
// on the 30240-line azimuth transform). Lengths whose 32-column tile admits only one workgroup per CU (N > 310,
// tile > 80 KB) take 48 columns where the 48-column tile fits: still one workgroup, but 12 waves instead of 8 (the
// 360-point passes of the 30240-line transform: A 0.598 -> 0.574 ms, A' 0.545 -> 0.524 ms; 40 columns: slower).
template <int N, class Seq>
inline hipError_t launch_cs(const ColFftArgs& a, bool inverse, hipStream_t s) {
    if constexpr ((size_t)N * 33 * sizeof(float2) > 80 * 1024) {
        const size_t lds48 = sizeof(float2) * (size_t)(N * 49 + tw_lds_elems(N) +
                                                       (a.tw_full ? tw_lds_elems(a.groups * N) : 0));
        if (a.col_width == 0 && a.in_m > 1 && lds48 <= max_lds_bytes()) return launch_cs_w<N, Seq, 48>(a, inverse, s);
    }
    const bool wide_fits = sizeof(float2) * (size_t)N * 33 <= max_lds_bytes();
    const int w = (a.col_width == 16 || a.col_width == 32) ? a.col_width : (a.in_m > 1 && wide_fits ? 32 : 16);
    return w == 32 ? launch_cs_w<N, Seq, 32>(a, inverse, s) : launch_cs_w<N, Seq, 16>(a, inverse, s);
}

// The per-unit dispatchers: true (and the launch status in *err) when the length is in that unit's list.
bool launch_col_static_0(const ColFftArgs& a, bool inverse, hipStream_t s, hipError_t* err);
bool launch_col_static_1(const ColFftArgs& a, bool inverse, hipStream_t s, hipError_t* err);
bool launch_col_static_2(const ColFftArgs& a, bool inverse, hipStream_t s, hipError_t* err);
bool launch_col_static_3(const ColFftArgs& a, bool inverse, hipStream_t s, hipError_t* err);
bool launch_col_static_4(const ColFftArgs& a, bool inverse, hipStream_t s, hipError_t* err);
bool launch_col_static_5(const ColFftArgs& a, bool inverse, hipStream_t s, hipError_t* err);

}  // namespace asarf
