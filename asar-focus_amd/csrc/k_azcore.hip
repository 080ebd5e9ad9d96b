// Fused azimuth core: forward pass B of the azimuth FFT, 16-tap sinc RCMC, azimuth matched filter and the
// inverse pass B', one workgroup per (k1 group, 64-column tile), out of place (pass-A output -> pass-B' output).
//
// With the four-step split N = N1*N2, pass A leaves group k1's length-N2 transform inputs in the contiguous rows
// k1*N2 + m; pass B turns them into the bins k1 + N1*k2 (k2 < N2) of those rows. RCMC interpolates along range
// within one bin and the matched filter is pointwise, so a workgroup that transforms its k1 group for its 64
// columns plus a range halo can correct the bins in LDS and run the inverse length-N2 transform before anything
// returns to HBM: the spectrum makes one round trip to HBM instead of three (B out, RCMC in and out, B' in).
//   staged columns:    [x0 - 8, x0 + 80)   (forward transforms of the 8 + 64 + 16 staged columns)
//   produced columns:  [x0, x0 + 64)       (RCMC reads x + p - 8 .. x + p + 7, p = round(shift) <= kAzCoreReach)
// The per-pixel arithmetic is k_rcmc_tile's (k_rda.hip; without the two zero-weight end taps of the window) and the
// transforms are the Stockham passes of the column kernels, so the core reproduces the separate kernels up to the
// rounding of a different N1 x N2 split.
//
// With N2 = 84 the tile (84 x 89 complex samples) and the twiddle tables take about 60 KB of LDS.
// Phases (512 lanes; L = 4, 8 or 16 lanes per column transform, see core_lanes):
//   1. forward transforms of the 88 staged columns of the group's N2 rows, packed L lanes per column into
//      batches of at most 512 / L columns (LDS exchange through each column's own tile column), bins left in the
//      tile;
//   2. RCMC + matched filter: lane (wave w, column c) produces bins w, w + 8, ... of column c (a wave reads 64
//      consecutive tile columns of one bin row); results held in registers until every tap is read, then
//      written over the produced columns;
//   3. inverse transforms of the 64 produced columns (same packing), inverse inter-pass twiddle, store.
#include "interp.h"
#include "kernels.h"

namespace asarf {

constexpr int kAcThreads = 512;
constexpr int kAcOut = 64;                                    // produced columns per tile
constexpr int kAcLeft = 8;                                    // first produced column inside the tile
constexpr int kAcCols = kAcLeft + kAcOut + 8 + kAzCoreReach;  // staged columns per tile
constexpr int kAcPitch = kAcCols + 1;
static_assert(kAcOut == 64, "phase 2 maps one wave onto the produced columns");

// Lanes per column transform. Fewer lanes than the 16 of the column kernels: a length-N2 pass of radix R has only
// N2 / R butterflies (7 for the radix-12 pass of 84), so 16 lanes leave most of them idle, and every wave of the
// workgroup issues each pass. With L lanes the staged columns are packed L lanes each into as few waves as they
// fill (wave slots past the last column skip the transform), at most 512 / L columns per batch. Chosen per length
// from {4, 8, 16}: the fewest wave-issued butterfly elements over the forward (kAcCols columns) and inverse (kAcOut)
// transforms, with at most 32 elements per lane (64 VGPRs of transform state).
template <int N2, class Seq, int L>
constexpr long core_issue_cost() {
    long rounds = 0;
    for (int p = 0; p < Seq::count; ++p) rounds += (long)((N2 / Seq::at(p) + L - 1) / L) * Seq::at(p);
    long waves = 0;
    for (int cols : {kAcCols, kAcOut}) {
        const int cb = cols < kAcThreads / L ? cols : kAcThreads / L;
        for (int base = 0; base < cols; base += cb) {
            const int n = cols - base < cb ? cols - base : cb;
            waves += (n * L + 63) / 64;
        }
    }
    return waves * rounds;
}

template <int N2, class Seq>
constexpr int core_lanes() {
    int best = 16;
    long best_cost = core_issue_cost<N2, Seq, 16>();
    if (seq_max_elems<N2, Seq, 8>() <= 32 && core_issue_cost<N2, Seq, 8>() < best_cost) {
        best = 8;
        best_cost = core_issue_cost<N2, Seq, 8>();
    }
    if (seq_max_elems<N2, Seq, 4>() <= 32 && core_issue_cost<N2, Seq, 4>() < best_cost) best = 4;
    return best;
}

// Length-N2 transforms of the tile columns [C0, C0 + NC) (tile column c at tl[i * kAcPitch + c]), L lanes each,
// packed in batches of at most 512 / L columns: thread t of a batch of n columns takes column t % n, lane t / n.
// INV = false: forward, from global memory (load(c, m) = row m of tile column c) into the tile; INV = true:
// inverse, from the tile into global memory (store(c, n, value) = row n of tile column c). Every thread reaches
// every barrier; threads past the batch's columns skip the arithmetic.
template <bool INV, int N2, class Seq, int L, int C0, int NC, class Load, class Store>
__device__ __forceinline__ void core_transforms(float2* tl, const TwLds& tw2, Load&& load, Store&& store) {
    constexpr int E = seq_max_elems<N2, Seq, L>();
    constexpr int CB = NC < kAcThreads / L ? NC : kAcThreads / L;
    constexpr int NB = (NC + CB - 1) / CB;
    const int tid = threadIdx.x;
    float2 v[E];
    static_for<0, NB>([&](auto bc) __attribute__((always_inline)) {
        constexpr int base = decltype(bc)::value * CB;
        constexpr int n = NC - base < CB ? NC - base : CB;
        const bool act = tid < n * L;
        const int c = C0 + base + tid % n;
        const int lane = tid / n;
        auto lds_store = [&](int i, float2 val) { tl[i * kAcPitch + c] = val; };
        auto lds_load = [&](int i) { return tl[i * kAcPitch + c]; };
        auto src_load = [&](int m) { return load(c, m); };
        constexpr int R0 = INV ? Seq::rat(0) : Seq::at(0);
        if constexpr (INV) {
            if (act) pass_load<R0>(v, lane, L, N2, lds_load);
            __syncthreads();  // every element of the column is read before the first exchange overwrites it
        } else {
            if (act) pass_load<R0>(v, lane, L, N2, src_load);
        }
        if (act) pass_compute<R0, E, INV>(v, lane, L, N2, 1, tw2);
        static_for<1, Seq::count>([&](auto pc) __attribute__((always_inline)) {
            constexpr int p = decltype(pc)::value;
            constexpr int Rp = INV ? Seq::rat(p - 1) : Seq::at(p - 1);
            constexpr int R = INV ? Seq::rat(p) : Seq::at(p);
            constexpr int ns_prev = INV ? Seq::rprefix(p - 1) : Seq::prefix(p - 1);
            constexpr int ns = INV ? Seq::rprefix(p) : Seq::prefix(p);
            if (act) pass_store<Rp>(v, lane, L, N2, ns_prev, lds_store);
            __syncthreads();
            if (act) pass_load<R>(v, lane, L, N2, lds_load);
            __syncthreads();
            if (act) pass_compute<R, E, INV>(v, lane, L, N2, ns, tw2);
        });
        constexpr int RL = INV ? Seq::rat(Seq::count - 1) : Seq::at(Seq::count - 1);
        if (act)
            pass_visit<RL>(v, lane, L, N2, [&](int k, float2& val, bool ok) {
                if (ok) store(c, k, val);
            });
    });
}

template <int N2, class Seq>
__global__ __launch_bounds__(kAcThreads) void k_az_core(AzCoreArgs a) {
    static_assert(Seq::product() == N2, "radix sequence must factor N2");
    extern __shared__ float2 lds[];
    if (a.gate && *a.gate > kAzCoreReach) return;
    // XCD-aware work order: workgroups are dealt to the 8 XCDs round robin, so consecutive work items
    // (neighbouring tiles of one k1 group, whose staged halos overlap) are given to the same XCD and its L2
    const int per_xcd = gridDim.x / 8;
    const int w = (blockIdx.x % 8) * per_xcd + blockIdx.x / 8;
    if (w >= a.items) return;
    const int k1 = w / a.ntiles;
    const int x0 = a.col0 + (w % a.ntiles) * kAcOut;
    const int tid = threadIdx.x;
    float2* tl = lds;  // [N2][kAcPitch]
    const TwLds tw2 = tw_fill(lds + N2 * kAcPitch, a.tw2, N2, tid, kAcThreads);
    const TwLds twf = tw_fill(lds + N2 * kAcPitch + tw_lds_elems(N2), a.tw_full, a.nbins, tid, kAcThreads);
    constexpr int L = core_lanes<N2, Seq>();
    const long row0 = (long)k1 * N2;

    // ---- 1. forward pass B of the staged columns (columns outside the image are zero, never read as taps)
    core_transforms<false, N2, Seq, L, 0, kAcCols>(
        tl, tw2,
        [&](int c, int m) -> float2 {
            const int x = x0 - kAcLeft + c;
            return (x >= 0 && x < a.nrv) ? a.in[(row0 + m) * a.stride + x] : make_float2(0.f, 0.f);
        },
        [&](int c, int k, float2 val) { tl[k * kAcPitch + c] = val; });
    __syncthreads();

    // ---- 2. RCMC + matched filter of the produced columns (the arithmetic of k_rcmc_tile)
    {
        const int cw = tid & 63;
        const int wv = tid >> 6;
        const int xc = x0 + cw;
        const bool xok = xc < a.nrv;
        const ColConst cc = a.col[xok ? xc : a.nrv - 1];
        const double step = a.prf / (double)a.nbins;
        const double inv_step = (double)a.nbins / a.prf;
        const double half_bins = 0.5 * a.beta * (double)a.nbins;
        const double m_dc = rint(cc.dc_off * inv_step);
        const double half = 0.5 * a.prf;
        float ws[16];  // signed window (-1)^i w[i]
#pragma unroll
        for (int i = 0; i < 16; ++i) ws[i] = (i & 1) ? -a.window[i] : a.window[i];
        constexpr int NR = (N2 + 7) / 8;
        float2 res[NR];
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int m = wv + 8 * i;
            if (m >= N2) break;
            const int k = k1 + a.n1 * m;
            const double fb = (k < a.nbins / 2) ? (double)k * step : (double)(k - a.nbins) * step;
            const double fn = alias_about_dc(fb, cc.dc, a.prf);
            float2 val = make_float2(0.f, 0.f);
            // outside the azimuth band the matched filter zeroes the pixel: skip its interpolation
            if (!outside_band(fn, m_dc, inv_step, half_bins)) {
                const double shift = cc.shift_k * fn * fn;
                const double p = floor(shift + 0.5);  // round half away from zero: shift >= 0
                const int pi = (int)fmin(p, 4096.0);
                const int cx = xc + pi;
                if (!(cx - 8 < 0 || cx + 8 >= a.nrv)) {
                    const float d = (float)(shift - p);
                    // the gate guarantees pi <= kAzCoreReach; the clamp only keeps the taps inside the tile
                    const float2* tap = tl + m * kAcPitch + min(kAcLeft + cw + pi - 8, kAcCols - 16);
                    val = (d == 0.0f) ? tap[8] : sinc16_interior(tap, ws, d);
                }
                double fc = fb - cc.dc_off;
                if (fc > half)
                    fc -= a.prf;
                else if (fc < -half)
                    fc += a.prf;
                const double turns = cc.inv_ka * (0.5 * fc * fc + cc.dc_off * fb);
                const float fr = (float)(turns - rint(turns));
                float sn, cs;
                sincos_turns(fr, &sn, &cs);
                val = make_float2(val.x * cs + val.y * sn, val.y * cs - val.x * sn);
            }
            res[i] = val;
        }
        __syncthreads();  // every tap is read before the produced columns are overwritten
#pragma unroll
        for (int i = 0; i < NR; ++i) {
            const int m = wv + 8 * i;
            if (m >= N2) break;
            tl[m * kAcPitch + kAcLeft + cw] = res[i];
        }
    }
    __syncthreads();

    // ---- 3. inverse pass B' of the produced columns, inverse inter-pass twiddle, store
    core_transforms<true, N2, Seq, L, kAcLeft, kAcOut>(
        tl, tw2, [&](int, int) { return make_float2(0.f, 0.f); },
        [&](int c, int n, float2 val) {
            const int x = x0 + c - kAcLeft;
            if (x >= a.nrv) return;
            const float2 wn = twf(k1 * n);  // k1 * n < n1 * n2
            a.out[(row0 + n) * a.stride + x] = cmulc(val, wn);
        });
}

size_t az_core_lds_bytes(int n2, int nbins) {
    return sizeof(float2) * (size_t)(n2 * kAcPitch + tw_lds_elems(n2) + tw_lds_elems(nbins));
}

int az_core_tile_columns() { return kAcOut; }

template <int N2, class Seq>
static hipError_t core_launch(const AzCoreArgs& a, size_t lds, bool prepare, hipStream_t s) {
    if (prepare)
        return hipFuncSetAttribute((const void*)k_az_core<N2, Seq>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                   (int)lds);
    const unsigned grid = (unsigned)((a.items + 7) / 8 * 8);
    hipLaunchKernelGGL((k_az_core<N2, Seq>), dim3(grid), dim3(kAcThreads), lds, s, a);
    return hipGetLastError();
}

// Second-pass lengths with a compiled core and their radix sequences, in the planner's order of preference (the
// split of the azimuth FFT takes the first whose first-pass length has a compile-time column kernel). 84, 90, 96
// and 98 were measured on the production scenes; the others cover the azimuth lengths those four do not divide
// (powers of two, 3^a 5^b, ...). With 128 the tile takes about 92 KB of LDS (one workgroup per CU).
#define ASARF_AZ_CORE_LENGTHS(X)                                                                                  \
    X(84, 12, 7) X(90, 9, 10) X(98, 7, 14) X(96, 8, 12) X(80, 10, 8) X(100, 10, 10) X(72, 9, 8) X(105, 15, 7)     \
    X(81, 9, 9) X(112, 16, 7) X(64, 8, 8) X(120, 12, 10) X(75, 15, 5) X(126, 14, 9) X(128, 16, 8) X(125, 25, 5)

bool az_core_supported(int n2) {
    switch (n2) {
#define ASARF_AZ_LABEL(N, ...) case N:
        ASARF_AZ_CORE_LENGTHS(ASARF_AZ_LABEL)
#undef ASARF_AZ_LABEL
            return true;
        default:
            return false;
    }
}

int az_core_lengths(int* out, int capacity) {
    static const int lengths[] = {
#define ASARF_AZ_ITEM(N, ...) N,
        ASARF_AZ_CORE_LENGTHS(ASARF_AZ_ITEM)
#undef ASARF_AZ_ITEM
    };
    const int n = (int)(sizeof(lengths) / sizeof(lengths[0]));
    for (int i = 0; i < n && i < capacity; ++i) out[i] = lengths[i];
    return n;
}

static hipError_t core_dispatch(const AzCoreArgs& a, int n2, bool prepare, hipStream_t s) {
    const size_t lds = az_core_lds_bytes(n2, a.nbins);
    if (lds > max_lds_bytes()) return hipErrorInvalidValue;
    switch (n2) {
#define ASARF_AZ_LAUNCH(N, ...) \
    case N: return core_launch<N, RadixSeq<__VA_ARGS__>>(a, lds, prepare, s);
        ASARF_AZ_CORE_LENGTHS(ASARF_AZ_LAUNCH)
#undef ASARF_AZ_LAUNCH
        default: return hipErrorInvalidValue;
    }
}

// Raises the kernel's dynamic LDS limit (once per plan, outside any graph capture).
hipError_t az_core_prepare(int n2, int nbins) {
    AzCoreArgs a{};
    a.nbins = nbins;
    return core_dispatch(a, n2, true, nullptr);
}

hipError_t launch_az_core(const AzCoreArgs& a, int n2, hipStream_t s) {
    if (!a.in || !a.out || a.in == a.out || a.items <= 0) return hipErrorInvalidValue;
    // the interpolation skips taps 0 and 15 (sinc16_interior): their window weights must be zero
    if (!(std::fabs(a.window[0]) < 1e-12f && std::fabs(a.window[15]) < 1e-12f)) return hipErrorInvalidValue;
    return core_dispatch(a, n2, false, s);
}

}  // namespace asarf
