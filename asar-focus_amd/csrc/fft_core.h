// Device-side FFT building blocks for gfx950: compile-time small DFTs held in registers and a Stockham pass
// engine that runs mixed-radix transforms over LDS with a runtime radix sequence.
//
// Stockham formulation (self-sorting, natural-order output): a pass of radix R after a prefix of passes whose
// radices multiply to Ns treats butterfly j (0 <= j < N/R) as
//   inputs   x[j + r*N/R],                         r = 0..R-1
//   twiddle  x_r *= W_N^{r*(j%Ns)*N/(Ns*R)}         (W_N = exp(-2 pi i / N))
//   outputs  y[(j/Ns)*Ns*R + j%Ns + r*Ns] = DFT_R(x)_r
// The first pass of a transform reads from global memory, the last pass leaves its outputs in registers at
// positions j + r*N/R, so producers and consumers fuse into the transform without an LDS round trip.
#pragma once

#include <hip/hip_runtime.h>

#include <type_traits>

namespace asarf {

// ------------------------------------------------------------------------------------------------ complex ----
__device__ __forceinline__ float2 cadd(float2 a, float2 b) { return make_float2(a.x + b.x, a.y + b.y); }
__device__ __forceinline__ float2 csub(float2 a, float2 b) { return make_float2(a.x - b.x, a.y - b.y); }
__device__ __forceinline__ float2 cmul(float2 a, float2 b) {
    return make_float2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}
// a * conj(b)
__device__ __forceinline__ float2 cmulc(float2 a, float2 b) {
    return make_float2(a.x * b.x + a.y * b.y, a.y * b.x - a.x * b.y);
}
__device__ __forceinline__ float2 cscale(float2 a, float s) { return make_float2(a.x * s, a.y * s); }
__device__ __forceinline__ float2 cconj(float2 a) { return make_float2(a.x, -a.y); }
// Multiply by -i (forward direction) or +i (inverse direction).
template <bool INV>
__device__ __forceinline__ float2 mul_mj(float2 a) {
    return INV ? make_float2(-a.y, a.x) : make_float2(a.y, -a.x);
}

// --------------------------------------------------------------------------------------- constexpr trig ----
constexpr double kPi = 3.14159265358979323846264338327950288;

constexpr double cx_sin_poly(double x) {
    // Taylor series on |x| <= pi/2; 16 terms are far below double rounding there.
    double term = x, sum = x;
    for (int k = 1; k < 16; ++k) {
        term *= -x * x / ((2.0 * k) * (2.0 * k + 1.0));
        sum += term;
    }
    return sum;
}
constexpr double cx_sin(double x) {
    while (x > kPi) x -= 2.0 * kPi;
    while (x < -kPi) x += 2.0 * kPi;
    if (x > kPi / 2) x = kPi - x;
    if (x < -kPi / 2) x = -kPi - x;
    return cx_sin_poly(x);
}
constexpr double cx_cos(double x) { return cx_sin(x + kPi / 2); }

// W_R^m in the transform direction: exp(-+2 pi i m / R).
template <int R, int M, bool INV>
struct Tw {
    static constexpr double ang = (INV ? 2.0 : -2.0) * kPi * (double)(M % R) / (double)R;
    static constexpr float re = (float)cx_cos(ang);
    static constexpr float im = (float)cx_sin(ang);
};

template <int B, int E, class F>
__device__ __forceinline__ void static_for(F&& f) {
    if constexpr (B < E) {
        f(std::integral_constant<int, B>{});
        static_for<B + 1, E>(f);
    }
}

constexpr int smallest_factor(int n) {
    for (int f = 2; f * f <= n; ++f)
        if (n % f == 0) return f;
    return n;
}
constexpr bool is_prime(int n) { return n > 1 && smallest_factor(n) == n; }

// --------------------------------------------------------------------------------------------- small DFTs ----
// In-register DFT of size R over v[0..R-1].
template <int R, bool INV, class Enable = void>
struct Dft;

template <bool INV>
struct Dft<1, INV> {
    static __device__ __forceinline__ void run(float2*) {}
};

template <bool INV>
struct Dft<2, INV> {
    static __device__ __forceinline__ void run(float2* v) {
        const float2 a = v[0], b = v[1];
        v[0] = cadd(a, b);
        v[1] = csub(a, b);
    }
};

template <bool INV>
struct Dft<4, INV> {
    static __device__ __forceinline__ void run(float2* v) {
        const float2 a0 = cadd(v[0], v[2]), a1 = csub(v[0], v[2]);
        const float2 b0 = cadd(v[1], v[3]), b1 = mul_mj<INV>(csub(v[1], v[3]));
        v[0] = cadd(a0, b0);
        v[2] = csub(a0, b0);
        v[1] = cadd(a1, b1);
        v[3] = csub(a1, b1);
    }
};

// Odd prime P: symmetric/antisymmetric pairs, X_m = v0 + sum_k c_km a_k -+ i sum_k s_km b_k.
template <int P, bool INV>
struct Dft<P, INV, typename std::enable_if<(P > 2) && is_prime(P)>::type> {
    static __device__ __forceinline__ void run(float2* v) {
        constexpr int H = (P - 1) / 2;
        float2 a[H], b[H];
        float2 x0 = v[0];
        float2 sum = x0;
        static_for<1, H + 1>([&](auto kk) {
            constexpr int k = decltype(kk)::value;
            a[k - 1] = cadd(v[k], v[P - k]);
            b[k - 1] = csub(v[k], v[P - k]);
            sum = cadd(sum, a[k - 1]);
        });
        float2 out[P];
        out[0] = sum;
        static_for<1, H + 1>([&](auto mm) {
            constexpr int m = decltype(mm)::value;
            float2 re = x0, im = make_float2(0.f, 0.f);
            static_for<1, H + 1>([&](auto kk) {
                constexpr int k = decltype(kk)::value;
                constexpr float c = (float)cx_cos(2.0 * kPi * (double)((k * m) % P) / (double)P);
                constexpr float s = (float)cx_sin(2.0 * kPi * (double)((k * m) % P) / (double)P);
                re.x += c * a[k - 1].x;
                re.y += c * a[k - 1].y;
                im.x += s * b[k - 1].x;
                im.y += s * b[k - 1].y;
            });
            // j*im = (-im.y, im.x); forward: X_m = re - j*im, X_{P-m} = re + j*im
            const float2 jim = make_float2(-im.y, im.x);
            if (INV) {
                out[m] = cadd(re, jim);
                out[P - m] = csub(re, jim);
            } else {
                out[m] = csub(re, jim);
                out[P - m] = cadd(re, jim);
            }
        });
        static_for<0, P>([&](auto ii) {
            constexpr int i = decltype(ii)::value;
            v[i] = out[i];
        });
    }
};

// Composite R = F*M (F = smallest factor, except R = 4^k uses F = 4): decimation in time,
// X[k1 + M k2] = sum_{n1<F} W_F^{n1 k2} W_R^{n1 k1} DFT_M(x[F n2 + n1])[k1].
template <int R>
constexpr int composite_factor() {
    return (R % 4 == 0 && R > 4) ? 4 : smallest_factor(R);
}

template <int R, bool INV>
struct Dft<R, INV, typename std::enable_if<(R > 4) && !is_prime(R)>::type> {
    static __device__ __forceinline__ void run(float2* v) {
        constexpr int F = composite_factor<R>();
        constexpr int M = R / F;
        float2 y[F][M];
        static_for<0, F>([&](auto n1c) {
            constexpr int n1 = decltype(n1c)::value;
            static_for<0, M>([&](auto n2c) {
                constexpr int n2 = decltype(n2c)::value;
                y[n1][n2] = v[F * n2 + n1];
            });
            Dft<M, INV>::run(y[n1]);
            static_for<1, M>([&](auto k1c) {
                constexpr int k1 = decltype(k1c)::value;
                if constexpr (n1 * k1 != 0) {
                    constexpr int e = (n1 * k1) % R;
                    if constexpr (e * 4 == R) {
                        y[n1][k1] = mul_mj<INV>(y[n1][k1]);
                    } else {
                        const float2 w = make_float2(Tw<R, e, INV>::re, Tw<R, e, INV>::im);
                        y[n1][k1] = cmul(y[n1][k1], w);
                    }
                }
            });
        });
        static_for<0, M>([&](auto k1c) {
            constexpr int k1 = decltype(k1c)::value;
            float2 u[F];
            static_for<0, F>([&](auto n1c) {
                constexpr int n1 = decltype(n1c)::value;
                u[n1] = y[n1][k1];
            });
            Dft<F, INV>::run(u);
            static_for<0, F>([&](auto k2c) {
                constexpr int k2 = decltype(k2c)::value;
                v[k1 + M * k2] = u[k2];
            });
        });
    }
};

// ----------------------------------------------------------------------------------------- pass engine ----
// Two-level twiddle table in LDS: W_n^e = hi[e >> bits] * lo[e & (2^bits - 1)], about 2 sqrt(n) entries, so every
// twiddle is two LDS reads and one complex multiply instead of a global-memory read.
struct TwLds {
    const float2* hi;
    const float2* lo;
    int bits;
    int mask;
    __device__ __forceinline__ float2 operator()(int e) const { return cmul(hi[e >> bits], lo[e & mask]); }
};

__host__ __device__ inline int tw_bits(int n) {
    int b = 0;
    while ((1 << (2 * b)) < n) ++b;
    return b;
}
__host__ __device__ inline int tw_lds_elems(int n) {
    const int b = tw_bits(n);
    return ((n - 1) >> b) + 1 + (1 << b);
}

// Fills the two-level table from the global table W_n^m (m < n) and returns its accessor. Caller syncs.
__device__ __forceinline__ TwLds tw_fill(float2* dst, const float2* tw_global, int n, int tid, int nthreads) {
    const int b = tw_bits(n);
    const int nhi = ((n - 1) >> b) + 1;
    const int nlo = 1 << b;
    for (int i = tid; i < nhi + nlo; i += nthreads) {
        const int m = i < nhi ? (i << b) : (i - nhi);
        dst[i] = tw_global[m < n ? m : 0];
    }
    return TwLds{dst, dst + nhi, b, nlo - 1};
}

// Radices compiled into the runtime dispatch. The host planner only emits these.
#ifndef ASARF_FOR_EACH_RADIX
#define ASARF_FOR_EACH_RADIX(X) X(2) X(3) X(4) X(5) X(7) X(8) X(16)
#endif

// Compile-time radix sequence of a transform (forward order; the inverse runs it reversed).
template <int... Rs>
struct RadixSeq {
    static constexpr int count = sizeof...(Rs);
    static constexpr int at(int p) {
        constexpr int a[] = {Rs...};
        return a[p];
    }
    static constexpr int rat(int p) { return at(count - 1 - p); }
    static constexpr int prefix(int p) {
        int x = 1;
        for (int i = 0; i < p; ++i) x *= at(i);
        return x;
    }
    static constexpr int rprefix(int p) {
        int x = 1;
        for (int i = 0; i < p; ++i) x *= rat(i);
        return x;
    }
    static constexpr int product() { return prefix(count); }
};


// Elements per lane: the largest R * ceil((N / R) / lanes) over the passes.
template <int N, class Seq, int LANES>
constexpr int seq_max_elems() {
    int m = 0;
    for (int p = 0; p < Seq::count; ++p) {
        const int r = Seq::at(p);
        const int nb = (N / r + LANES - 1) / LANES;
        m = nb * r > m ? nb * r : m;
    }
    return m;
}

constexpr int kMaxPasses = 8;

struct FftPlanDev {
    int n;                      // transform length
    int npass;                  // number of Stockham passes
    unsigned long long radices; // forward radix of pass p in bits [5p, 5p+5); the inverse runs them in reverse
    const float2* tw;           // W_n^m, m in [0, n), FP64-accurate values rounded to float
};

__host__ __device__ inline int plan_radix(const FftPlanDev& p, int pass) {
    return (int)((p.radices >> (5 * pass)) & 31ull);
}

// Load phase: v[i*R + r] = src(j + r*n/R) for this thread's butterflies j = s + i*S. Lanes past the last
// butterfly load the last butterfly again: the register array is written unconditionally, which keeps it in
// registers (a predicated write invites the compiler to merge writes with a computed index and demote the array).
// A butterfly that is past the end for every lane of the wave is skipped (a wave vote and a scalar branch)
// instead of computed as a clamped copy; its registers keep stale values that no store or visit uses.
// VOTE0 = false: the caller guarantees that the first butterfly of every lane is valid (n / R >= S), and the
// vote is only taken for the later ones.
__device__ __forceinline__ bool wave_any(bool c) { return __builtin_amdgcn_ballot_w64(c) != 0; }

template <int R, int MAXE, bool VOTE0 = true, class Src>
__device__ __forceinline__ void pass_load(float2 (&v)[MAXE], int s, int S, int n, Src&& src) {
    constexpr int NB = MAXE / R;
    const int nr = n / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        if ((VOTE0 || i > 0) && !wave_any(s + i * S < nr)) break;
        const int j = min(s + i * S, nr - 1);
#pragma unroll
        for (int r = 0; r < R; ++r) v[i * R + r] = src(j + r * nr);
    }
}

// j / ns and j % ns without integer division: q = trunc((j + 0.5) / ns) evaluated in float is exact while
// n * 2^-23 < 0.5 (n < 4M), since (j + 0.5)/ns stays at least 0.5/ns away from an integer.
__device__ __forceinline__ int fast_div(int j, int ns, float inv_ns) {
    return __float2int_rz(((float)j + 0.5f) * inv_ns);
}

// Twiddle + DFT of every butterfly of the thread (pass with prefix product ns); clamped lanes compute a copy.
// TW(e) returns W_n^e for 0 <= e < n.
// TW1: one table lookup per butterfly (w = W^base) and the other powers by products (w^(2a) = (w^a)^2,
// w^r = w^hb w^(r - hb)), instead of one lookup per power of two: fewer LDS reads for the same number of complex
// products; w^r then carries up to about log2(R) + 2 roundings more than the table (a few 1e-7 for R <= 16).
template <int R, int MAXE, bool INV, bool VOTE0 = true, bool TW1 = false, class TwAt>
__device__ __forceinline__ void pass_compute(float2 (&v)[MAXE], int s, int S, int n, int ns, TwAt&& tw_at) {
    constexpr int NB = MAXE / R;
    const int nr = n / R;
    if (ns > 1) {
        const int tstride = n / (ns * R);
        const float inv_ns = 1.0f / (float)ns;
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            if ((VOTE0 || i > 0) && !wave_any(s + i * S < nr)) break;
            const int j = min(s + i * S, nr - 1);
            const int k = j - fast_div(j, ns, inv_ns) * ns;
            const int base = k * tstride;
            // w^r from one table lookup per power of two and products of at most two factors: w^r = w^(2^a) w^b
            // with b < 2^a, so every twiddle carries at most a few ulp more than the table.
            float2 pw[R];
            pw[1] = tw_at(base);
#pragma unroll
            for (int r = 2; r < R; ++r) {
                if ((r & (r - 1)) == 0) {
                    pw[r] = TW1 ? cmul(pw[r / 2], pw[r / 2]) : tw_at(base * r);
                } else {
                    int hb = 1;
                    while (hb * 2 <= r) hb *= 2;
                    pw[r] = cmul(pw[hb], pw[r - hb]);
                }
            }
#pragma unroll
            for (int r = 1; r < R; ++r) v[i * R + r] = INV ? cmulc(v[i * R + r], pw[r]) : cmul(v[i * R + r], pw[r]);
            Dft<R, INV>::run(&v[i * R]);
            // keep butterflies in sequence: hoisting every butterfly's twiddles at once costs registers, not time
            __builtin_amdgcn_sched_barrier(0);
        }
        return;
    }
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        if ((VOTE0 || i > 0) && !wave_any(s + i * S < nr)) break;
        Dft<R, INV>::run(&v[i * R]);
    }
}

// Store phase: dst((j/ns)*ns*R + j%ns + r*ns, value).
template <int R, int MAXE, class Dst>
__device__ __forceinline__ void pass_store(float2 (&v)[MAXE], int s, int S, int n, int ns, Dst&& dst) {
    constexpr int NB = MAXE / R;
    const int nr = n / R;
    const float inv_ns = 1.0f / (float)ns;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int j = s + i * S;
        if (j < nr) {
            const int k = j - fast_div(j, ns, inv_ns) * ns;
            const int base = (j - k) * R + k;
#pragma unroll
            for (int r = 0; r < R; ++r) dst(base + r * ns, v[i * R + r]);
        }
    }
}

// Visit the thread's elements after a final pass (positions j + r*n/R) with f(position, value&, valid).
// Every lane visits (clamped lanes see their copy with valid == false) so f may update the value unconditionally.
template <int R, int MAXE, class F>
__device__ __forceinline__ void pass_visit(float2 (&v)[MAXE], int s, int S, int n, F&& f) {
    constexpr int NB = MAXE / R;
    const int nr = n / R;
#pragma unroll
    for (int i = 0; i < NB; ++i) {
        const int jj = s + i * S;
        const bool valid = jj < nr;
        const int j = min(jj, nr - 1);
#pragma unroll
        for (int r = 0; r < R; ++r) f(j + r * nr, v[i * R + r], valid);
    }
}

// Radix dispatch is expanded in the kernel body, not in helper functions that take the register array by
// reference: the array must be promoted to registers before control-flow simplification sees the per-radix
// accesses, otherwise writes from different cases get merged behind a computed index and the array is demoted
// to scratch memory.
//
// Kernels using these macros have a template parameter RSET and name their state: `v` (float2[kMaxE]), `lane`,
// `nlanes`, `fft_n`, the twiddle accessor `tw_at(int e)` = W_n^e, and the LDS accessors `lds_store(int, float2)`
// / `lds_load(int)`.
// Kernels are instantiated per radix set (a bit mask, see radix_bit); cases outside the set compile to nothing,
// so each instantiation only carries the register pressure of the radices it can run.
constexpr int radix_bit(int r) {
    return r == 2 ? 1 : r == 3 ? 2 : r == 4 ? 4 : r == 5 ? 8 : r == 7 ? 16 : r == 8 ? 32 : r == 16 ? 64 : 0;
}
constexpr int kRadixAll = 127;

#define ASARF_CASE_(RR, CALL)                            \
    case RR:                                             \
        if constexpr ((RSET & radix_bit(RR)) != 0) {     \
            CALL(RR);                                    \
        }                                                \
        break;
#define ASARF_DISPATCH(RVAL, CALL)                                                              \
    switch (RVAL) {                                                                           \
        ASARF_CASE_(2, CALL) ASARF_CASE_(3, CALL) ASARF_CASE_(4, CALL) ASARF_CASE_(5, CALL)   \
        ASARF_CASE_(7, CALL) ASARF_CASE_(8, CALL) ASARF_CASE_(16, CALL) default : break;     \
    }

#define ASARF_STORE_LDS_(RR) pass_store<RR>(v, lane, nlanes, fft_n, prev_ns_, lds_store)
#define ASARF_LOAD_LDS_(RR) pass_load<RR>(v, lane, nlanes, fft_n, lds_load)
#define ASARF_COMPUTE_FWD_(RR) pass_compute<RR, kMaxE, false>(v, lane, nlanes, fft_n, ns_, tw_at)
#define ASARF_COMPUTE_INV_(RR) pass_compute<RR, kMaxE, true>(v, lane, nlanes, fft_n, ns_, tw_at)
#define ASARF_COMPUTE1_FWD(RR) pass_compute<RR, kMaxE, false>(v, lane, nlanes, fft_n, 1, tw_at)
#define ASARF_COMPUTE1_INV(RR) pass_compute<RR, kMaxE, true>(v, lane, nlanes, fft_n, 1, tw_at)

// Passes 1..NPASS-1 of one transform direction (pass 0 already loaded and computed). RADIX_AT(p) gives the
// radix of pass p in this direction. On exit the registers hold the final outputs.
#define ASARF_RUN_PASSES(INVB, NPASS, RADIX_AT)                            \
    {                                                                      \
        int ns_ = RADIX_AT(0);                                             \
        int prev_ns_ = 1;                                                  \
        for (int p_ = 1; p_ < (NPASS); ++p_) {                             \
            {                                                              \
                const int rs_ = RADIX_AT(p_ - 1);                          \
                ASARF_DISPATCH(rs_, ASARF_STORE_LDS_)                      \
            }                                                              \
            __syncthreads();                                               \
            const int r_ = RADIX_AT(p_);                                   \
            ASARF_DISPATCH(r_, ASARF_LOAD_LDS_)                            \
            __syncthreads();                                               \
            if (INVB) {                                                    \
                ASARF_DISPATCH(r_, ASARF_COMPUTE_INV_)                     \
            } else {                                                       \
                ASARF_DISPATCH(r_, ASARF_COMPUTE_FWD_)                     \
            }                                                              \
            prev_ns_ = ns_;                                                \
            ns_ *= r_;                                                     \
        }                                                                  \
    }

}  // namespace asarf
