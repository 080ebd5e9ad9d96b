// Learned RCMC interpolator: training-pair generation from the plain 16-tap sinc RCMC, the network's forward and
// backward pass with per-workgroup gradient reduction, Adam, the tap table over (d, slope) cells and the error
// judge that decides where the table may replace the sinc kernel.
#include "rcmc_net.h"

namespace asarf {

// ------------------------------------------------------------------------------------------------ RNG ------
// Philox-4x32-10 (Salmon et al., SC'11), the generator of the scene simulator, evaluated per pair index.
__device__ inline void philox10(unsigned c[4], unsigned k0, unsigned k1) {
#pragma unroll
    for (int r = 0; r < 10; ++r) {
        const unsigned long long p0 = (unsigned long long)0xD2511F53u * c[0];
        const unsigned long long p1 = (unsigned long long)0xCD9E8D57u * c[2];
        const unsigned hi0 = (unsigned)(p0 >> 32), lo0 = (unsigned)p0;
        const unsigned hi1 = (unsigned)(p1 >> 32), lo1 = (unsigned)p1;
        const unsigned n0 = hi1 ^ c[1] ^ k0, n2 = hi0 ^ c[3] ^ k1;
        c[0] = n0;
        c[1] = lo1;
        c[2] = n2;
        c[3] = lo0;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

__device__ inline double alias_dc(double f, double dc, double prf) {
    const double half = 0.5 * prf;
    if (dc < 0.0) {
        if (f > half + dc) f -= prf;
    } else if (dc > 0.0) {
        if (f < -half + dc) f += prf;
    }
    return f;
}

// ------------------------------------------------------------------------------------------ training pairs --
// One thread per pair: a random (bin row, column) whose 16 taps lie inside [0, nrv); the target is the 16-tap
// sinc RCMC output, computed exactly as k_rcmc_mf does.
__global__ __launch_bounds__(256) void k_rcmc_pairs(PairArgs a) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n_pairs) return;
    const double step = a.prf / (double)a.nbins;
    int q = 0, x = 0, cx = 0;
    double shift = 0.0, p = 0.0, fn = 0.0;
    const ColConst* cc = nullptr;
    for (int attempt = 0; attempt < 16; ++attempt) {
        unsigned c[4] = {(unsigned)i, (unsigned)(i >> 32), (unsigned)attempt, 0x52434D43u};
        philox10(c, (unsigned)a.seed, (unsigned)(a.seed >> 32));
        q = (int)(((unsigned long long)c[0] * (unsigned)a.nbins) >> 32);
        x = 8 + (int)(((unsigned long long)c[1] * (unsigned)(a.nrv - 16)) >> 32);
        const int k = (a.n2 == 1) ? q : (q / a.n2 + a.n1 * (q % a.n2));
        const double fb = (k < a.nbins / 2) ? (double)k * step : (double)(k - a.nbins) * step;
        cc = a.col + x;
        fn = alias_dc(fb, cc->dc, a.prf);
        shift = cc->shift_k * fn * fn;
        p = round(shift);
        cx = x + (int)fmin(p, 4096.0);
        if (!(cx - 8 < 0 || cx + 8 >= a.nrv)) break;
        cx = -1;
    }
    float* out = a.pairs + i * kPairFloats;
    if (cx < 0) {  // no admissible position found: an all-zero pair (weight 0 in the loss)
        for (int j = 0; j < kPairFloats; ++j) out[j] = 0.f;
        return;
    }
    const float2* row = a.img + (long)q * a.stride;
    const float d = (float)(shift - p);
    float2 t;
    if (d == 0.0f) {
        t = row[cx];
    } else {
        float2 acc = make_float2(0.f, 0.f);
        float den = 0.f;
#pragma unroll
        for (int j = 0; j < 16; ++j) {
            const float wj = ((j & 1) ? -a.window[j] : a.window[j]) * __builtin_amdgcn_rcpf(d + (float)(8 - j));
            const float2 s = row[cx - 8 + j];
            acc.x += wj * s.x;
            acc.y += wj * s.y;
            den += wj;
        }
        const float inv = __builtin_amdgcn_rcpf(den);
        t = (den == 0.f) ? make_float2(0.f, 0.f) : make_float2(acc.x * inv, acc.y * inv);
    }
    float2 sm[kNetTaps];
    float e = 0.f;
#pragma unroll
    for (int j = 0; j < kNetTaps; ++j) {
        sm[j] = row[cx - 4 + j];
        e += sm[j].x * sm[j].x + sm[j].y * sm[j].y;
    }
    const float r = sqrtf(e / (float)kNetTaps);
    const float inv = r > 1e-30f ? 1.0f / r : 0.f;
    out[0] = 2.0f * d;
    out[1] = net_slope(cc->shift_k, fn, a.prf);
#pragma unroll
    for (int j = 0; j < kNetTaps; ++j) {
        out[2 + 2 * j] = sm[j].x * inv;
        out[3 + 2 * j] = sm[j].y * inv;
    }
    out[18] = t.x * inv;
    out[19] = t.y * inv;
}

hipError_t launch_rcmc_pairs(const PairArgs& a, hipStream_t s) {
    if (a.n_pairs <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_rcmc_pairs, dim3((unsigned)((a.n_pairs + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

// --------------------------------------------------------------------------------------- forward/backward --
// One wave per workgroup, one sample per lane. The bf16-rounded weights are broadcast from LDS; every lane keeps
// its sample's hidden vectors in its own LDS row (pitch 65: a lane's own row and a row shared by all lanes are
// both read without bank conflicts). The gradient of the wave's 64 samples is reduced in LDS, each lane owning
// one hidden unit's row (or column) of each dW, and written as this workgroup's partial sum.
__device__ __forceinline__ float tanh_f(float x) {
    const float ax = fminf(fabsf(x), 15.0f);
    const float e = __expf(-2.0f * ax);
    const float t = (1.0f - e) / (1.0f + e);
    return copysignf(t, x);
}

constexpr int kNetPitch = kNetH + 1;

__global__ __launch_bounds__(kNetGroupSamples) void k_net_grad(NetGradArgs a) {
    const float* __restrict__ wq = a.wq;
    const float* __restrict__ pairs = a.pairs;
    const long n = a.n;
    float* __restrict__ partials = a.partials;
    float* __restrict__ loss_part = a.loss_part;
    __shared__ float W[kNetP];
    __shared__ float H1[kNetGroupSamples][kNetPitch];  // h1, later dz1
    __shared__ float H2[kNetGroupSamples][kNetPitch];  // h2, later dz2
    __shared__ float DO[kNetGroupSamples][kNetOut + 1];
    __shared__ float X[kNetGroupSamples][kNetIn];
    const int l = threadIdx.x;
    for (int i = l; i < kNetP; i += kNetGroupSamples) W[i] = wq[i];
    const long si = (long)blockIdx.x * kNetGroupSamples + l;
    const bool valid = si < n;
    float pr[kPairFloats];
#pragma unroll
    for (int j = 0; j < kPairFloats; ++j) pr[j] = valid ? pairs[si * kPairFloats + j] : 0.f;
    X[l][0] = pr[0];
    X[l][1] = pr[1];
    __syncthreads();
    float* h1 = H1[l];
    float* h2 = H2[l];
    for (int j = 0; j < kNetH; ++j)
        h1[j] = tanh_f(W[kNetOW1 + 2 * j] * pr[0] + W[kNetOW1 + 2 * j + 1] * pr[1] + W[kNetOB1 + j]);
    for (int i = 0; i < kNetH; ++i) {
        float acc = W[kNetOB2 + i];
        const float* w2 = W + kNetOW2 + i * kNetH;
#pragma unroll 16
        for (int j = 0; j < kNetH; ++j) acc += w2[j] * h1[j];
        h2[i] = tanh_f(acc);
    }
    float o[kNetOut];
#pragma unroll
    for (int k = 0; k < kNetOut; ++k) o[k] = W[kNetOB3 + k];
    for (int i = 0; i < kNetH; ++i) {
        const float hv = h2[i];
#pragma unroll
        for (int k = 0; k < kNetOut; ++k) o[k] += W[kNetOW3 + k * kNetH + i] * hv;
    }
    {
        float base[kNetTaps];
        sinc8_weights(0.5f * pr[0], a.window8, base);
#pragma unroll
        for (int j = 0; j < kNetTaps; ++j) o[2 * j] += base[j];
    }
    // prediction sum_j (a_j + i b_j) s_j and its squared error against the target
    float er = -pr[18], ei = -pr[19];
#pragma unroll
    for (int j = 0; j < kNetTaps; ++j) {
        const float sr = pr[2 + 2 * j], sim = pr[3 + 2 * j];
        er += o[2 * j] * sr - o[2 * j + 1] * sim;
        ei += o[2 * j] * sim + o[2 * j + 1] * sr;
    }
    float loss = valid ? er * er + ei * ei : 0.f;
    float dout[kNetOut];
#pragma unroll
    for (int j = 0; j < kNetTaps; ++j) {
        const float sr = pr[2 + 2 * j], sim = pr[3 + 2 * j];
        dout[2 * j] = valid ? 2.0f * (er * sr + ei * sim) : 0.f;
        dout[2 * j + 1] = valid ? 2.0f * (ei * sr - er * sim) : 0.f;
        DO[l][2 * j] = dout[2 * j];
        DO[l][2 * j + 1] = dout[2 * j + 1];
    }
    __syncthreads();
    float* part = partials + (long)blockIdx.x * kNetP;
    {
        // lane l owns column l of W3: dW3[k][l] = sum_s dout_s[k] h2_s[l]
        float acc[kNetOut];
#pragma unroll
        for (int k = 0; k < kNetOut; ++k) acc[k] = 0.f;
        for (int s = 0; s < kNetGroupSamples; ++s) {
            const float hv = H2[s][l];
#pragma unroll
            for (int k = 0; k < kNetOut; ++k) acc[k] += DO[s][k] * hv;
        }
#pragma unroll
        for (int k = 0; k < kNetOut; ++k) part[kNetOW3 + k * kNetH + l] = acc[k];
        if (l < kNetOut) {
            float b = 0.f;
            for (int s = 0; s < kNetGroupSamples; ++s) b += DO[s][l];
            part[kNetOB3 + l] = b;
        }
    }
    __syncthreads();
    // dz2 = (W3^T dout) * (1 - h2^2), in place of h2
    for (int i = 0; i < kNetH; ++i) {
        float acc = 0.f;
#pragma unroll
        for (int k = 0; k < kNetOut; ++k) acc += W[kNetOW3 + k * kNetH + i] * dout[k];
        const float hv = h2[i];
        h2[i] = acc * (1.0f - hv * hv);
    }
    __syncthreads();
    {
        // lane l owns row l of W2: dW2[l][j] = sum_s dz2_s[l] h1_s[j]
        float acc[kNetH];
#pragma unroll
        for (int j = 0; j < kNetH; ++j) acc[j] = 0.f;
        float b = 0.f;
        for (int s = 0; s < kNetGroupSamples; ++s) {
            const float g = H2[s][l];
            b += g;
#pragma unroll
            for (int j = 0; j < kNetH; ++j) acc[j] += g * H1[s][j];
        }
#pragma unroll
        for (int j = 0; j < kNetH; ++j) part[kNetOW2 + l * kNetH + j] = acc[j];
        part[kNetOB2 + l] = b;
    }
    __syncthreads();
    // dz1 = (W2^T dz2) * (1 - h1^2), in place of h1
    for (int j = 0; j < kNetH; ++j) {
        float acc = 0.f;
#pragma unroll 16
        for (int i = 0; i < kNetH; ++i) acc += W[kNetOW2 + i * kNetH + j] * h2[i];
        const float hv = h1[j];
        h1[j] = acc * (1.0f - hv * hv);
    }
    __syncthreads();
    {
        // lane l owns row l of W1: dW1[l][k] = sum_s dz1_s[l] x_s[k]
        float a0 = 0.f, a1 = 0.f, b = 0.f;
        for (int s = 0; s < kNetGroupSamples; ++s) {
            const float g = H1[s][l];
            a0 += g * X[s][0];
            a1 += g * X[s][1];
            b += g;
        }
        part[kNetOW1 + 2 * l] = a0;
        part[kNetOW1 + 2 * l + 1] = a1;
        part[kNetOB1 + l] = b;
    }
    for (int off = 32; off > 0; off >>= 1) loss += __shfl_xor(loss, off);
    if (l == 0) loss_part[blockIdx.x] = loss;
}

hipError_t launch_net_grad(const NetGradArgs& a, hipStream_t s) {
    const long groups = (a.n + kNetGroupSamples - 1) / kNetGroupSamples;
    if (groups <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_net_grad, dim3((unsigned)groups), dim3(kNetGroupSamples), 0, s, a);
    return hipGetLastError();
}

__global__ void k_net_reduce(const float* partials, int groups, float* out) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (p >= kNetP) return;
    float acc = 0.f;
    for (int g = 0; g < groups; ++g) acc += partials[(long)g * kNetP + p];
    out[p] = acc;
}

hipError_t launch_net_reduce(const float* partials, int groups, float* out, hipStream_t s) {
    hipLaunchKernelGGL(k_net_reduce, dim3((kNetP + 255) / 256), dim3(256), 0, s, partials, groups, out);
    return hipGetLastError();
}

// Adam (Kingma & Ba) on the FP32 master parameters; the forward pass uses their bf16 rounding.
__global__ void k_net_adam(float* master, float* m, float* v, float* wq, const float* partials, int groups,
                           float inv_batch, float lr, float bc1, float bc2, const float* loss_part,
                           float* loss_slot) {
    const int p = blockIdx.x * blockDim.x + threadIdx.x;
    if (loss_slot && p == 0) {
        float l = 0.f;
        for (int k = 0; k < groups; ++k) l += loss_part[k];
        *loss_slot = l * inv_batch;
    }
    if (p >= kNetP) return;
    float g = 0.f;
    for (int k = 0; k < groups; ++k) g += partials[(long)k * kNetP + p];
    g *= inv_batch;
    const float mm = 0.9f * m[p] + 0.1f * g;
    const float vv = 0.999f * v[p] + 0.001f * g * g;
    m[p] = mm;
    v[p] = vv;
    const float w = master[p] - lr * (mm / bc1) / (sqrtf(vv / bc2) + 1e-8f);
    master[p] = w;
    wq[p] = bf16_round(w);
}

hipError_t launch_net_adam(float* master, float* m, float* v, float* wq, const float* partials, int groups,
                           float inv_batch, float lr, int step, const float* loss_part, float* loss_slot,
                           hipStream_t s) {
    const float bc1 = 1.0f - powf(0.9f, (float)step);
    const float bc2 = 1.0f - powf(0.999f, (float)step);
    hipLaunchKernelGGL(k_net_adam, dim3((kNetP + 255) / 256), dim3(256), 0, s, master, m, v, wq, partials, groups,
                       inv_batch, lr, bc1, bc2, loss_part, loss_slot);
    return hipGetLastError();
}

// ------------------------------------------------------------------------------------------ tap table ------
// One thread per (d, slope) cell: the network evaluated at the cell centre.
__global__ __launch_bounds__(256) void k_net_lut(NetLutArgs a) {
    __shared__ float W[kNetP];
    for (int i = threadIdx.x; i < kNetP; i += blockDim.x) W[i] = a.wq[i];
    __syncthreads();
    const int c = blockIdx.x * blockDim.x + threadIdx.x;
    if (c >= a.nd * a.ns) return;
    const int id = c % a.nd, is = c / a.nd;
    const float d = ((float)id + 0.5f) / (float)a.nd - 0.5f;
    const float slope = ((float)is + 0.5f) * 2.0f / (float)a.ns - 1.0f;
    const float x0 = 2.0f * d, x1 = slope;
    float h1[kNetH];
#pragma unroll
    for (int j = 0; j < kNetH; ++j)
        h1[j] = tanh_f(W[kNetOW1 + 2 * j] * x0 + W[kNetOW1 + 2 * j + 1] * x1 + W[kNetOB1 + j]);
    float o[kNetOut];
#pragma unroll
    for (int k = 0; k < kNetOut; ++k) o[k] = W[kNetOB3 + k];
    for (int i = 0; i < kNetH; ++i) {
        float acc = W[kNetOB2 + i];
#pragma unroll
        for (int j = 0; j < kNetH; ++j) acc += W[kNetOW2 + i * kNetH + j] * h1[j];
        const float h = tanh_f(acc);
#pragma unroll
        for (int k = 0; k < kNetOut; ++k) o[k] += W[kNetOW3 + k * kNetH + i] * h;
    }
    float base[kNetTaps];
    sinc8_weights(d, a.window8, base);
#pragma unroll
    for (int j = 0; j < kNetTaps; ++j) a.lut[(long)c * kNetTaps + j] = make_float2(base[j] + o[2 * j], o[2 * j + 1]);
}

hipError_t launch_net_lut(const NetLutArgs& a, hipStream_t s) {
    const int cells = a.nd * a.ns;
    hipLaunchKernelGGL(k_net_lut, dim3((cells + 255) / 256), dim3(256), 0, s, a);
    return hipGetLastError();
}

// One thread per validation pair: squared error of the table taps and of the plain 8-tap Blackman sinc against
// the 16-tap target, accumulated per judge cell.
__global__ __launch_bounds__(256) void k_net_judge(NetJudgeArgs a) {
    const long i = (long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= a.n) return;
    const float* pr = a.pairs + i * kPairFloats;
    const float d = 0.5f * pr[0], slope = pr[1];
    float e2 = 0.f;
#pragma unroll
    for (int j = 0; j < kPairFloats; ++j) e2 += pr[j] * pr[j];
    if (e2 == 0.f) return;  // empty pair
    const int cell = net_cell(d, slope, a.nd, a.ns);
    const float2* w = a.lut + (long)cell * kNetTaps;
    float er = -pr[18], ei = -pr[19];
    float sr8 = 0.f, si8 = 0.f, den = 0.f;
#pragma unroll
    for (int j = 0; j < kNetTaps; ++j) {
        const float sr = pr[2 + 2 * j], sim = pr[3 + 2 * j];
        const float2 wj = w[j];
        er += wj.x * sr - wj.y * sim;
        ei += wj.x * sim + wj.y * sr;
        const float ws = ((j & 1) ? -a.window8[j] : a.window8[j]) * __builtin_amdgcn_rcpf(d + (float)(4 - j));
        sr8 += ws * sr;
        si8 += ws * sim;
        den += ws;
    }
    float e8r, e8i;
    if (d == 0.0f) {
        e8r = pr[2 + 2 * 4] - pr[18];
        e8i = pr[3 + 2 * 4] - pr[19];
    } else {
        e8r = sr8 / den - pr[18];
        e8i = si8 / den - pr[19];
    }
    const int jd = (cell % a.nd) * a.judge_nd / a.nd;
    const int jc = (cell / a.nd) * a.judge_nd + jd;
    atomicAdd(a.err_sum + jc, er * er + ei * ei);
    atomicAdd(a.sinc8_err_sum + jc, e8r * e8r + e8i * e8i);
    atomicAdd(a.count + jc, 1.0f);
}

hipError_t launch_net_judge(const NetJudgeArgs& a, hipStream_t s) {
    if (a.n <= 0) return hipSuccess;
    hipLaunchKernelGGL(k_net_judge, dim3((unsigned)((a.n + 255) / 256)), dim3(256), 0, s, a);
    return hipGetLastError();
}

}  // namespace asarf
