// Host side of the learned RCMC interpolator: parameter initialisation and transfer, gradient evaluation,
// training with Adam, and the (d, slope) tap table with its error judge.
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

#include "asarf.h"
#include "internal.h"
#include "rcmc_net.h"

using namespace asarf;

#define HIP_CHECK(expr)                                                                                    \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return set_error(e_ == hipErrorOutOfMemory ? ASARF_ENOMEM : ASARF_EHIP, "%s: %s (%s:%d)", #expr, \
                             hipGetErrorString(e_), __FILE__, __LINE__);                                   \
    } while (0)

static void net_free(asarf_rcmc_net* n) {
    void* bufs[] = {n->master, n->m, n->v, n->wq, n->partials, n->loss_part, n->grad, n->lut, n->lut_ok, n->judge};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
}

static uint64_t splitmix64(uint64_t& s) {
    uint64_t z = (s += 0x9E3779B97F4A7C15ull);
    z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
    z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
    return z ^ (z >> 31);
}

static int upload_params(asarf_rcmc_net* n, const float* host) {
    std::vector<float> q(kNetP);
    for (int i = 0; i < kNetP; ++i) q[i] = bf16_round(host[i]);
    HIP_CHECK(hipMemcpyAsync(n->master, host, sizeof(float) * kNetP, hipMemcpyHostToDevice, n->ctx->stream));
    HIP_CHECK(hipMemcpyAsync(n->wq, q.data(), sizeof(float) * kNetP, hipMemcpyHostToDevice, n->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(n->ctx->stream));
    return ASARF_OK;
}

static int ensure_groups(asarf_rcmc_net* n, long groups) {
    if (groups <= n->groups_cap) return ASARF_OK;
    if (n->partials) (void)hipFree(n->partials);
    if (n->loss_part) (void)hipFree(n->loss_part);
    n->partials = nullptr;
    n->loss_part = nullptr;
    n->groups_cap = 0;
    HIP_CHECK(hipMalloc(&n->partials, sizeof(float) * (size_t)groups * kNetP));
    HIP_CHECK(hipMalloc(&n->loss_part, sizeof(float) * (size_t)groups));
    n->groups_cap = groups;
    return ASARF_OK;
}

extern "C" int asarf_rcmc_net_create(asarf_ctx* c, unsigned long long seed, asarf_rcmc_net** out) {
    if (!c || !out) return set_error(ASARF_EINVAL, "asarf_rcmc_net_create: null argument");
    int st = ctx_init(c);
    if (st != ASARF_OK) return st;
    auto* n = new asarf_rcmc_net();
    n->ctx = c;
    auto fail = [&](int s) {
        net_free(n);
        delete n;
        return s;
    };
    for (float** b : {&n->master, &n->m, &n->v, &n->wq, &n->grad})
        if (hipMalloc(b, sizeof(float) * kNetP) != hipSuccess)
            return fail(set_error(ASARF_ENOMEM, "asarf_rcmc_net_create: parameters"));
    (void)hipMemsetAsync(n->m, 0, sizeof(float) * kNetP, c->stream);
    (void)hipMemsetAsync(n->v, 0, sizeof(float) * kNetP, c->stream);
    // Glorot-uniform weights, zero biases
    std::vector<float> p(kNetP, 0.f);
    uint64_t s = seed;
    auto uni = [&](float lim) { return (float)(((double)(splitmix64(s) >> 11) * 0x1.0p-53 * 2.0 - 1.0) * lim); };
    const float l1 = std::sqrt(6.f / (kNetIn + kNetH)), l2 = std::sqrt(6.f / (kNetH + kNetH)),
                l3 = std::sqrt(6.f / (kNetH + kNetOut));
    for (int i = 0; i < kNetH * kNetIn; ++i) p[kNetOW1 + i] = uni(l1);
    for (int i = 0; i < kNetH * kNetH; ++i) p[kNetOW2 + i] = uni(l2);
    for (int i = 0; i < kNetOut * kNetH; ++i) p[kNetOW3 + i] = uni(l3);
    if ((st = upload_params(n, p.data())) != ASARF_OK) return fail(st);
    for (int i = 0; i < 8; ++i) n->window8[i] = blackman_value(i, 8);
    *out = n;
    return clear_error();
}

extern "C" int asarf_rcmc_net_destroy(asarf_rcmc_net* n) {
    if (!n) return set_error(ASARF_EINVAL, "asarf_rcmc_net_destroy: null");
    (void)hipStreamSynchronize(n->ctx->stream);
    net_free(n);
    delete n;
    return clear_error();
}

extern "C" int asarf_rcmc_net_num_params(void) { return kNetP; }

extern "C" int asarf_rcmc_net_get_params(asarf_rcmc_net* n, float* out, int count) {
    if (!n || !out || count != kNetP) return set_error(ASARF_EINVAL, "asarf_rcmc_net_get_params: need %d floats", kNetP);
    HIP_CHECK(hipMemcpyAsync(out, n->master, sizeof(float) * kNetP, hipMemcpyDeviceToHost, n->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(n->ctx->stream));
    return clear_error();
}

extern "C" int asarf_rcmc_net_set_params(asarf_rcmc_net* n, const float* in, int count) {
    if (!n || !in || count != kNetP) return set_error(ASARF_EINVAL, "asarf_rcmc_net_set_params: need %d floats", kNetP);
    for (int i = 0; i < kNetP; ++i)
        if (!std::isfinite(in[i])) return set_error(ASARF_EINVAL, "asarf_rcmc_net_set_params: non-finite value");
    const int st = upload_params(n, in);
    if (st != ASARF_OK) return st;
    n->step = 0;
    (void)hipMemsetAsync(n->m, 0, sizeof(float) * kNetP, n->ctx->stream);
    (void)hipMemsetAsync(n->v, 0, sizeof(float) * kNetP, n->ctx->stream);
    return clear_error();
}

extern "C" int asarf_rcmc_net_grad(asarf_rcmc_net* n, const float* pairs, long n_pairs, float* grad_out,
                                   double* loss_out) {
    if (!n || !pairs || !grad_out || n_pairs <= 0) return set_error(ASARF_EINVAL, "asarf_rcmc_net_grad: argument");
    const long groups = (n_pairs + kNetGroupSamples - 1) / kNetGroupSamples;
    int st = ensure_groups(n, groups);
    if (st != ASARF_OK) return st;
    hipStream_t s = n->ctx->stream;
    NetGradArgs ga{n->wq, pairs, n_pairs, n->partials, n->loss_part, {}};
    for (int i = 0; i < 8; ++i) ga.window8[i] = n->window8[i];
    HIP_CHECK(launch_net_grad(ga, s));
    HIP_CHECK(launch_net_reduce(n->partials, (int)groups, n->grad, s));
    HIP_CHECK(hipMemcpyAsync(grad_out, n->grad, sizeof(float) * kNetP, hipMemcpyDeviceToHost, s));
    std::vector<float> lp(groups);
    HIP_CHECK(hipMemcpyAsync(lp.data(), n->loss_part, sizeof(float) * groups, hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    if (loss_out) {
        double l = 0.0;
        for (float v : lp) l += v;
        *loss_out = l;
    }
    return clear_error();
}

extern "C" int asarf_rcmc_net_train(asarf_rcmc_net* n, const float* pairs, long n_pairs, int batch, int steps,
                                    float lr, float* loss_history) {
    if (!n || !pairs || n_pairs <= 0 || steps <= 0 || !(lr > 0))
        return set_error(ASARF_EINVAL, "asarf_rcmc_net_train: argument");
    if (batch <= 0 || batch % kNetGroupSamples || batch > n_pairs)
        return set_error(ASARF_EINVAL, "asarf_rcmc_net_train: batch %d must be a positive multiple of %d <= n_pairs",
                         batch, kNetGroupSamples);
    const long groups = batch / kNetGroupSamples;
    int st = ensure_groups(n, groups);
    if (st != ASARF_OK) return st;
    hipStream_t s = n->ctx->stream;
    const long nb = n_pairs / batch;  // whole batches; the pairs are already in random order
    float* d_hist = nullptr;
    if (loss_history) HIP_CHECK(hipMalloc(&d_hist, sizeof(float) * (size_t)steps));
    for (int k = 0; k < steps; ++k) {
        const float* b = pairs + (long)(k % nb) * batch * kPairFloats;
        NetGradArgs ga{n->wq, b, batch, n->partials, n->loss_part, {}};
        for (int i = 0; i < 8; ++i) ga.window8[i] = n->window8[i];
        hipError_t e = launch_net_grad(ga, s);
        // cosine decay to 1 % of the initial rate
        const float lrk = lr * (0.01f + 0.99f * 0.5f * (1.0f + std::cos(3.14159265f * (float)k / (float)steps)));
        n->step += 1;
        if (e == hipSuccess)
            e = launch_net_adam(n->master, n->m, n->v, n->wq, n->partials, (int)groups, 1.0f / (float)batch, lrk,
                                n->step, n->loss_part, d_hist ? d_hist + k : nullptr, s);
        if (e != hipSuccess) {
            if (d_hist) (void)hipFree(d_hist);
            return set_error(ASARF_EHIP, "asarf_rcmc_net_train: step %d: %s", k, hipGetErrorString(e));
        }
    }
    if (d_hist) {
        const hipError_t e1 = hipMemcpyAsync(loss_history, d_hist, sizeof(float) * (size_t)steps,
                                             hipMemcpyDeviceToHost, s);
        const hipError_t e2 = hipStreamSynchronize(s);
        (void)hipFree(d_hist);
        if (e1 != hipSuccess || e2 != hipSuccess) return set_error(ASARF_EHIP, "asarf_rcmc_net_train: loss copy");
    }
    HIP_CHECK(hipStreamSynchronize(s));
    return clear_error();
}

extern "C" int asarf_rcmc_net_build_table(asarf_rcmc_net* n, const float* val_pairs, long n_val, int nd, int ns,
                                          int judge_nd, float threshold, float* judge_rms, float* judge_sinc8_rms,
                                          int* n_judge_ok) {
    if (!n || !val_pairs || n_val <= 0 || nd <= 0 || ns <= 0 || judge_nd <= 0 || nd % judge_nd || !(threshold > 0))
        return set_error(ASARF_EINVAL, "asarf_rcmc_net_build_table: argument");
    if ((long)nd * ns > (1L << 22)) return set_error(ASARF_EINVAL, "asarf_rcmc_net_build_table: table too large");
    hipStream_t s = n->ctx->stream;
    for (void* b : {(void*)n->lut, (void*)n->lut_ok, (void*)n->judge})
        if (b) (void)hipFree(b);
    n->lut = nullptr;
    n->lut_ok = nullptr;
    n->judge = nullptr;
    const int nj = judge_nd * ns;
    HIP_CHECK(hipMalloc(&n->lut, sizeof(float2) * (size_t)nd * ns * kNetTaps));
    HIP_CHECK(hipMalloc(&n->lut_ok, (size_t)nj));
    HIP_CHECK(hipMalloc(&n->judge, sizeof(float) * 3 * (size_t)nj));
    n->nd = nd;
    n->ns = ns;
    n->judge_nd = judge_nd;
    NetLutArgs la{n->wq, nd, ns, n->lut, {}};
    for (int i = 0; i < 8; ++i) la.window8[i] = n->window8[i];
    HIP_CHECK(launch_net_lut(la, s));
    HIP_CHECK(hipMemsetAsync(n->judge, 0, sizeof(float) * 3 * (size_t)nj, s));
    NetJudgeArgs ja{};
    ja.lut = n->lut;
    ja.nd = nd;
    ja.ns = ns;
    ja.judge_nd = judge_nd;
    ja.pairs = val_pairs;
    ja.n = n_val;
    for (int i = 0; i < 8; ++i) ja.window8[i] = n->window8[i];
    ja.err_sum = n->judge;
    ja.count = n->judge + nj;
    ja.sinc8_err_sum = n->judge + 2 * nj;
    HIP_CHECK(launch_net_judge(ja, s));
    std::vector<float> h(3 * (size_t)nj);
    HIP_CHECK(hipMemcpyAsync(h.data(), n->judge, sizeof(float) * h.size(), hipMemcpyDeviceToHost, s));
    HIP_CHECK(hipStreamSynchronize(s));
    std::vector<unsigned char> ok(nj);
    int n_ok = 0;
    for (int j = 0; j < nj; ++j) {
        const float cnt = h[nj + j];
        const float rms = cnt > 0 ? std::sqrt(h[j] / cnt) : INFINITY;
        const float rms8 = cnt > 0 ? std::sqrt(h[2 * nj + j] / cnt) : INFINITY;
        if (judge_rms) judge_rms[j] = rms;
        if (judge_sinc8_rms) judge_sinc8_rms[j] = rms8;
        // a cell is trusted when enough validation pairs fell into it and their RMS error is below threshold
        ok[j] = (cnt >= 16.f && rms <= threshold) ? 1 : 0;
        n_ok += ok[j];
    }
    HIP_CHECK(hipMemcpyAsync(n->lut_ok, ok.data(), (size_t)nj, hipMemcpyHostToDevice, s));
    HIP_CHECK(hipStreamSynchronize(s));
    n->n_ok = n_ok;
    if (n_judge_ok) *n_judge_ok = n_ok;
    return clear_error();
}
