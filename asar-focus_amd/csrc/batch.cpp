// Batched focusing: many blocks of uneven size, each with its own plan, executed either block by block (launch,
// then wait, as a host loop over the reference's per-scene flow would) or as one captured graph that forks the
// blocks over several streams and is replayed with a single launch.
#include <vector>

#include "asarf.h"
#include "internal.h"

using namespace asarf;

#define HIP_CHECK(expr)                                                                                    \
    do {                                                                                                   \
        hipError_t e_ = (expr);                                                                            \
        if (e_ != hipSuccess)                                                                              \
            return set_error(e_ == hipErrorOutOfMemory ? ASARF_ENOMEM : ASARF_EHIP, "%s: %s (%s:%d)", #expr, \
                             hipGetErrorString(e_), __FILE__, __LINE__);                                   \
    } while (0)
#define ST_CHECK(expr)                 \
    do {                               \
        int s_ = (expr);               \
        if (s_ != ASARF_OK) return s_; \
    } while (0)

struct asarf_batch {
    asarf_ctx* ctx = nullptr;
    std::vector<hipStream_t> streams;
    std::vector<asarf_ctx*> sctx;  // one context per stream (plans launch on their context's stream)
    struct Block {
        asarf_plan* plan = nullptr;
        const void* raw = nullptr;
        float* imp = nullptr;
        asarf_image slc{};
        int stream = 0;
    };
    std::vector<Block> blocks;
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    hipEvent_t fork = nullptr;
    std::vector<hipEvent_t> joins;
};

static void batch_free(asarf_batch* b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    for (auto& bl : b->blocks)
        if (bl.plan) asarf_plan_destroy(bl.plan);
    for (auto* c : b->sctx) asarf_ctx_destroy(c);
    for (auto s : b->streams) (void)hipStreamDestroy(s);
    if (b->fork) (void)hipEventDestroy(b->fork);
    for (auto e : b->joins) (void)hipEventDestroy(e);
}

extern "C" int asarf_batch_create(asarf_ctx* ctx, int n_streams, asarf_batch** out) {
    if (!ctx || !out || n_streams < 1 || n_streams > 16)
        return set_error(ASARF_EINVAL, "asarf_batch_create: argument (1..16 streams)");
    ST_CHECK(ctx_init(ctx));
    auto* b = new asarf_batch();
    b->ctx = ctx;
    auto fail = [&](int st) {
        batch_free(b);
        delete b;
        return st;
    };
    if (hipSetDevice(ctx->device) != hipSuccess) return fail(set_error(ASARF_EHIP, "asarf_batch_create: device"));
    for (int i = 0; i < n_streams; ++i) {
        hipStream_t s;
        if (hipStreamCreateWithFlags(&s, hipStreamNonBlocking) != hipSuccess)
            return fail(set_error(ASARF_EHIP, "asarf_batch_create: stream"));
        b->streams.push_back(s);
        asarf_ctx* c = nullptr;
        int st = asarf_ctx_create(ctx->device, s, &c);
        if (st != ASARF_OK) return fail(st);
        b->sctx.push_back(c);
        st = ctx_init(c);
        if (st != ASARF_OK) return fail(st);
        // the parent context's autotuner choices apply to every stream's plans
        c->tuned_col_n1 = ctx->tuned_col_n1;
        for (const auto& kv : ctx->tuned_rows) {
            asarf_ctx::RowTune rt;
            rt.use_static = kv.second.use_static;
            rt.static_variant = kv.second.static_variant;
            if (kv.second.plan) {
                auto p = std::make_unique<FftPlanHost>();
                p->n = kv.second.plan->n;
                p->radix = kv.second.plan->radix;
                p->d_tw = get_twiddles(c, p->n, &st);
                if (!p->d_tw) return fail(st);
                rt.plan = p.get();
                c->plans[std::make_pair(p->n, -1000 - c->tune_key++)] = std::move(p);
            }
            c->tuned_rows[kv.first] = rt;
        }
        hipEvent_t e;
        if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
            return fail(set_error(ASARF_EHIP, "asarf_batch_create: event"));
        b->joins.push_back(e);
    }
    if (hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess)
        return fail(set_error(ASARF_EHIP, "asarf_batch_create: event"));
    *out = b;
    return clear_error();
}

extern "C" int asarf_batch_add(asarf_batch* b, const asarf_focus_params* p, const float* chirp_iq, const void* raw,
                               float* imp_out, int* index) {
    if (!b || !p || !chirp_iq || !raw) return set_error(ASARF_EINVAL, "asarf_batch_add: null argument");
    if (p->imp && !imp_out) return set_error(ASARF_EINVAL, "asarf_batch_add: IMP output required");
    if (b->exec) return set_error(ASARF_EINVAL, "asarf_batch_add: the batch is already captured");
    asarf_batch::Block bl;
    bl.stream = (int)(b->blocks.size() % b->streams.size());
    ST_CHECK(asarf_plan_create(b->sctx[bl.stream], p, chirp_iq, &bl.plan));
    plan_single_stream(bl.plan);
    bl.raw = raw;
    bl.imp = imp_out;
    b->blocks.push_back(bl);
    if (index) *index = (int)b->blocks.size() - 1;
    return clear_error();
}

static int enqueue_all(asarf_batch* b) {
    for (auto& bl : b->blocks) ST_CHECK(asarf_plan_execute(bl.plan, bl.raw, &bl.slc, bl.imp));
    return ASARF_OK;
}

extern "C" int asarf_batch_execute(asarf_batch* b, int mode) {
    if (!b) return set_error(ASARF_EINVAL, "asarf_batch_execute: null");
    if (b->blocks.empty()) return set_error(ASARF_EINVAL, "asarf_batch_execute: empty batch");
    if (mode == ASARF_BATCH_SEQUENTIAL) {
        for (auto& bl : b->blocks) {
            ST_CHECK(asarf_plan_execute(bl.plan, bl.raw, &bl.slc, bl.imp));
            HIP_CHECK(hipStreamSynchronize(b->streams[bl.stream]));
        }
        return clear_error();
    }
    if (mode != ASARF_BATCH_GRAPH) return set_error(ASARF_EINVAL, "asarf_batch_execute: mode %d", mode);
    hipStream_t s0 = b->streams[0];
    if (!b->exec) {
        // one eager pass settles every lazily initialised table before capture, then the fork-join capture
        ST_CHECK(enqueue_all(b));
        for (auto s : b->streams) HIP_CHECK(hipStreamSynchronize(s));
        HIP_CHECK(hipStreamBeginCapture(s0, hipStreamCaptureModeRelaxed));
        int st = ASARF_OK;
        hipError_t e = hipEventRecord(b->fork, s0);
        for (size_t i = 1; i < b->streams.size() && e == hipSuccess; ++i) e = hipStreamWaitEvent(b->streams[i], b->fork, 0);
        if (e == hipSuccess) st = enqueue_all(b);
        for (size_t i = 1; i < b->streams.size() && e == hipSuccess; ++i) {
            e = hipEventRecord(b->joins[i], b->streams[i]);
            if (e == hipSuccess) e = hipStreamWaitEvent(s0, b->joins[i], 0);
        }
        hipGraph_t g = nullptr;
        const hipError_t ee = hipStreamEndCapture(s0, &g);
        if (e != hipSuccess || ee != hipSuccess || st != ASARF_OK) {
            if (g) (void)hipGraphDestroy(g);
            if (st != ASARF_OK) return st;
            return set_error(ASARF_EHIP, "asarf_batch_execute: capture: %s",
                             hipGetErrorString(e != hipSuccess ? e : ee));
        }
        b->graph = g;
        HIP_CHECK(hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0));
    }
    HIP_CHECK(hipGraphLaunch(b->exec, s0));
    return clear_error();
}

extern "C" int asarf_batch_sync(asarf_batch* b) {
    if (!b) return set_error(ASARF_EINVAL, "asarf_batch_sync: null");
    for (auto s : b->streams) HIP_CHECK(hipStreamSynchronize(s));
    return clear_error();
}

extern "C" int asarf_batch_size(const asarf_batch* b) { return b ? (int)b->blocks.size() : 0; }

extern "C" int asarf_batch_slc(asarf_batch* b, int index, asarf_image* slc) {
    if (!b || !slc || index < 0 || index >= (int)b->blocks.size())
        return set_error(ASARF_EINVAL, "asarf_batch_slc: argument");
    if (!b->blocks[index].slc.data) return set_error(ASARF_EINVAL, "asarf_batch_slc: block %d not executed", index);
    *slc = b->blocks[index].slc;
    return clear_error();
}

extern "C" int asarf_batch_results(asarf_batch* b, int index, asarf_focus_results* r) {
    if (!b || !r || index < 0 || index >= (int)b->blocks.size())
        return set_error(ASARF_EINVAL, "asarf_batch_results: argument");
    return asarf_plan_results(b->blocks[index].plan, r);
}

extern "C" int asarf_batch_destroy(asarf_batch* b) {
    if (!b) return set_error(ASARF_EINVAL, "asarf_batch_destroy: null");
    for (auto s : b->streams) (void)hipStreamSynchronize(s);
    batch_free(b);
    delete b;
    return clear_error();
}
