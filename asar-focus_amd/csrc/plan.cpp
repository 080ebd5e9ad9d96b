// Fused focusing plan: the GPU stage of main.cc:233-327 as one stream-ordered sequence with no host round trip.
//
//   raw bytes --(1 pass: I/Q + Doppler moments)--> finalize (FP64, 1 lane) --> column constants
//   raw bytes --(row kernel: correct, FFT, x conj(H)/n, IFFT)--> image rows < Na, Nrv columns
//   image --(column pass A, B: azimuth FFT, permuted bins)--> RCMC + matched filter + column pass B' (one
//         tile kernel, out of place) --(column pass A': azimuth IFFT, 1/Nap, rows < Na)--> SLC
//   [IMP] SLC --(crop, multilook x RSL)--> --(SRGR)--> detected ground-range image
//
// Device memory: one Nap x stride cf32 image (stride = Nrv rounded up to 16 for 128 B rows) plus small tables.
#include <algorithm>
#include <cmath>
#include <cstdlib>
#include <cstring>
#include <vector>

#include "asarf.h"
#include "internal.h"
#include "poly.h"
#include "rcmc_net.h"

using namespace asarf;

#define HIP_CHECK(expr)                                                                                   \
    do {                                                                                                  \
        hipError_t e_ = (expr);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return set_error(e_ == hipErrorOutOfMemory ? ASARF_ENOMEM : ASARF_EHIP, "%s: %s (%s:%d)", #expr, \
                             hipGetErrorString(e_), __FILE__, __LINE__);                                  \
    } while (0)
#define ST_CHECK(expr)                 \
    do {                               \
        int s_ = (expr);               \
        if (s_ != ASARF_OK) return s_; \
    } while (0)

struct asarf_plan {
    asarf_ctx* ctx = nullptr;
    asarf_focus_params p{};
    int nrp = 0, nap = 0, nrv = 0;
    long stride = 0;
    int nblocks = 0, step = 0, nstrips = 0, raw_vec = 1;  // raw_vec: widest column count per lane the rows allow
    const FftPlanHost* rplan = nullptr;
    ColSplit cs;
    // device buffers
    float2* img = nullptr;
    float2* img2 = nullptr;  // second image: out-of-place target of the column passes and of the RCMC stage
    bool core = false;       // fused azimuth core: pass B + RCMC + matched filter + pass B' in one kernel
    // Column chunks of the azimuth tail (core, then in-place pass A', per chunk) alternating between the context's
    // stream and a second stream, each chunk's core after the previous chunk's: one chunk's pass A' (HBM-bound)
    // can overlap the next chunk's core (VALU-bound). Opt-in (ASARF_AZ_CHUNKS).
    int az_chunks = 1;
    hipStream_t s1 = nullptr;
    hipEvent_t ev_fork = nullptr, ev_join = nullptr;
    std::vector<hipEvent_t> ev_core;
    float2* hconj = nullptr;
    ColConst* col = nullptr;
    long long* partials = nullptr;
    long long* totals = nullptr;
    long long* corr = nullptr;  // straddle-column lag sums of the eight-column statistics pass
    DevResults* res = nullptr;
    float* ml = nullptr;
    float* rsl = nullptr;
    // fused multilook + SRGR (k_ml_srgr): per ground column source index and fraction, ownership ranges
    bool ml_srgr = false;
    int* gr_r0 = nullptr;
    float* gr_fx = nullptr;
    int* gr_start = nullptr;
    size_t bytes = 0;
    // results window
    int crop_pos[2] = {0, 0}, crop_neg[2] = {0, 0};  // {front, back}
    int near = 8, far = 8, nr_s = 0;
    float ml_ratio = 0.f;
    int imp_cap = 0;
    // RCMC interpolator (0 = 16-tap sinc, 1 = learned table with sinc fallback, 2 = 8-tap sinc)
    int rcmc_interp = 0;
    bool rcmc_row_only = false;
    const asarf_rcmc_net* net = nullptr;
    int fft_mfma = 0;  // split-bfloat16 MFMA transforms
    LinesDesc lines{};  // level-0 lines of the current execution (asarf_plan_execute_lines / _fbaq4)
    float2* fbaq_lut = nullptr;  // RAW_ASAR_FBAQ4: block-major copy of the reconstruction tables (4096 entries)
    // timing
    bool timing = false;
    std::vector<hipEvent_t> ev;
    std::vector<const char*> ev_names;
    std::vector<float> last_ms;
};

// Tiling of the raw statistics pass at `vec` columns per lane: tiles of 256*vec columns, and the Doppler blocks
// one tile can touch.
static void raw_tiles(const asarf_plan* pl, int vec, int* ntiles, int* nslots) {
    const int tw = kThreads * vec;
    *ntiles = (pl->p.nr + tw - 1) / tw;
    *nslots = (tw - 1) / pl->step + 2;
}

static int plan_alloc(asarf_plan* pl, void** p, size_t bytes) {
    HIP_CHECK(hipMalloc(p, bytes));
    pl->bytes += bytes;
    return ASARF_OK;
}

static void plan_free(asarf_plan* pl) {
    void* bufs[] = {pl->img,    pl->img2, pl->hconj, pl->col, pl->partials, pl->totals,
                    pl->corr,   pl->res,  pl->ml,    pl->rsl, pl->fbaq_lut, pl->gr_r0, pl->gr_fx, pl->gr_start};
    for (void* b : bufs)
        if (b) (void)hipFree(b);
    for (hipEvent_t e : pl->ev) (void)hipEventDestroy(e);
    for (hipEvent_t e : pl->ev_core) (void)hipEventDestroy(e);
    if (pl->ev_fork) (void)hipEventDestroy(pl->ev_fork);
    if (pl->ev_join) (void)hipEventDestroy(pl->ev_join);
    if (pl->s1) (void)hipStreamDestroy(pl->s1);
}

extern "C" int asarf_plan_create(asarf_ctx* c, const asarf_focus_params* p, const float* chirp_iq, asarf_plan** out) {
    if (!c || !p || !chirp_iq || !out) return set_error(ASARF_EINVAL, "asarf_plan_create: null argument");
    if (p->na < 2 || p->nr < 32 || p->nc <= 0 || p->nc >= p->nr)
        return set_error(ASARF_ESHAPE, "asarf_plan_create: na=%d nr=%d nc=%d", p->na, p->nr, p->nc);
    if (p->raw_format != ASARF_RAW_U8 && p->raw_format != ASARF_RAW_I8 && p->raw_format != ASARF_RAW_ERS_LINES &&
        p->raw_format != ASARF_RAW_ASAR_FBAQ4)
        return set_error(ASARF_EINVAL, "asarf_plan_create: raw_format %d (u8, i8, ERS lines or ASAR FBAQ-4 records)",
                         p->raw_format);
    if (p->raw_format == ASARF_RAW_ASAR_FBAQ4 && (p->line_bytes < 3 || p->swst_mult <= 0))
        return set_error(ASARF_EINVAL,
                         "asarf_plan_create: FBAQ-4 records need line_bytes (%ld) >= 3 and swst_mult (%d) > 0",
                         p->line_bytes, p->swst_mult);
    if (p->raw_format == ASARF_RAW_ERS_LINES &&
        (p->samples_per_line <= 0 || p->samples_per_line > p->nr || p->line_bytes < 2L * p->samples_per_line ||
         p->swst_mult <= 0))
        return set_error(ASARF_EINVAL,
                         "asarf_plan_create: ERS lines need 0 < samples_per_line (%d) <= nr (%d), line_bytes (%ld) >= "
                         "2 samples_per_line and swst_mult (%d) > 0",
                         p->samples_per_line, p->nr, p->line_bytes, p->swst_mult);
    if (!(p->prf > 0) || !(p->wavelength > 0) || !(p->range_spacing > 0) || !(p->slant_range_first > 0) ||
        !(p->az_bandwidth_fraction > 0))
        return set_error(ASARF_EINVAL, "asarf_plan_create: non-positive radar parameter");
    if (p->quirks & ~ASARF_QUIRKS_REFERENCE) return set_error(ASARF_EINVAL, "asarf_plan_create: unknown quirk bits");
    const int nrp = p->nrp > 0 ? p->nrp : asarf_fft_padding_size(p->nr + p->nc);
    const int nap = p->nap > 0 ? p->nap : asarf_fft_padding_size(p->na);
    if (nrp < p->nr + p->nc || nap < p->na)
        return set_error(ASARF_ESHAPE, "asarf_plan_create: FFT sizes %d x %d too small", nrp, nap);
    const int nrv = p->nr - p->nc;
    const double vr0 = polyval(p->vr_poly, 3, 0.0), vr1 = polyval(p->vr_poly, 3, (double)nrv);
    if (!(vr0 > 0) || !(vr1 > 0)) return set_error(ASARF_EINVAL, "asarf_plan_create: Vr must be > 0");
    const int step = p->nr / 32;
    int nblocks = 0;
    for (int xb = 0; xb < p->nr; xb += step) {
        if (!(p->quirks & ASARF_Q4_DC_EXTRA_BLOCK) && xb + step > p->nr) break;
        ++nblocks;
    }
    if (nblocks > kMaxDcBlocks) return set_error(ASARF_EUNSUPPORTED, "asarf_plan_create: %d Doppler blocks", nblocks);
    if (p->imp) {
        if (!(p->azimuth_spacing > 0) || !(p->ground_spacing > 0) || p->gr_range_size <= 0 || !(p->reference_range > 0))
            return set_error(ASARF_EINVAL, "asarf_plan_create: IMP parameters");
        if (nrv <= 16) return set_error(ASARF_ESHAPE, "asarf_plan_create: too few range samples for IMP");
    }
    ST_CHECK(ctx_init(c));
    if ((size_t)nrv * sizeof(float2) > c->max_lds)
        return set_error(ASARF_EUNSUPPORTED, "asarf_plan_create: %d range samples exceed LDS", nrv);

    auto* pl = new asarf_plan();
    pl->ctx = c;
    pl->p = *p;
    pl->nrp = nrp;
    pl->nap = nap;
    pl->nrv = nrv;
    pl->stride = (nrv + kColTile - 1) / kColTile * kColTile;
    pl->step = step;
    pl->nblocks = nblocks;
    // 8 columns per lane when the rows allow 16-byte loads and a lane meets at most one Doppler block boundary,
    // 4 when they allow 8-byte loads, else 1 (the width of a call also depends on the raw buffer's alignment)
    pl->raw_vec = (p->nr % 8 == 0 && step >= 8) ? 8 : (p->nr % 4 == 0) ? 4 : 1;
    if (p->raw_format == ASARF_RAW_ERS_LINES)  // lanes of 4 samples valid or NaN together, 8-byte loads
        pl->raw_vec = (p->nr % 4 == 0 && p->samples_per_line % 4 == 0 && p->swst_mult % 4 == 0 &&
                       p->line_bytes % 8 == 0)
                          ? 4
                          : 1;
    if (p->raw_format == ASARF_RAW_ASAR_FBAQ4) pl->raw_vec = 1;  // FP64 moments, one column per lane
    pl->nstrips = (p->na + kStripRows - 1) / kStripRows;
    auto fail = [&](int st) {
        plan_free(pl);
        delete pl;
        return st;
    };
    int st = ASARF_OK;
    pl->rplan = get_plan(c, nrp, kRowThreads, &st);
    if (!pl->rplan) return fail(st);
    st = get_col_split(c, nap, &pl->cs);
    if (st != ASARF_OK) return fail(st);
    // ASARF_RCMC_ROW=1: always the row RCMC kernel (the column-tiled kernel is not launched); for comparisons
    const char* row_env = std::getenv("ASARF_RCMC_ROW");
    pl->rcmc_row_only = row_env && row_env[0] == '1';
    // Fused azimuth core (pass B + RCMC + matched filter + pass B' in one kernel, k_azcore.hip): the default when
    // the azimuth length splits into n1 x n2 with a compiled second-pass length n2 (a short n2 keeps the
    // workgroup's n2 x 89 spectrum tile within a workgroup's LDS). ASARF_AZ_FUSED=0 keeps the separate kernels (and
    // the balanced split). Whether the core or the separate kernels run is decided on the device by the RCMC
    // shift bound (the core's range halo covers shifts up to kAzCoreReach pixels).
    const char* fz_env = std::getenv("ASARF_AZ_FUSED");
    // The compiled second-pass lengths are tried in their order of preference: first those whose first-pass
    // length n1 = nap / n2 has a compile-time column kernel and where both the core's tile and the 32-column tile of
    // the strided passes A and A' leave room for two workgroups per CU, then those with a compile-time first pass,
    // then any. (On the 30240-line ASAR scene: 105 x 288 runs the plan in 3.13 ms against 3.21 ms for 84 x 360,
    // whose 360-point passes hold one workgroup per CU.) ASARF_AZ_CORE_N2=<n2> tries only that one.
    if (!(fz_env && fz_env[0] == '0') && pl->cs.n2 > 1) {
        int lens[32];
        const int nl = std::min(az_core_lengths(lens, 32), 32);
        const char* n2_env = std::getenv("ASARF_AZ_CORE_N2");
        const int n2_only = n2_env ? std::atoi(n2_env) : 0;
        const size_t half_lds = c->max_lds / 2;
        auto two_per_cu = [&](int n1, int n2) {
            const size_t pass_lds = sizeof(float2) * (size_t)(n1 * 33 + tw_lds_elems(n1) + tw_lds_elems(nap));
            return pass_lds <= half_lds && az_core_lds_bytes(n2, nap) <= half_lds;
        };
        for (int round = 0; round < 3 && !pl->core; ++round) {
            for (int i = 0; i < nl; ++i) {
                const int n2 = lens[i];
                if (nap % n2 != 0 || (n2_only > 0 && n2 != n2_only)) continue;
                if (round < 2 && !(c->col_static && col_static_length(nap / n2))) continue;
                if (round == 0 && !two_per_cu(nap / n2, n2)) continue;
                ColSplit alt;
                if (get_col_split(c, nap, &alt, n2) != ASARF_OK || alt.n2 != n2) continue;
                if (!az_core_supported(n2) || az_core_lds_bytes(n2, nap) > c->max_lds) continue;
                if (az_core_prepare(n2, nap) != hipSuccess) continue;
                pl->cs = alt;
                pl->core = true;
                break;
            }
        }
        clear_error();
    }
    if ((st = plan_alloc(pl, (void**)&pl->img, sizeof(float2) * (size_t)nap * pl->stride))) return fail(st);
    // second image: the passes ping-pong between the two (A img -> img2, B img2 -> img, RCMC img -> img2,
    // B' img2 -> img, A' img -> img2), since out-of-place passes stream faster than in-place ones
    if ((st = plan_alloc(pl, (void**)&pl->img2, sizeof(float2) * (size_t)nap * pl->stride))) return fail(st);
    if ((st = plan_alloc(pl, (void**)&pl->hconj, sizeof(float2) * (size_t)nrp))) return fail(st);
    if ((st = plan_alloc(pl, (void**)&pl->col, sizeof(ColConst) * (size_t)nrv))) return fail(st);
    size_t partial_vals = 0;  // the larger of the two tilings (the width is chosen per call from the raw alignment)
    for (int v : {1, pl->raw_vec >= 4 ? 4 : 1, pl->raw_vec}) {
        int nt = 0, ns = 0;
        raw_tiles(pl, v, &nt, &ns);
        partial_vals = std::max(partial_vals, (size_t)pl->nstrips * nt * (kRawStats + ns * kRawLag));
    }
    if ((st = plan_alloc(pl, (void**)&pl->partials, sizeof(long long) * partial_vals))) return fail(st);
    if ((st = plan_alloc(pl, (void**)&pl->totals, sizeof(long long) * (size_t)(kRawStats + nblocks * kRawLag))))
        return fail(st);
    if ((st = plan_alloc(pl, (void**)&pl->corr, sizeof(long long) * (size_t)pl->nstrips * nblocks * kRawLag)))
        return fail(st);
    if ((st = plan_alloc(pl, (void**)&pl->res, sizeof(DevResults)))) return fail(st);
    if (p->raw_format == ASARF_RAW_ASAR_FBAQ4 &&
        (st = plan_alloc(pl, (void**)&pl->fbaq_lut, sizeof(float2) * 4096)))
        return fail(st);

    // chirp spectrum conj(FFT(chirp))/n: inverse-direction transform of the conjugated chirp
    {
        std::vector<float2> h(nrp);
        for (int i = 0; i < nrp; ++i) h[i] = make_float2(chirp_iq[2 * i], -chirp_iq[2 * i + 1]);
        if (hipMemcpy(pl->hconj, h.data(), sizeof(float2) * (size_t)nrp, hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_error(ASARF_EHIP, "asarf_plan_create: chirp upload"));
        RowFftArgs a{};
        a.in = pl->hconj;
        a.in_stride = nrp;
        a.out = pl->hconj;
        a.out_stride = nrp;
        a.rows = 1;
        a.scale = 1.0f / (float)nrp;
        a.plan = pl->rplan->dev();
        if (launch_row_fft(a, true, c->stream) != hipSuccess)
            return fail(set_error(ASARF_EHIP, "asarf_plan_create: chirp transform"));
    }
    // results window (main_flow.cc:155-201) for both signs of the Doppler constant term
    const int max_ap = aperture_pixels_max(p->wavelength, p->prf, p->az_bandwidth_fraction, p->vr_poly,
                                           p->slant_range_first, p->range_spacing, p->nr, p->nc);
    const int lb = std::max(0, p->lines_before), la = std::max(0, p->lines_after);
    pl->crop_pos[0] = lb + std::max(0, max_ap - lb);
    pl->crop_pos[1] = la;
    pl->crop_neg[0] = lb;
    pl->crop_neg[1] = la + std::max(0, max_ap - la);
    pl->near = asarf_rcmc_window_size() / 2;
    pl->far = asarf_rcmc_window_size() / 2;
    pl->nr_s = nrv - pl->near - pl->far;
    if (p->imp) {
        pl->ml_ratio = (float)(12.5 / p->azimuth_spacing);
        const int rows_pos = p->na - pl->crop_pos[0] - pl->crop_pos[1];
        const int rows_neg = p->na - pl->crop_neg[0] - pl->crop_neg[1];
        if (rows_pos <= 4 || rows_neg <= 4) return fail(set_error(ASARF_ESHAPE, "asarf_plan_create: crop leaves no lines"));
        pl->imp_cap = std::max((int)std::round((float)rows_pos / pl->ml_ratio),
                               (int)std::round((float)rows_neg / pl->ml_ratio));
        if ((st = plan_alloc(pl, (void**)&pl->rsl, sizeof(float) * (size_t)pl->nr_s))) return fail(st);
        // range spreading loss table over the cropped range axis (range_spreading_loss.cu:46-52)
        std::vector<float> tab(pl->nr_s);
        const double r0 = p->slant_range_first + (double)pl->near * p->range_spacing;
        for (int i = 0; i < pl->nr_s; ++i) {
            const double r = r0 + (double)i * p->range_spacing;
            tab[i] = (float)(1.0 / std::sqrt(std::pow(p->reference_range / r, 3.0)));
        }
        if (hipMemcpy(pl->rsl, tab.data(), sizeof(float) * (size_t)pl->nr_s, hipMemcpyHostToDevice) != hipSuccess)
            return fail(set_error(ASARF_EHIP, "asarf_plan_create: RSL table upload"));
        // Multilook and SRGR fused into one kernel (the detected slant-range image stays in LDS) where the ground
        // columns' source columns, computed on the device as k_srgr computes them, never decrease; otherwise (or
        // with ASARF_IMP_FUSED=0) the two kernels with the slant-range image between them.
        const char* imp_env = std::getenv("ASARF_IMP_FUSED");
        const int gr = p->gr_range_size;
        if (!(imp_env && imp_env[0] == '0') && gr > 0 && ml_srgr_prepare(pl->ml_ratio) == hipSuccess) {
            const int tiles = ml_srgr_tiles(pl->nr_s);
            if ((st = plan_alloc(pl, (void**)&pl->gr_r0, sizeof(int) * (size_t)gr))) return fail(st);
            if ((st = plan_alloc(pl, (void**)&pl->gr_fx, sizeof(float) * (size_t)gr))) return fail(st);
            if ((st = plan_alloc(pl, (void**)&pl->gr_start, sizeof(int) * (size_t)(tiles + 1)))) return fail(st);
            SrgrArgs sa{};
            sa.in_nr = pl->nr_s;
            for (int i = 0; i < 5; ++i) sa.grsr[i] = p->grsr_poly[i];
            sa.dr = p->range_spacing;
            sa.gs = p->ground_spacing;
            sa.out_nr = gr;
            std::vector<int> r0(gr);
            if (launch_srgr_index(sa, pl->gr_r0, pl->gr_fx, c->stream) != hipSuccess ||
                hipStreamSynchronize(c->stream) != hipSuccess ||
                hipMemcpy(r0.data(), pl->gr_r0, sizeof(int) * (size_t)gr, hipMemcpyDeviceToHost) != hipSuccess)
                return fail(set_error(ASARF_EHIP, "asarf_plan_create: slant-to-ground index"));
            if (std::is_sorted(r0.begin(), r0.end())) {
                std::vector<int> start(tiles + 1);
                for (int t = 0; t < tiles; ++t)
                    start[t] = (int)(std::lower_bound(r0.begin(), r0.end(), t * ml_srgr_owned_columns()) - r0.begin());
                start[tiles] = gr;
                if (hipMemcpy(pl->gr_start, start.data(), sizeof(int) * (size_t)(tiles + 1), hipMemcpyHostToDevice) !=
                    hipSuccess)
                    return fail(set_error(ASARF_EHIP, "asarf_plan_create: slant-to-ground ranges"));
                pl->ml_srgr = true;
            }
        }
        if (!pl->ml_srgr &&
            (st = plan_alloc(pl, (void**)&pl->ml, sizeof(float) * (size_t)pl->imp_cap * pl->nr_s)))
            return fail(st);
    }
    if (pl->core) {
        // One chunk by default: with the 48-column tiles of the long strided passes (141 KB of LDS) a pass-A'
        // workgroup cannot share a CU with a core workgroup (60 KB), and the two-stream chunks measured slower on
        // the ASAR scene (3.20 ms unsplit against 3.26-3.28 ms with 2-5 chunks; with 32-column tiles they had
        // gained 0.04 ms). ASARF_AZ_CHUNKS=<n> splits the tail into n chunks.
        const int tiles = (nrv + az_core_tile_columns() - 1) / az_core_tile_columns();
        int chunks = 1;
        const char* ch_env = std::getenv("ASARF_AZ_CHUNKS");
        if (ch_env && std::atoi(ch_env) > 0) chunks = std::min(std::atoi(ch_env), tiles);
        pl->az_chunks = chunks;
        if (chunks > 1) {
            if (hipStreamCreateWithFlags(&pl->s1, hipStreamNonBlocking) != hipSuccess ||
                hipEventCreateWithFlags(&pl->ev_fork, hipEventDisableTiming) != hipSuccess ||
                hipEventCreateWithFlags(&pl->ev_join, hipEventDisableTiming) != hipSuccess)
                return fail(set_error(ASARF_EHIP, "asarf_plan_create: azimuth chunk stream"));
            pl->ev_core.assign(chunks, nullptr);
            for (hipEvent_t& e : pl->ev_core)
                if (hipEventCreateWithFlags(&e, hipEventDisableTiming) != hipSuccess)
                    return fail(set_error(ASARF_EHIP, "asarf_plan_create: azimuth chunk events"));
        }
    }
    if (hipStreamSynchronize(c->stream) != hipSuccess) return fail(set_error(ASARF_EHIP, "asarf_plan_create: sync"));
    *out = pl;
    return clear_error();
}

extern "C" int asarf_plan_destroy(asarf_plan* pl) {
    if (!pl) return set_error(ASARF_EINVAL, "asarf_plan_destroy: null");
    (void)hipStreamSynchronize(pl->ctx->stream);
    plan_free(pl);
    delete pl;
    return clear_error();
}

extern "C" size_t asarf_plan_device_bytes(const asarf_plan* pl) { return pl ? pl->bytes : 0; }

extern "C" int asarf_plan_imp_capacity(const asarf_plan* pl, int* rows, int* cols) {
    if (!pl || !rows || !cols) return set_error(ASARF_EINVAL, "asarf_plan_imp_capacity: null argument");
    *rows = pl->imp_cap;
    *cols = pl->p.imp ? pl->p.gr_range_size : 0;
    return clear_error();
}

extern "C" int asarf_plan_set_timing(asarf_plan* pl, int enable) {
    if (!pl) return set_error(ASARF_EINVAL, "asarf_plan_set_timing: null");
    pl->timing = enable != 0;
    return clear_error();
}

static int mark(asarf_plan* pl, const char* name) {
    if (!pl->timing) return ASARF_OK;
    const size_t i = pl->ev_names.size();
    if (i >= pl->ev.size()) {
        hipEvent_t e;
        HIP_CHECK(hipEventCreate(&e));
        pl->ev.push_back(e);
    }
    pl->ev_names.push_back(name);
    HIP_CHECK(hipEventRecord(pl->ev[i], pl->ctx->stream));
    return ASARF_OK;
}

// Statistics, Doppler, column constants, range compression and the forward azimuth transform (both passes):
// on return pl->img holds the range-Doppler image in permuted bin order.
// Statistics, range compression and the forward azimuth passes. With core_path, pass B is gated on the RCMC shift
// bound: it runs only where the fused azimuth core (which includes pass B) will not.
static int enqueue_to_spectrum(asarf_plan* pl, const void* raw, bool core_path) {
    asarf_ctx* c = pl->ctx;
    const asarf_focus_params& p = pl->p;
    hipStream_t s = c->stream;
    // 1. one pass over the raw bytes: I/Q and Doppler-block moments
    RawMomentsArgs rm{};
    rm.raw = raw;
    rm.nr = p.nr;
    rm.na = p.na;
    rm.step = pl->step;
    rm.nblocks = pl->nblocks;
    const uintptr_t ra_bits = reinterpret_cast<uintptr_t>(raw);
    rm.vec = (pl->raw_vec == 8 && (ra_bits & 15) == 0) ? 8 : (pl->raw_vec >= 4 && (ra_bits & 7) == 0) ? 4 : 1;
    rm.corr = pl->corr;
    rm.lines = pl->lines;
    raw_tiles(pl, rm.vec, &rm.ntiles, &rm.nslots);
    rm.nstrips = pl->nstrips;
    rm.partials = pl->partials;
    HIP_CHECK(launch_raw_moments(rm, p.raw_format, s));
    ST_CHECK(mark(pl, "raw_moments"));
    RawReduceArgs rr{};
    rr.partials = pl->partials;
    rr.nstrips = pl->nstrips;
    rr.ntiles = rm.ntiles;
    rr.tile_cols = kThreads * rm.vec;
    rr.corr = rm.vec == 8 ? pl->corr : nullptr;
    rr.nslots = rm.nslots;
    rr.step = pl->step;
    rr.nblocks = pl->nblocks;
    rr.totals = pl->totals;
    rr.f64 = p.raw_format == ASARF_RAW_ASAR_FBAQ4;
    HIP_CHECK(launch_reduce_raw(rr, s));
    ST_CHECK(mark(pl, "reduce"));
    FinalizeArgs fa{};
    fa.totals = pl->totals;
    fa.f64 = rr.f64;
    fa.nblocks = pl->nblocks;
    fa.step = pl->step;
    fa.nr = p.nr;
    fa.na = p.na;
    fa.x_stride = pl->nrp;
    fa.n_total = p.n_total_samples ? (double)p.n_total_samples : (double)p.na * (double)p.nr;
    fa.quirks = p.quirks;
    fa.prf = p.prf;
    fa.crop_dc_pos = pl->crop_pos[0];
    fa.crop_dc_neg = pl->crop_neg[0];
    fa.imp_in_rows_pos = p.na - pl->crop_pos[0] - pl->crop_pos[1];
    fa.imp_in_rows_neg = p.na - pl->crop_neg[0] - pl->crop_neg[1];
    fa.ml_ratio = p.imp ? pl->ml_ratio : 0.f;
    fa.res = pl->res;
    HIP_CHECK(launch_finalize(fa, s));
    ST_CHECK(mark(pl, "finalize"));
    ColConstArgs ca{};
    ca.col = pl->col;
    ca.nrv = pl->nrv;
    ca.dc_poly = pl->res->dc_poly;
    for (int i = 0; i < 3; ++i) ca.vr_poly[i] = p.vr_poly[i];
    ca.wavelength = p.wavelength;
    ca.range_spacing = p.range_spacing;
    ca.r_first = p.slant_range_first;
    ca.prf = p.prf;
    ca.nbins = pl->nap;
    ca.reach = &pl->res->rcmc_reach;
    HIP_CHECK(launch_col_const(ca, s));
    ST_CHECK(mark(pl, "column_constants"));
    // 2. fused range compression
    RcArgs ra{};
    ra.in = raw;
    ra.in_stride = p.nr;
    ra.nr_in = p.nr;
    ra.out = pl->img;
    ra.out_stride = pl->stride;
    ra.rows = p.na;
    ra.nrv = pl->nrv;
    ra.zero_to = pl->nrv;
    ra.coef = &pl->res->coef;
    ra.hconj = pl->hconj;
    ra.plan = pl->rplan->dev();
    ra.mfma = pl->fft_mfma;
    ra.lines = pl->lines;
    {
        const auto t = c->tuned_rows.find(pl->nrp);
        ra.no_static = (t != c->tuned_rows.end() && !t->second.use_static) ? 1 : 0;
        ra.static_variant = (t != c->tuned_rows.end()) ? t->second.static_variant : 0;
        // ASARF_RC_VARIANT=<v> forces compile-time variant v (comparison runs and tests)
        const char* v_env = std::getenv("ASARF_RC_VARIANT");
        if (v_env && range_compress_static_variants(pl->nrp) > 0) {
            ra.no_static = 0;
            ra.static_variant = std::atoi(v_env) % range_compress_static_variants(pl->nrp);
        }
    }
    HIP_CHECK(launch_range_compress(ra, p.raw_format, s));
    ST_CHECK(mark(pl, "range_compression"));
    // 3. azimuth FFT (permuted bins), rows >= Na read as zero; with two passes, A img -> img2 and B img2 -> img
    // (out of place streams faster than in place)
    if (pl->cs.n2 > 1) {
        ST_CHECK(col_forward_pass(c, pl->cs, 0, pl->img, pl->stride, p.na, pl->img2, pl->stride, pl->nrv,
                                  pl->fft_mfma));
        ST_CHECK(mark(pl, "azimuth_fft_a"));
        ST_CHECK(col_forward_pass(c, pl->cs, 1, pl->img2, pl->stride, pl->nap, pl->img, pl->stride, pl->nrv,
                                  pl->fft_mfma, core_path ? &pl->res->rcmc_reach : nullptr));
        ST_CHECK(mark(pl, "azimuth_fft_b"));
    } else {
        ST_CHECK(col_forward_pass(c, pl->cs, 0, pl->img, pl->stride, p.na, pl->img, pl->stride, pl->nrv,
                                  pl->fft_mfma));
        ST_CHECK(mark(pl, "azimuth_fft_a"));
    }
    return ASARF_OK;
}

static int plan_execute(asarf_plan* pl, const void* raw, asarf_image* slc, float* imp_out);

extern "C" int asarf_plan_execute(asarf_plan* pl, const void* raw, asarf_image* slc, float* imp_out) {
    if (!pl || !raw || !slc) return set_error(ASARF_EINVAL, "asarf_plan_execute: null argument");
    if (pl->p.imp && !imp_out) return set_error(ASARF_EINVAL, "asarf_plan_execute: IMP output required");
    if (pl->p.raw_format == ASARF_RAW_ERS_LINES)
        return set_error(ASARF_EINVAL, "asarf_plan_execute: a plan of ERS lines runs with asarf_plan_execute_lines");
    if (pl->p.raw_format == ASARF_RAW_ASAR_FBAQ4)
        return set_error(ASARF_EINVAL, "asarf_plan_execute: a plan of FBAQ-4 records runs with asarf_plan_execute_fbaq4");
    return plan_execute(pl, raw, slc, imp_out);
}

extern "C" int asarf_plan_execute_fbaq4(asarf_plan* pl, const void* records, const short* swst, int min_swst,
                                        const float* i_lut, const float* q_lut, int lut_items, asarf_image* slc,
                                        float* imp_out) {
    if (!pl || !records || !swst || !i_lut || !q_lut || !slc)
        return set_error(ASARF_EINVAL, "asarf_plan_execute_fbaq4: null argument");
    if (pl->p.imp && !imp_out) return set_error(ASARF_EINVAL, "asarf_plan_execute_fbaq4: IMP output required");
    if (pl->p.raw_format != ASARF_RAW_ASAR_FBAQ4)
        return set_error(ASARF_EINVAL, "asarf_plan_execute_fbaq4: the plan was created for raw format %d",
                         pl->p.raw_format);
    if (lut_items < 4096)
        return set_error(ASARF_ESHAPE, "asarf_plan_execute_fbaq4: the tables need 4096 entries, not %d", lut_items);
    HIP_CHECK(launch_fbaq4_lut_t(i_lut, q_lut, pl->fbaq_lut, pl->ctx->stream));
    pl->lines.swst = swst;
    pl->lines.min_swst = min_swst;
    pl->lines.mult = pl->p.swst_mult;
    pl->lines.spl = 0;
    pl->lines.line_bytes = pl->p.line_bytes;
    pl->lines.lut = pl->fbaq_lut;
    return plan_execute(pl, records, slc, imp_out);
}

extern "C" int asarf_plan_execute_lines(asarf_plan* pl, const void* lines, const short* swst, int min_swst,
                                        asarf_image* slc, float* imp_out) {
    if (!pl || !lines || !swst || !slc) return set_error(ASARF_EINVAL, "asarf_plan_execute_lines: null argument");
    if (pl->p.imp && !imp_out) return set_error(ASARF_EINVAL, "asarf_plan_execute_lines: IMP output required");
    if (pl->p.raw_format != ASARF_RAW_ERS_LINES)
        return set_error(ASARF_EINVAL, "asarf_plan_execute_lines: the plan was created for raw format %d",
                         pl->p.raw_format);
    pl->lines.swst = swst;
    pl->lines.min_swst = min_swst;
    pl->lines.mult = pl->p.swst_mult;
    pl->lines.spl = pl->p.samples_per_line;
    pl->lines.line_bytes = pl->p.line_bytes;
    return plan_execute(pl, lines, slc, imp_out);
}

// The interpolator a plan execution applies: a learned table whose judge accepted no cell is the 16-tap sinc
// everywhere, so it runs as such (and may take the fused azimuth core).
static int effective_interp(const asarf_plan* pl) {
    if (pl->rcmc_interp == ASARF_RCMC_LEARNED && pl->net && pl->net->n_ok == 0) return ASARF_RCMC_SINC16;
    return pl->rcmc_interp;
}

// Whether an execution takes the fused azimuth core path (the device-side shift bound still decides whether the
// core or the gated separate kernels do the work).
static bool core_path_of(const asarf_plan* pl) {
    return pl->core && effective_interp(pl) == ASARF_RCMC_SINC16 && !pl->fft_mfma && !pl->rcmc_row_only;
}

namespace asarf {
void plan_single_stream(asarf_plan* pl) {
    if (pl) pl->az_chunks = 1;
}
}  // namespace asarf

static int plan_execute(asarf_plan* pl, const void* raw, asarf_image* slc, float* imp_out) {
    asarf_ctx* c = pl->ctx;
    const asarf_focus_params& p = pl->p;
    hipStream_t s = c->stream;
    pl->ev_names.clear();
    ST_CHECK(mark(pl, "start"));
    const bool core_path = core_path_of(pl);
    ST_CHECK(enqueue_to_spectrum(pl, raw, core_path));
    float2* slc_buf = pl->img;  // where the SLC ends (img2 after an out-of-place last pass)
    RcmcArgs rc{};
    rc.img = pl->img;
    rc.stride = pl->stride;
    rc.nrv = pl->nrv;
    rc.nbins = pl->nap;
    rc.n1 = pl->cs.n1;
    rc.n2 = pl->cs.n2;
    rc.col = pl->col;
    rc.prf = p.prf;
    rc.beta = p.az_bandwidth_fraction;
    rc.mode = 0;
    // a table whose judge accepted no cell runs as the 16-tap sinc directly, without the per-pixel cell lookups
    rc.interp = effective_interp(pl);
    if (pl->net) {
        rc.lut = pl->net->lut;
        rc.lut_ok = pl->net->lut_ok;
        rc.lut_nd = pl->net->nd;
        rc.lut_ns = pl->net->ns;
        rc.judge_nd = pl->net->judge_nd;
    }
    for (int i = 0; i < 8; ++i) rc.window8[i] = blackman_value(i, 8);
    if (core_path) {
        const ColSplit& cs = pl->cs;
        const int n = cs.n;
        // 4. fused azimuth core: pass B + RCMC + matched filter + pass B' (img2 -> img) when the RCMC shift bound
        // computed on the device fits the core's halo; otherwise pass B (enqueued above), RCMC + matched filter
        // and pass B' as separate kernels. Both variants are enqueued and the device-side bound makes exactly
        // one of them do the work.
        const int* gate = &pl->res->rcmc_reach;
        // the separate kernels for shift bounds beyond the core's halo first (pass B ran above; RCMC img -> img2,
        // then B' img2 -> img), so that the core chunks below read an untouched pass-A output in img2
        rc.out = pl->img2;
        rc.fused_gate = gate;
        rc.tile_gate = gate;
        HIP_CHECK(launch_rcmc_tile(rc, s));
        HIP_CHECK(launch_rcmc_mf(rc, s));
        ColFftArgs ib = col_args(pl->ctx, pl->img2, pl->stride, pl->img, pl->stride, pl->nrv, cs.n1, cs.n2, 1, n,
                                 cs.n2, 1, n, cs.full->d_tw, 1.f, cs.p2);
        ib.fused_gate = gate;
        HIP_CHECK(launch_col_fft(ib, true, s));
        ST_CHECK(mark(pl, "azimuth_core_wide"));
        AzCoreArgs ac{};
        ac.in = pl->img2;
        ac.out = pl->img;
        ac.stride = pl->stride;
        ac.nrv = pl->nrv;
        ac.nbins = pl->nap;
        ac.n1 = cs.n1;
        ac.col = pl->col;
        ac.prf = p.prf;
        ac.beta = p.az_bandwidth_fraction;
        ac.tw2 = cs.p2->d_tw;
        ac.tw_full = cs.full->d_tw;
        ac.gate = gate;
        rcmc_window(ac.window);
        // 4. fused azimuth core (img2 -> img) and 5. inverse pass A' in place (img), 1/Nap, rows < Na, per column
        // chunk. Chunk k's core waits for chunk k - 1's core (on the other stream), so it runs beside chunk
        // k - 1's pass A'; the core reads pass-A columns up to 16 past its chunk from img2, which no pass A'
        // writes. With per-kernel timing the chunks run unsplit on one stream.
        const int tw = az_core_tile_columns();
        const int tiles = (pl->nrv + tw - 1) / tw;
        // one chunk under stream capture: the second stream's fork and join are not taken into a caller's graph
        hipStreamCaptureStatus cap = hipStreamCaptureStatusNone;
        HIP_CHECK(hipStreamIsCapturing(s, &cap));
        const int chunks = (pl->timing || cap != hipStreamCaptureStatusNone) ? 1 : pl->az_chunks;
        if (chunks > 1) {
            HIP_CHECK(hipEventRecord(pl->ev_fork, s));
            HIP_CHECK(hipStreamWaitEvent(pl->s1, pl->ev_fork, 0));
        }
        for (int k = 0; k < chunks; ++k) {
            hipStream_t sk = (k % 2 == 0) ? s : pl->s1;
            const int t0 = (int)((long)tiles * k / chunks), t1 = (int)((long)tiles * (k + 1) / chunks);
            const int c0 = t0 * tw, c1 = std::min(t1 * tw, pl->nrv);
            if (k > 0) HIP_CHECK(hipStreamWaitEvent(sk, pl->ev_core[k - 1], 0));
            ac.col0 = c0;
            ac.ntiles = t1 - t0;
            ac.items = cs.n1 * ac.ntiles;
            HIP_CHECK(launch_az_core(ac, cs.n2, sk));
            if (chunks > 1) HIP_CHECK(hipEventRecord(pl->ev_core[k], sk));
            if (chunks == 1) ST_CHECK(mark(pl, "azimuth_core_fused"));
            ColFftArgs ia = col_args(pl->ctx, pl->img + c0, pl->stride, pl->img + c0, pl->stride, c1 - c0, cs.n2, 1,
                                     cs.n2, n, 1, cs.n2, p.na, nullptr, 1.0f / (float)pl->nap, cs.p1);
            HIP_CHECK(launch_col_fft(ia, true, sk));
        }
        if (chunks > 1) {
            HIP_CHECK(hipEventRecord(pl->ev_join, pl->s1));
            HIP_CHECK(hipStreamWaitEvent(s, pl->ev_join, 0));
        }
        ST_CHECK(mark(pl, "azimuth_ifft_a"));
        slc_buf = pl->img;
    } else {
        // 4. RCMC (any interpolator) + azimuth matched filter (img -> img2): the column-tiled kernel when the
        // device-side shift bound fits its halo, else the row kernel; both are enqueued and the bound makes
        // exactly one of them work
        const ColSplit& cs = pl->cs;
        const int n = cs.n;
        rc.out = pl->img2;
        if (!pl->rcmc_row_only) {
            rc.tile_gate = &pl->res->rcmc_reach;
            HIP_CHECK(launch_rcmc_tile(rc, s));
        }
        HIP_CHECK(launch_rcmc_mf(rc, s));
        ST_CHECK(mark(pl, "rcmc_matched_filter"));
        // 5. azimuth IFFT (img2 -> img), 1/Nap, rows < Na
        if (cs.n2 == 1) {
            ColFftArgs ia = col_args(pl->ctx, pl->img2, pl->stride, pl->img, pl->stride, pl->nrv, 1, 0, 1, n, 0, 1, p.na,
                                     nullptr, 1.0f / (float)pl->nap, cs.p1);
            ia.mfma = pl->fft_mfma;
            HIP_CHECK(launch_col_fft(ia, true, s));
            ST_CHECK(mark(pl, "azimuth_ifft_b"));
        } else {
            ColFftArgs ib = col_args(pl->ctx, pl->img2, pl->stride, pl->img, pl->stride, pl->nrv, cs.n1, cs.n2, 1, n, cs.n2,
                                     1, n, cs.full->d_tw, 1.f, cs.p2);
            ib.mfma = pl->fft_mfma;
            HIP_CHECK(launch_col_fft(ib, true, s));
            ST_CHECK(mark(pl, "azimuth_ifft_b"));
            // inverse pass A' (img -> img2, out of place), 1/Nap, rows < Na: the SLC ends in img2
            ColFftArgs ia = col_args(pl->ctx, pl->img, pl->stride, pl->img2, pl->stride, pl->nrv, cs.n2, 1, cs.n2, n, 1, cs.n2,
                                     p.na, nullptr, 1.0f / (float)pl->nap, cs.p1);
            ia.mfma = pl->fft_mfma;
            HIP_CHECK(launch_col_fft(ia, true, s));
            ST_CHECK(mark(pl, "azimuth_ifft_a"));
            slc_buf = pl->img2;
        }
    }
    *slc = asarf_image{slc_buf, pl->nrv, p.na, (int)pl->stride, pl->nap, ASARF_CF32};
    if (p.imp) {
        // 6. crop (results window), multilook x RSL, SRGR
        MultilookArgs ma{};
        ma.slc = slc_buf;
        ma.stride = pl->stride;
        ma.row0_dev = &pl->res->crop_front;
        ma.col0 = pl->near;
        ma.in_rows_dev = &pl->res->pad;
        ma.nr = pl->nr_s;
        ma.out_rows = pl->imp_cap;
        ma.out_rows_dev = &pl->res->imp_rows;
        ma.ratio = pl->ml_ratio;
        ma.rsl = pl->rsl;
        ma.out = pl->ml;
        if (pl->ml_srgr) {
            MlSrgrArgs fa{};
            fa.ml = ma;
            fa.r0 = pl->gr_r0;
            fa.fx = pl->gr_fx;
            fa.g_start = pl->gr_start;
            fa.out_nr = p.gr_range_size;
            fa.out = imp_out;
            HIP_CHECK(launch_ml_srgr(fa, s));
            ST_CHECK(mark(pl, "multilook_srgr"));
            return clear_error();
        }
        HIP_CHECK(launch_multilook(ma, s));
        ST_CHECK(mark(pl, "multilook_rsl"));
        SrgrArgs sa{};
        sa.in = pl->ml;
        sa.in_nr = pl->nr_s;
        sa.rows = pl->imp_cap;
        sa.rows_dev = &pl->res->imp_rows;
        for (int i = 0; i < 5; ++i) sa.grsr[i] = p.grsr_poly[i];
        sa.dr = p.range_spacing;
        sa.gs = p.ground_spacing;
        sa.out_nr = p.gr_range_size;
        sa.out = imp_out;
        HIP_CHECK(launch_srgr(sa, s));
        ST_CHECK(mark(pl, "srgr"));
    }
    return clear_error();
}

extern "C" int asarf_plan_kernel_times(asarf_plan* pl, const char** names, float* ms, int cap, int* count) {
    if (!pl || !count) return set_error(ASARF_EINVAL, "asarf_plan_kernel_times: null argument");
    if (!pl->timing || pl->ev_names.size() < 2) {
        *count = 0;
        return clear_error();
    }
    HIP_CHECK(hipEventSynchronize(pl->ev[pl->ev_names.size() - 1]));
    const int n = (int)pl->ev_names.size() - 1;
    *count = n;
    for (int i = 0; i < n && i < cap; ++i) {
        float t = 0.f;
        HIP_CHECK(hipEventElapsedTime(&t, pl->ev[i], pl->ev[i + 1]));
        if (ms) ms[i] = t;
        if (names) names[i] = pl->ev_names[i + 1];
    }
    return clear_error();
}

extern "C" int asarf_plan_results(asarf_plan* pl, asarf_focus_results* r) {
    if (!pl || !r) return set_error(ASARF_EINVAL, "asarf_plan_results: null argument");
    DevResults d;
    HIP_CHECK(hipMemcpyAsync(&d, pl->res, sizeof(d), hipMemcpyDeviceToHost, pl->ctx->stream));
    HIP_CHECK(hipStreamSynchronize(pl->ctx->stream));
    std::memset(r, 0, sizeof(*r));
    r->iq.dc_i = d.dc_i;
    r->iq.dc_q = d.dc_q;
    r->iq.iq_gain = d.iq_gain;
    r->iq.phase_error = d.phase_error;
    r->iq.stddev_i = d.stddev_i;
    r->iq.stddev_q = d.stddev_q;
    for (int i = 0; i < 3; ++i) r->dc_poly[i] = d.dc_poly[i];
    r->n_blocks = d.nblocks;
    for (int i = 0; i < d.nblocks && i < 64; ++i) {
        r->block_freq[i] = d.block_freq[i];
        r->block_x[i] = d.block_x[i];
    }
    r->nrv = pl->nrv;
    const bool pos = d.dc_poly[2] > 0.0;
    r->crop_front = pos ? pl->crop_pos[0] : pl->crop_neg[0];
    r->crop_back = pos ? pl->crop_pos[1] : pl->crop_neg[1];
    r->crop_near = pl->near;
    r->crop_far = pl->far;
    r->imp_rows = pl->p.imp ? d.imp_rows : 0;
    r->rcmc_shift_bound = d.rcmc_reach;
    const bool fused_ran = core_path_of(pl) && d.rcmc_reach <= kAzCoreReach;
    r->azimuth_core_fused = fused_ran ? 1 : 0;
    return clear_error();
}

extern "C" int asarf_plan_set_rcmc(asarf_plan* pl, int interp, const asarf_rcmc_net* net) {
    if (!pl) return set_error(ASARF_EINVAL, "asarf_plan_set_rcmc: null plan");
    if (interp < ASARF_RCMC_SINC16 || interp > ASARF_RCMC_SINC8)
        return set_error(ASARF_EINVAL, "asarf_plan_set_rcmc: interpolator %d", interp);
    if (interp == ASARF_RCMC_LEARNED && (!net || !net->lut || !net->lut_ok))
        return set_error(ASARF_EINVAL, "asarf_plan_set_rcmc: the learned interpolator needs a net with a built table");
    if (net && net->ctx != pl->ctx) return set_error(ASARF_EINVAL, "asarf_plan_set_rcmc: net of another context");
    pl->rcmc_interp = interp;
    pl->net = interp == ASARF_RCMC_LEARNED ? net : nullptr;
    return clear_error();
}

extern "C" int asarf_plan_rcmc_pairs(asarf_plan* pl, const void* raw, long n_pairs, unsigned long long seed,
                                     float* pairs) {
    if (!pl || !raw || !pairs || n_pairs <= 0) return set_error(ASARF_EINVAL, "asarf_plan_rcmc_pairs: argument");
    if (pl->p.raw_format != ASARF_RAW_U8 && pl->p.raw_format != ASARF_RAW_I8)
        return set_error(ASARF_EUNSUPPORTED, "asarf_plan_rcmc_pairs: plans of u8 or i8 raw samples only");
    if (pl->nrv < 32) return set_error(ASARF_ESHAPE, "asarf_plan_rcmc_pairs: too few range samples");
    pl->ev_names.clear();
    ST_CHECK(enqueue_to_spectrum(pl, raw, false));
    PairArgs a{};
    a.img = pl->img;
    a.stride = pl->stride;
    a.nrv = pl->nrv;
    a.nbins = pl->nap;
    a.n1 = pl->cs.n1;
    a.n2 = pl->cs.n2;
    a.col = pl->col;
    a.prf = pl->p.prf;
    rcmc_window(a.window);
    a.n_pairs = n_pairs;
    a.seed = seed;
    a.pairs = pairs;
    HIP_CHECK(launch_rcmc_pairs(a, pl->ctx->stream));
    return clear_error();
}

extern "C" int asarf_plan_set_fft(asarf_plan* pl, int precision) {
    if (!pl) return set_error(ASARF_EINVAL, "asarf_plan_set_fft: null plan");
    if (precision == ASARF_FFT_FP32) {
        pl->fft_mfma = 0;
        return clear_error();
    }
    if (precision != ASARF_FFT_SPLIT_BF16) return set_error(ASARF_EINVAL, "asarf_plan_set_fft: precision %d", precision);
    if (pl->p.raw_format != ASARF_RAW_U8 && pl->p.raw_format != ASARF_RAW_I8)
        return set_error(ASARF_EUNSUPPORTED, "asarf_plan_set_fft: split-bf16 transforms take u8 or i8 raw samples");
    auto col_ok = [](int n) { return n == 128 || n == 256; };
    if (pl->nrp != 8192)
        return set_error(ASARF_EUNSUPPORTED, "asarf_plan_set_fft: split-bf16 range transforms need nrp = 8192, not %d",
                         pl->nrp);
    if (!col_ok(pl->cs.n1) || (pl->cs.n2 != 1 && !col_ok(pl->cs.n2)))
        return set_error(ASARF_EUNSUPPORTED,
                         "asarf_plan_set_fft: split-bf16 azimuth passes need a %d = n1 x n2 split into 128/256, got "
                         "%d x %d", pl->nap, pl->cs.n1, pl->cs.n2);
    pl->fft_mfma = 1;
    return clear_error();
}
