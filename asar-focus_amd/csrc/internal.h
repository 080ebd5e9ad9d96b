// Internal declarations shared by the C ABI translation units.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <memory>
#include <string>
#include <vector>

#include "asarf.h"
#include "kernels.h"

namespace asarf {

int set_error(int status, const char* fmt, ...) __attribute__((format(printf, 2, 3)));
int clear_error();

int aperture_pixels_max(double wavelength, double prf, double beta, const double* vr_poly, double r_first, double dr,
                        int range_size, int chirp_samples);

// Radix plan + device twiddle table for one transform length.
struct FftPlanHost {
    int n = 0;
    std::vector<int> radix;
    float2* d_tw = nullptr;
    FftPlanDev dev() const {
        FftPlanDev d{};
        d.n = n;
        d.npass = (int)radix.size();
        d.radices = 0;
        for (size_t i = 0; i < radix.size() && i < (size_t)kMaxPasses; ++i)
            d.radices |= (unsigned long long)radix[i] << (5 * i);
        d.tw = d_tw;
        return d;
    }
};

// Four-step split of a column transform of length n = n1 * n2 (n2 == 1: single pass).
struct ColSplit {
    int n = 0, n1 = 0, n2 = 0;
    const FftPlanHost* p1 = nullptr;
    const FftPlanHost* p2 = nullptr;
    const FftPlanHost* full = nullptr;  // twiddle table of length n (inter-pass twiddle)
};

}  // namespace asarf

struct asarf_ctx {
    int device = 0;
    hipStream_t stream = nullptr;
    bool own_stream = false;
    bool initialized = false;
    size_t max_lds = 65536;
    // column-transform knobs, read from the environment at initialisation (ASARF_COL_STATIC, ASARF_COL_WIDTH)
    bool col_static = true;
    int col_width = 0;
    // plan caches keyed by (length, lanes per transform)
    std::map<std::pair<int, int>, std::unique_ptr<asarf::FftPlanHost>> plans;
    std::map<int, float2*> tw_tables;
    // grow-only scratch
    void* d_scratch = nullptr;
    size_t scratch_bytes = 0;
    void* h_pinned = nullptr;
    size_t pinned_bytes = 0;
    // autotuner choices (asarf_autotune_fft): fastest column split (n1) per column length, and per row length
    // whether the compile-time range-compression kernel wins and otherwise the fastest runtime radix plan
    std::map<int, int> tuned_col_n1;
    struct RowTune {
        bool use_static = true;
        int static_variant = 0;
        const asarf::FftPlanHost* plan = nullptr;
    };
    std::map<int, RowTune> tuned_rows;
    int tune_key = 0;  // next private plan-cache key for tuner candidates
};

namespace asarf {
int ctx_init(asarf_ctx* ctx);
int ctx_scratch(asarf_ctx* ctx, size_t bytes, void** out);
int ctx_pinned(asarf_ctx* ctx, size_t bytes, void** out);
const FftPlanHost* get_plan(asarf_ctx* ctx, int n, int lanes, int* status);
float2* get_twiddles(asarf_ctx* ctx, int n, int* status);
int get_col_split(asarf_ctx* ctx, int n, ColSplit* out, int n2_max = 0);
// Column-transform launch arguments: element m of group g is read from row g*in_g + m*in_m, element k written
// to row g*out_g + k*out_m; twf (nullptr = none) is the inter-pass twiddle table.
ColFftArgs col_args(const asarf_ctx* c, const float2* in, long in_stride, float2* out, long out_stride, int ncols,
                    int groups, long in_g, long in_m, int in_rows, long out_g, long out_m, int out_rows,
                    const float2* twf, float scale, const FftPlanHost* sub);
// Doppler-centroid lag sums (CalculateDopplerCentroid, fractional_doppler_centroid.cu:98-162): the block layout
// of an image (validated; `who` names the entry point in errors) and the run that fills d_partials (device) with
// nvals doubles, copies them into h_partials (host) and fits the polynomial. dc_run allocates nothing.
constexpr int kDcMaxPartials = 2048;  // strips x blocks, so nvals <= 2 * kDcMaxPartials
struct DcLayout {
    int step = 0, nb = 0, nstrips = 0;
    size_t nvals = 0;
};
int dc_layout(const asarf_image* img, double prf, int quirks, const char* who, DcLayout* out);
int dc_run(asarf_ctx* c, const asarf_image* img, double prf, const DcLayout& layout, double* d_partials,
           double* h_partials, double* poly3, double* block_freq, double* block_x, int max_blocks, int* n_blocks);
int check_cf32(const asarf_image* img, const char* who);
bool radix_plan(int n, int lanes, std::vector<int>& out);
bool radix_plan_relaxed(int n, int lanes, std::vector<int>& out);
bool radix_plan_fits(int n, int lanes, const std::vector<int>& radix);
// Column transform passes of a four-step split (api.cpp); mfma selects the split-bfloat16 kernels.
int col_forward_pass(asarf_ctx* c, const ColSplit& cs, int pass, const float2* in, long in_stride, int in_rows,
                     float2* img, long stride, int ncols, int mfma = 0, const int* fused_gate = nullptr);
int col_inverse_pass_to(asarf_ctx* c, const ColSplit& cs, int pass, const float2* in, long in_stride, float2* out,
                        long out_stride, int ncols, int out_rows, float scale);
// Runs the plan's azimuth tail as one chunk on its context's stream (batches: their blocks overlap already).
void plan_single_stream(asarf_plan* plan);
int col_inverse_pass(asarf_ctx* c, const ColSplit& cs, int pass, float2* img, long stride, int ncols, int out_rows,
                     float scale, int mfma = 0);
}  // namespace asarf

// Learned RCMC interpolator state (device buffers; the host API is in rcmc_net.cpp).
struct asarf_rcmc_net {
    asarf_ctx* ctx = nullptr;
    float* master = nullptr;  // FP32 parameters
    float* m = nullptr;       // Adam moments
    float* v = nullptr;
    float* wq = nullptr;      // bf16-rounded copy used by the forward pass
    float* partials = nullptr;
    float* loss_part = nullptr;
    float* grad = nullptr;
    long groups_cap = 0;
    int step = 0;
    // tap table and judge
    float2* lut = nullptr;
    unsigned char* lut_ok = nullptr;
    float* judge = nullptr;   // err_sum | count | sinc8_err_sum, each ns * judge_nd
    int nd = 0, ns = 0, judge_nd = 0;
    int n_ok = 0;             // judge cells accepted by the last table build
    float window8[8] = {};
};
