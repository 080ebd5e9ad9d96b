// Kernel argument blocks and host-side launchers (internal to libasarf.so).
#pragma once

#include <hip/hip_runtime.h>

#include <cstddef>
#include <cstdint>

#include "fft_core.h"

namespace asarf {

constexpr int kThreads = 256;      // generic kernels and column transforms
constexpr int kRowThreads = 512;   // row transforms: one row per workgroup
constexpr int kColTile = 16;   // columns per workgroup in column transforms (16 x 8 B = one 128 B line)
constexpr int kMaxE = 16;      // complex elements per lane in a Stockham pass
constexpr int kStripRows = 128; // rows per workgroup in the raw statistics pass
constexpr int kRawStats = 6;    // SI, SQ, SII, SQQ, SIQ, N
constexpr int kRawLag = 9;      // LII, LQQ, LIQ, LQI, LI1, LQ1, LI2, LQ2, M per Doppler block
constexpr int kMaxDcBlocks = 64;
// Largest integer RCMC shift (pixels) the fused azimuth core's range halo covers; above it the plan runs the
// separate pass-B, RCMC/matched-filter and pass-B' kernels instead.
constexpr int kAzCoreReach = 8;
// Largest integer RCMC shift (pixels) covered by the column-tiled RCMC kernel's right halo.
constexpr int kRcmcTileReach = 24;

enum RawFormat { RAW_CF32 = 0, RAW_U8 = 1, RAW_I8 = 2, RAW_ERS_LINES = 3, RAW_ASAR_FBAQ4 = 4 };

// Level-0 lines. RAW_ERS_LINES: line y holds spl u8 (I, Q) pairs at bytes + y * line_bytes. RAW_ASAR_FBAQ4: record y
// at bytes + y * line_bytes holds a 2-byte little-endian length L, then L bytes in blocks of 64 (a quantiser block
// id and 63 code bytes, I code in the high nibble); sample s is code byte 64 (s / 63) + 1 + s % 63 when that is
// below L, reconstructed from the block-major table lut (16 entries per block id). Sample s of line y lands on
// range sample s + mult * (swst[y] - min_swst); range samples below the row's logical width that no sample lands
// on are NaN (zero after the I/Q correction, ZeroNaNs).
struct LinesDesc {
    const short* swst;
    int min_swst, mult, spl;
    long line_bytes;
    const float2* lut;  // RAW_ASAR_FBAQ4: lut[16 b + c] = (I, Q) of block id b and converted code c
    __device__ __forceinline__ int shift(int y) const { return mult * ((int)swst[y] - min_swst); }
};

// FBAQ-4 sample s of the record rec (length len from its first two bytes) into *v; false where the record has no
// such sample (Fbaq4IndexCalc, envisat_parse_util.cuh:75-84: codes 8..15 -> 7..0, 0..7 -> 8..15).
__device__ __forceinline__ int fbaq4_record_length(const unsigned char* rec) { return (int)rec[0] | ((int)rec[1] << 8); }
__device__ __forceinline__ bool fbaq4_sample(const unsigned char* rec, int len, long rec_bytes, const float2* lut,
                                             int s, float2* v) {
    if (s < 0) return false;
    const int blk = s / 63;
    const int k = 64 * blk + 1 + (s - 63 * blk);
    if (k >= len || 2 + (long)k >= rec_bytes) return false;
    const int bid = rec[2 + 64 * blk], code = rec[2 + k];
    const int ci = code >> 4, cq = code & 15;
    const float2* lt = lut + 16 * bid;
    *v = make_float2(lt[ci > 7 ? 15 - ci : 8 + ci].x, lt[cq > 7 ? 15 - cq : 8 + cq].y);
    return true;
}

// One row of N range samples from an FBAQ-4 record, lane `lane` of T taking x = lane, lane + T, ...: the record's
// sample x - shift where one lands below nr_in (through correct()), else 0 (NaN after ZeroNaNs, or padding), handed
// to store(x, value). Decoded in phases (every block id and code byte of the lane, then every table entry, then the
// stores), so that all of a lane's dependent loads are in flight together.
template <int N, int T, class Correct, class Store>
__device__ __forceinline__ void fbaq4_row_decode(const unsigned char* rec, long rec_bytes, const float2* lut, int shift,
                                                 int nr_in, int lane, Correct correct, Store store) {
    constexpr int IT = (N + T - 1) / T;
    const int len = fbaq4_record_length(rec);
    int bid[IT], code[IT];
    bool ok[IT];
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int x = lane + it * T;
        const int s = x - shift;
        const int blk = s >= 0 ? s / 63 : 0;
        const int k = 64 * blk + 1 + (s - 63 * blk);
        ok[it] = x < N && x < nr_in && s >= 0 && k < len && 2 + (long)k < rec_bytes;
        bid[it] = code[it] = 0;
        if (ok[it]) {
            bid[it] = rec[2 + 64 * blk];
            code[it] = rec[2 + k];
        }
    }
#pragma unroll
    for (int it = 0; it < IT; ++it) {
        const int x = lane + it * T;
        float2 v = make_float2(0.f, 0.f);
        if (ok[it]) {
            const int ci = code[it] >> 4, cq = code[it] & 15;
            const float2* lt = lut + 16 * bid[it];
            v = correct(lt[ci > 7 ? 15 - ci : 8 + ci].x, lt[cq > 7 ? 15 - cq : 8 + cq].y);
        }
        if (x < N) store(x, v);
    }
}

// Affine I/Q correction: I' = (I - dc_i) * inv_gain, Q' = (Q - dc_q - I' * sin_corr) * inv_cos.
struct IqCoef {
    float dc_i, dc_q, inv_gain, sin_corr, inv_cos;
    int identity;
};

// Device-resident results of the fused statistics/Doppler stage.
struct DevResults {
    double dc_i, dc_q, iq_gain, phase_error, stddev_i, stddev_q;
    double dc_poly[3];
    double block_freq[kMaxDcBlocks];
    double block_x[kMaxDcBlocks];
    int nblocks;
    int crop_front;  // lines removed at the start of the SLC for the IMP tail
    int imp_rows;    // rows of the multilooked product
    int pad;
    int rcmc_reach;  // bound on the integer RCMC shift over all columns and Doppler bins (pixels)
    IqCoef coef;
};

// Per range column constants of the range-Doppler stage (FP64).
struct ColConst {
    double dc;        // Doppler centroid at this column [Hz]
    double dc_off;    // bin-quantised centroid [Hz]
    double shift_k;   // RCMC shift in pixels per Hz^2: lambda^2 R0 / (8 Vr^2 dr)
    double inv_ka;    // 1 / azimuth FM rate: lambda R0 / (2 Vr^2)
};

struct RowFftArgs {
    const float2* in;
    long in_stride;
    float2* out;
    long out_stride;
    int rows;
    float scale;
    FftPlanDev plan;
    int mfma;           // split-bfloat16 MFMA transform (length 8192 only)
};

struct RcArgs {
    const void* in;     // raw I/Q bytes (row pitch nr_in samples) or cf32 rows (row pitch in_stride)
    long in_stride;     // elements
    int nr_in;          // valid input samples per row (raw); columns beyond are zero
    float2* out;
    long out_stride;
    int rows;
    int nrv;            // valid output columns
    int zero_to;        // columns [nrv, zero_to) are written as zero
    const IqCoef* coef; // device, nullptr = identity
    const float2* hconj; // conj(FFT(chirp)) / n
    FftPlanDev plan;
    int mfma;           // split-bfloat16 MFMA transforms (length 8192 only)
    int no_static;      // use the runtime-radix kernel even where a compile-time kernel exists (autotuner)
    int static_variant; // which compile-time variant of the length (autotuner; 0 = default)
    LinesDesc lines;    // RAW_ERS_LINES, RAW_ASAR_FBAQ4
};

struct ColFftArgs {
    const float2* in;
    long in_stride;
    float2* out;
    long out_stride;
    int ncols;
    int ntiles;
    int groups;            // number of sub-transforms per column
    long in_g, in_m;       // input row of element m of group g: g*in_g + m*in_m
    int in_rows;           // rows >= in_rows read as zero
    long out_g, out_m;     // output row of element k of group g: g*out_g + k*out_m
    int out_rows;          // rows >= out_rows are not written
    const float2* tw_full; // post-twiddle W_full^{g*k} (conjugated for the inverse); nullptr = none
    float scale;
    FftPlanDev sub;
    const int* fused_gate; // non-null: do nothing when *fused_gate <= kAzCoreReach (the fused kernel ran)
    int mfma;              // split-bfloat16 MFMA sub-transform (lengths 128 and 256 only)
    int no_static;         // 1: runtime-radix kernel even where a compile-time kernel exists (comparison runs)
    int col_width;         // columns per workgroup of the compile-time kernels: 16, 32, or 0 = by access pattern
};

struct RcmcArgs {
    float2* img;
    long stride;
    int nrv;
    int nbins;        // azimuth FFT length
    int n1, n2;       // storage row q holds frequency bin (q / n2) + n1 * (q % n2)
    const ColConst* col;
    double prf;
    double beta;      // azimuth bandwidth fraction
    int mode;         // 0 = RCMC + matched filter, 1 = RCMC only, 2 = matched filter only
    const int* fused_gate;  // non-null: do nothing when *fused_gate <= kAzCoreReach (the fused kernel ran)
    // interpolator: 0 = 16-tap Blackman sinc (the reference's), 1 = learned 8-tap table where its judge cell
    // passed (16-tap sinc elsewhere), 2 = plain 8-tap Blackman sinc
    int interp;
    float2* out;                  // nullptr: in place on img
    const int* tile_gate;         // non-null: the row kernel does nothing when *tile_gate <= kRcmcTileReach
    const float2* lut;            // [ns][nd][8] complex taps
    const unsigned char* lut_ok;  // [ns][judge_nd]
    int lut_nd, lut_ns, judge_nd;
    float window8[8];
    float window[16];  // Blackman window of the 16-tap interpolator (filled by the launchers)
};

// Fused azimuth core: forward pass B + RCMC + matched filter + inverse pass B' (out of place; split n1 x n2).
struct AzCoreArgs {
    const float2* in;      // pass-A output: row k1*n2 + m holds element m of group k1's length-n2 transform
    float2* out;           // inverse pass-B' output with the inverse inter-pass twiddle applied
    long stride;
    int nrv;
    int nbins;             // n1 * n2
    int n1;
    int col0;              // first produced column (a multiple of az_core_tile_columns())
    int ntiles;            // produced tiles, from col0 (the last may end at nrv)
    int items;             // n1 * ntiles
    const ColConst* col;
    double prf, beta;
    const float2* tw2;     // W_n2^m
    const float2* tw_full; // W_nbins^m
    const int* gate;       // non-null: do nothing when *gate > kAzCoreReach
    float window[16];      // Blackman window of the RCMC interpolator
};

struct RawMomentsArgs {
    const void* raw;
    int nr, na;
    int step;         // DC block width
    int nblocks;      // Doppler blocks (columns at or beyond nblocks*step only enter the statistics)
    int vec;          // adjacent columns per lane: 8 (one 16-byte load per row, packed dot products), 4 (one
                      // 8-byte load) or 1
    int ntiles;       // tiles of 256*vec columns
    int nslots;       // Doppler blocks a tile can touch: (256*vec - 1)/step + 2
    int nstrips;
    long long* partials; // [strip][tile][kRawStats + nslots*kRawLag]
    long long* corr;     // vec 8: [nstrips][nblocks][kRawLag] lag sums of the odd block-start columns (written for
                         // odd starts only)
    LinesDesc lines;     // RAW_ERS_LINES, RAW_ASAR_FBAQ4: rows are shifted lines, NaN where no sample lands
};

struct RawReduceArgs {
    const long long* partials;
    int nstrips, ntiles, nslots, step, nblocks;
    int tile_cols;      // 256 * vec
    const long long* corr;  // vec 8: straddle-column lag sums ([nstrips][nblocks][kRawLag]) to move between
                            // neighbouring blocks (else null)
    long long* totals;  // [kRawStats + nblocks*kRawLag]
    int f64;            // partials and totals hold FP64 sums (bit patterns in the int64 slots): RAW_ASAR_FBAQ4
};

struct FinalizeArgs {
    const long long* totals; // [kRawStats + nblocks*kRawLag]
    int f64;                 // totals are FP64 bit patterns (RAW_ASAR_FBAQ4) instead of exact integer sums
    int nblocks, step, nr, na, x_stride;
    double n_total;
    int quirks;
    double prf;
    // crop decision for the IMP tail
    int crop_dc_pos, crop_dc_neg;
    int imp_in_rows_pos, imp_in_rows_neg;
    float ml_ratio;
    DevResults* res;
};

struct ColConstArgs {
    ColConst* col;
    int nrv;
    const double* dc_poly;  // device; nullptr: dc_val
    double dc_val[3];       // the Doppler polynomial by value (host callers)
    double vr_poly[3];
    double wavelength, range_spacing, r_first, prf;
    int nbins;
    int* reach;             // non-null: atomic max of the per-column RCMC shift bound (cleared beforehand)
    int clear_reach;        // launch_col_const clears *reach first (when no earlier kernel of the stream does)
};

struct MultilookArgs {
    const float2* slc;
    long stride;
    int row0;               // first input row (crop)
    const int* row0_dev;    // device override of row0 (fused plan), nullptr = use row0
    int col0;
    int in_rows;            // input rows after the crop
    const int* in_rows_dev; // device override
    int nr;                 // range samples
    int out_rows;           // rows to produce (upper bound when device override is used)
    const int* out_rows_dev;
    float ratio;
    const float* rsl;       // optional per-column factor
    float* out;             // out_rows x nr
};

struct SrgrArgs {
    const float* in;
    int in_nr;
    int rows;
    const int* rows_dev;
    double grsr[5];
    double dr, gs;
    int out_nr;
    float* out;
};

// Multilook x RSL fused with slant-to-ground resampling (k_imp.hip): workgroup t multilooks the slant columns
// [t * 255, t * 255 + 256) of a tile of output rows into LDS and resamples the ground columns
// [g_start[t], g_start[t + 1]), those whose first source column r0 lies in [t * 255, t * 255 + 255).
struct MlSrgrArgs {
    MultilookArgs ml;     // ml.out unused
    const int* r0;        // per ground column: first source slant column (clamped as k_srgr clamps it)
    const float* fx;      // per ground column: interpolation fraction
    const int* g_start;   // ml_srgr_tiles(ml.nr) + 1 entries
    int out_nr;           // ground columns
    float* out;           // out_rows x out_nr
};

// Level-0 sample decode into a padded cf32 image (k_ingest.hip).
struct IngestArgs {
    const uint8_t* bytes;     // line / record y at bytes + y * line_bytes
    long line_bytes;
    int lines;                // decoded lines (rows of the logical region)
    int samples_per_line;     // ERS: I/Q byte pairs per line
    const uint16_t* swst;     // per-line sampling-window-start codes
    int min_swst;
    int swst_mult;            // columns per SWST code step (ERS 4, ASAR 1)
    const float* i_lut;       // FBAQ-4 reconstruction tables
    const float* q_lut;
    int lut_items;
    const float2* lut_t;      // 4096-entry scratch for the block-major table copy (nullptr: read i_lut / q_lut)
    float2* out;
    int x_size, x_stride, y_stride;  // logical width, row pitch and padded height of the image
    int vec16;                // set by the launcher: column pairs stored with 16-byte stores
    int w32;                  // set by the launcher: ERS sample pairs read with aligned 4-byte loads
};

// Output conditioning into MDS records (k_output.hip).
struct OutputArgs {
    const float2* slc;        // SLC (cf32), or
    const float* det;         // detected image (f32)
    long x_stride;            // elements per input row
    int x_size, y_size;
    float calibration;
    int header;               // record header bytes (zeroed)
    long record_bytes;        // header + x_size * bytes per pixel
    unsigned char* dest;
};

// ---- launchers (k_*.hip) ----
hipError_t launch_row_fft(const RowFftArgs& a, bool inverse, hipStream_t s);
hipError_t launch_range_compress(const RcArgs& a, int raw_format, hipStream_t s);
bool launch_range_compress_static(const RcArgs& a, int raw_format, hipStream_t s, hipError_t* err);
int range_compress_static_variants(int n);
// Padded-layout range compression (k_rc_pad.hip): variants per length and their launcher.
int range_compress_pad_variants(int n);
bool launch_range_compress_pad(const RcArgs& a, int fmt, int variant, hipStream_t s, hipError_t* err);
hipError_t launch_col_fft(const ColFftArgs& a, bool inverse, hipStream_t s);
hipError_t launch_rc_mfma(const RcArgs& a, int raw_format, hipStream_t s);
hipError_t launch_row_mfma(const RowFftArgs& a, bool inverse, hipStream_t s);
hipError_t launch_col_mfma(const ColFftArgs& a, bool inverse, hipStream_t s);
bool launch_col_static(const ColFftArgs& a, bool inverse, hipStream_t s, hipError_t* err);
bool col_static_length(int n);
hipError_t launch_rcmc_mf(const RcmcArgs& a, hipStream_t s);
hipError_t launch_rcmc_tile(const RcmcArgs& a, hipStream_t s);
// Fused azimuth core (k_azcore.hip): LDS bytes of a workgroup, produced columns per tile, the second-pass
// lengths it is compiled for, the once-per-plan LDS limit, the launch.
size_t az_core_lds_bytes(int n2, int nbins);
int az_core_tile_columns();
bool az_core_supported(int n2);
int az_core_lengths(int* out, int capacity);  // the compiled n2 in order of preference; returns their count
hipError_t az_core_prepare(int n2, int nbins);
hipError_t launch_az_core(const AzCoreArgs& a, int n2, hipStream_t s);
void rcmc_window(float w[16]);
hipError_t launch_col_const(const ColConstArgs& a, hipStream_t s);
hipError_t launch_moments_cf32(const float2* img, long stride, int xs, int ys, int xstr, double* partials,
                               int nwg, hipStream_t s);
hipError_t launch_iq_apply_cf32(float2* img, long stride, int xs, int ys, int xstr, int ystr, int width,
                                IqCoef c, hipStream_t s);
hipError_t launch_dc_cf32(const float2* img, long stride, int ys, int x_begin0, int step, int nblocks, int xstr,
                          int nstrips, double* partials, hipStream_t s);
hipError_t launch_reduce_f64(const double* partials, int nparts, int nvals, double* totals, hipStream_t s);
hipError_t launch_raw_moments(const RawMomentsArgs& a, int raw_format, hipStream_t s);
hipError_t launch_reduce_raw(const RawReduceArgs& a, hipStream_t s);
hipError_t launch_finalize(const FinalizeArgs& a, hipStream_t s);
hipError_t launch_multilook(const MultilookArgs& a, hipStream_t s);
hipError_t launch_rsl(float* img, int nr, int rows, const float* tab, hipStream_t s);
hipError_t launch_srgr(const SrgrArgs& a, hipStream_t s);
// The source column and fraction of every ground column, as k_srgr computes them (out_nr entries each).
hipError_t launch_srgr_index(const SrgrArgs& a, int* r0, float* fx, hipStream_t s);
int ml_srgr_tiles(int nr);                 // workgroup columns of the fused kernel
int ml_srgr_owned_columns();               // slant columns per workgroup that own ground columns (255)
hipError_t ml_srgr_prepare(float ratio);   // raises the kernel's LDS limit for the ratio's row tile
hipError_t launch_ml_srgr(const MlSrgrArgs& a, hipStream_t s);
hipError_t launch_zero_cols(float2* img, long stride, int rows, int c0, int c1, hipStream_t s);
hipError_t launch_ingest(const IngestArgs& a, int kind, hipStream_t s);  // kind 0 = ERS IM, 1 = ASAR FBAQ-4
hipError_t launch_condition(const OutputArgs& a, int detected, hipStream_t s);
int fbaq4_index_host(int block, int code);
// Block-major copy of the two 4096-entry FBAQ-4 tables: lut_t[16 b + c] = (i_lut[256 c + b], q_lut[256 c + b]).
hipError_t launch_fbaq4_lut_t(const float* i_lut, const float* q_lut, float2* lut_t, hipStream_t s);

// Max dynamic LDS bytes per workgroup the launchers may request (set by the context at initialisation).
size_t max_lds_bytes();

}  // namespace asarf
