/*
 * asarf — MI355X-native range-Doppler SAR focusser: C ABI.
 *
 * This is the boundary a host flow with the reference's interfaces (cgi-estonia-space/asar-focus 0.3.2) calls in
 * place of its CUDA stage. Each entry point names the reference function it replaces. All device pointers are
 * HIP device memory on the context's device; calls are ordered on the context's stream and are asynchronous
 * unless documented as returning host results. Every call returns an asarf_status; the reason for a failure is
 * available from asarf_last_error() on the calling thread. Arguments are validated before anything is launched.
 * The library never frees memory it did not allocate.
 */
#ifndef ASARF_H
#define ASARF_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ------------------------------------------------------------------------------------------ status codes */
typedef enum {
    ASARF_OK = 0,
    ASARF_EINVAL = 1,       /* null pointer, bad flag or parameter value */
    ASARF_ESHAPE = 2,       /* image size/stride inconsistent (reference: std::logic_error) */
    ASARF_EWORKSPACE = 3,   /* workspace smaller than required (cufft_checks.h:92-98, device_padded_image.cuh:147-151) */
    ASARF_EHIP = 4,         /* HIP runtime error */
    ASARF_ENOMEM = 5,       /* device allocation failed */
    ASARF_EUNSUPPORTED = 6  /* size outside what the kernels support (e.g. FFT length beyond LDS) */
} asarf_status;

/* Reference quirks reproduced by default (ASARF_QUIRKS_REFERENCE). */
enum {
    ASARF_Q1_IQ_PADDED_WIDTH = 1 << 0, /* I/Q correction over the padded row width, iq_correction.cu:197-198 */
    ASARF_Q2_COS_NO_SQRT = 1 << 1,     /* cos_corr = 1 - d^2, iq_correction.cu:259-262 */
    ASARF_Q4_DC_EXTRA_BLOCK = 1 << 2   /* 33rd Doppler block over the padding, fractional_doppler_centroid.cu:107-112 */
};
#define ASARF_QUIRKS_REFERENCE (ASARF_Q1_IQ_PADDED_WIDTH | ASARF_Q2_COS_NO_SQRT | ASARF_Q4_DC_EXTRA_BLOCK)

enum { ASARF_CF32 = 0, ASARF_F32 = 1 };                     /* image element types */
/* raw sample formats (I,Q interleaved). ASARF_RAW_ERS_LINES: ERS IM level-0 lines as the parser leaves them
 * (asarf_parse_ers_im_packet): samples_per_line u8 (I, Q) pairs per line, line_bytes apart, each line placed
 * swst_mult x (SWST code - minimum) range samples to the right, as ConvertErsImSamplesToComplex does
 * (envisat_format_kernels.cu:116-133); range samples no line sample lands on are NaN. ASARF_RAW_ASAR_FBAQ4: ASAR IM
 * FBAQ-4 records as asarf_parse_asar_im_packet writes them (2-byte little-endian length, then blocks of a
 * quantiser id and 63 code bytes), line_bytes apart, decoded through the reconstruction tables and placed
 * swst_mult x (SWST code - minimum) range samples to the right, as ConvertAsarImBlocksToComplex does
 * (envisat_format_kernels.cu:160-210). */
enum { ASARF_RAW_CF32 = 0, ASARF_RAW_U8 = 1, ASARF_RAW_I8 = 2, ASARF_RAW_ERS_LINES = 3, ASARF_RAW_ASAR_FBAQ4 = 4 };
enum { ASARF_FORWARD = 0, ASARF_INVERSE = 1 };
enum { ASARF_FFT_ROWS = 0, ASARF_FFT_COLUMNS = 1 };

/* ------------------------------------------------------------------------------------------ memory model */
/* Padded image: logical (x_size, y_size) inside (x_stride, y_stride); row-major, range fastest.
 * Replaces DevicePaddedImage (cuda_util/device_padded_image.cuh:28-158). */
typedef struct {
    void* data;
    int x_size, y_size;
    int x_stride, y_stride;
    int dtype; /* ASARF_CF32 (8 B) or ASARF_F32 (4 B) */
} asarf_image;

/* Owning device buffer with its byte size. Replaces CudaWorkspace (cuda_util/cuda_workplace.h:28-90). */
typedef struct {
    void* data;
    size_t bytes;
} asarf_workspace;

typedef struct asarf_ctx asarf_ctx;

/* Context: device, stream and caches (FFT twiddles, scratch). One context per host thread.
 * stream: the hipStream_t to run on (NULL = the device's default stream). Device initialisation is deferred
 * to the first call that launches work, so argument validation works on hosts without a GPU.
 * Replaces CudaInit (cuda_util/cuda_device_init.h:26-43). */
int asarf_ctx_create(int device, void* stream, asarf_ctx** out);
int asarf_ctx_destroy(asarf_ctx* ctx);
int asarf_sync(asarf_ctx* ctx);
void* asarf_ctx_stream(asarf_ctx* ctx);
const char* asarf_last_error(void);
const char* asarf_version(void);

/* InitPadded + ZeroMemory (device_padded_image.cuh:30-41,116). */
int asarf_image_alloc(asarf_ctx* ctx, int x_size, int y_size, int x_stride, int y_stride, int dtype, asarf_image* out);
int asarf_image_free(asarf_ctx* ctx, asarf_image* img);
int asarf_image_zero(asarf_ctx* ctx, asarf_image* img);
size_t asarf_image_bytes(const asarf_image* img); /* TotalByteSize (device_padded_image.cuh:72-74) */
/* CudaWorkspace(byte_size) / destructor (cuda_workplace.h:35-38,89). */
int asarf_workspace_alloc(asarf_ctx* ctx, size_t bytes, asarf_workspace* out);
int asarf_workspace_free(asarf_ctx* ctx, asarf_workspace* ws);
/* InitExtPtr (device_padded_image.cuh:49-61): the image adopts the workspace memory; ws becomes empty. */
int asarf_image_adopt(asarf_image* img, int x_size, int y_size, int x_stride, int y_stride, int dtype,
                      asarf_workspace* ws);
/* ReleaseMemory (device_padded_image.cuh:104-111): the image gives its memory to ws (which must be empty). */
int asarf_image_release(asarf_image* img, asarf_workspace* ws);

/* Strided copy of the sub-rectangle (x0, y0, width, height) of img's logical region to dst (device or host
 * memory, dst_pitch bytes per row); asynchronous on the context stream (synchronises when dst is pageable host
 * memory). Replaces the cudaMemcpy2D subset of SubsetResultsAndReassembleMeta (main_flow.cc:615-620) and
 * CopyToHostLogicalSize (device_padded_image.cuh:81-87). */
int asarf_image_subset_copy(asarf_ctx* ctx, const asarf_image* img, int x0, int y0, int width, int height, void* dst,
                            size_t dst_pitch);

/* ------------------------------------------------------------------------------------------ host helpers */
/* Smallest 2^a 3^b 5^c 7^d strictly greater than n (<= 131072), or -1.
 * Replaces GetOptimalFFTSizes().upper_bound(n) (cuda_util/cufft_plan.cc:147-181, main.cc:217,223). */
int asarf_fft_padding_size(int n);

/* ChirpInfo (sar/sar_metadata.h:34-44). */
typedef struct {
    double range_sampling_rate; /* Hz */
    double pulse_duration;      /* s */
    int n_samples;
    double pulse_bandwidth;     /* Hz */
    double kr;                  /* Hz/s */
    int apply_window;           /* periodic Hamming */
} asarf_chirp_info;
/* GenerateChirpData (sar/sar_chirp.cc:70-104), FP64 on the host: writes `padding` complex floats (I,Q). */
int asarf_generate_chirp(const asarf_chirp_info* chirp, int padding, float* out_iq);

/* Polyfit / Polyval (util/src/math_utils.cc:23-44, util/include/math_utils.h:55-63); highest power first. */
int asarf_polyfit(const double* x, const double* y, int n, int degree, double* coeffs);
double asarf_polyval(const double* p, int n, double x);

/* ------------------------------------------------------------------------------------------ stages */
/* SARResults subset written by the I/Q correction (sar/sar_metadata.h:46-55). */
typedef struct {
    double dc_i, dc_q;
    double iq_gain;
    double phase_error; /* degrees */
    double stddev_i, stddev_q;
} asarf_iq_results;

/* RawDataCorrection + ZeroNaNs (sar/iq_correction.cu:188-273, main.cc:244-245), in place on a cf32 image.
 * n_total_samples: divisor of the mean and deviation (0 -> x_size*y_size). Returns host results (synchronises). */
int asarf_iq_correction(asarf_ctx* ctx, asarf_image* img, size_t n_total_samples, int quirks,
                        asarf_iq_results* results);

/* CalculateDopplerCentroid (sar/fractional_doppler_centroid.cu:98-162): poly[3] highest power first, in range
 * index. Optional block_freq/block_x (capacity max_blocks) receive the per-block estimates. Synchronises. */
int asarf_doppler_centroid(asarf_ctx* ctx, const asarf_image* img, double prf, int quirks, double* poly3,
                           double* block_freq, double* block_x, int max_blocks, int* n_blocks);

/* No counterpart in the reference: asarf_doppler_centroid for callers that cannot allocate (tight serving loops,
 * harnesses that forbid hipMalloc). Same arguments, results and errors; it calls no hipMalloc, hipFree or other
 * allocating routine. Its lag sums go to one fixed device array per GPU (2 x 2048 doubles, 32 KiB, held until the
 * process ends) and return through the stack; layouts needing more are rejected with ASARF_ESHAPE (the block
 * layout of asarf_doppler_centroid never does). Calls that overlap on one GPU share that array and race. */
int asarf_doppler_centroid_noalloc(asarf_ctx* ctx, const asarf_image* img, double prf, int quirks, double* poly3,
                                   double* block_freq, double* block_x, int max_blocks, int* n_blocks);

/* RangeCompression (sar/range_compression.cu:57-111), in place on a cf32 image whose x_stride equals the chirp
 * length. chirp_iq: host, n_chirp complex floats (the padded, windowed chirp). On return x_size = x_size - nc.
 * ws may be NULL (the fused row kernel needs no FFT work area). */
int asarf_range_compression(asarf_ctx* ctx, asarf_image* img, const float* chirp_iq, int n_chirp, int chirp_samples,
                            asarf_workspace* ws);

/* Parameters of RangeDopplerAlgorithm taken from SARMetadata (range_doppler_algorithm.cu:337-353). */
typedef struct {
    double wavelength;
    double range_spacing;
    double slant_range_first;
    double prf;
    double az_bandwidth_fraction;
    double dc_poly[3]; /* highest power first */
    double vr_poly[3]; /* highest power first */
} asarf_rda_params;

/* RangeDopplerAlgorithm (sar/range_doppler_algorithm.cu:312-407): azimuth FFT, 16-tap sinc RCMC, azimuth
 * matched filter, azimuth IFFT, 1/Nap. src is the range-compressed image. If out->data is NULL, out adopts ws
 * (as the reference does) and afterwards ws holds src's memory; otherwise out is used as given and src is left
 * as scratch. out receives the SLC in its logical region (x < src.x_size, y < src.y_size). */
int asarf_range_doppler(asarf_ctx* ctx, const asarf_rda_params* params, asarf_image* src, asarf_image* out,
                        asarf_workspace* ws);

/* GetRcmcParameters (sar/focussing_details.h:24-28): window size of the RCMC interpolator (16). */
int asarf_rcmc_window_size(void);

/* AzimuthMultiLook (sar/az_multilook.cu:82-101): slc cf32 (x_size == x_stride not required), output f32 rows x
 * slc.x_size into out (capacity checked). Writes the output row count and the new azimuth spacing. */
int asarf_az_multilook(asarf_ctx* ctx, const asarf_image* slc, double azimuth_spacing, asarf_workspace* out,
                       int* out_rows, double* new_azimuth_spacing);

/* RangeSpreadingLossCorrection (sar/range_spreading_loss.cu:42-69), in place on f32 rows. */
int asarf_range_spreading_loss(asarf_ctx* ctx, float* img, int range_size, int azimuth_size, double slant_range_first,
                               double range_spacing, double reference_range);

/* Kernel part of SlantRangeToGroundRangeInterpolation (sar/srgr.cu:49-71,278-296): grsr[5] maps ground distance
 * [m] to slant distance from the first sample [m] (highest power first), evaluated in FP64. */
int asarf_srgr(asarf_ctx* ctx, const float* in, int in_range_size, int azimuth_size, const double* grsr5,
               double slant_range_spacing, double ground_spacing, int out_range_size, float* out);

/* Host part of SlantRangeToGroundRangeInterpolation (sar/srgr.cu:73-121) for a spherical earth: samples 100 slant
 * ranges R_i = R_first + i*(range_size/100)*dr, ground distance along the sphere from the first one, degree-4
 * fit. Writes grsr[5] and the output range size floor(g_99 / ground_spacing). */
int asarf_srgr_fit_spherical(double slant_range_first, double range_spacing, int range_size, double orbit_radius,
                             double earth_radius, double ground_spacing, double* grsr5, int* out_range_size);

/* ----------------------------------------------------------------------------------- level-0 sample decode */
/* ConvertErsImSamplesToComplex (envisat_format/src/envisat_format_kernels.cu:116-158) fused with the NaN fill and
 * the padded copy that follow it (main_flow.cc:537-540, main.cc:233-239), in one pass over `out` (cf32, device):
 * line y of `lines` holds samples_per_line (I, Q) u8 pairs at samples + y * line_bytes (device); sample s lands at
 * column s + 4 * (swst_codes[y] - min_swst) (swst_codes on the device). Logical elements that no sample reaches
 * are NaN, padding elements 0; samples shifted past the logical row are dropped. out->y_size must equal lines. */
int asarf_ingest_ers_im(asarf_ctx* ctx, const unsigned char* samples, long line_bytes, int samples_per_line,
                        int lines, const unsigned short* swst_codes, int min_swst, asarf_image* out);
/* ConvertAsarImBlocksToComplex + Fbaq4IndexCalc (envisat_format_kernels.cu:160-247,
 * envisat_format/include/envisat_parse_util.cuh:75-84), fused likewise: record y at records + y * record_bytes
 * is a 2-byte little-endian length L and L bytes of 64-byte blocks (a quantiser block id, then 63 codes with the I
 * code in the high nibble); code c of block b reconstructs to lut[256 * (c > 7 ? 15 - c : 8 + c) + b] from
 * i_lut / q_lut (lut_items floats each, device; 4096 in the reference). Sample s (the s-th code of the record)
 * lands at column s + swst_codes[y] - min_swst. */
int asarf_ingest_asar_fbaq4(asarf_ctx* ctx, const unsigned char* records, long record_bytes, int records_n,
                            const unsigned short* swst_codes, int min_swst, const float* i_lut, const float* q_lut,
                            int lut_items, asarf_image* out);
/* Fbaq4IndexCalc (envisat_parse_util.cuh:75-84) on the host: the reconstruction-table index of code (0..15) in
 * quantiser block `block` (0..255). */
int asarf_fbaq4_index(int block, int code);

/* ------------------------------------------------------------------------------ FP64 host geometry (host) */
/* Orbit state vector (sar/orbit_state_vector.h): ECEF position [m] and velocity [m/s] at time t [s]. Times are on
 * any epoch near the data (a double keeps 0.1 us at 1e8 s) and are used at the reference's whole-microsecond
 * resolution. */
typedef struct {
    double t;
    double pos[3];
    double vel[3];
} asarf_osv;

/* What the geometry functions read of SARMetadata (sar/sar_metadata.h): orbit, line timing, image and radar. */
typedef struct {
    const asarf_osv* osv;        /* state vectors, increasing t */
    int n_osv;
    double first_line_time;      /* s, same epoch as osv[].t */
    double line_time_interval;   /* s */
    int azimuth_size, range_size, chirp_samples;
    double slant_range_first, range_spacing;                   /* m */
    double carrier_frequency, prf, az_bandwidth_fraction;      /* Hz, Hz, fraction */
    double vr_poly[3], dc_poly[3];  /* current Vr and Doppler polynomials over the range index (aperture length) */
    double center_point[3];         /* ECEF initial guess of the geolocation */
} asarf_geometry;

/* InterpolateOrbit (sar/orbit_state_vector.cc:23-48): Lagrange interpolation through all n state vectors. */
int asarf_interpolate_orbit(const asarf_osv* osv, int n, double t, asarf_osv* out);
/* RangeDopplerGeoLocate (util/src/geo_tools.cc:27-85): the zero-Doppler point at slant_range on the WGS84 ellipsoid
 * (Newton from init_guess until the step is < 1e-6 m; ASARF_EINVAL after 101 steps). */
int asarf_geolocate(const double* sensor_pos, const double* sensor_vel, const double* init_guess, double slant_range,
                    double* out_xyz);
/* CalcIncidenceAngle (util/src/geo_tools.cc:87-106), degrees. */
double asarf_incidence_angle(const double* earth_point, const double* sensor_pos);
/* EstimateProcessingVelocity (sar/processing_velocity_estimation.cc:56-121): Vr at range indices 0, step, 2 step, ...
 * below range_size - chirp_samples (step = that / 10; 10 or 11 positions) from the orbit (hyperbolic range fit over
 * 9 aperture samples) and their degree-2 fit, highest power first. Optional outputs: the samples (up to 16) and
 * their range indices. */
int asarf_estimate_processing_velocity(const asarf_geometry* geometry, double* vr_poly3, double* vr_samples,
                                       double* range_samples, int* n_samples);
/* Host part of SlantRangeToGroundRangeInterpolation (sar/srgr.cu:73-107) from the orbit: grsr[5] (degree 4, highest
 * power first, slant - R_first as a function of ground distance) and the output range size. */
int asarf_srgr_fit(const asarf_geometry* geometry, double ground_spacing, double* grsr5, int* out_range_size);
/* Geo2xyzWgs84 (util/include/geo_tools.h:52-74): geodetic latitude/longitude [deg] and height [m] to ECEF. */
int asarf_geo_to_xyz(double latitude, double longitude, double altitude, double* out_xyz);
/* xyz2geoWGS84 (util/include/geo_tools.h:76-95): ECEF to geodetic [latitude, longitude] in degrees (Bowring). */
int asarf_xyz_to_geo(const double* xyz, double* out_lat_lon);
/* Scene centre and azimuth pixel spacing (main_flow.cc:499-522): the mid-line, mid-swath zero-Doppler point from
 * guess_xyz, and its distance to the same range's point one line later (azimuth_spacing may be NULL). Reads every
 * field of geometry except vr_poly, dc_poly and center_point. */
int asarf_scene_center(const asarf_geometry* geometry, const double* guess_xyz, double* out_center_xyz,
                       double* azimuth_spacing);
/* Tie points of one line of the geolocation grid ADS (envisat_lvl1_ads.h:640-646). */
typedef struct {
    unsigned samp_numbers[11];
    float slant_range_times[11]; /* two-way, ns */
    float angles[11];            /* incidence, deg */
    int lats[11], longs[11];     /* microdegrees */
} asarf_tie_points;
/* FillGeoLocationAds (envisat_lvl1_writer.cc:277-324) for the line az_idx: 11 points at range samples
 * i * range_size / 10, slant range from range_spacing, or (grsr5 != NULL) from the slant-to-ground polynomial of
 * the ground range sample * range_spacing. */
int asarf_geolocation_tie_points(const asarf_geometry* geometry, int az_idx, const double* grsr5,
                                 asarf_tie_points* out);
/* CalcAperturePixels (sar/sar_metadata.h:117-141): the synthetic aperture in lines at the range indices 0, step, ...
 * below range_size - chirp_samples (step = that / 10) from vr_poly and dc_poly; *n_out entries (10 or 11) are
 * written, ASARF_EINVAL when capacity is smaller. */
int asarf_aperture_pixels(const asarf_geometry* geometry, int* out, int capacity, int* n_out);

/* AzimuthRangeWindow (main_flow.h): lines and range pixels removed from the focused image before output. */
typedef struct {
    long lines_before_sensing_start;  /* lines read before the requested start */
    long lines_after_sensing_start;   /* aperture shortfall removed after the start */
    long lines_before_sensing_stop;   /* aperture shortfall removed before the stop */
    long lines_after_sensing_stop;    /* lines read after the requested stop */
    int near_range_pixels, far_range_pixels;
} asarf_results_window_t;
/* CalcResultsWindow (main_flow.cc:155-201). The subset itself is asarf_image_subset_copy with
 * y0 = lines_before_sensing_start + lines_after_sensing_start, x0 = near_range_pixels. */
int asarf_results_window(double dc_constant, long lines_before_start, long lines_after_stop, long max_aperture,
                         int rcmc_window, asarf_results_window_t* out);

/* ------------------------------------------------------------------------------------- output conditioning */
/* ConditionResultsSLC / CalibrateClampToBigEndian (envisat_format/src/envisat_format_kernels.cu:28-68, 100-111):
 * writes slc->y_size records to dest (device) of header_bytes zero bytes then slc->x_size pixels of big-endian
 * int16 I and Q = clamp(value * calibration) truncated toward zero. dest_bytes >= y_size * (header + 4 x_size). */
int asarf_condition_slc(asarf_ctx* ctx, const asarf_image* slc, float calibration, int header_bytes, void* dest,
                        size_t dest_bytes);
/* ConditionResultsDetected / CalibrateClampToBigEndianDetected (envisat_format_kernels.cu:70-98, 113-121): records
 * of header_bytes zero bytes then range_size uint16 of 10 log10(value * calibration) (single precision), saturated
 * to [0, 65535] with negative and NaN values 0 (survey quirk Q17 made defined), stored low byte first as the
 * reference's kernel does. img is azimuth_size x range_size floats (device, contiguous). */
int asarf_condition_detected(asarf_ctx* ctx, const float* img, int range_size, int azimuth_size, float calibration,
                             int header_bytes, void* dest, size_t dest_bytes);
/* CalculateStatistics / CalculateStatisticsImaged (main_flow.cc:666-676): out4 = {mean I, mean Q, std I, std Q}
 * of the logical region (NaN samples skipped, sums divided by x_size * y_size), or with imaged != 0
 * {(mean I + mean Q) / 2, 0, (std I + std Q) / 2, 0}. Synchronises. */
int asarf_image_statistics(asarf_ctx* ctx, const asarf_image* img, int imaged, double* out4);

/* Batched complex FFT (test hook; replaces the cuFFT plans of cufft_plan.cc:34-91). ROWS: `batch` transforms of
 * length n at data + i*dist, contiguous. COLUMNS: transforms of length n down `batch` columns of a row-major
 * array with row stride `dist` elements. Unnormalised unless scale != 1. In place. */
int asarf_fft_c2c(asarf_ctx* ctx, void* data, int n, int batch, long dist, int layout, int direction, float scale);

/* ------------------------------------------------------------------------------------------ fused plan */
/* The whole GPU stage of main.cc:233-327 as one stream-ordered sequence with no host round trip:
 * raw bytes -> statistics + Doppler (one pass) -> fused range compression -> azimuth FFT -> RCMC + matched filter
 * -> azimuth IFFT -> [IMP: crop, multilook x RSL, SRGR]. */
typedef struct {
    int na, nr, nc;          /* azimuth lines, range samples, chirp samples */
    int nrp, nap;            /* FFT sizes, 0 -> reference padding rule */
    int raw_format;          /* ASARF_RAW_U8, ASARF_RAW_I8, ASARF_RAW_ERS_LINES or ASARF_RAW_ASAR_FBAQ4 */
    double prf, wavelength, range_spacing, slant_range_first, az_bandwidth_fraction;
    double vr_poly[3];
    size_t n_total_samples;  /* 0 -> na*nr */
    int quirks;
    int imp;                 /* also produce the detected ground-range product */
    double azimuth_spacing;  /* m, for the multilook ratio 12.5/azimuth_spacing */
    double reference_range;  /* RSL reference range [m] (asar_constants.h:37: 800 km) */
    double grsr_poly[5];     /* ground -> slant polynomial (see asarf_srgr_fit_spherical) */
    int gr_range_size;       /* ground-range output samples */
    double ground_spacing;   /* 12.5 m */
    int lines_before, lines_after; /* extra lines around the requested sensing window (main.cc:137-142) */
    /* ASARF_RAW_ERS_LINES: I/Q pairs per line (nr = samples_per_line + swst_mult x (max - min SWST code)), bytes
     * from one line to the next, range samples per SWST code step (4 for ERS IM). ASARF_RAW_ASAR_FBAQ4: line_bytes
     * is the record pitch and swst_mult 1 (samples_per_line unused). */
    int samples_per_line;
    long line_bytes;
    int swst_mult;
} asarf_focus_params;

typedef struct {
    asarf_iq_results iq;
    double dc_poly[3];
    int n_blocks;
    double block_freq[64];
    double block_x[64];
    int nrv;                 /* SLC range samples */
    int crop_front, crop_back, crop_near, crop_far; /* results window (main_flow.cc:155-201) */
    int imp_rows;
    /* Azimuth core of the last execute: 1 = fused forward pass / RCMC / matched filter / inverse pass kernel,
     * 0 = separate kernels (no second column pass, tile beyond LDS, or RCMC shift bound above the halo). */
    int azimuth_core_fused;
    int rcmc_shift_bound; /* device bound on the integer RCMC shift (pixels); 0 when not computed */
} asarf_focus_results;

typedef struct asarf_plan asarf_plan;

/* chirp_iq: host, nrp complex floats. Allocates the plan's device buffers (image, tables, scratch). */
int asarf_plan_create(asarf_ctx* ctx, const asarf_focus_params* params, const float* chirp_iq, asarf_plan** out);
int asarf_plan_destroy(asarf_plan* plan);
/* Device bytes held by the plan. */
size_t asarf_plan_device_bytes(const asarf_plan* plan);
/* Runs the stage. raw: device, na*nr I/Q byte pairs. On return (asynchronously) *slc views the plan-owned SLC
 * (x_size = nrv, y_size = na); imp_out (device, may be NULL when params.imp == 0) receives up to
 * asarf_plan_imp_capacity rows x gr_range_size floats. */
int asarf_plan_execute(asarf_plan* plan, const void* raw, asarf_image* slc, float* imp_out);
/* asarf_plan_execute for a plan of ASARF_RAW_ERS_LINES: lines (device) as described there, swst (device, na int16
 * SWST codes, one per line) and their minimum; the NaN fill and the padded copy of main_flow.cc:530-583 and
 * main.cc:233-239 happen in the statistics and range-compression loads, with no intermediate image. */
int asarf_plan_execute_lines(asarf_plan* plan, const void* lines, const short* swst, int min_swst, asarf_image* slc,
                             float* imp_out);
/* asarf_plan_execute for a plan of ASARF_RAW_ASAR_FBAQ4: records (device) as described there, swst (device, na int16
 * codes) and their minimum, i_lut / q_lut (device, the instrument file's 4096-entry FBAQ-4 tables, see
 * asarf_ins_fbaq4_tables). ConvertAsarImBlocksToComplex with the NaN fill and padded copy of main_flow.cc:530-583
 * and main.cc:233-239, fused into the statistics (FP64 sums: the samples are table values) and range-compression
 * loads, with no intermediate image. */
int asarf_plan_execute_fbaq4(asarf_plan* plan, const void* records, const short* swst, int min_swst, const float* i_lut,
                             const float* q_lut, int lut_items, asarf_image* slc, float* imp_out);
int asarf_plan_imp_capacity(const asarf_plan* plan, int* rows, int* cols);
/* Copies the estimator outputs to the host (synchronises). */
int asarf_plan_results(asarf_plan* plan, asarf_focus_results* results);
/* Per-kernel timing of the last execute (HIP events on the plan's stream; enable before executing). */
int asarf_plan_set_timing(asarf_plan* plan, int enable);
int asarf_plan_kernel_times(asarf_plan* plan, const char** names, float* ms, int capacity, int* count);

/* ---------------------------------------------------------------------------- split-precision transforms */
/* No counterpart in the reference (whose transforms are cuFFT FP32, cufft_plan.cc:34-91): FFTs whose radix-16
 * passes run as batched 16-point DFT GEMMs on the matrix cores with bfloat16 hi/lo operands and FP32 sums. */
enum { ASARF_FFT_FP32 = 0, ASARF_FFT_SPLIT_BF16 = 1 };
/* As asarf_fft_c2c with split-bf16 kernels: rows of 8192; columns of 128, 256 or a product of the two. */
int asarf_fft_c2c_split_bf16(asarf_ctx* ctx, void* data, int n, int batch, long dist, int layout, int direction,
                             float scale);
/* Transform precision of a plan: ASARF_FFT_SPLIT_BF16 needs nrp = 8192 and an azimuth length that splits into
 * 128/256 passes (e.g. 32768 = 128 x 256); otherwise ASARF_EUNSUPPORTED. Default ASARF_FFT_FP32. */
int asarf_plan_set_fft(asarf_plan* plan, int precision);

/* ------------------------------------------------------------------------------- autotuner and batches */
/* No counterpart in the reference. Measures the candidate schedules of one transform length on the device and
 * records the fastest in the context; plans created afterwards use it. ROWS: range compression of `batch`
 * lines of length n (the compile-time variants where the length has them, runtime radix orders); *choice = -1 - v
 * for compile-time variant v, else the runtime candidate index. COLUMNS: forward azimuth passes of n x batch; *choice = n1 of
 * the chosen split (n = single pass). *best_ms = its time, *n_candidates = candidates measured. Synchronises. */
int asarf_autotune_fft(asarf_ctx* ctx, int n, int layout, int batch, int reps, int* choice, float* best_ms,
                       int* n_candidates);
/* Autotuner cache file (no counterpart in the reference, whose transforms are cuFFT plans): keeps the tuner's
 * choices across processes. Save writes the context's choices (text: a header with the device architecture, then
 * one line per row or column length); load reads such a file into a context on a device of the same architecture
 * and returns the number of choices in *n_loaded. A malformed file, or one tuned on another architecture, is
 * ASARF_EINVAL and changes nothing (a malformed file is rejected before any device call). */
int asarf_tune_save(asarf_ctx* ctx, const char* path);
int asarf_tune_load(asarf_ctx* ctx, const char* path, int* n_loaded);
/* Transform lengths with compile-time kernels (no counterpart in the reference). 1 when azimuth sub-transforms of
 * length n (a single column pass, or either pass of a two-pass split) run a compile-time kernel: every
 * 2^a 3^b 5^c 7^d from 16 to 400; else 0. */
int asarf_fft_static_column_length(int n);
/* A range transform length for nr + nc samples (pass as asarf_focus_params.nrp): the smallest length with a
 * compile-time range-compression kernel (6144, 6400, 8192) that is >= asarf_fft_padding_size(n) and at most 1.5
 * times it, else asarf_fft_padding_size(n). The range-compressed lines are the same as with the reference's
 * padding up to FP32 rounding: the transform is a zero-padded linear correlation, and the kept samples never see
 * the wrap. The azimuth length is not open to this (its padding decides which lines wrap into the aperture). */
int asarf_fft_fast_range_size(int n);

/* Many blocks of uneven size in one batch: each block gets its own plan on one of n_streams streams (round
 * robin; each stream has its own context that inherits ctx's autotuner choices). SEQUENTIAL runs the blocks
 * one after another, waiting for each; GRAPH captures all blocks once (fork-join over the streams) and replays
 * the graph with one launch per call (asynchronous: asarf_batch_sync waits). Raw inputs and IMP outputs are
 * device buffers the caller keeps alive. */
enum { ASARF_BATCH_SEQUENTIAL = 0, ASARF_BATCH_GRAPH = 1 };
typedef struct asarf_batch asarf_batch;
int asarf_batch_create(asarf_ctx* ctx, int n_streams, asarf_batch** out);
int asarf_batch_add(asarf_batch* batch, const asarf_focus_params* params, const float* chirp_iq, const void* raw,
                    float* imp_out, int* index);
int asarf_batch_execute(asarf_batch* batch, int mode);
int asarf_batch_sync(asarf_batch* batch);
int asarf_batch_size(const asarf_batch* batch);
int asarf_batch_slc(asarf_batch* batch, int index, asarf_image* slc);
int asarf_batch_results(asarf_batch* batch, int index, asarf_focus_results* results);
int asarf_batch_destroy(asarf_batch* batch);

/* ------------------------------------------------------------------------------ learned RCMC interpolator */
/* No counterpart in the reference: an 8-tap interpolator whose complex taps are produced by a small network
 * (MLP 2 -> 64 -> 64 -> 16, tanh, bfloat16 weights, FP32 sums) from the fractional RCMC shift d and the local
 * migration slope, trained by this library's own kernels on pairs produced by the 16-tap sinc RCMC
 * (range_doppler_algorithm.cu:107-195), with the 16-tap sinc kept as fallback and error judge.
 * A training pair is ASARF_RCMC_PAIR_FLOATS floats: 2 d, slope, 8 complex samples at cx-4..cx+3 and the
 * complex 16-tap target, samples and target divided by the samples' RMS. */
#define ASARF_RCMC_PAIR_FLOATS 20
enum { ASARF_RCMC_SINC16 = 0, ASARF_RCMC_LEARNED = 1, ASARF_RCMC_SINC8 = 2 };
typedef struct asarf_rcmc_net asarf_rcmc_net;

/* Glorot-uniform weights from the seed, zero biases, fresh Adam state. */
int asarf_rcmc_net_create(asarf_ctx* ctx, unsigned long long seed, asarf_rcmc_net** out);
int asarf_rcmc_net_destroy(asarf_rcmc_net* net);
int asarf_rcmc_net_num_params(void);
/* FP32 master parameters (host arrays of asarf_rcmc_net_num_params() floats). set resets the Adam state. */
int asarf_rcmc_net_get_params(asarf_rcmc_net* net, float* out, int count);
int asarf_rcmc_net_set_params(asarf_rcmc_net* net, const float* in, int count);
/* Sum over the pairs (device) of the gradient of |prediction - target|^2 with respect to the parameters, the
 * forward pass using the bfloat16-rounded parameters; *loss = sum of the squared errors. Synchronises. */
int asarf_rcmc_net_grad(asarf_rcmc_net* net, const float* pairs, long n_pairs, float* grad_out, double* loss);
/* Adam (beta 0.9/0.999), learning rate lr with cosine decay to lr/100 over `steps` steps of `batch` pairs
 * (a multiple of 64) taken in order from the device pairs. loss_history (host, `steps` floats, may be NULL)
 * receives each step's mean batch loss. Synchronises. */
int asarf_rcmc_net_train(asarf_rcmc_net* net, const float* pairs, long n_pairs, int batch, int steps, float lr,
                         float* loss_history);
/* Evaluates the network at the centres of nd x ns (d, slope) cells into the tap table, then judges it on the
 * validation pairs: per judge cell (judge_nd x ns, nd a multiple of judge_nd) the RMS error of the table taps
 * and of a plain 8-tap Blackman sinc against the 16-tap target; cells with at least 16 pairs and an RMS error
 * <= threshold use the table, the rest fall back to the 16-tap sinc. judge_rms / judge_sinc8_rms (host,
 * judge_nd*ns floats, may be NULL) receive the errors; *n_judge_ok the number of accepted cells. */
int asarf_rcmc_net_build_table(asarf_rcmc_net* net, const float* val_pairs, long n_val, int nd, int ns, int judge_nd,
                               float threshold, float* judge_rms, float* judge_sinc8_rms, int* n_judge_ok);
/* Runs the plan's front half (statistics, range compression, azimuth FFT) on raw and writes n_pairs training
 * pairs (device, n_pairs x ASARF_RCMC_PAIR_FLOATS floats) drawn at seeded random (bin, column) positions. */
int asarf_plan_rcmc_pairs(asarf_plan* plan, const void* raw, long n_pairs, unsigned long long seed, float* pairs);
/* Selects the plan's RCMC interpolator (ASARF_RCMC_*); the learned one needs a net with a built table, which
 * must outlive the plan's use of it. When the table's last build accepted no judge cell the learned path is the
 * 16-tap sinc everywhere, and the plan runs that kernel without the per-pixel table lookups. */
int asarf_plan_set_rcmc(asarf_plan* plan, int interpolator, const asarf_rcmc_net* net);

/* --------------------------------------------------------------------------- level-0 host utilities (host) */
/* mjd (envisat_types.h:37-75): days since 2000-01-01 (negative before), whole seconds of the day, microseconds. */
typedef struct {
    int days;
    unsigned seconds;
    unsigned micros;
} asarf_mjd;
/* mjd ordering (envisat_types.h:42-73): -1, 0 or 1 as a is before, equal to or after b (both non-NULL; NULL
 * compares as equal and sets the last error). */
int asarf_mjd_compare(const asarf_mjd* a, const asarf_mjd* b);
/* MjdAddMicros (envisat_utils.h:74-78); micros may be negative. */
int asarf_mjd_add_micros(const asarf_mjd* in, long long micros, asarf_mjd* out);
/* CounterGap (parse_util.h:33-45): steps from last to current of a counter that wraps after limit
 * (MaxValueForBits: limit = 2^bits - 1). */
unsigned long long asarf_counter_gap(unsigned long long last, unsigned long long current, unsigned long long limit);

/* Patc (patc.h): the ERS time-correlation record. */
typedef struct {
    unsigned short cyclic_counter;
    char mission[2];
    unsigned orbit_number;
    int epoch_1950;      /* days since 1950-01-01 */
    int milliseconds;    /* of that day */
    unsigned sbt_counter;
    unsigned sbt_period; /* ns per SBT tick */
} asarf_patc;
/* ParsePatc (patc.cc:25-57) from the 54 bytes of a PATC file. */
int asarf_parse_patc(const unsigned char* bytes, long n_bytes, asarf_patc* out);
/* A registered PATC: its time point in microseconds since 2000-01-01, SBT counter and period. */
typedef struct {
    long long time_us;
    unsigned sbt_counter;
    unsigned sbt_period_ns;
} asarf_patc_reference;
/* RegisterPatc (ers_sbt.cc:40-63): validates (no negative fields, not before 1991) into a reference the caller
 * keeps (the reference holds it in a process-wide variable). */
int asarf_patc_register(const asarf_patc* patc, asarf_patc_reference* out);
/* AdjustIspSensingTime (ers_sbt.cc:65-86): ISP time from the SBT counter; ref == NULL leaves it unchanged (no PATC
 * registered). */
int asarf_adjust_isp_sensing_time(const asarf_patc_reference* ref, asarf_mjd* isp_time, unsigned sbt, int overflow);

/* ProductTypes / Instrument (asar_constants.h:33-35), same order. */
enum {
    ASARF_PRODUCT_SAR_IM0 = 0,
    ASARF_PRODUCT_ASA_IM0 = 1,
    ASARF_PRODUCT_SAR_IMS = 2,
    ASARF_PRODUCT_SAR_IMP = 3,
    ASARF_PRODUCT_ASA_IMS = 4,
    ASARF_PRODUCT_ASA_IMP = 5,
    ASARF_PRODUCT_UNIDENTIFIED = 6
};
enum { ASARF_INSTRUMENT_SAR = 0, ASARF_INSTRUMENT_ASAR = 1 };
/* GetProductTypeFrom (asar_constants.cc:58-66): by the product name's prefix. */
int asarf_product_type_from_name(const char* product_name);
/* TryDetermineTargetProductFrom (asar_constants.cc:68-90): target "IMS" / "IMP" (any case); ASARF_EINVAL otherwise. */
int asarf_target_product(int in_product, const char* target, int* out_product);
/* GetInstrumentFrom (asar_constants.cc:92-102); ASARF_EINVAL for detected and unidentified products. */
int asarf_instrument(int product, int* out_instrument);

/* Echo metadata of one ASAR IM level-0 packet (EchoMeta, envisat_im_parse.cc:54-80, plus the packet header). */
typedef struct {
    asarf_mjd isp_sensing_time, fep_time;
    unsigned short isp_length, packet_id, sequence_count, packet_length;
    unsigned char segmentation_flags, mode_id, antenna_beam_set_no, comp_ratio;
    unsigned char echo_flag, noise_flag, cal_flag, cal_type;
    unsigned short cycle_count;
    unsigned mode_count;
    unsigned long long onboard_time;
    unsigned short pri_code, swst_code, echo_window_code;
    unsigned char upconverter_raw, downconverter_raw, tx_pol, rx_pol, cal_row_number;
    unsigned short tx_pulse_code;
    unsigned char beam_adj_delta, chirp_pulse_bw_code, aux_tx_monitor_level;
    unsigned short resampling_factor;
    int sample_bytes; /* packet length - 29 */
    int samples;      /* 63 per whole 64-byte block, length - 1 of a partial one */
    int dsr_bytes;    /* 39 + FEP ISP length: the offset of the next data set record */
} asarf_asar_echo;
/* One ASAR IM level-0 data set record (envisat_im_parse.cc:433-600): its echo metadata, and for an echo packet
 * (record != NULL) its sample blocks as asarf_ingest_asar_fbaq4's record (2-byte little-endian length, bytes,
 * zeros to record_bytes). ASARF_EINVAL for a data field header length other than 29 or a mode other than IM. */
int asarf_parse_asar_im_packet(const unsigned char* packet, long n_bytes, asarf_asar_echo* meta,
                               unsigned char* record, long record_bytes);

/* Echo metadata of one ERS IM level-0 packet (EchoMeta of ers_im_parse.cc). */
typedef struct {
    asarf_mjd isp_sensing_time; /* as recorded, before the SBT correction */
    unsigned short isp_length;
    unsigned data_record_number;
    unsigned char packet_counter, subcommutation_counter, obrc_orbit, activity_task, sample_flags;
    unsigned sbt;               /* satellite binary time (1/256 s ticks) */
    unsigned image_format_counter;
    unsigned short swst_code, pri_code;
} asarf_ers_echo;
/* One 11498-byte ERS IM level-0 data set record (ers_im_parse.cc:366-548): its metadata, and (line != NULL) the
 * 11232 bytes of its 5616 (I, Q) pairs as asarf_ingest_ers_im's line. ASARF_EINVAL for an ISP length other than
 * 11465 or an auxiliary section not starting with 0xAA. */
int asarf_parse_ers_im_packet(const unsigned char* packet, long n_bytes, asarf_ers_echo* meta, unsigned char* line);

/* ProductHeader::Get (envisat_mph_sph_parser.h:41-77): the value of KEY in '\n'-separated KEY=value text (MPH,
 * SPH), one pair of surrounding quotes removed, NUL-terminated into value (truncated to capacity - 1). */
int asarf_header_get(const char* text, long n_bytes, const char* key, char* value, int capacity);
/* DSD_lvl0 (envisat_mph_sph_parser.h:30-38). Names are the descriptor's lines as they stand. */
typedef struct {
    char ds_name[64], ds_type[16], filename[96];
    unsigned long long ds_offset, ds_size, num_dsr, dsr_size;
} asarf_dsd;
/* ParseMdsr (envisat_lvl0_parser.cc:98-102): the measurement data set descriptor of a level-0 file (the first
 * complete one of the four at the end of the SPH). Numbers are read as the reference reads them: from the first
 * digit 1-9 after '=', so a sign is dropped ("-0000000001" reads as 1). */
int asarf_level0_mdsr(const unsigned char* file, long n_bytes, asarf_dsd* out);

/* ------------------------------------------------------------------------- auxiliary files (host) */
/* L1ProductMetadata of an orbit source (doris_orbit.h, filled by doris_orbit.cc:260-304). */
typedef struct {
    char orbit_name[64];       /* MPH PRODUCT */
    char vector_source[3];     /* "DP" (DORIS PRECISE), "RO" (REAPER ORBIT) or "??", from SPH DS_NAME */
    char phase;                /* first character of MPH PHASE */
    unsigned cycle;            /* MPH CYCLE */
    double delta_ut1;          /* s */
    unsigned abs_orbit;
    unsigned quality;
    asarf_mjd state_vector_time;
    double pos[3], vel[3];     /* m, m/s */
} asarf_doris_meta;
/* Parsable::CreateFromFile + CreateOrbitInfo (doris_orbit.cc:82-176): the state vectors of a DOR_VOR_AX file (its
 * bytes, 206606 for ENVISAT or 233954 for ERS) from 5 minutes before start to 5 minutes after stop, as asarf_osv
 * with t in seconds since 2000-01-01 (and, times != NULL, the exact times); fewer than 10 is an error (a corrupt
 * file for the window). *meta (may be NULL): the L1 metadata of doris_orbit.cc:260-304. ASARF_EWORKSPACE when more
 * than capacity vectors are kept (*n_out stays 0). */
int asarf_doris_orbit(const unsigned char* file, long n_bytes, const asarf_mjd* start, const asarf_mjd* stop,
                      asarf_osv* osv, asarf_mjd* times, int capacity, int* n_out, asarf_doris_meta* meta);
/* The validity window in a DORIS orbit file name (doris_orbit.cc:178-246): DOR_VOR_AXVF-P<created>_<hhmmss>_
 * <start ymd>_<hhmmss>_<stop ymd>_<hhmmss>, or the ERS REAPER form with one more HHMMSS item. ASARF_EINVAL when the
 * name is not such a file name (the reference skips those). */
int asarf_doris_filename_window(const char* name, asarf_mjd* valid_start, asarf_mjd* valid_stop);
/* FindCONFile (envisat_aux_file.cc:41-65): process_gain_ims[7] and process_gain_imp[7] (swaths IS1-IS7) of a
 * 5721-byte ASAR processor configuration file. */
int asarf_con_process_gains(const unsigned char* con, long n_bytes, float* gain_ims, float* gain_imp);
/* GetXca (envisat_aux_file.cc:99-131): the image-mode VV scaling factors (SLC and precision, swaths IS1-IS7) of a
 * 28177-byte external calibration file; ASARF_EINVAL for a GADS DSR length other than 26552 or 0. */
int asarf_xca_scaling_factors(const unsigned char* xca, long n_bytes, float* slc_vv, float* precision_vv);
/* The product calibration constant of main.cc:336-337 (detected: gain_imp / sqrt(precision_vv)) and main.cc:384-388
 * (SLC: gain_ims / sqrt(slc_vv), / 8 for ERS), in FP64; it scales the statistics and the conditioned records. */
int asarf_calibration_constant(int detected, int ers, float process_gain, float scaling_factor, double* out);
/* What the level-0 flow reads of an instrument characterisation file for one image-mode swath (main_flow.cc:299-300,
 * 389-434): radar frequency and sampling rate, the nominal chirp, the range gate bias and the PRI count before the
 * sampling window (mode_timelines[0].r_values[swath]). */
typedef struct {
    float radar_frequency, radar_sampling_rate, range_gate_bias;
    float chirp_amplitude[4], chirp_phase[4], chirp_duration;
    unsigned short swst_rank;
} asarf_ins_swath;
int asarf_ins_swath_params(const unsigned char* ins, long n_bytes, int swath_idx, asarf_ins_swath* out);
/* The radar timing of main_flow.cc:389-460 (no range recalibration): chirp rate, duration, samples and bandwidth;
 * PRF and line time; two-way slant range time and slant range to the first sample; wavelength and range spacing. */
typedef struct {
    double prf, line_time_interval, two_way_slant_range_time, slant_range_first, wavelength, range_spacing;
    double chirp_kr, chirp_duration, chirp_bandwidth;
    int chirp_samples;
} asarf_timing;
int asarf_radar_timing(const asarf_ins_swath* ins, int ers, int pri_code, int min_swst, int swst_mult,
                       asarf_timing* out);

/* Summary-quality thresholds and expected values of the processor configuration file (envisat_lvl0_parser.cc:211-233). */
typedef struct {
    float thresh_chirp_broadening, thresh_chirp_sidelobe, thresh_chirp_islr;
    float thresh_input_mean, exp_input_mean, thresh_input_std_dev, exp_input_std_dev;
    float thresh_dop_cen, thresh_dop_amb;
    float thresh_output_mean, exp_output_mean, thresh_output_std_dev, exp_output_std_dev;
    float thresh_input_missing_lines, thresh_input_gaps;
    unsigned lines_per_gap;
} asarf_sq_thresholds;
/* The thresholds of a 5721-byte configuration file; the expected output mean / std dev are the IMP ones when
 * detected, else the IMS ones. */
int asarf_con_summary_quality(const unsigned char* con, long n_bytes, int detected, asarf_sq_thresholds* out);

/* ----------------------------------------------------------------------------------- level-1 header (host) */
/* What ConstructEnvisatFileHeader (envisat_lvl1_writer.cc:326-570) reads of the product's metadata. Times are mjd
 * (days since 2000-01-01); geometry->first_line_time and the orbit times are seconds since 2000-01-01, the same
 * instants as first_line_time. For a detected product the geometry describes the ground-range image (range_size
 * ground samples, range_spacing the ground spacing) and grsr5 is its slant-to-ground polynomial. */
typedef struct {
    const char* l0_product_name;   /* the level-0 product's name: characters 7 and 9 become S / P and 1 */
    int detected;                  /* 1: precision image (IMP, uint16 records), 0: SLC (IMS, int16 I/Q records) */
    const char* acquisition_station, *processing_station, *software_version;
    asarf_mjd processing_time;     /* the reference takes the wall clock; here the caller's, for repeatable files */
    asarf_mjd sensing_start, sensing_stop, first_line_time, last_line_time;
    const asarf_doris_meta* orbit; /* MPH orbit fields (NULL: blank) */
    int rel_orbit;
    asarf_mjd utc_sbt_time;
    unsigned sat_binary_time, clock_step;
    asarf_mjd leap_utc;
    int leap_sign;
    char leap_err;
    int product_err;
    const char* swath;             /* "IS1" .. "IS7" */
    int ascending;
    const char* polarization;      /* "V/V", ... */
    const char* echo_comp_method, *echo_comp_ratio, *init_cal_method, *init_cal_ratio, *per_cal_method,
        *per_cal_ratio, *noise_method, *noise_ratio;  /* NULL: "NONE" / blank */
    const char* configuration_file, *instrument_file, *external_characterization_file, *external_calibration_file,
        *orbit_dataset_name;       /* referenced auxiliary files ("" or NULL characterization file: not used) */
    const asarf_geometry* geometry;
    double azimuth_spacing, two_way_slant_range_time, range_sampling_rate;
    int chirp_window;              /* the range chirp was Hamming windowed */
    double chirp_alpha, chirp_kr, azimuth_window_alpha;
    const double* grsr5;           /* detected products: ground-to-slant polynomial (NULL for SLC) */
    unsigned first_sbt;
    asarf_mjd first_mjd;
    unsigned short first_swst_code, last_swst_code, pri_code, tx_pulse_len_code, tx_bw_code, echo_win_len_code,
        up_code, down_code, resamp_code, beam_adj_code, beam_set_num_code, tx_monitor_code;
    unsigned swst_changes, swst_rank;
    float tx_monitor_value;        /* the converted auxiliary TX monitor level (the reference's 260-entry table) */
    asarf_iq_results iq;
    double statistics[4];          /* output mean I, Q (or amplitude), std dev I, Q, calibrated */
    asarf_sq_thresholds sq;
    int mds_records, mds_record_size;
} asarf_l1_inputs;
/* Bytes of the header (MPH through the geolocation grid) for an SLC (0) or detected (1) product. */
long asarf_l1_header_size(int detected);
/* ConstructEnvisatFileHeader: MPH (1247 bytes), SPH with 18 DSDs (6099), summary quality ADS (170), main processing
 * parameters ADS (2009), Doppler centroid ADS (55), [slant-to-ground ADS (55)], 12 geolocation grid ADSRs (521 each),
 * as the bytes before the MDS records. ASARF_EWORKSPACE when capacity is short; ASARF_EINVAL when a value does not
 * fit its header field. */
int asarf_l1_header(const asarf_l1_inputs* in, unsigned char* out, long capacity, long* n_out);

/* FetchEnvisatL0ImForecastMeta (envisat_im_parse.cc:297-384): the ASAR IM records to focus between the sensing
 * times start and stop. *before_start / *after_stop: in, the lines allowed before the start and after the stop;
 * out, the lines kept there. offsets (relative to mds) and ISP times of the n_out kept records, the first an echo.
 * The leading calibration records are dropped by their own flags (the reference tests the last record read). */
int asarf_asar_forecast(const unsigned char* mds, long mds_bytes, long num_dsr, const asarf_mjd* start,
                        const asarf_mjd* stop, long* before_start, long* after_stop, long* offsets, asarf_mjd* times,
                        long capacity, long* n_out);

/* FetchErsL0ImForecastMeta (ers_im_parse.cc:228-331): the ERS IM records (11498 bytes each) to focus between
 * start and stop, ISP times SBT-corrected with patc (NULL: as recorded), a gap in the data record number filled by
 * repeating the previous entry (at most 900 missing records; offsets repeat). In/out counts as in
 * asarf_asar_forecast. */
int asarf_ers_forecast(const unsigned char* mds, long mds_bytes, long num_dsr, const asarf_mjd* start,
                       const asarf_mjd* stop, const asarf_patc_reference* patc, long* before_start, long* after_stop,
                       long* offsets, asarf_mjd* times, long capacity, long* n_out);

/* The FBAQ-4 reconstruction tables of an ASAR instrument characterisation file (ASA_INS_AX, 173273 bytes;
 * envisat_aux_file.cc:67-95, envisat_aux_file.h:143-149): 4096 I and 4096 Q floats for asarf_ingest_asar_fbaq4. */
int asarf_ins_fbaq4_tables(const unsigned char* ins, long n_bytes, float* i_lut, float* q_lut);

#ifdef __cplusplus
}
#endif
#endif
