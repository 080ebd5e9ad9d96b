/* Seeded synthetic raw-scene generator (host). See scene.c. */
#ifndef ASARF_SCENE_H
#define ASARF_SCENE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

enum { ASARF_SCENE_U8_ERS = 0, ASARF_SCENE_I8 = 1 };

typedef struct {
    int na, nr, nc;           /* azimuth lines, range samples, chirp samples */
    int format;               /* ASARF_SCENE_U8_ERS (offset binary, bias 15.5, 0..31) or ASARF_SCENE_I8 */
    double fs, kr;            /* range sampling rate [Hz], chirp FM rate [Hz/s] */
    double prf, wavelength;   /* [Hz], [m] */
    double vr, r_first;       /* effective velocity [m/s], slant range of sample 0 [m] */
    double dc_near, dc_far;   /* Doppler centroid at range pixel 0 and at pixel nr-nc-1 [Hz], linear between */
    double az_band_frac;      /* simulated Doppler extent as a fraction of PRF */
    int n_targets_az, n_targets_rg; /* regular grid of unit point targets */
    int n_targets_random;     /* additional randomly placed targets */
    double tgt_rg_lo, tgt_rg_hi, tgt_az_lo, tgt_az_hi; /* placement fractions of (nr-nc) and na */
    double snr_db;            /* white noise: mean target power / noise power; <= -900 disables */
    double clutter_power;     /* compound-Gaussian clutter power per complex sample; 0 disables */
    double clutter_shape;     /* gamma texture shape (mean 1) */
    double clutter_rho;       /* lag-1 azimuth correlation of the clutter (0 = white); the lag-1 phase is the local
                                 Doppler centroid, 2 pi f_dc / PRF, as for echoes seen through the antenna beam */
    double quant_std_lsb;     /* quantiser gain chosen so that std(I) is this many LSB */
    uint64_t seed;
} asarf_scene_params;

void asarf_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
/* Target positions (range pixel after range compression, azimuth line of zero-Doppler time). Returns count. */
int asarf_scene_targets(const asarf_scene_params* p, double* tx, double* ty, int max_targets);
/* raw_out: na*nr*2 bytes (I,Q) or NULL; float_out: na*nr*2 floats before quantisation or NULL. Returns #targets. */
int asarf_scene_generate(const asarf_scene_params* p, void* raw_out, float* float_out);

#ifdef __cplusplus
}
#endif
#endif
