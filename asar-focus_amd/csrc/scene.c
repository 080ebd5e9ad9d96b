/*
 * Seeded synthetic raw-scene generator for ERS/ASAR-like Imaging Mode blocks (host C, OpenMP).
 *
 * The reference focuses real Level-0 products only; this generator supplies the synthetic raw blocks that stand in
 * for them (uint8 5-bit offset-binary I/Q as produced by the ERS ingest, envisat_format_kernels.cu:116-158, or
 * complex int8). The echo model follows the reference's sign conventions:
 *   transmitted pulse exp(+i pi Kr t^2), t = (i - Nc/2)/fs (sar/sar_chirp.cc:80),
 *   azimuth phase exp(-i 4 pi (R(eta) - R0)/lambda), R(eta) = sqrt(R0^2 + Vr^2 (eta - eta0)^2),
 *   echo leading edge at range sample (R(eta) - R_first)/dr, so a target at range pixel x focuses at x,
 *   Doppler extent az_band_frac*PRF centred on the local Doppler centroid (range_doppler_algorithm.cu:122-140).
 * Distributed clutter is compound Gaussian (gamma texture) with an AR(1) azimuth correlation whose lag-1 phase is
 * the local Doppler centroid, so the ACCC estimator sees the same Doppler signature as for real echoes.
 * Random numbers come from Philox-4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11), a counter-based generator, so every
 * sample is a pure function of (seed, line, sample) regardless of threading.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "scene.h"

/* ---------------- Philox-4x32-10 ---------------- */
static inline void philox_round(uint32_t* c, const uint32_t* k) {
    const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
    const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
    const uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
    const uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
    const uint32_t n0 = hi1 ^ c[1] ^ k[0];
    const uint32_t n1 = lo1;
    const uint32_t n2 = hi0 ^ c[3] ^ k[1];
    const uint32_t n3 = lo0;
    c[0] = n0;
    c[1] = n1;
    c[2] = n2;
    c[3] = n3;
}

void asarf_philox4x32_10(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c[4] = {ctr[0], ctr[1], ctr[2], ctr[3]};
    uint32_t k[2] = {key[0], key[1]};
    for (int r = 0; r < 10; r++) {
        philox_round(c, k);
        k[0] += 0x9E3779B9u;
        k[1] += 0xBB67AE85u;
    }
    memcpy(out, c, sizeof(c));
}

static inline double u01(uint32_t v) { return ((double)v + 0.5) * (1.0 / 4294967296.0); }

/* Two independent N(0,1) values for (stream, a, b) by Box-Muller on one Philox block. */
static inline void gauss2(uint64_t seed, uint32_t stream, uint32_t a, uint32_t b, double* g0, double* g1,
                          double* e0, double* e1) {
    const uint32_t ctr[4] = {a, b, stream, 0x5A5Au};
    const uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    asarf_philox4x32_10(ctr, key, o);
    const double r = sqrt(-2.0 * log(u01(o[0])));
    const double t = 2.0 * M_PI * u01(o[1]);
    *g0 = r * cos(t);
    *g1 = r * sin(t);
    if (e0) *e0 = u01(o[2]);
    if (e1) *e1 = u01(o[3]);
}

static double dc_at(const asarf_scene_params* p, double x) {
    const int nrv = p->nr - p->nc;
    const double span = nrv > 1 ? (double)(nrv - 1) : 1.0;
    return p->dc_near + (p->dc_far - p->dc_near) * x / span;
}

int asarf_scene_targets(const asarf_scene_params* p, double* tx, double* ty, int max_targets) {
    const int nrv = p->nr - p->nc;
    int n = 0;
    if (p->n_targets_az > 0 && p->n_targets_rg > 0) {
        for (int j = 0; j < p->n_targets_az; j++)
            for (int i = 0; i < p->n_targets_rg; i++) {
                const double fr = p->n_targets_rg > 1 ? (double)i / (p->n_targets_rg - 1) : 0.5;
                const double fa = p->n_targets_az > 1 ? (double)j / (p->n_targets_az - 1) : 0.5;
                if (n < max_targets) {
                    tx[n] = round((p->tgt_rg_lo + (p->tgt_rg_hi - p->tgt_rg_lo) * fr) * nrv);
                    ty[n] = round((p->tgt_az_lo + (p->tgt_az_hi - p->tgt_az_lo) * fa) * p->na);
                }
                n++;
            }
    }
    for (int t = 0; t < p->n_targets_random; t++) {
        double g0, g1, e0, e1;
        gauss2(p->seed, 7u, (uint32_t)t, 0xFFFFFFFFu, &g0, &g1, &e0, &e1);
        if (n < max_targets) {
            tx[n] = round((p->tgt_rg_lo + (p->tgt_rg_hi - p->tgt_rg_lo) * e0) * nrv);
            ty[n] = round((p->tgt_az_lo + (p->tgt_az_hi - p->tgt_az_lo) * e1) * p->na);
        }
        n++;
    }
    return n;
}

/* Adds the echoes of all targets to line y (complex double, length nr). */
static void add_targets_line(const asarf_scene_params* p, const double* tx, const double* ty, int nt, int y,
                             double* line) {
    const double dr = 299792458.0 / (2.0 * p->fs);
    const double eta = (double)y / p->prf;
    const int half_nc = p->nc / 2;
    for (int t = 0; t < nt; t++) {
        const double r0 = p->r_first + tx[t] * dr;
        const double ka = 2.0 * p->vr * p->vr / (p->wavelength * r0);
        const double fdc = dc_at(p, tx[t]);
        const double eta0 = ty[t] / p->prf;
        const double bw = 0.5 * p->az_band_frac * p->prf;
        /* Doppler f = -Ka (eta - eta0); keep f in [fdc - bw, fdc + bw] */
        const double de = eta - eta0;
        const double f = -ka * de;
        if (f < fdc - bw || f > fdc + bw) continue;
        const double v2d2 = p->vr * p->vr * de * de;
        const double rr = sqrt(r0 * r0 + v2d2);
        const double dR = v2d2 / (rr + r0); /* R - R0, cancellation-free */
        const double az_ph = -4.0 * M_PI * dR / p->wavelength;
        const double x0 = (rr - p->r_first) / dr;
        int k0 = (int)ceil(x0);
        int k1 = (int)floor(x0 + (double)p->nc - 1e-9);
        if (k0 < 0) k0 = 0;
        if (k1 > p->nr - 1) k1 = p->nr - 1;
        for (int k = k0; k <= k1; k++) {
            const double u = ((double)k - x0 - (double)half_nc) / p->fs;
            const double ph = az_ph + M_PI * p->kr * u * u;
            line[2 * k] += cos(ph);
            line[2 * k + 1] += sin(ph);
        }
    }
}

int asarf_scene_generate(const asarf_scene_params* p, void* raw_out, float* float_out) {
    if (!p || p->na <= 0 || p->nr <= 0 || p->nc <= 0 || p->nc >= p->nr) return -1;
    const int na = p->na, nr = p->nr;
    const int nt = asarf_scene_targets(p, NULL, NULL, 0);
    double* tx = (double*)malloc(sizeof(double) * (nt > 0 ? nt : 1));
    double* ty = (double*)malloc(sizeof(double) * (nt > 0 ? nt : 1));
    asarf_scene_targets(p, tx, ty, nt);

    /* target power per sample (unit amplitude echoes), used to set the noise level for the SNR */
    const double dr = 299792458.0 / (2.0 * p->fs);
    double tgt_energy = 0.0;
    for (int t = 0; t < nt; t++) {
        const double r0 = p->r_first + tx[t] * dr;
        const double ka = 2.0 * p->vr * p->vr / (p->wavelength * r0);
        tgt_energy += (p->az_band_frac * p->prf / ka) * p->prf * (double)p->nc;
    }
    const double mean_tgt_power = tgt_energy / ((double)na * nr);
    const double noise_power = (p->snr_db > -900.0 && nt > 0) ? mean_tgt_power / pow(10.0, p->snr_db / 10.0) : 0.0;
    const double sig_n = sqrt(0.5 * noise_power);
    const double sig_c = sqrt(0.5 * p->clutter_power);
    const double shape = p->clutter_shape > 0 ? p->clutter_shape : 1.0;

    float* buf = float_out;
    float* own = NULL;
    if (!buf) {
        own = (float*)malloc(sizeof(float) * 2 * (size_t)na * nr);
        if (!own) {
            free(tx);
            free(ty);
            return -2;
        }
        buf = own;
    }
    /* clutter: compound-Gaussian, AR(1) along azimuth with lag-1 phase 2 pi f_dc(x) / PRF; column-parallel */
    if (sig_c > 0.0) {
        const double rho = p->clutter_rho > 0.0 && p->clutter_rho < 1.0 ? p->clutter_rho : 0.0;
        const double innov = sqrt(1.0 - rho * rho);
#pragma omp parallel for schedule(static)
        for (int x = 0; x < nr; x++) {
            const double ph = 2.0 * M_PI * dc_at(p, (double)x) / p->prf;
            const double rc = rho * cos(ph), rs = rho * sin(ph);
            double cr = 0.0, ci = 0.0;
            for (int y = 0; y < na; y++) {
                double g0, g1;
                gauss2(p->seed, 2u, (uint32_t)y, (uint32_t)x, &g0, &g1, NULL, NULL);
                if (y == 0 || rho == 0.0) {
                    cr = g0;
                    ci = g1;
                } else {
                    const double nr_ = rc * cr - rs * ci + innov * g0;
                    const double ni_ = rs * cr + rc * ci + innov * g1;
                    cr = nr_;
                    ci = ni_;
                }
                /* texture ~ Gamma(shape, 1/shape) per 64x64 cell (Erlang sum of exponentials) */
                const uint32_t cell = ((uint32_t)(y >> 6) << 16) ^ (uint32_t)(x >> 6);
                double tex = 0.0;
                const int ks = (int)shape;
                for (int k = 0; k < ks; k += 2) {
                    double a, b, e0, e1;
                    gauss2(p->seed, 3u, cell, (uint32_t)k, &a, &b, &e0, &e1);
                    tex += -log(e0);
                    if (k + 1 < ks) tex += -log(e1);
                }
                tex = ks > 0 ? tex / (double)ks : 1.0;
                const double s = sig_c * sqrt(tex);
                const size_t o = ((size_t)y * nr + x) * 2;
                buf[o] = (float)(s * cr);
                buf[o + 1] = (float)(s * ci);
            }
        }
    } else {
        memset(buf, 0, sizeof(float) * 2 * (size_t)na * nr);
    }
    double sum2 = 0.0;
#pragma omp parallel reduction(+ : sum2)
    {
        double* line = (double*)malloc(sizeof(double) * 2 * (size_t)nr);
#pragma omp for schedule(dynamic, 16)
        for (int y = 0; y < na; y++) {
            for (int x = 0; x < nr; x++) {
                const size_t o = ((size_t)y * nr + x) * 2;
                line[2 * x] = buf[o];
                line[2 * x + 1] = buf[o + 1];
            }
            add_targets_line(p, tx, ty, nt, y, line);
            for (int x = 0; x < nr; x++) {
                if (sig_n > 0.0) {
                    double g0, g1;
                    gauss2(p->seed, 1u, (uint32_t)y, (uint32_t)x, &g0, &g1, NULL, NULL);
                    line[2 * x] += sig_n * g0;
                    line[2 * x + 1] += sig_n * g1;
                }
                const size_t o = ((size_t)y * nr + x) * 2;
                buf[o] = (float)line[2 * x];
                buf[o + 1] = (float)line[2 * x + 1];
                sum2 += line[2 * x] * line[2 * x] + line[2 * x + 1] * line[2 * x + 1];
            }
        }
        free(line);
    }
    const double std_comp = sqrt(sum2 / (2.0 * (double)na * nr));
    const double k = std_comp > 0 ? p->quant_std_lsb / std_comp : 1.0;
    if (raw_out) {
        uint8_t* o8 = (uint8_t*)raw_out;
#pragma omp parallel for schedule(static)
        for (long i = 0; i < 2L * na * nr; i++) {
            double v = (double)buf[i] * k;
            if (p->format == ASARF_SCENE_U8_ERS) {
                v = round(15.5 + v);
                if (v < 0) v = 0;
                if (v > 31) v = 31;
                o8[i] = (uint8_t)v;
            } else {
                v = round(v);
                if (v < -127) v = -127;
                if (v > 127) v = 127;
                ((int8_t*)o8)[i] = (int8_t)v;
            }
        }
    }
    free(own);
    free(tx);
    free(ty);
    return nt;
}
