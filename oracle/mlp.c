/*
 * FP64 oracle of the learned RCMC interpolator's network (TEST INFRASTRUCTURE ONLY; see oracle.h).
 *
 * The network has no counterpart in the reference: it maps (2 d, slope) to corrections of 8 complex taps (MLP
 * 2 -> 64 -> 64 -> 16, tanh hidden units, linear output) added to the plain 8-tap Blackman sinc weights; the
 * prediction is sum_j (a_j + i b_j) s_j over the pair's 8 samples and the per-pair loss |prediction - target|^2. Its training pairs come from the 16-tap sinc RCMC of
 * sar/range_doppler_algorithm.cu:107-195. Parameters are used as given (the caller passes the bfloat16-rounded
 * values the device forward pass uses); everything here is double precision.
 *
 * Pinned by: central finite differences of or_mlp_loss_grad's loss (tests/test_oracle_mlp.py), the closed form
 * of a network whose only non-zero parameters are output biases, and the sinc weights' identity at d = 0, unit
 * sum and agreement with numpy's Blackman window times numpy's sinc.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

enum { IN = 2, H = 64, OUT = 16, TAPS = 8, PAIR = 20 };
enum { OW1 = 0, OB1 = OW1 + H * IN, OW2 = OB1 + H, OB2 = OW2 + H * H, OW3 = OB2 + H, OB3 = OW3 + OUT * H,
       NP = OB3 + OUT };

int or_mlp_num_params(void) { return NP; }

/* Forward pass of one input: the 16 network outputs (tap corrections, real/imaginary interleaved). */
void or_mlp_forward(const double* w, double u0, double u1, double* out16) {
    double h1[H], h2[H];
    for (int j = 0; j < H; ++j) h1[j] = tanh(w[OW1 + 2 * j] * u0 + w[OW1 + 2 * j + 1] * u1 + w[OB1 + j]);
    for (int i = 0; i < H; ++i) {
        double a = w[OB2 + i];
        for (int j = 0; j < H; ++j) a += w[OW2 + i * H + j] * h1[j];
        h2[i] = tanh(a);
    }
    for (int k = 0; k < OUT; ++k) {
        double a = w[OB3 + k];
        for (int i = 0; i < H; ++i) a += w[OW3 + k * H + i] * h2[i];
        out16[k] = a;
    }
}

/* Plain 8-tap Blackman sinc weights at fractional shift d (taps cx-4..cx+3), unit sum; the interpolation weights
 * are these plus the network outputs. */
void or_sinc8_weights(double d, double* out8) {
    if (d == 0.0) {
        for (int j = 0; j < TAPS; ++j) out8[j] = j == 4 ? 1.0 : 0.0;
        return;
    }
    double den = 0.0;
    for (int j = 0; j < TAPS; ++j) {
        const double a = 2.0 * M_PI * (double)j / 7.0;
        const double w = (double)(float)(0.42 - 0.5 * cos(a) + 0.08 * cos(2.0 * a));
        out8[j] = ((j & 1) ? -w : w) / (d + 4.0 - (double)j);
        den += out8[j];
    }
    for (int j = 0; j < TAPS; ++j) out8[j] /= den;
}

/* Sum over n pairs of the loss and (grad != NULL) of its gradient with respect to the NP parameters. */
double or_mlp_loss_grad(const double* w, const float* pairs, long n, double* grad) {
    double loss = 0.0;
    if (grad) memset(grad, 0, sizeof(double) * NP);
    for (long s = 0; s < n; ++s) {
        const float* p = pairs + s * PAIR;
        const double u0 = p[0], u1 = p[1];
        double h1[H], h2[H], o[OUT];
        for (int j = 0; j < H; ++j) h1[j] = tanh(w[OW1 + 2 * j] * u0 + w[OW1 + 2 * j + 1] * u1 + w[OB1 + j]);
        for (int i = 0; i < H; ++i) {
            double a = w[OB2 + i];
            for (int j = 0; j < H; ++j) a += w[OW2 + i * H + j] * h1[j];
            h2[i] = tanh(a);
        }
        for (int k = 0; k < OUT; ++k) {
            double a = w[OB3 + k];
            for (int i = 0; i < H; ++i) a += w[OW3 + k * H + i] * h2[i];
            o[k] = a;
        }
        double base[TAPS];
        or_sinc8_weights(0.5 * u0, base);
        for (int j = 0; j < TAPS; ++j) o[2 * j] += base[j];
        double er = -(double)p[18], ei = -(double)p[19];
        for (int j = 0; j < TAPS; ++j) {
            const double sr = p[2 + 2 * j], si = p[3 + 2 * j];
            er += o[2 * j] * sr - o[2 * j + 1] * si;
            ei += o[2 * j] * si + o[2 * j + 1] * sr;
        }
        loss += er * er + ei * ei;
        if (!grad) continue;
        double dout[OUT], dz2[H], dz1[H];
        for (int j = 0; j < TAPS; ++j) {
            const double sr = p[2 + 2 * j], si = p[3 + 2 * j];
            dout[2 * j] = 2.0 * (er * sr + ei * si);
            dout[2 * j + 1] = 2.0 * (ei * sr - er * si);
        }
        for (int k = 0; k < OUT; ++k) {
            grad[OB3 + k] += dout[k];
            for (int i = 0; i < H; ++i) grad[OW3 + k * H + i] += dout[k] * h2[i];
        }
        for (int i = 0; i < H; ++i) {
            double a = 0.0;
            for (int k = 0; k < OUT; ++k) a += w[OW3 + k * H + i] * dout[k];
            dz2[i] = a * (1.0 - h2[i] * h2[i]);
            grad[OB2 + i] += dz2[i];
            for (int j = 0; j < H; ++j) grad[OW2 + i * H + j] += dz2[i] * h1[j];
        }
        for (int j = 0; j < H; ++j) {
            double a = 0.0;
            for (int i = 0; i < H; ++i) a += w[OW2 + i * H + j] * dz2[i];
            dz1[j] = a * (1.0 - h1[j] * h1[j]);
            grad[OB1 + j] += dz1[j];
            grad[OW1 + 2 * j] += dz1[j] * u0;
            grad[OW1 + 2 * j + 1] += dz1[j] * u1;
        }
    }
    return loss;
}
