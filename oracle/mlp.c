/*
 * ORACLE (test infrastructure only) -- A21: the refinement-indicator MLP
 * 128 -> 64 -> 64 -> 1 with GELU (tanh form, Hendrycks & Gimpel 2016:
 * 0.5 z (1 + tanh(sqrt(2/pi) (z + 0.044715 z^3)))), mean-squared-error
 * loss, forward and backward.  No counterpart in the reference.
 *
 * Parameters, one flat vector of 12,481: W1[64][128] | b1[64] | W2[64][64]
 * | b2[64] | W3[64] | b3.
 *
 * quantize = 0: plain FP64 throughout (pinned by central finite differences
 * and by torch autograd on the CPU, tests/test_oracle_mlp.py).
 * quantize = 1: the GPU kernels' arithmetic model -- every GEMM operand
 * rounded to BF16 (round-to-nearest-even from FP32): inputs, weights, both
 * hidden activations and the two back-propagated deltas; biases, bias
 * gradients, pre-activations and the output stay unrounded; sums in FP64
 * here (FP32 on the GPU).
 */
#include "oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NI ORC_MLP_IN
#define NH ORC_MLP_HIDDEN

enum { OW1 = 0, OB1 = NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH, OB3 = OW3 + NH };

double orc_bf16_round(double x)
{
    float f = (float)x;
    uint32_t u;
    memcpy(&u, &f, 4);
    if ((u & 0x7fffffffu) > 0x7f800000u)
        return x; /* NaN */
    u += 0x7fffu + ((u >> 16) & 1u);
    u &= 0xffff0000u;
    memcpy(&f, &u, 4);
    return (double)f;
}

static double q(double x, int on) { return on ? orc_bf16_round(x) : x; }

static const double kC = 0.7978845608028654; /* sqrt(2/pi) */
static const double kA = 0.044715;

static double gelu(double z)
{
    const double t = tanh(kC * (z + kA * z * z * z));
    return 0.5 * z * (1.0 + t);
}

static double gelu_d(double z)
{
    const double t = tanh(kC * (z + kA * z * z * z));
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * kC * (1.0 + 3.0 * kA * z * z);
}

double orc_mlp_loss_grad(const double *params, int64_t B, const double *X, const double *y,
                         int quantize, double *pred, double *grad)
{
    double *Wq = (double *)malloc(sizeof(double) * ORC_MLP_NPARAM);
    for (int i = 0; i < ORC_MLP_NPARAM; ++i)
        Wq[i] = params[i];
    if (quantize) { /* GEMM operands: the three weight matrices */
        for (int i = OW1; i < OB1; ++i)
            Wq[i] = q(params[i], 1);
        for (int i = OW2; i < OB2; ++i)
            Wq[i] = q(params[i], 1);
        for (int i = OW3; i < OB3; ++i)
            Wq[i] = q(params[i], 1);
    }
    if (grad)
        memset(grad, 0, sizeof(double) * ORC_MLP_NPARAM);
    double loss = 0.0;
    double x[NI], z1[NH], a1[NH], z2[NH], a2[NH], d2[NH], d1[NH];
    for (int64_t r = 0; r < B; ++r) {
        for (int i = 0; i < NI; ++i)
            x[i] = q(X[r * NI + i], quantize);
        for (int o = 0; o < NH; ++o) {
            double s = 0.0;
            for (int i = 0; i < NI; ++i)
                s += Wq[OW1 + o * NI + i] * x[i];
            z1[o] = s + Wq[OB1 + o];
            a1[o] = q(gelu(z1[o]), quantize);
        }
        for (int o = 0; o < NH; ++o) {
            double s = 0.0;
            for (int i = 0; i < NH; ++i)
                s += Wq[OW2 + o * NH + i] * a1[i];
            z2[o] = s + Wq[OB2 + o];
            a2[o] = q(gelu(z2[o]), quantize);
        }
        double p = 0.0;
        for (int i = 0; i < NH; ++i)
            p += Wq[OW3 + i] * a2[i];
        p += Wq[OB3];
        if (pred)
            pred[r] = p;
        const double e = p - y[r];
        loss += e * e;
        if (!grad)
            continue;
        const double d3 = 2.0 * e / (double)B;
        for (int i = 0; i < NH; ++i)
            grad[OW3 + i] += d3 * a2[i];
        grad[OB3] += d3;
        for (int o = 0; o < NH; ++o) {
            const double d = d3 * Wq[OW3 + o] * gelu_d(z2[o]);
            grad[OB2 + o] += d;
            d2[o] = q(d, quantize);
        }
        for (int o = 0; o < NH; ++o)
            for (int i = 0; i < NH; ++i)
                grad[OW2 + o * NH + i] += d2[o] * a1[i];
        for (int i = 0; i < NH; ++i) {
            double s = 0.0;
            for (int o = 0; o < NH; ++o)
                s += Wq[OW2 + o * NH + i] * d2[o];
            const double d = s * gelu_d(z1[i]);
            grad[OB1 + i] += d;
            d1[i] = q(d, quantize);
        }
        for (int o = 0; o < NH; ++o)
            for (int i = 0; i < NI; ++i)
                grad[OW1 + o * NI + i] += d1[o] * x[i];
    }
    free(Wq);
    return B ? loss / (double)B : 0.0;
}
