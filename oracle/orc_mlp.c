/*
 * ORACLE -- test infrastructure, not product code.
 *
 * FP64 forward and backward of the warm-start network (5 -> 128 -> 128 -> 128 -> 5, GELU in
 * its tanh form on the three hidden layers, linear output) with a mean-squared-error loss.
 * The GPU trains the same network with BF16 MFMA kernels; this is the reference its gradients
 * are checked against.  The backward pass is itself pinned by central finite differences of
 * orc_mlp_loss_grad's loss (tests/test_oracle_mlp.py).
 */
#include "photic_oracle.h"

#include <math.h>
#include <stdlib.h>
#include <string.h>

#define NI ORC_MLP_IN
#define NH ORC_MLP_HID
#define NO ORC_MLP_OUT

static const double GELU_C = 0.7978845608028654; /* sqrt(2/pi) */
static const double GELU_A = 0.044715;

static double gelu(double z)
{
    return 0.5 * z * (1.0 + tanh(GELU_C * (z + GELU_A * z * z * z)));
}

static double gelu_grad(double z)
{
    double t = tanh(GELU_C * (z + GELU_A * z * z * z));
    return 0.5 * (1.0 + t) + 0.5 * z * (1.0 - t * t) * GELU_C * (1.0 + 3.0 * GELU_A * z * z);
}

typedef struct {
    const double *W1, *b1, *W2, *b2, *W3, *b3, *W4, *b4;
} mlp_view;

static mlp_view view(const double *w)
{
    mlp_view v;
    v.W1 = w;
    v.b1 = v.W1 + NH * NI;
    v.W2 = v.b1 + NH;
    v.b2 = v.W2 + NH * NH;
    v.W3 = v.b2 + NH;
    v.b3 = v.W3 + NH * NH;
    v.W4 = v.b3 + NH;
    v.b4 = v.W4 + NO * NH;
    return v;
}

/* One sample: pre-activations z1..z3, activations h1..h3, output o. */
static void fwd1(const mlp_view *v, const double *x, double *z1, double *h1, double *z2,
                 double *h2, double *z3, double *h3, double *o)
{
    for (int j = 0; j < NH; j++) {
        double s = v->b1[j];
        for (int i = 0; i < NI; i++)
            s += v->W1[j * NI + i] * x[i];
        z1[j] = s;
        h1[j] = gelu(s);
    }
    for (int j = 0; j < NH; j++) {
        double s = v->b2[j];
        for (int i = 0; i < NH; i++)
            s += v->W2[j * NH + i] * h1[i];
        z2[j] = s;
        h2[j] = gelu(s);
    }
    for (int j = 0; j < NH; j++) {
        double s = v->b3[j];
        for (int i = 0; i < NH; i++)
            s += v->W3[j * NH + i] * h2[i];
        z3[j] = s;
        h3[j] = gelu(s);
    }
    for (int j = 0; j < NO; j++) {
        double s = v->b4[j];
        for (int i = 0; i < NH; i++)
            s += v->W4[j * NH + i] * h3[i];
        o[j] = s;
    }
}

void orc_mlp_forward(const double *w, int n, const double *in, double *out)
{
    mlp_view v = view(w);
    double z1[NH], h1[NH], z2[NH], h2[NH], z3[NH], h3[NH];
    for (int s = 0; s < n; s++)
        fwd1(&v, in + s * NI, z1, h1, z2, h2, z3, h3, out + s * NO);
}

double orc_mlp_loss_grad(const double *w, int n, const double *in, const double *target,
                         double *grad)
{
    mlp_view v = view(w);
    double z1[NH], h1[NH], z2[NH], h2[NH], z3[NH], h3[NH], o[NO];
    double d1[NH], d2[NH], d3[NH], d4[NO];
    double loss = 0.0, scale = 1.0 / ((double)n * NO);
    if (grad)
        memset(grad, 0, sizeof(double) * ORC_MLP_NPARAM);
    double *gW1 = grad, *gb1 = gW1 ? gW1 + NH * NI : NULL, *gW2 = gb1 ? gb1 + NH : NULL,
           *gb2 = gW2 ? gW2 + NH * NH : NULL, *gW3 = gb2 ? gb2 + NH : NULL,
           *gb3 = gW3 ? gW3 + NH * NH : NULL, *gW4 = gb3 ? gb3 + NH : NULL,
           *gb4 = gW4 ? gW4 + NO * NH : NULL;
    for (int s = 0; s < n; s++) {
        const double *x = in + s * NI, *t = target + s * NO;
        fwd1(&v, x, z1, h1, z2, h2, z3, h3, o);
        for (int j = 0; j < NO; j++) {
            double e = o[j] - t[j];
            loss += e * e;
            d4[j] = 2.0 * e * scale;
        }
        if (!grad)
            continue;
        for (int j = 0; j < NO; j++) {
            gb4[j] += d4[j];
            for (int i = 0; i < NH; i++)
                gW4[j * NH + i] += d4[j] * h3[i];
        }
        for (int i = 0; i < NH; i++) {
            double s3 = 0.0;
            for (int j = 0; j < NO; j++)
                s3 += v.W4[j * NH + i] * d4[j];
            d3[i] = s3 * gelu_grad(z3[i]);
        }
        for (int j = 0; j < NH; j++) {
            gb3[j] += d3[j];
            for (int i = 0; i < NH; i++)
                gW3[j * NH + i] += d3[j] * h2[i];
        }
        for (int i = 0; i < NH; i++) {
            double s2 = 0.0;
            for (int j = 0; j < NH; j++)
                s2 += v.W3[j * NH + i] * d3[j];
            d2[i] = s2 * gelu_grad(z2[i]);
        }
        for (int j = 0; j < NH; j++) {
            gb2[j] += d2[j];
            for (int i = 0; i < NH; i++)
                gW2[j * NH + i] += d2[j] * h1[i];
        }
        for (int i = 0; i < NH; i++) {
            double s1 = 0.0;
            for (int j = 0; j < NH; j++)
                s1 += v.W2[j * NH + i] * d2[j];
            d1[i] = s1 * gelu_grad(z1[i]);
        }
        for (int j = 0; j < NH; j++) {
            gb1[j] += d1[j];
            for (int i = 0; i < NI; i++)
                gW1[j * NI + i] += d1[j] * x[i];
        }
    }
    return loss * scale;
}
