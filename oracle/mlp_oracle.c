/*
 * CPU oracle of the configuration picker's perceptron (the HIP forward,
 * backward and Adam kernels are checked against it).
 *
 * TEST INFRASTRUCTURE ONLY: used by tests/ as the checker, never by the
 * shipped library.  No reference counterpart (the reference tunes by
 * compile-time macros, sgemm_tune.h:6-15 / dgemm_tune.h:6-15); this is the
 * textbook two-hidden-layer ReLU network with mean-squared loss, in double.
 *
 * Layout (identical to the library's): one flat parameter array
 *   W1[in][h] b1[h] W2[h][h] b2[h] W3[h][out] b3[out]
 * with row-major weights ([input][output]); activations row-major [batch][width].
 */
#include "oracle.h"
#include <math.h>
#include <stdlib.h>
#include <string.h>

#define IN ORACLE_MLP_IN
#define HID ORACLE_MLP_HID
#define OUT ORACLE_MLP_OUT
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + IN * HID)
#define OFF_W2 (OFF_B1 + HID)
#define OFF_B2 (OFF_W2 + HID * HID)
#define OFF_W3 (OFF_B2 + HID)
#define OFF_B3 (OFF_W3 + HID * OUT)

static void layer(const double *x, int B, int in, const double *w, const double *b, int out, int relu, double *z) {
    for (int r = 0; r < B; ++r)
        for (int o = 0; o < out; ++o) {
            double acc = b[o];
            for (int i = 0; i < in; ++i) acc += x[r * in + i] * w[i * out + o];
            z[r * out + o] = (relu && acc < 0.0) ? 0.0 : acc;
        }
}

void oracle_mlp_forward(const double *p, const double *x, int B, double *y) {
    double *h1 = malloc(sizeof(double) * B * HID), *h2 = malloc(sizeof(double) * B * HID);
    layer(x, B, IN, p + OFF_W1, p + OFF_B1, HID, 1, h1);
    layer(h1, B, HID, p + OFF_W2, p + OFF_B2, HID, 1, h2);
    layer(h2, B, HID, p + OFF_W3, p + OFF_B3, OUT, 0, y);
    free(h1);
    free(h2);
}

double oracle_mlp_loss(const double *p, const double *x, const double *t, const double *mask, int B) {
    double *y = malloc(sizeof(double) * B * OUT);
    oracle_mlp_forward(p, x, B, y);
    double s = 0.0;
    for (int i = 0; i < B * OUT; ++i) {
        const double d = mask[i] * (y[i] - t[i]);
        s += d * d;
    }
    free(y);
    return s / (double)(B * OUT);
}

/* Backpropagation: dL/dz of a layer's output, then dW = x^T dz, db = sum dz,
 * dx = dz W^T masked by the ReLU that produced x. */
void oracle_mlp_grad(const double *p, const double *x, const double *t, const double *mask, int B, double *g,
                     double *loss) {
    double *h1 = malloc(sizeof(double) * B * HID), *h2 = malloc(sizeof(double) * B * HID);
    double *y = malloc(sizeof(double) * B * OUT), *dy = malloc(sizeof(double) * B * OUT);
    double *dh2 = malloc(sizeof(double) * B * HID), *dh1 = malloc(sizeof(double) * B * HID);
    layer(x, B, IN, p + OFF_W1, p + OFF_B1, HID, 1, h1);
    layer(h1, B, HID, p + OFF_W2, p + OFF_B2, HID, 1, h2);
    layer(h2, B, HID, p + OFF_W3, p + OFF_B3, OUT, 0, y);
    const double inv = 1.0 / (double)(B * OUT);
    double s = 0.0;
    for (int i = 0; i < B * OUT; ++i) {
        const double d = mask[i] * (y[i] - t[i]);
        s += d * d;
        dy[i] = 2.0 * mask[i] * d * inv;   /* mask is 0/1, so mask*d == d */
    }
    *loss = s * inv;
    memset(g, 0, sizeof(double) * ORACLE_MLP_NPARAMS);
    const double *layer_in[3] = {x, h1, h2};
    const int ins[3] = {IN, HID, HID}, outs[3] = {HID, HID, OUT};
    const int offw[3] = {OFF_W1, OFF_W2, OFF_W3}, offb[3] = {OFF_B1, OFF_B2, OFF_B3};
    double *dz[3] = {dh1, dh2, dy};
    for (int l = 2; l >= 0; --l) {
        const int in = ins[l], out = outs[l];
        for (int r = 0; r < B; ++r)
            for (int o = 0; o < out; ++o) {
                const double d = dz[l][r * out + o];
                g[offb[l] + o] += d;
                for (int i = 0; i < in; ++i) g[offw[l] + i * out + o] += layer_in[l][r * in + i] * d;
            }
        if (l > 0) {
            /* gradient into the ReLU output feeding layer l */
            const double *h = layer_in[l];
            double *dh = dz[l - 1];
            for (int r = 0; r < B; ++r)
                for (int i = 0; i < in; ++i) {
                    double acc = 0.0;
                    if (h[r * in + i] > 0.0)
                        for (int o = 0; o < out; ++o) acc += dz[l][r * out + o] * p[offw[l] + i * out + o];
                    dh[r * in + i] = acc;
                }
        }
    }
    free(h1);
    free(h2);
    free(y);
    free(dy);
    free(dh2);
    free(dh1);
}

/* Adam (Kingma & Ba 2015) with bias correction, one step. */
void oracle_adam(double *p, const double *g, double *m, double *v, int n, double lr, double b1, double b2,
                 double eps, int step) {
    const double bc1 = 1.0 - pow(b1, step), bc2 = 1.0 - pow(b2, step);
    for (int i = 0; i < n; ++i) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        p[i] -= lr * (m[i] / bc1) / (sqrt(v[i] / bc2) + eps);
    }
}
