/*
 * oracle_mlp.c -- CPU restatement, in double precision, of the Krylov-schedule
 * proposer network that libreapers_hip.so trains and evaluates on the GPU.
 * Test infrastructure only, like the rest of oracle/: the GPU forward,
 * gradient and AdamW kernels are checked against these loops, and these
 * loops are checked against central finite differences and closed forms.
 *
 * The network has no counterpart in the reference (cl91/REAPERS evolves with
 * one fixed Krylov dimension, inc/algo.h:266-269).  Layout of the 5122
 * parameters: W1[64][12], b1[64], W2[64][64], b2[64], W3[2][64], b3[2];
 * y = W3 tanh(W2 tanh(W1 x + b1) + b2) + b3.
 * Loss: sum_{s,o} w[s][o] (y[s][o] - t[s][o])^2 / n.
 * AdamW: Loshchilov and Hutter, "Decoupled weight decay regularization",
 * ICLR 2019, Alg. 2 with the bias corrections of Kingma and Ba, ICLR 2015.
 */
#include <math.h>
#include <string.h>

#include "oracle.h"

enum { NI = 12, NH = 64, NO = 2 };
enum { OW1 = 0, OB1 = OW1 + NH * NI, OW2 = OB1 + NH, OB2 = OW2 + NH * NH, OW3 = OB2 + NH,
       OB3 = OW3 + NO * NH, NP = OB3 + NO };

static void fwd(const double *p, const double *x, double *h1, double *h2, double *y)
{
    for (int j = 0; j < NH; j++) {
        double z = p[OB1 + j];
        for (int i = 0; i < NI; i++) z += p[OW1 + j * NI + i] * x[i];
        h1[j] = tanh(z);
    }
    for (int j = 0; j < NH; j++) {
        double z = p[OB2 + j];
        for (int k = 0; k < NH; k++) z += p[OW2 + j * NH + k] * h1[k];
        h2[j] = tanh(z);
    }
    for (int o = 0; o < NO; o++) {
        double z = p[OB3 + o];
        for (int j = 0; j < NH; j++) z += p[OW3 + o * NH + j] * h2[j];
        y[o] = z;
    }
}

int or_mlp_nparams(void) { return NP; }

void or_mlp_forward_f64(const double *p, const double *X, int64_t n, double *Y)
{
    double h1[NH], h2[NH];
    for (int64_t s = 0; s < n; s++) fwd(p, X + s * NI, h1, h2, Y + s * NO);
}

double or_mlp_loss_grad_f64(const double *p, const double *X, const double *T, const double *W,
                            int64_t n, double *g)
{
    double h1[NH], h2[NH], y[NO], dy[NO], dz2[NH], dz1[NH], loss = 0;
    memset(g, 0, sizeof(double) * NP);
    for (int64_t s = 0; s < n; s++) {
        const double *x = X + s * NI;
        fwd(p, x, h1, h2, y);
        for (int o = 0; o < NO; o++) {
            double w = W ? W[s * NO + o] : 1.0;
            double e = y[o] - T[s * NO + o];
            loss += w * e * e;
            dy[o] = 2.0 * w * e / (double)n;
            g[OB3 + o] += dy[o];
            for (int j = 0; j < NH; j++) g[OW3 + o * NH + j] += dy[o] * h2[j];
        }
        for (int j = 0; j < NH; j++) {
            double d = 0;
            for (int o = 0; o < NO; o++) d += p[OW3 + o * NH + j] * dy[o];
            dz2[j] = d * (1.0 - h2[j] * h2[j]);
            g[OB2 + j] += dz2[j];
            for (int k = 0; k < NH; k++) g[OW2 + j * NH + k] += dz2[j] * h1[k];
        }
        for (int k = 0; k < NH; k++) {
            double d = 0;
            for (int j = 0; j < NH; j++) d += p[OW2 + j * NH + k] * dz2[j];
#if defined(OR_PLANT) && OR_PLANT == 4
            dz1[k] = d;                          /* planted: tanh' dropped */
#else
            dz1[k] = d * (1.0 - h1[k] * h1[k]);
#endif
            g[OB1 + k] += dz1[k];
            for (int i = 0; i < NI; i++) g[OW1 + k * NI + i] += dz1[k] * x[i];
        }
    }
    return loss / (double)n;
}

void or_mlp_adam_f64(double *p, const double *g, double *m, double *v, int64_t np, double lr,
                     double b1, double b2, double eps, double wd, int64_t step)
{
    double c1 = 1.0 - pow(b1, (double)step), c2 = 1.0 - pow(b2, (double)step);
    for (int64_t e = 0; e < np; e++) {
        m[e] = b1 * m[e] + (1.0 - b1) * g[e];
        v[e] = b2 * v[e] + (1.0 - b2) * g[e] * g[e];
        double mh = m[e] / c1, vh = v[e] / c2;
        p[e] = p[e] - lr * (mh / (sqrt(vh) + eps) + wd * p[e]);
    }
}
