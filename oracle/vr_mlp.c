/*
 * vr_mlp.c - binary64 restatement of the light-cluster proposal network.
 *
 * TEST INFRASTRUCTURE ONLY (the checker for the product's BF16 MFMA kernels;
 * loaded by tests/ and never by the product).  The reference has no learned
 * proposal: this file follows the network definition of the benchmark
 * configuration (inputs: 3 normalised position values + 6-way face one-hot;
 * widths 9-64-64-256; ReLU; softmax over clusters) and Adam (Kingma & Ba,
 * 2015, Algorithm 1).  Pinned by central finite differences and Adam's
 * closed-form first step (tests/test_oracle_mlp.py).
 *
 * Parameter layout (flat): W1[64][9], b1[64], W2[64][64], b2[64],
 * W3[256][64], b3[256]; weights row-major (output, input).
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#define NI 9
#define NH 64
#define NO 256
#define OFF_W1 0
#define OFF_B1 (OFF_W1 + NH * NI)
#define OFF_W2 (OFF_B1 + NH)
#define OFF_B2 (OFF_W2 + NH * NH)
#define OFF_W3 (OFF_B2 + NH)
#define OFF_B3 (OFF_W3 + NO * NH)
#define NPARAM (OFF_B3 + NO)

uint32_t or_mlp_num_params(void) { return NPARAM; }

/* one sample's forward pass; h1, h2: NH, z: NO */
static void forward1(const double* P, const double* x, double* h1, double* h2, double* z)
{
    for (int o = 0; o < NH; o++) {
        double s = P[OFF_B1 + o];
        for (int i = 0; i < NI; i++) s += P[OFF_W1 + o * NI + i] * x[i];
        h1[o] = s > 0 ? s : 0;
    }
    for (int o = 0; o < NH; o++) {
        double s = P[OFF_B2 + o];
        for (int i = 0; i < NH; i++) s += P[OFF_W2 + o * NH + i] * h1[i];
        h2[o] = s > 0 ? s : 0;
    }
    for (int o = 0; o < NO; o++) {
        double s = P[OFF_B3 + o];
        for (int i = 0; i < NH; i++) s += P[OFF_W3 + o * NH + i] * h2[i];
        z[o] = s;
    }
}

static void softmax(const double* z, double* q)
{
    double m = z[0];
    for (int o = 1; o < NO; o++) m = z[o] > m ? z[o] : m;
    double s = 0;
    for (int o = 0; o < NO; o++) q[o] = exp(z[o] - m), s += q[o];
    for (int o = 0; o < NO; o++) q[o] /= s;
}

/* logits z[B][NO] and probabilities q[B][NO] (either may be NULL) */
void or_mlp_forward(const double* P, const double* x, uint32_t B, double* z, double* q)
{
    double h1[NH], h2[NH], zz[NO], qq[NO];
    for (uint32_t b = 0; b < B; b++) {
        forward1(P, x + (size_t)b * NI, h1, h2, zz);
        if (z) memcpy(z + (size_t)b * NO, zz, sizeof(zz));
        if (q) {
            softmax(zz, qq);
            memcpy(q + (size_t)b * NO, qq, sizeof(qq));
        }
    }
}

/* loss = (1/B) sum_b w_b * -log softmax(z_b)[t_b]; grad (NPARAM) = dloss/dP */
double or_mlp_loss_grad(const double* P, const double* x, const uint32_t* t, const double* w, uint32_t B,
                        double* grad)
{
    double h1[NH], h2[NH], z[NO], q[NO], dz[NO], d2[NH], d1[NH];
    double loss = 0;
    if (grad) memset(grad, 0, sizeof(double) * NPARAM);
    for (uint32_t b = 0; b < B; b++) {
        const double* xb = x + (size_t)b * NI;
        forward1(P, xb, h1, h2, z);
        softmax(z, q);
        const double wb = w ? w[b] : 1.0;
        loss += wb * -log(q[t[b]]);
        if (!grad) continue;
        for (int o = 0; o < NO; o++) dz[o] = wb * (q[o] - (o == (int)t[b] ? 1.0 : 0.0)) / (double)B;
        for (int i = 0; i < NH; i++) d2[i] = 0;
        for (int o = 0; o < NO; o++) {
            grad[OFF_B3 + o] += dz[o];
            for (int i = 0; i < NH; i++) {
                grad[OFF_W3 + o * NH + i] += dz[o] * h2[i];
                d2[i] += P[OFF_W3 + o * NH + i] * dz[o];
            }
        }
        for (int i = 0; i < NH; i++) d2[i] = h2[i] > 0 ? d2[i] : 0;
        for (int i = 0; i < NH; i++) d1[i] = 0;
        for (int o = 0; o < NH; o++) {
            grad[OFF_B2 + o] += d2[o];
            for (int i = 0; i < NH; i++) {
                grad[OFF_W2 + o * NH + i] += d2[o] * h1[i];
                d1[i] += P[OFF_W2 + o * NH + i] * d2[o];
            }
        }
        for (int i = 0; i < NH; i++) d1[i] = h1[i] > 0 ? d1[i] : 0;
        for (int o = 0; o < NH; o++) {
            grad[OFF_B1 + o] += d1[o];
            for (int i = 0; i < NI; i++) grad[OFF_W1 + o * NI + i] += d1[o] * xb[i];
        }
    }
    return loss / (double)B;
}

/* Adam (Kingma & Ba 2015, Algorithm 1), step t >= 1 */
void or_adam_step(double* P, const double* g, double* m, double* v, uint32_t t, double lr, double b1, double b2,
                  double eps, uint32_t n)
{
    const double c1 = 1.0 - pow(b1, (double)t), c2 = 1.0 - pow(b2, (double)t);
    for (uint32_t i = 0; i < n; i++) {
        m[i] = b1 * m[i] + (1.0 - b1) * g[i];
        v[i] = b2 * v[i] + (1.0 - b2) * g[i] * g[i];
        const double mh = m[i] / c1, vh = v[i] / c2;
        P[i] -= lr * mh / (sqrt(vh) + eps);
    }
}
