/* e13mlp_ref.c -- TEST INFRASTRUCTURE ONLY (the oracle): the subtree cost
 * model's perceptron in double precision, straight from the definitions, to
 * check the library's fp32 GPU kernels.  Nothing in the product links this.
 *
 * Network (parameter layout as in the library, 6402 values):
 *   h1 = relu(W1 x + b1)   W1[64][32] at 0,    b1 at 2048
 *   h2 = relu(W2 h1 + b2)  W2[64][64] at 2112, b2 at 6208
 *   y  = W3 h2 + b3        W3[2][64]  at 6272, b3 at 6400
 * Loss L = (1/n) sum_s sum_o (y[s][o] - t[s][o])^2.
 * The gradient is checked against central finite differences in the tests.
 */
#include <string.h>

#define NI 32
#define NH 64
#define NO 2
#define OW1 0
#define OB1 2048
#define OW2 2112
#define OB2 6208
#define OW3 6272
#define OB3 6400
#define NP 6402

static void fwd(const double *P, const double *x, double *h1, double *h2, double *y) {
  for (int i = 0; i < NH; i++) {
    double a = P[OB1 + i];
    for (int j = 0; j < NI; j++) a += P[OW1 + i * NI + j] * x[j];
    h1[i] = a > 0 ? a : 0;
  }
  for (int k = 0; k < NH; k++) {
    double a = P[OB2 + k];
    for (int j = 0; j < NH; j++) a += P[OW2 + k * NH + j] * h1[j];
    h2[k] = a > 0 ? a : 0;
  }
  for (int o = 0; o < NO; o++) {
    double a = P[OB3 + o];
    for (int k = 0; k < NH; k++) a += P[OW3 + o * NH + k] * h2[k];
    y[o] = a;
  }
}

void eo_mlp_forward(const double *P, const double *x, int n, double *y) {
  double h1[NH], h2[NH];
  for (int s = 0; s < n; s++) fwd(P, x + (size_t)s * NI, h1, h2, y + (size_t)s * NO);
}

/* Loss and its gradient (backpropagation written out). */
double eo_mlp_grad(const double *P, const double *x, const double *t, int n, double *g) {
  memset(g, 0, NP * sizeof(double));
  double loss = 0, h1[NH], h2[NH], y[NO], dy[NO], d2[NH], d1[NH];
  for (int s = 0; s < n; s++) {
    const double *xs = x + (size_t)s * NI;
    fwd(P, xs, h1, h2, y);
    for (int o = 0; o < NO; o++) {
      const double d = y[o] - t[(size_t)s * NO + o];
      loss += d * d;
      dy[o] = 2.0 * d / n;
    }
    for (int o = 0; o < NO; o++) {
      g[OB3 + o] += dy[o];
      for (int k = 0; k < NH; k++) g[OW3 + o * NH + k] += dy[o] * h2[k];
    }
    for (int k = 0; k < NH; k++) {
      double a = 0;
      for (int o = 0; o < NO; o++) a += P[OW3 + o * NH + k] * dy[o];
      d2[k] = h2[k] > 0 ? a : 0;
      g[OB2 + k] += d2[k];
      for (int j = 0; j < NH; j++) g[OW2 + k * NH + j] += d2[k] * h1[j];
    }
    for (int j = 0; j < NH; j++) {
      double a = 0;
      for (int k = 0; k < NH; k++) a += P[OW2 + k * NH + j] * d2[k];
      d1[j] = h1[j] > 0 ? a : 0;
      g[OB1 + j] += d1[j];
      for (int i = 0; i < NI; i++) g[OW1 + j * NI + i] += d1[j] * xs[i];
    }
  }
  return loss / n;
}
