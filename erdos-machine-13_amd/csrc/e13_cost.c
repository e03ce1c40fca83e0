/* e13_cost.c -- the subtree cost model around the persistent search: frontier
 * features, the cost log, and the host side of the perceptron's GPU
 * inference and training (kernels in e13_mlp.hip).
 *
 * No reference counterpart.  The reference balances its OpenMP loop over a
 * static front (src/e13search.c:357-420, dynamic schedule in natural order) and
 * estimates subtree sizes by random descents (probe(), :256-313); here a small
 * network trained on the search's own logs predicts the leaves and steps of a
 * frontier node's subtree, and the persistent search splits and orders its
 * frontier by that prediction.  Nothing is pruned.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "e13_ctx.h"

void e13_cost_targets(uint64_t leaves, uint64_t steps, float *t2) {
  t2[0] = (float)(log1p((double)leaves) / 16.0);
  t2[1] = (float)(log1p((double)steps) / 24.0);
}

/* 32 features of a frontier node (all O(1) in magnitude). */
void e13i_node_features(const e13_problem *p, const E13Node *nd, float *f) {
  memset(f, 0, E13_NFEAT * sizeof(float));
  const int nslot = 4 * p->nq + 1;
  const int first = nd->next_slot & 0xFFFF, end = (nd->next_slot >> 16) ? (nd->next_slot >> 16) : nslot;
  f[0] = 1.f;
  f[1] = (float)nd->depth / 16.f;
  f[2] = (float)nd->nbound / (float)(2 * p->nq);
  f[3] = (float)nd->budget / 4.f;
  f[4] = (float)(p->X + 1) / 15.f;
  f[5] = (float)(log1p((double)nd->t) / 14.0);
  f[6] = (float)nd->stage / 9.f;
  f[7] = (float)nd->nt / 9.f;
  f[8] = (float)(nd->q - E13_QOFF) / 1300.f;
  f[9] = (float)(log1p((double)nd->ones) / 8.0);
  f[10] = (float)nd->st / 16.f;
  f[11] = (float)(nd->pend + 1) / 33.f;
  f[12] = (float)(end - first) / (float)nslot;
  f[13] = (float)nd->nintro / 4.f;
  f[14] = (float)(nd->flags & 1);
  f[15] = (float)((nd->flags >> 1) & 1);
  for (int s = 0; s < 16 && s < p->nq; s++) { /* cells of state s already decided */
    int d = 0;
    for (int b = 0; b < 2; b++) {
      const int code = nd->tab[2 * s + b];
      d += code != E13_C_EDIT && code != E13_C_FREE;
    }
    f[16 + s] = (float)d / 2.f;
  }
}

void e13_mlp_init(uint64_t seed, float *P) {
  uint64_t st = seed;
  const int fan[3] = {32, 64, 64}, off[3] = {0, 2112, 6272}, cnt[3] = {2048, 4096, 128};
  const int boff[3] = {2048, 6208, 6400}, bcnt[3] = {64, 64, 2};
  for (int l = 0; l < 3; l++) {
    const double r = sqrt(6.0 / fan[l]);
    for (int k = 0; k < cnt[l]; k++) {
      const double u = (double)(e13i_splitmix(&st) >> 11) * (1.0 / 9007199254740992.0);
      P[off[l] + k] = (float)((2.0 * u - 1.0) * r);
    }
    for (int k = 0; k < bcnt[l]; k++) P[boff[l] + k] = 0.f;
  }
}

/* Sample buffers for n rows (and the per-tile gradient partials). */
static int mlp_buffers(e13_ctx *c, size_t n) {
  if (!c->d_mlp && hipMalloc((void **)&c->d_mlp, 4 * E13_MLP_PARAMS * sizeof(float)) != hipSuccess)
    return E13_ENOMEM;
  if (c->mlp_rows >= n) return E13_OK;
  hipFree(c->d_mlp_x);
  hipFree(c->d_mlp_t);
  hipFree(c->d_mlp_y);
  hipFree(c->d_mlp_part);
  hipFree(c->d_mlp_sse);
  hipFree(c->d_mlp_idx);
  c->d_mlp_x = c->d_mlp_t = c->d_mlp_y = c->d_mlp_part = c->d_mlp_sse = NULL;
  c->d_mlp_idx = NULL;
  c->mlp_rows = 0;
  const size_t tiles = (n + 31) / 32;
  if (hipMalloc((void **)&c->d_mlp_x, n * E13_NFEAT * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&c->d_mlp_t, n * 2 * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&c->d_mlp_y, n * 2 * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&c->d_mlp_part, tiles * E13_MLP_PARAMS * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&c->d_mlp_sse, tiles * sizeof(float)) != hipSuccess ||
      hipMalloc((void **)&c->d_mlp_idx, n * sizeof(int32_t)) != hipSuccess)
    return E13_ENOMEM;
  c->mlp_rows = n;
  return E13_OK;
}

int e13_mlp_forward(e13_ctx *c, const float *P, const float *x, int64_t n, float *y) {
  if (!c || !P || !x || !y || n < 1 || n > (1 << 26)) return E13_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  int rc = mlp_buffers(c, (size_t)n);
  if (rc) return rc;
  HIPCHK(hipMemcpyAsync(c->d_mlp, P, E13_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_mlp_x, x, (size_t)n * E13_NFEAT * sizeof(float), hipMemcpyHostToDevice, c->stream));
  if (e13k_mlp_forward(c->d_mlp, c->d_mlp_x, NULL, (int)n, c->d_mlp_y, c->stream)) return E13_EHIP;
  HIPCHK(hipMemcpyAsync(y, c->d_mlp_y, (size_t)n * 2 * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  return E13_OK;
}

int e13i_mlp_predict_dev(e13_ctx *c, const float *x, int n, float *y) {
  if (!c->have_model) return E13_EINVAL;
  return e13_mlp_forward(c, c->model, x, n, y);
}

static double sum_sse(e13_ctx *c, int tiles, int *rc) {
  float *h = (float *)malloc((size_t)tiles * sizeof(float));
  double s = 0;
  if (!h) {
    *rc = E13_ENOMEM;
    return 0;
  }
  if (hipMemcpyAsync(h, c->d_mlp_sse, (size_t)tiles * sizeof(float), hipMemcpyDeviceToHost, c->stream) !=
          hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess)
    *rc = E13_EHIP;
  else
    for (int k = 0; k < tiles; k++) s += h[k];
  free(h);
  return s;
}

int e13_mlp_grad(e13_ctx *c, const float *P, const float *x, const float *t, int64_t n, float *grad, double *loss) {
  if (!c || !P || !x || !t || !grad || n < 1 || n > (1 << 26)) return E13_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  int rc = mlp_buffers(c, (size_t)n);
  if (rc) return rc;
  float *dP = c->d_mlp, *dG = c->d_mlp + E13_MLP_PARAMS;
  HIPCHK(hipMemcpyAsync(dP, P, E13_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_mlp_x, x, (size_t)n * E13_NFEAT * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_mlp_t, t, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, c->stream));
  if (e13k_mlp_grad(dP, c->d_mlp_x, c->d_mlp_t, NULL, 0, (int)n, 1.f / (float)n, c->d_mlp_part, c->d_mlp_sse, dG,
                    c->stream))
    return E13_EHIP;
  HIPCHK(hipMemcpyAsync(grad, dG, E13_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, c->stream));
  const double sse = sum_sse(c, (int)((n + 31) / 32), &rc);
  if (rc) return rc;
  if (loss) *loss = sse / (double)n;
  return E13_OK;
}

int e13_mlp_train(e13_ctx *c, float *P, const float *x, const float *t, int64_t n, int epochs, int batch, float lr,
                  uint64_t seed, float *loss_per_epoch) {
  if (!c || !P || !x || !t || n < 1 || n > (1 << 26) || epochs < 1 || batch < 1 || !(lr > 0)) return E13_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  int rc = mlp_buffers(c, (size_t)n);
  if (rc) return rc;
  float *dP = c->d_mlp, *dG = dP + E13_MLP_PARAMS, *dM = dG + E13_MLP_PARAMS, *dV = dM + E13_MLP_PARAMS;
  HIPCHK(hipMemcpyAsync(dP, P, E13_MLP_PARAMS * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(dM, 0, 2 * E13_MLP_PARAMS * sizeof(float), c->stream));
  HIPCHK(hipMemcpyAsync(c->d_mlp_x, x, (size_t)n * E13_NFEAT * sizeof(float), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemcpyAsync(c->d_mlp_t, t, (size_t)n * 2 * sizeof(float), hipMemcpyHostToDevice, c->stream));
  int32_t *perm = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  float *sse = (float *)malloc(((size_t)n / 32 + 2) * sizeof(float));
  if (!perm || !sse) {
    free(perm);
    free(sse);
    return E13_ENOMEM;
  }
  for (int64_t i = 0; i < n; i++) perm[i] = (int32_t)i;
  uint64_t st = seed;
  const double b1 = 0.9, b2 = 0.999;
  int64_t step = 0;
  for (int e = 0; e < epochs && !rc; e++) {
    for (int64_t i = n - 1; i > 0; i--) { /* Fisher-Yates with splitmix64 */
      const int64_t j = (int64_t)(e13i_splitmix(&st) % (uint64_t)(i + 1));
      const int32_t tmp = perm[i];
      perm[i] = perm[j];
      perm[j] = tmp;
    }
    if (hipMemcpyAsync(c->d_mlp_idx, perm, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream) !=
        hipSuccess) {
      rc = E13_EHIP;
      break;
    }
    double ep_loss = 0;
    int nb = 0;
    for (int64_t r0 = 0; r0 < n; r0 += batch, nb++) {
      const int m = (int)(n - r0 < batch ? n - r0 : batch);
      step++;
      const float c1 = (float)(1.0 - pow(b1, (double)step)), c2 = (float)(1.0 - pow(b2, (double)step));
      if (e13k_mlp_grad(dP, c->d_mlp_x, c->d_mlp_t, c->d_mlp_idx, (int)r0, m, 1.f / (float)m, c->d_mlp_part,
                        c->d_mlp_sse, dG, c->stream) ||
          e13k_mlp_adam(dP, dG, dM, dV, lr, (float)b1, (float)b2, 1e-8f, c1, c2, c->stream)) {
        rc = E13_EHIP;
        break;
      }
      if (loss_per_epoch) {
        const int tiles = (m + 31) / 32;
        if (hipMemcpyAsync(sse, c->d_mlp_sse, (size_t)tiles * sizeof(float), hipMemcpyDeviceToHost, c->stream) !=
                hipSuccess ||
            hipStreamSynchronize(c->stream) != hipSuccess) {
          rc = E13_EHIP;
          break;
        }
        double s = 0;
        for (int k = 0; k < tiles; k++) s += sse[k];
        ep_loss += s / (double)m;
      }
    }
    if (loss_per_epoch) loss_per_epoch[e] = (float)(ep_loss / (nb ? nb : 1));
  }
  if (!rc && (hipMemcpyAsync(P, dP, E13_MLP_PARAMS * sizeof(float), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipStreamSynchronize(c->stream) != hipSuccess))
    rc = E13_EHIP;
  free(perm);
  free(sse);
  return rc;
}

int e13_set_cost_model(e13_ctx *c, const float *P) {
  if (!c) return E13_EINVAL;
  if (!P) {
    c->have_model = 0;
    return E13_OK;
  }
  memcpy(c->model, P, sizeof c->model);
  c->have_model = 1;
  return E13_OK;
}

int e13_get_cost_log(e13_ctx *c, e13_cost_rec *out, int64_t cap, int64_t *n) {
  if (!c) return E13_EINVAL;
  if (n) *n = c->cost_n;
  if (!out) return E13_OK;
  const int64_t k = c->cost_n < cap ? c->cost_n : cap;
  if (k > 0) memcpy(out, c->cost_log, (size_t)k * sizeof *out);
  return E13_OK;
}

void e13i_free_cost(e13_ctx *c) {
  hipFree(c->d_mlp);
  hipFree(c->d_mlp_x);
  hipFree(c->d_mlp_t);
  hipFree(c->d_mlp_y);
  hipFree(c->d_mlp_part);
  hipFree(c->d_mlp_sse);
  hipFree(c->d_mlp_idx);
  hipFree(c->d_seed_cnt);
  free(c->cost_log);
}
