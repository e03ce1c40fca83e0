/* e13_estimate.c -- the probe estimator: the reference's estimate mode
 * (src/e13search.c:473-510) over its probe() (:256-313), run by k_probe.
 *
 * Probe i draws from its own splitmix64 stream (state splitmix(0x5EED0000 +
 * (seed << 40) + i)), so its four sums -- leaves, nodes, steps, survivors,
 * each child weighted by the product of the live-child counts above it --
 * depend on (problem, seed, i) alone and not on which wave ran it.  The host
 * adds the per-probe sums in probe order and prints the mean and standard
 * error the way the reference does.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>

#include "e13_ctx.h"

#define PROBES_PER_LAUNCH (1 << 22) /* 128 MiB of per-probe sums */

void e13i_free_estimate(e13_ctx *c) {
  hipFree(c->d_probe_out);
  c->d_probe_out = NULL;
  c->probe_cap = 0;
}

int e13_estimate(e13_ctx *c, const e13_problem *p, uint64_t seed, int64_t probes, double *per_probe,
                 e13_estimate_sum *out) {
  if (!c || !p || !out || probes < 1 || p->nq < 1 || p->nq > E13_MAXNQ || p->budget < 0) return E13_EINVAL;
  for (int i = 0; i < 2 * p->nq; i++)
    if (p->tab[i] < E13_FORCED || p->tab[i] >= 4 * p->nq) return E13_EINVAL;
  memset(out, 0, sizeof *out);
  HIPCHK(hipSetDevice(c->device));
  const int gate = c->tune.gate == 2;
  const int wg = c->tune.wg_threads >= 64 && c->tune.wg_threads <= 256 ? c->tune.wg_threads / 64 * 64 : 256;
  if (e13k_probe_lds_bytes(wg, gate) > c->lds_max) return E13_EINVAL;
  /* every workgroup resident at once: waves take probes from one cursor */
  int per_cu = e13k_probe_max_wgs(wg, gate);
  if (per_cu < 1) per_cu = 1;
  const int grid = c->ncu * per_cu;
  const size_t chunk = probes < PROBES_PER_LAUNCH ? (size_t)probes : PROBES_PER_LAUNCH;
  int rc = e13i_ensure_buf((void **)&c->d_probe_out, &c->probe_cap, chunk, 4 * sizeof(double));
  if (rc) return rc;
  double *host = per_probe ? NULL : (double *)malloc(chunk * 4 * sizeof(double));
  if (!per_probe && !host) return E13_ENOMEM;

  E13Problem pb;
  e13i_fill_problem(p, &pb);
  E13ProbeArgs a;
  memset(&a, 0, sizeof a);
  a.pb = c->d_pb;
  a.key = c->d_key;
  a.tgt = c->d_tgt;
  a.out = c->d_probe_out;
  a.ctr = c->d_ctr;
  a.seed = seed;
  a.nq = p->nq;
  a.slice = c->tune.probe_slice;
  a.budget = p->budget;
  a.nintro = p->sym_n0;
  a.solo = c->tune.solo_steps;
  rc = E13_OK;
  if (hipMemcpyAsync(c->d_pb, &pb, sizeof pb, hipMemcpyHostToDevice, c->stream) != hipSuccess) rc = E13_EHIP;
  for (int64_t p0 = 0; rc == E13_OK && p0 < probes; p0 += (int64_t)chunk) {
    const size_t n = (size_t)(probes - p0) < chunk ? (size_t)(probes - p0) : chunk;
    a.p0 = (uint64_t)p0;
    a.n = (uint32_t)n;
    E13Counters ctr;
    float ms = 0.f;
    double *dst = per_probe ? per_probe + 4 * p0 : host;
    if (hipMemsetAsync(c->d_ctr, 0, sizeof(E13Counters), c->stream) != hipSuccess ||
        hipEventRecord(c->ev0, c->stream) != hipSuccess) {
      rc = E13_EHIP;
      break;
    }
    int lr = e13k_launch_probe(&a, grid, wg, gate, c->stream);
    if (lr != 0) {
      fprintf(stderr, "libe13: probe launch failed: %s\n", hipGetErrorString((hipError_t)lr));
      rc = E13_EHIP;
      break;
    }
    if (hipEventRecord(c->ev1, c->stream) != hipSuccess ||
        hipMemcpyAsync(dst, c->d_probe_out, n * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
            hipSuccess ||
        hipMemcpyAsync(&ctr, c->d_ctr, sizeof ctr, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess || hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) {
      rc = E13_EHIP;
      break;
    }
    if (ctr.overflow) { /* 4: LDS area away from address 0; 8, 16: a probe's bookkeeping went wrong */
      fprintf(stderr, "libe13: probe kernel error flags %u\n", ctr.overflow);
      rc = E13_EHIP;
      break;
    }
    out->kernel_ms += ms;
    out->sim_steps += ctr.steps;
    out->rerun_steps += ctr.cyc[0];
    out->rounds += (int64_t)ctr.rounds;
    out->lane_slots += (int64_t)ctr.lane_slots;
    out->launches++;
    for (size_t k = 0; k < n; k++) /* in probe order */
      for (int j = 0; j < 4; j++) {
        const double v = dst[4 * k + j];
        out->sum[j] += v;
        out->sq[j] += v * v;
      }
  }
  free(host);
  if (rc) return rc;
  out->probes = probes;
  for (int j = 0; j < 4; j++) { /* e13search.c:503-506 */
    const double m = out->sum[j] / (double)probes, var = out->sq[j] / (double)probes - m * m;
    out->mean[j] = m;
    out->se[j] = var > 0 ? sqrt(var / (double)probes) : 0.0;
  }
  c->stats.kernel_ms = out->kernel_ms;
  c->stats.launches = out->launches;
  c->stats.rounds = out->rounds;
  c->stats.lane_slots = out->lane_slots;
  c->stats.grid = grid;
  c->stats.wg_threads = wg;
  c->stats.slice = a.slice;
  c->stats.gate = c->tune.gate;
  return E13_OK;
}
