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

_Static_assert(E13_STATIC_PROBES == E13_SCRATCH_PROBES, "the static estimator's limit is its scratch's size");
_Static_assert(sizeof(E13Counters) <= E13_SCRATCH_PB - E13_SCRATCH_CTR && sizeof(E13Counters) % 4 == 0,
               "the counters fit their scratch slot and clear as words");
_Static_assert(sizeof(E13Problem) <= E13_SCRATCH_ROOT - E13_SCRATCH_PB, "the problem record fits its scratch slot");
_Static_assert(sizeof(E13Node) <= E13_SCRATCH_BYTES - E13_SCRATCH_ROOT, "the root node fits its scratch slot");

/* The estimate's root as a node: the blank tape stopped before the read of
   cell 0 at t = 0, every cell unread (e13search.c:265-271: probe() starts with
   fixed[] all zero, so editable cells branch and free cells take any move). */
void e13i_probe_root(const e13_problem *p, E13Node *nd) {
  memset(nd, 0, sizeof *nd);
  nd->q = E13_QOFF;
  nd->q0 = E13_QOFF;
  nd->nt = 1;
  nd->st0 = -1;
  nd->flags = 3; /* abs_ok, rel_ok */
  nd->budget = p->budget;
  nd->nintro = p->sym_n0;
  nd->pend = 0;
  for (int i = 0; i < E13_MAXCELLS; i++)
    nd->tab[i] = i < 2 * p->nq && p->tab[i] != E13_FORCED ? E13_C_EDIT : E13_C_FREE;
}

void e13i_free_estimate(e13_ctx *c) {
  hipFree(c->d_probe_out);
  c->d_probe_out = NULL;
  c->probe_cap = 0;
}

/* Device buffers one estimate runs over: per-probe sums for chunk probes, the
   launch counters and the problem record. */
typedef struct {
  double *out;
  E13Counters *ctr;
  E13Problem *pb;
  E13Node *root;
  size_t chunk;
  int static_scratch; /* counters cleared by a fill kernel (no memset node) */
} est_bufs;

static int check_problem(const e13_ctx *c, const e13_problem *p, int64_t probes, const e13_estimate_sum *out) {
  if (!c || !p || !out || probes < 1 || p->nq < 1 || p->nq > E13_MAXNQ || p->budget < 0) return E13_EINVAL;
  for (int i = 0; i < 2 * p->nq; i++)
    if (p->tab[i] < E13_FORCED || p->tab[i] >= 4 * p->nq) return E13_EINVAL;
  return E13_OK;
}

/* Launch shape: every workgroup resident at once, waves take probes from one cursor. */
static int probe_shape(const e13_ctx *c, int *gate, int *wg, int *grid) {
  *gate = c->tune.gate == 2;
  *wg = c->tune.wg_threads >= 64 && c->tune.wg_threads <= 256 ? c->tune.wg_threads / 64 * 64 : 256;
  if (e13k_probe_lds_bytes(*wg, *gate) > c->lds_max) return E13_EINVAL;
  int per_cu = e13k_probe_max_wgs(*wg, *gate);
  if (per_cu < 1) per_cu = 1;
  *grid = c->ncu * per_cu;
  return E13_OK;
}

/* The probes in launches of at most b->chunk, each launch's per-probe sums copied
   to per_probe + 4*p0 (or to host, reused, when per_probe is NULL) and added in
   probe order. */
static int estimate_run(e13_ctx *c, const e13_problem *p, uint64_t seed, int64_t probes, double *per_probe,
                        double *host, const est_bufs *b, e13_estimate_sum *out) {
  int gate, wg, grid;
  int rc = probe_shape(c, &gate, &wg, &grid);
  if (rc) return rc;
  E13Problem pb;
  e13i_fill_problem(p, &pb);
  E13ProbeArgs a;
  memset(&a, 0, sizeof a);
  a.pb = b->pb;
  a.key = c->d_key;
  a.tgt = c->d_tgt;
  a.out = b->out;
  a.ctr = b->ctr;
  a.seed = seed;
  a.nq = p->nq;
  a.slice = c->tune.probe_slice;
  a.solo = c->tune.solo_steps;
  a.roots = b->root;
  a.per_root = (uint64_t)probes; /* every probe starts at the one root */
  E13Node root;
  e13i_probe_root(p, &root);
  if (hipMemcpyAsync(b->pb, &pb, sizeof pb, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(b->root, &root, sizeof root, hipMemcpyHostToDevice, c->stream) != hipSuccess)
    return E13_EHIP;
  for (int64_t p0 = 0; rc == E13_OK && p0 < probes; p0 += (int64_t)b->chunk) {
    const size_t n = (size_t)(probes - p0) < b->chunk ? (size_t)(probes - p0) : b->chunk;
    a.p0 = (uint64_t)p0;
    a.n = (uint32_t)n;
    E13Counters ctr;
    float ms = 0.f;
    double *dst = per_probe ? per_probe + 4 * p0 : host;
    const int zr = b->static_scratch
                       ? e13k_launch_zero(b->ctr, (uint32_t)(sizeof(E13Counters) / 4), c->stream)
                       : (int)hipMemsetAsync(b->ctr, 0, sizeof(E13Counters), c->stream);
    if (zr != 0 || hipEventRecord(c->ev0, c->stream) != hipSuccess) {
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
        hipMemcpyAsync(dst, b->out, n * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipMemcpyAsync(&ctr, b->ctr, sizeof ctr, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
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

int e13_estimate(e13_ctx *c, const e13_problem *p, uint64_t seed, int64_t probes, double *per_probe,
                 e13_estimate_sum *out) {
  int rc = check_problem(c, p, probes, out);
  if (rc) return rc;
  memset(out, 0, sizeof *out);
  HIPCHK(hipSetDevice(c->device));
  int gate, wg, grid;
  if ((rc = probe_shape(c, &gate, &wg, &grid)) != 0) return rc;
  const size_t chunk = probes < PROBES_PER_LAUNCH ? (size_t)probes : PROBES_PER_LAUNCH;
  rc = e13i_ensure_buf((void **)&c->d_probe_out, &c->probe_cap, chunk, 4 * sizeof(double));
  if (rc) return rc;
  double *host = per_probe ? NULL : (double *)malloc(chunk * 4 * sizeof(double));
  if (!per_probe && !host) return E13_ENOMEM;
  const est_bufs b = {c->d_probe_out, c->d_ctr, c->d_pb, c->d_proot, chunk, 0};
  rc = estimate_run(c, p, seed, probes, per_probe, host, &b, out);
  free(host);
  return rc;
}

/* e13_estimate over the device's static scratch: one launch, nothing allocated. */
int e13_estimate_static(e13_ctx *c, const e13_problem *p, uint64_t seed, int64_t probes, double *per_probe,
                        e13_estimate_sum *out) {
  int rc = check_problem(c, p, probes, out);
  if (rc) return rc;
  if (!per_probe) return E13_EINVAL;
  if (probes > E13_STATIC_PROBES) {
    fprintf(stderr, "libe13: e13_estimate_static: %lld probes exceed the static scratch's %d\n", (long long)probes,
            E13_STATIC_PROBES);
    return E13_ESCRATCH;
  }
  if (!c->scratch) {
    fprintf(stderr, "libe13: e13_estimate_static: no static scratch address on device %d\n", c->device);
    return E13_EHIP;
  }
  memset(out, 0, sizeof *out);
  HIPCHK(hipSetDevice(c->device));
  unsigned char *s = (unsigned char *)c->scratch;
  const est_bufs b = {(double *)s, (E13Counters *)(s + E13_SCRATCH_CTR), (E13Problem *)(s + E13_SCRATCH_PB),
                      (E13Node *)(s + E13_SCRATCH_ROOT), (size_t)E13_SCRATCH_PROBES, 1};
  return estimate_run(c, p, seed, probes, per_probe, NULL, &b, out);
}

/* Subtree steps below n device nodes of one problem (pb_dev its device
   record): the mean step sum of per_node probes from each node, the
   reference's probe() started at that node instead of the root; -1 for a node
   with no pending cell (a machine paused at a cap). */
int e13i_probe_nodes(e13_ctx *c, const e13_problem *p, const E13Problem *pb_dev, const E13Node *nodes, size_t n,
                     int per_node, uint64_t seed, double *est) {
  if (n == 0) return E13_OK;
  if (per_node < 1 || (uint64_t)n * (uint64_t)per_node > 0xFFFFFFFFull) return E13_EINVAL;
  int gate, wg, grid;
  int rc = probe_shape(c, &gate, &wg, &grid);
  if (rc) return rc;
  const size_t total = n * (size_t)per_node;
  if ((rc = e13i_ensure_buf((void **)&c->d_probe_out, &c->probe_cap, total, 4 * sizeof(double)))) return rc;
  double *host = (double *)malloc(total * 4 * sizeof(double));
  if (!host) return E13_ENOMEM;
  E13ProbeArgs a;
  memset(&a, 0, sizeof a);
  a.pb = pb_dev;
  a.key = c->d_key;
  a.tgt = c->d_tgt;
  a.out = c->d_probe_out;
  a.ctr = c->d_pctr;
  a.seed = seed;
  a.n = (uint32_t)total;
  a.nq = p->nq;
  a.slice = c->tune.probe_slice;
  a.solo = c->tune.solo_steps;
  a.roots = nodes;
  a.per_root = (uint64_t)per_node;
  E13Counters ctr;
  rc = E13_OK;
  if (hipMemsetAsync(c->d_pctr, 0, sizeof(E13Counters), c->stream) != hipSuccess) rc = E13_EHIP;
  int lr = rc ? 0 : e13k_launch_probe(&a, grid, wg, gate, c->stream);
  if (lr != 0) {
    fprintf(stderr, "libe13: probe launch failed: %s\n", hipGetErrorString((hipError_t)lr));
    rc = E13_EHIP;
  }
  if (!rc && (hipMemcpyAsync(host, c->d_probe_out, total * 4 * sizeof(double), hipMemcpyDeviceToHost, c->stream) !=
                  hipSuccess ||
              hipMemcpyAsync(&ctr, c->d_pctr, sizeof ctr, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipStreamSynchronize(c->stream) != hipSuccess))
    rc = E13_EHIP;
  if (!rc && ctr.overflow) {
    fprintf(stderr, "libe13: probe kernel error flags %u\n", ctr.overflow);
    rc = E13_EHIP;
  }
  for (size_t i = 0; !rc && i < n; i++) {
    double s = 0.0;
    int paused = 0;
    for (int j = 0; j < per_node; j++) {
      const double v = host[4 * (i * (size_t)per_node + (size_t)j) + 2];
      paused |= v < 0.0;
      s += v;
    }
    est[i] = paused ? -1.0 : s / (double)per_node;
  }
  free(host);
  return rc;
}
