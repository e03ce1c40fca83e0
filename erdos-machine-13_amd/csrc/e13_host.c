/* e13_host.c -- C host of libe13.so: problem setup, the search driver around
 * the gfx950 expansion kernel, complete-table batches, the plain survivor
 * re-check and the reference's output lines.
 *
 * References: src/e13search.c (setup :317-355, set_sym :423-426, run_delete
 * :357-420, report :184-200, run_full :428-433) and the GPU driver loop of
 * src/e13gpu.py:381-437.
 */
#define __HIP_PLATFORM_AMD__ 1
#include <hip/hip_runtime_api.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <pthread.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "e13.h"
#include "e13_dev.h"

#include "e13_ctx.h"

static double host_ms(void);
static const long T14S[E13_NS + 1] = {0, 2, 22, 90, 268, 944, 3434, 13450, 52068, 208864}; /* e13search.c:70 */

const char *e13_strerror(int code) {
  switch (code) {
  case E13_OK: return "ok";
  case E13_EINVAL: return "invalid argument";
  case E13_EHIP: return "HIP runtime error";
  case E13_ENOMEM: return "out of memory";
  case E13_EFORCED: return "forced cells exceed fmax";
  case E13_ESYM: return "symmetric state is named by a fixed cell or has one";
  case E13_ESYMSTART: return "start state in the symmetric states";
  case E13_ESTACK: return "node stack overflow";
  case E13_ENODEV: return "no HIP device";
  case E13_ESTOPPED: return "stopped cleanly; the checkpoint is saved";
  case E13_ECKPT: return "checkpoint file unreadable, damaged or for another search";
  case E13_ELOST: return "persistent search: a mailbox delivery never arrived complete";
  case E13_EWATCHDOG: return "persistent search: idle watchdog (no work and no end mark)";
  case E13_ELDS: return "dynamic LDS does not start at LDS address 0";
  case E13_ESCRATCH: return "the request exceeds the static scratch (more than E13_STATIC_PROBES probes)";
  case E13_ERESIDENT: return "persistent search: not every workgroup became resident at once";
  default: return "unknown error";
  }
}


/* ------------------------------------------------------------------------ */
/* Zobrist tables (e13search.c:98-103, 437-444).                             */
/* ------------------------------------------------------------------------ */
uint64_t e13i_splitmix(uint64_t *s) {
  uint64_t z = (*s += 0x9E3779B97F4A7C15ULL);
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

int e13_keys(uint64_t *key, uint64_t *tgt) {
  uint64_t seed = 20261028ULL;
  for (int i = 0; i < 1364; i++) key[i] = e13i_splitmix(&seed);
  tgt[0] = 0;
  for (int n = 1; n <= E13_NS; n++) {
    uint64_t h = 0;
    for (int c = -1; c <= (1 << n) - 2; c++) h ^= key[c + 64];
    tgt[n] = h;
  }
  return 0;
}

/* ------------------------------------------------------------------------ */
/* Machine text and setup.                                                   */
/* ------------------------------------------------------------------------ */
static int parse_cell(const char *s, int len) { /* e13search.c:105-108 */
  if (len >= 1 && s[0] == '-') return E13_HALT;
  if (len < 3 || (s[0] != '0' && s[0] != '1') || (s[1] != 'L' && s[1] != 'R')) return -100;
  int n = 0;
  for (int i = 2; i < len; i++) {
    if (s[i] < '0' || s[i] > '9') return -100;
    n = 10 * n + (s[i] - '0');
  }
  return (n << 2) | ((s[1] == 'R') << 1) | (s[0] - '0');
}

int e13_parse_machine(const char *text, int32_t *base, int max_states) {
  int ncell = 0;
  const char *p = text;
  while (*p) {
    while (*p == ' ' || *p == '\t' || *p == '\n' || *p == '\r') p++;
    if (!*p) break;
    const char *s = p;
    while (*p && *p != ' ' && *p != '\t' && *p != '\n' && *p != '\r') p++;
    int code = parse_cell(s, (int)(p - s));
    if (code == -100) return E13_EINVAL;
    if (ncell >= 2 * max_states) break;
    base[ncell++] = code;
  }
  if (ncell % 2) return E13_EINVAL;
  return ncell / 2;
}

int e13_setup(const int32_t *base, int nq_base, int X, int fmax, const uint8_t *freebase, int sym_lo, int sym_hi,
              e13_problem *out) {
  if (!base || !out || nq_base < 1 || nq_base > 16 || X < -1 || X >= nq_base) return E13_EINVAL;
  memset(out, 0, sizeof *out);
  int nforced = 0, row = 0;
  for (int q = 0; q < nq_base; q++) {
    if (q == X) continue;
    for (int a = 0; a < 2; a++) {
      int code = base[2 * q + a];
      if (freebase && freebase[2 * q + a]) {
        code = E13_FORCED;
      } else if (code >= 0 && (code >> 2) == X) {
        code = E13_FORCED;
        nforced++;
      } else if (code >= 0) {
        int nx = code >> 2;
        if (X >= 0 && nx > X) nx--;
        code = (nx << 2) | (code & 3);
      }
      out->tab[2 * row + a] = (int8_t)code;
    }
    row++;
  }
  for (int i = 2 * row; i < E13_MAXCELLS; i++) out->tab[i] = E13_FORCED;
  out->nq = row;
  out->X = X;
  out->nforced = nforced;
  out->sym_lo = out->sym_hi = -1;
  out->sym_n0 = 0;
  /* set_sym (e13search.c:423-426): LO and HI are base states, shifted down
     when above X.  With X = -1 both move down by one (lo > -1 holds for every
     lo >= 0), and a LO that lands below 0 switches symmetry off. */
  const int lo = sym_lo >= 0 ? sym_lo - (sym_lo > X) : -1, hi = sym_hi - (sym_hi >= X);
  if (lo >= 0) { /* the checks at e13search.c:336-351 */
    if (hi >= row) return E13_EINVAL; /* the reference would read past its table */
    uint8_t named[16] = {0};
    /* the scan covers 2*(NQ-1) cells, NQ the base states (e13search.c:341): the
       whole table when a state is deleted, all but the last row when none is */
    for (int i = 0; i < 2 * (nq_base - 1); i++) {
      int c = out->tab[i];
      if (c >= 0 && (c >> 2) >= lo && (c >> 2) <= hi) named[(c >> 2) - lo] = 1;
    }
    int n0 = 0;
    while (n0 <= hi - lo && named[n0]) n0++;
    for (int q = lo + n0; q <= hi; q++)
      if (named[q - lo] || out->tab[2 * q] != E13_FORCED || out->tab[2 * q + 1] != E13_FORCED) return E13_ESYM;
    if (lo == 0) return E13_ESYMSTART;
    out->sym_lo = lo;
    out->sym_hi = hi;
    out->sym_n0 = n0;
  }
  if (freebase) fmax = nforced; /* e13search.c:352 */
  if (nforced > fmax) {
    out->budget = -1;
    return E13_EFORCED;
  }
  out->budget = fmax - nforced;
  return E13_OK;
}

/* ------------------------------------------------------------------------ */
/* Context.                                                                  */
/* ------------------------------------------------------------------------ */
static void tuning_defaults(e13_tuning *t) {
  t->form = 3;
  t->slice = 32;
  t->wg_threads = 256;
  t->wgs_per_cu = 2;
  t->cap = 4096;
  t->batch = 1 << 17;
  t->gate = 1;
  t->refill_tries = 2;
  t->stack_nodes = 1LL << 25;
  t->seed_per_wave = 0;
  t->sched_every = 32;
  t->donate_max = 4;
  t->idle_naps = 4;
  t->mail_every = 1;
  t->trace_events = 0;
  t->graph_chain = 0;
  t->solo_steps = 4096;
  t->evict_after = 16384;
  t->probe_slice = 32;
}

static void tuning_merge(e13_tuning *dst, const e13_tuning *src) {
  if (!src) return;
  if (src->form > 0) dst->form = src->form;
  if (src->slice > 0) dst->slice = src->slice;
  if (src->wg_threads > 0) dst->wg_threads = src->wg_threads;
  if (src->wgs_per_cu > 0) dst->wgs_per_cu = src->wgs_per_cu;
  if (src->cap > 0) dst->cap = src->cap;
  if (src->batch > 0) dst->batch = src->batch;
  if (src->gate > 0) dst->gate = src->gate;
  if (src->refill_tries > 0) dst->refill_tries = src->refill_tries;
  if (src->stack_nodes > 0) dst->stack_nodes = src->stack_nodes;
  if (src->seed_per_wave > 0) dst->seed_per_wave = src->seed_per_wave;
  if (src->seed_per_wave < 0) dst->seed_per_wave = 0; /* negative: no seeding */
  if (src->sched_every > 0) dst->sched_every = src->sched_every;
  if (src->donate_max > 0) dst->donate_max = src->donate_max;
  if (src->idle_naps > 0) dst->idle_naps = src->idle_naps;
  if (src->mail_every > 0) dst->mail_every = src->mail_every;
  if (src->trace_events > 0) dst->trace_events = src->trace_events;
  if (src->trace_events < 0) dst->trace_events = 0; /* negative: tracing off */
  if (src->graph_chain > 0) dst->graph_chain = src->graph_chain;
  if (src->graph_chain < 0) dst->graph_chain = 0; /* negative: host loop */
  if (src->cost_log > 0) dst->cost_log = 1;
  if (src->cost_log < 0) dst->cost_log = 0;
  if (src->learned > 0) dst->learned = src->learned > 3 ? 3 : src->learned;
  if (src->solo_steps > 0) dst->solo_steps = src->solo_steps;
  if (src->solo_steps < 0) dst->solo_steps = 0;
  if (src->evict_after > 0) dst->evict_after = src->evict_after;
  if (src->evict_after < 0) dst->evict_after = 0;
  if (src->learned < 0) dst->learned = 0;
  if (src->probe_slice > 0) dst->probe_slice = src->probe_slice;
  if (src->cut_live > 0) dst->cut_live = src->cut_live;
  if (src->cut_live < 0) dst->cut_live = 0;
  if (src->phase_cycles > 0) dst->phase_cycles = 1;
  if (src->phase_cycles < 0) dst->phase_cycles = 0;
  if (src->nt1_skip > 0) dst->nt1_skip = src->nt1_skip;
  if (src->nt1_skip < 0) dst->nt1_skip = 0;
  if (src->tape_words > 0) dst->tape_words = src->tape_words;
  if (src->tape_words < 0) dst->tape_words = 0;
  if (src->wide_nodes > 0) dst->wide_nodes = src->wide_nodes;
}

static int tuning_valid(const e13_tuning *t) {
  if (t->form < 1 || t->form > 3) return 0;
  if (t->slice < 1 || t->slice > 1 << 16) return 0;
  if (t->probe_slice < 1 || t->probe_slice > 1 << 16) return 0;
  if (t->wg_threads < 64 || t->wg_threads > 256 || t->wg_threads % 64) return 0;
  if (t->wgs_per_cu < 1 || t->wgs_per_cu > 16) return 0;
  if (t->cap < 1) return 0;
  if (t->batch < 1 || t->batch > (1 << 22)) return 0;
  if (t->gate < 1 || t->gate > 2) return 0;
  if (t->refill_tries < 1 || t->refill_tries > 64) return 0;
  if (t->stack_nodes < 1024) return 0;
  if (t->seed_per_wave < 0 || t->seed_per_wave > 64) return 0;
  if (t->sched_every < 1 || t->donate_max < 1 || t->idle_naps < 1 || t->mail_every < 1) return 0;
  if (t->trace_events < 0 || t->trace_events > 1 << 16) return 0;
  if (t->graph_chain < 0 || t->graph_chain > 4096) return 0;
  if (t->cut_live < 0 || t->cut_live > 64) return 0;
  if (t->nt1_skip < 0 || t->nt1_skip > 2) return 0;
  if (t->tape_words != 0 && (t->tape_words < E13_TW_MIN || t->tape_words > E13_TW)) return 0;
  if (t->wide_nodes < 0 || t->wide_nodes > (1 << 24)) return 0;
  return 1;
}

int e13_open(int device, const e13_tuning *t, e13_ctx **out) {
  if (!out) return E13_EINVAL;
  *out = NULL;
  int ndev = 0;
  if (hipGetDeviceCount(&ndev) != hipSuccess || ndev < 1) return E13_ENODEV;
  if (device < 0 || device >= ndev) return E13_EINVAL;
  e13_ctx *c = (e13_ctx *)calloc(1, sizeof *c);
  if (!c) return E13_ENOMEM;
  c->device = device;
  if (hipSetDevice(device) != hipSuccess) {
    free(c);
    return E13_EHIP;
  }
  hipDeviceProp_t prop;
  if (hipGetDeviceProperties(&prop, device) != hipSuccess) {
    free(c);
    return E13_EHIP;
  }
  c->ncu = prop.multiProcessorCount;
  c->lds_max = prop.sharedMemPerBlock;
  {
    int lds_cu = 0;
    if (hipDeviceGetAttribute(&lds_cu, hipDeviceAttributeMaxSharedMemoryPerMultiprocessor, device) != hipSuccess ||
        lds_cu <= 0)
      lds_cu = (int)prop.sharedMemPerBlock;
    c->lds_cu = (size_t)lds_cu;
  }
  snprintf(c->arch, sizeof c->arch, "%s", prop.gcnArchName);
  /* defaults, then this device's stored (autotuned) settings, then the caller's;
     the small-search profile likewise, over the stored main settings */
  tuning_defaults(&c->tune);
  char path[1024];
  e13_tuning stored;
  if (e13i_tuning_path(c->arch, c->ncu, 0, path, sizeof path) && e13_load_tuning(path, &stored) == E13_OK) {
    e13_tuning tmp = c->tune;
    tuning_merge(&tmp, &stored);
    if (tuning_valid(&tmp)) c->tune = tmp;
  }
  c->tune_small = c->tune;
  if (e13i_tuning_path(c->arch, c->ncu, 1, path, sizeof path) && e13_load_tuning(path, &stored) == E13_OK) {
    e13_tuning tmp = c->tune_small;
    tuning_merge(&tmp, &stored);
    if (tuning_valid(&tmp)) {
      c->tune_small = tmp;
      c->has_small = 1;
    }
  }
  c->auto_profile = 1;
  tuning_merge(&c->tune, t);
  tuning_merge(&c->tune_small, t);
  if (!tuning_valid(&c->tune) || !tuning_valid(&c->tune_small)) {
    free(c);
    return E13_EINVAL;
  }
  c->tune_open = c->tune;
  c->tune_small_open = c->tune_small;
  /* the static scratch of e13_estimate_static: its address on this device */
  if (e13k_scratch_addr(device, &c->scratch) != 0) c->scratch = NULL;
  if (hipStreamCreate(&c->stream) != hipSuccess || hipEventCreate(&c->ev0) != hipSuccess ||
      hipEventCreate(&c->ev1) != hipSuccess) {
    free(c);
    return E13_EHIP;
  }
  uint64_t key[1364], tgt[E13_NS + 1];
  e13_keys(key, tgt);
  int32_t dl[E13_NS + 1];
  dl[0] = 0;
  for (int n = 1; n <= E13_NS; n++) dl[n] = (int32_t)(4 * T14S[n] + 64); /* BUDGET_MULT*T14S[n]+64 */
  if (hipMalloc((void **)&c->d_ctr, sizeof(E13Counters)) != hipSuccess ||
      hipMalloc((void **)&c->d_key, sizeof key) != hipSuccess ||
      hipMalloc((void **)&c->d_tgt, sizeof tgt) != hipSuccess ||
      hipMalloc((void **)&c->d_pb, sizeof(E13Problem)) != hipSuccess ||
      hipMalloc((void **)&c->d_proot, sizeof(E13Node)) != hipSuccess ||
      hipMalloc((void **)&c->d_pctr, sizeof(E13Counters)) != hipSuccess ||
      hipMalloc((void **)&c->d_dl, sizeof dl) != hipSuccess) {
    e13_close(c);
    return E13_ENOMEM;
  }
  if (hipMemcpy(c->d_key, key, sizeof key, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_tgt, tgt, sizeof tgt, hipMemcpyHostToDevice) != hipSuccess ||
      hipMemcpy(c->d_dl, dl, sizeof dl, hipMemcpyHostToDevice) != hipSuccess) {
    e13_close(c);
    return E13_EHIP;
  }
  *out = c;
  return E13_OK;
}

void e13_close(e13_ctx *c) {
  if (!c) return;
  hipSetDevice(c->device);
  if (c->stream) hipStreamSynchronize(c->stream);
  hipFree(c->d_ctr);
  hipFree(c->d_key);
  hipFree(c->d_tgt);
  hipFree(c->d_pb);
  hipFree(c->d_proot);
  hipFree(c->d_pctr);
  free(c->seed_est);
  hipFree(c->d_dl);
  hipFree(c->d_stack);
  hipFree(c->d_out);
  hipFree(c->d_surv);
  hipFree(c->d_tabs);
  hipFree(c->d_caps);
  hipFree(c->d_leaves);
  hipFree(c->d_stacks);
  hipFree(c->d_mail);
  hipFree(c->d_hungry);
  hipFree(c->d_sch);
  hipFree(c->d_pbs);
  hipFree(c->d_pcnt);
  hipFree(c->d_seeds);
  hipFree(c->d_trace);
  e13i_free_cost(c);
  e13i_free_estimate(c);
  hipFree(c->d_wave_state);
  hipFree(c->d_wide);
  if (c->h_stop) hipHostFree(c->h_stop);
  hipFree(c->d_chain);
  if (c->ev0) hipEventDestroy(c->ev0);
  if (c->ev1) hipEventDestroy(c->ev1);
  if (c->stream) hipStreamDestroy(c->stream);
  free(c);
}

int e13_set_tuning(e13_ctx *c, const e13_tuning *t) {
  if (!c) return E13_EINVAL;
  if (!t) { /* back to what e13_open installed, with the automatic profile choice */
    c->tune = c->tune_open;
    c->tune_small = c->tune_small_open;
    c->auto_profile = 1;
    return E13_OK;
  }
  e13_tuning n = c->tune;
  tuning_merge(&n, t);
  if (!tuning_valid(&n)) return E13_EINVAL;
  c->tune = n;
  c->auto_profile = 0; /* installed settings hold for every search from now on */
  return E13_OK;
}

/* Edit depth of a problem: the cells a search binds with a free choice of
   values (E13_FORCED: free cells and rules into the deleted state) plus the
   edits it may make to the others. */
static int edit_depth(const e13_problem *p) {
  int d = p->budget;
  for (int i = 0; i < 2 * p->nq && i < E13_MAXCELLS; i++) d += p->tab[i] == E13_FORCED;
  return d;
}

int e13_search_profile(e13_ctx *c, const e13_problem *ps, int np, e13_tuning *t) {
  if (!c || !ps || np < 1) return E13_EINVAL;
  int small = c->has_small && c->auto_profile;
  for (int k = 0; k < np && small; k++)
    if (edit_depth(&ps[k]) > E13_SMALL_DEPTH) small = 0;
  if (t) *t = small ? c->tune_small : c->tune;
  return small ? E13_PROFILE_SMALL : E13_PROFILE_MAIN;
}

int e13_get_tuning(e13_ctx *c, e13_tuning *t) {
  if (!c || !t) return E13_EINVAL;
  *t = c->tune;
  return E13_OK;
}

int e13_get_stats(e13_ctx *c, e13_stats *s) {
  if (!c || !s) return E13_EINVAL;
  *s = c->stats;
  return E13_OK;
}

int e13_sync(e13_ctx *c) {
  if (!c) return E13_EINVAL;
  HIPCHK(hipStreamSynchronize(c->stream));
  return E13_OK;
}

/* Workgroup size that fits the LDS limit (drops waves, never below one). */
static int fit_wg(const e13_ctx *c, int wg, int gate_hash) {
  while (wg > 64 && e13k_lds_bytes(wg, gate_hash) > c->lds_max) wg -= 64;
  return wg;
}

int e13i_ensure_buf(void **p, size_t *cap, size_t need, size_t elem) {
  if (*cap >= need) return E13_OK;
  hipFree(*p);
  *p = NULL;
  *cap = 0;
  if (hipMalloc(p, need * elem) != hipSuccess) return E13_ENOMEM;
  *cap = need;
  return E13_OK;
}

/* ------------------------------------------------------------------------ */
/* Search.                                                                   */
/* ------------------------------------------------------------------------ */
void e13i_fill_problem(const e13_problem *p, E13Problem *pb) {
  memset(pb, 0, sizeof *pb);
  pb->nq = p->nq;
  pb->nslot = 4 * p->nq + 1;
  pb->X = p->X;
  pb->sym_lo = p->sym_lo;
  pb->sym_hi = p->sym_hi;
  for (int n = 1; n <= E13_NS; n++) pb->dl[n] = (int32_t)(4 * T14S[n] + 64); /* BUDGET_MULT*T14S+64 */
  for (int i = 0; i < E13_MAXCELLS; i++) pb->base[i] = i < 2 * p->nq ? p->tab[i] : E13_FORCED;
}

/* The blank-tape root (sim_init, e13search.c:123-129; init_root, e13gpu.py:276-283). */
void e13i_root_node(const e13_problem *p, E13Node *nd) {
  memset(nd, 0, sizeof *nd);
  nd->q = E13_QOFF;
  nd->st = 0;
  nd->t = 0;
  nd->nt = 1;
  nd->stage = 0;
  nd->st0 = -1;
  nd->q0 = E13_QOFF;
  nd->flags = 3;
  nd->budget = p->budget;
  nd->nintro = p->sym_n0;
  nd->pend = -1;
  for (int i = 0; i < E13_MAXCELLS; i++) {
    int c = i < 2 * p->nq ? p->tab[i] : E13_FORCED;
    uint8_t code;
    if (c == E13_FORCED) code = E13_C_FREE;
    else if (p->budget > 0) code = E13_C_EDIT; /* every unread cell branches while edits are left */
    else code = c < 0 ? E13_C_HALT : (uint8_t)c;
    nd->tab[i] = code;
  }
}

static void finish_survivor(const e13_problem *p, const E13SurvRec *r, e13_survivor *s, int prob);

/* Survivor records to survivor lines, each re-checked by the plain simulation:
   independent machines of up to ~835k steps each, so on host threads. */
typedef struct {
  const e13_problem *ps;
  int np;
  const E13SurvRec *rec;
  e13_survivor *surv;
  uint64_t n, first, stride;
} finish_job;

static void *finish_worker(void *arg) {
  const finish_job *j = (const finish_job *)arg;
  for (uint64_t k = j->first; k < j->n; k += j->stride) {
    int pr = j->rec[k].prob;
    if (pr < 0 || pr >= j->np) pr = 0;
    finish_survivor(&j->ps[pr], &j->rec[k], &j->surv[k], pr);
  }
  return NULL;
}

static void finish_all(const e13_problem *ps, int np, const E13SurvRec *rec, e13_survivor *surv, uint64_t n) {
  long cpus = sysconf(_SC_NPROCESSORS_ONLN);
  uint64_t nt = cpus > 1 ? (uint64_t)cpus : 1;
  if (nt > 16) nt = 16;
  if (nt > n) nt = n;
  pthread_t th[16];
  finish_job jobs[16];
  uint64_t started = 0;
  for (uint64_t t = 0; t < nt; t++) {
    jobs[t] = (finish_job){ps, np, rec, surv, n, t, nt};
    if (t > 0 && pthread_create(&th[t], NULL, finish_worker, &jobs[t]) == 0) started |= 1ull << t;
  }
  finish_worker(&jobs[0]);
  for (uint64_t t = 1; t < nt; t++) {
    if (started >> t & 1) pthread_join(th[t], NULL);
    else finish_worker(&jobs[t]); /* no thread: do its share here */
  }
}


static int check_problem(const e13_problem *p) {
  if (!p || p->nq < 1 || p->nq > E13_MAXNQ || p->budget < 0) return 0;
  for (int i = 0; i < 2 * p->nq; i++) {
    int c = p->tab[i];
    if (c < E13_FORCED || c >= 4 * p->nq) return 0;
  }
  if (p->sym_lo >= 0 && (p->sym_hi < p->sym_lo || p->sym_hi >= p->nq)) return 0;
  return 1;
}

/* The host-loop search as hipGraph launch chains: graph_chain iterations of
   (pop a batch, expand it, push the children) captured once, then the graph is
   launched until the device cursor finds the stack empty.  One host round
   trip per graph instead of per launch; the results are those of the plain
   host loop (same batches in the same order). */
static int run_chain(e13_ctx *c, E13ExpandArgs *a, size_t batch, int grid, int wg, int refill, size_t *peak) {
  const e13_tuning *tu = &c->tune;
  if (!c->d_chain && hipMalloc((void **)&c->d_chain, sizeof(E13Chain)) != hipSuccess) return E13_ENOMEM;
  E13Chain ch;
  memset(&ch, 0, sizeof ch);
  ch.top = 1; /* the root, already on the stack */
  ch.peak = 1;
  ch.batch = (uint32_t)batch;
  HIPCHK(hipMemcpyAsync(c->d_chain, &ch, sizeof ch, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  a->chain = c->d_chain;
  a->stack = c->d_stack;
  a->in = NULL;
  a->nin = 0;
  const int push_grid = 4 * c->ncu;
  hipGraph_t g = NULL;
  hipGraphExec_t ge = NULL;
  HIPCHK(hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal));
  int lr = 0;
  for (int i = 0; i < tu->graph_chain && lr == 0; i++) {
    lr = e13k_launch_chain_pop(c->d_chain, c->d_ctr, c->stream);
    if (lr == 0) lr = e13k_launch_expand(a, grid, wg, tu->gate == 2, refill, c->stream);
    if (lr == 0)
      lr = e13k_launch_chain_push(c->d_chain, c->d_stack, (uint64_t)c->stack_cap, c->d_out, a->out_cap, c->d_ctr,
                                  push_grid, c->stream);
  }
  hipError_t ce = hipStreamEndCapture(c->stream, &g);
  if (lr != 0 || ce != hipSuccess) {
    fprintf(stderr, "libe13: graph capture failed: %s\n", hipGetErrorString(lr ? (hipError_t)lr : ce));
    if (g) hipGraphDestroy(g);
    return E13_EHIP;
  }
  if (hipGraphInstantiate(&ge, g, NULL, NULL, 0) != hipSuccess) {
    hipGraphDestroy(g);
    return E13_EHIP;
  }
  int rc = E13_OK;
  for (;;) {
    float ms = 0;
    if (hipEventRecord(c->ev0, c->stream) != hipSuccess || hipGraphLaunch(ge, c->stream) != hipSuccess ||
        hipEventRecord(c->ev1, c->stream) != hipSuccess ||
        hipMemcpyAsync(&ch, c->d_chain, sizeof ch, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess || hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) {
      rc = E13_EHIP;
      break;
    }
    c->stats.kernel_ms += ms;
    c->stats.launches += 3 * tu->graph_chain;
    c->stats.graphs += 1;
    if (ch.error) {
      if (ch.error == 2) fprintf(stderr, "libe13: dynamic LDS does not start at LDS address 0\n");
      rc = ch.error == 2 ? E13_ELDS : E13_ESTACK;
      break;
    }
    if (ch.top == 0) break;
  }
  *peak = (size_t)ch.peak;
  hipGraphExecDestroy(ge);
  hipGraphDestroy(g);
  a->chain = NULL;
  a->stack = NULL;
  return rc;
}

int e13_search(e13_ctx *c, const e13_problem *p, e13_counts *out, e13_survivor *surv, uint64_t surv_cap,
               uint64_t *n_surv) {
  if (!c || !out || !check_problem(p)) return E13_EINVAL;
  const int ckpt = c->ckpt_path[0] != 0;
  if (c->tune.form == 3) return e13_search_many(c, p, 1, out, surv, surv_cap, n_surv);
  HIPCHK(hipSetDevice(c->device));
  const e13_tuning *tu = &c->tune;
  const int form = tu->form == 1 ? 1 : 2;
  const int nslot = 4 * p->nq + 1;
  const size_t batch = (size_t)tu->batch;
  c->stats.frontier = 1; /* the host loop starts from the root */
  int rc;
  if ((rc = e13i_ensure_buf((void **)&c->d_stack, &c->stack_cap, (size_t)tu->stack_nodes, sizeof(E13Node))))
    return rc;
  if ((rc = e13i_ensure_buf((void **)&c->d_out, &c->out_cap, batch * (size_t)nslot, sizeof(E13Node)))) return rc;
  size_t want_surv = surv_cap > 4096 ? (size_t)surv_cap : 4096;
  if ((rc = e13i_ensure_buf((void **)&c->d_surv, &c->surv_cap, want_surv, sizeof(E13SurvRec)))) return rc;

  E13ExpandArgs a;
  memset(&a, 0, sizeof a);
  E13Problem pb;
  e13i_fill_problem(p, &pb);
  HIPCHK(hipMemcpyAsync(c->d_pb, &pb, sizeof pb, hipMemcpyHostToDevice, c->stream));
  a.pb = c->d_pb;
  a.nslot = (uint32_t)nslot;
  a.nq = p->nq;
  a.out = c->d_out;
  a.ctr = c->d_ctr;
  a.surv = c->d_surv;
  a.key = c->d_key;
  a.tgt = c->d_tgt;
  a.out_cap = (uint32_t)(batch * (size_t)nslot);
  a.surv_cap = (uint32_t)c->surv_cap;
  a.slice = tu->slice;
  a.cap = tu->cap;
  a.refill_tries = tu->refill_tries;
  a.chunk = 256;

  E13Node root;
  e13i_root_node(p, &root);
  HIPCHK(hipMemsetAsync(c->d_ctr, 0, sizeof(E13Counters), c->stream));
  HIPCHK(hipMemcpyAsync(c->d_stack, &root, sizeof root, hipMemcpyHostToDevice, c->stream));
  size_t top = 1, peak = 1;
  const int refill = form == 2;
  const int wg = fit_wg(c, tu->wg_threads, tu->gate == 2);
  if (e13k_lds_bytes(wg, tu->gate == 2) > c->lds_max) return E13_EINVAL;
  int grid_used = 0;
  /* checkpoint, clean stop and resume (e13gpu.py:349-409) */
  int64_t launches = 0, call_launches = 0;
  double prior_ms = 0, t_start = host_ms(), last_ckpt = t_start, last_stop_check = 0;
  char stop_path[1100];
  stop_path[0] = 0;
  if (ckpt) {
    snprintf(stop_path, sizeof stop_path, "%s.stop", c->ckpt_path);
    if ((rc = e13i_ckpt_load(c, p, &top, &peak, &launches, &prior_ms)) < 0) return rc;
    if (rc == 1 && getenv("E13_TIMING"))
      fprintf(stderr, "libe13: resumed from %s: stack=%zu launches=%lld\n", c->ckpt_path, top, (long long)launches);
    remove(stop_path); /* left over from the last clean stop */
  }
  if (tu->graph_chain > 0 && !ckpt) {
    grid_used = refill ? c->ncu * tu->wgs_per_cu : (int)((batch * (size_t)nslot + (size_t)wg - 1) / (size_t)wg);
    if ((rc = run_chain(c, &a, batch, grid_used, wg, refill, &peak))) return rc;
    top = 0;
  }
  while (top > 0) {
    if (ckpt) {
      const double now = host_ms();
      int stop = c->ckpt_max_launches > 0 && call_launches >= c->ckpt_max_launches;
      if (now - last_stop_check > 100.0) { /* the stop file, at most ten times a second */
        last_stop_check = now;
        stop |= access(stop_path, F_OK) == 0;
      }
      if (stop || now - last_ckpt > 1000.0 * c->ckpt_every_s) {
        if ((rc = e13i_ckpt_save(c, p, top, peak, launches, prior_ms + now - t_start))) return rc;
        last_ckpt = host_ms();
        if (stop) {
          remove(stop_path);
          c->stats.peak_stack = (int64_t)peak;
          return E13_ESTOPPED;
        }
      }
    }
    size_t nin = top < batch ? top : batch;
    size_t base_i = top - nin;
    a.in = c->d_stack + base_i;
    a.nin = (uint32_t)nin;
    /* item cursor and child count restart at 0 each launch */
    HIPCHK(hipMemsetAsync(&c->d_ctr->item, 0, 2 * sizeof(uint32_t), c->stream));
    int grid;
    if (refill) {
      grid = c->ncu * tu->wgs_per_cu;
    } else {
      size_t items = nin * (size_t)nslot;
      grid = (int)((items + (size_t)wg - 1) / (size_t)wg);
    }
    grid_used = grid;
    HIPCHK(hipEventRecord(c->ev0, c->stream));
    int lr = e13k_launch_expand(&a, grid, wg, tu->gate == 2, refill, c->stream);
    if (lr != 0) {
      fprintf(stderr, "libe13: expand launch failed: %s\n", hipGetErrorString((hipError_t)lr));
      return E13_EHIP;
    }
    HIPCHK(hipEventRecord(c->ev1, c->stream));
    uint32_t io[4];
    HIPCHK(hipMemcpyAsync(io, &c->d_ctr->item, sizeof io, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    float ms = 0;
    HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
    c->stats.kernel_ms += ms;
    c->stats.launches++;
    launches++;
    call_launches++;
    uint32_t nout = io[1];
    if (io[3] & 4u) {
      fprintf(stderr, "libe13: dynamic LDS does not start at LDS address 0\n");
      return E13_ELDS;
    }
    if (io[3]) return E13_ESTACK; /* output buffer overflow (cannot happen: one child per item) */
    if (base_i + nout > c->stack_cap) return E13_ESTACK;
    if (nout)
      HIPCHK(hipMemcpyAsync(c->d_stack + base_i, c->d_out, (size_t)nout * sizeof(E13Node), hipMemcpyDeviceToDevice,
                            c->stream));
    top = base_i + nout;
    if (top > peak) peak = top;
  }
  E13Counters h;
  HIPCHK(hipMemcpyAsync(&h, c->d_ctr, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  out->leaves = h.leaves;
  out->steps = h.steps;
  out->survivors = h.survivors;
  for (int j = 0; j <= E13_NS; j++) out->dead_at_stage[j] = h.hist[j];
  c->stats.rounds += (int64_t)h.rounds;
  c->stats.lane_slots += (int64_t)h.lane_slots;
  for (int j = 0; j < 4; j++) c->stats.cycles[j] += (int64_t)h.cyc[j];
  if ((int64_t)peak > c->stats.peak_stack) c->stats.peak_stack = (int64_t)peak;
  c->stats.slice = tu->slice;
  c->stats.wg_threads = wg;
  c->stats.grid = grid_used;
  c->stats.gate = tu->gate;
  c->stats.form = form;

  uint64_t ns = h.nsurv;
  if (n_surv) *n_surv = ns;
  uint64_t keep = ns < c->surv_cap ? ns : c->surv_cap;
  if (surv && surv_cap && keep) {
    uint64_t nk = keep < surv_cap ? keep : surv_cap;
    E13SurvRec *rec = (E13SurvRec *)malloc(nk * sizeof(E13SurvRec));
    if (!rec) return E13_ENOMEM;
    HIPCHK(hipMemcpy(rec, c->d_surv, nk * sizeof(E13SurvRec), hipMemcpyDeviceToHost));
    finish_all(p, 1, rec, surv, nk);
    free(rec);
  }
  return E13_OK;
}

/* Survivor records -> e13_survivor, with the plain re-check and read mask. */
static void finish_survivor(const e13_problem *p, const E13SurvRec *r, e13_survivor *s, int prob) {
  memset(s, 0, sizeof *s);
  for (int i = 0; i < E13_MAXCELLS; i++) {
    int code = r->tab[i], v;
    if (i >= 2 * p->nq) v = E13_FORCED;
    else if (code == E13_C_EDIT) v = p->tab[i]; /* never read: still the base value */
    else if (code == E13_C_HALT) v = E13_HALT;
    else if (code == E13_C_FREE) v = E13_FORCED;
    else v = code;
    s->tab[i] = (int8_t)v;
  }
  s->st0 = r->st0;
  s->head0 = r->head0;
  s->abs_ok = r->abs_ok;
  s->rel_ok = r->rel_ok;
  s->steps = r->steps;
  s->ch[0] = r->ch[0];
  s->ch[1] = r->ch[1];
  s->ch[2] = r->ch[2];
  s->prob = prob;
  uint32_t mask = 0;
  s->recheck_ok = e13_recheck(s->tab, p->nq, s->st0, s->head0, s->abs_ok, s->rel_ok, s->steps, &mask);
  s->read_mask = mask;
}

/* ---- the frontier and its cost estimates ---- */
/* Estimated subtree steps below the frontier nodes d[0..n) (host copy hn): the
   installed network's prediction (mode 1) or the mean of E13_FRONTIER_PROBES
   probes from each node (mode 3, the non-learned estimate: the reference's
   probe() started at the node).  A node paused at a cap (no pending cell) is
   ranked costliest: it has run past the cap already. */
static int frontier_estimate(e13_ctx *c, const e13_problem *ps, int mode, const E13Node *d, const E13Node *hn,
                             size_t n, double *est) {
  if (mode == 3) {
    for (size_t i0 = 0; i0 < n;) { /* one launch per run of nodes of one problem */
      size_t i1 = i0 + 1;
      while (i1 < n && hn[i1].prob == hn[i0].prob) i1++;
      const int k = hn[i0].prob;
      const int rc = e13i_probe_nodes(c, &ps[k], c->d_pbs + k, d + i0, i1 - i0, E13_FRONTIER_PROBES,
                                      0xF20E5EEDull + (uint64_t)i0, est + i0);
      if (rc) return rc;
      i0 = i1;
    }
    for (size_t i = 0; i < n; i++)
      if (est[i] < 0.0) est[i] = 1e300;
    return E13_OK;
  }
  float *x = (float *)malloc(n * E13_NFEAT * sizeof(float)), *y = (float *)malloc(n * 2 * sizeof(float));
  int rc = E13_ENOMEM;
  if (x && y) {
    for (size_t i = 0; i < n; i++) e13i_node_features(&ps[hn[i].prob], &hn[i], x + i * E13_NFEAT);
    rc = e13i_mlp_predict_dev(c, x, (int)n, y);
    if (!rc) /* the network's second output is log1p(subtree steps) / 24 */
      for (size_t i = 0; i < n; i++) est[i] = expm1(24.0 * (double)y[2 * i + 1]);
  }
  free(x);
  free(y);
  return rc;
}

static const double *g_sort_est;
static int by_est_desc(const void *a, const void *b) {
  const double ca = g_sort_est[*(const int32_t *)a], cb = g_sort_est[*(const int32_t *)b];
  if (ca != cb) return ca > cb ? -1 : 1;
  return *(const int32_t *)a - *(const int32_t *)b;
}

/* Nodes d[0..n) by estimated cost, costliest first: host copy *hn_out, order
   *ord_out, estimates *est_out (indexed like hn). */
static int frontier_sorted(e13_ctx *c, const e13_problem *ps, int mode, const E13Node *d, size_t n,
                           E13Node **hn_out, int32_t **ord_out, double **est_out) {
  E13Node *hn = (E13Node *)malloc(n * sizeof(E13Node));
  double *est = (double *)malloc(n * sizeof(double));
  int32_t *ord = (int32_t *)malloc(n * sizeof(int32_t));
  int rc = E13_ENOMEM;
  if (hn && est && ord) {
    rc = E13_OK;
    if (hipMemcpyAsync(hn, d, n * sizeof(E13Node), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = E13_EHIP;
    if (!rc) rc = frontier_estimate(c, ps, mode, d, hn, n, est);
    if (!rc) {
      for (size_t i = 0; i < n; i++) ord[i] = (int32_t)i;
      g_sort_est = est;
      qsort(ord, n, sizeof(int32_t), by_est_desc);
    }
  }
  if (rc) {
    free(hn);
    free(est);
    free(ord);
    return rc;
  }
  *hn_out = hn;
  *ord_out = ord;
  *est_out = est;
  return E13_OK;
}

/* Split step of the ordered seeding: the costlier half of the frontier d[0..n)
   stays in d (to be expanded), the rest goes to `seeds`; *ne = nodes kept. */
static int learned_split(e13_ctx *c, const e13_problem *ps, int mode, E13Node *d, size_t n, E13Node *seeds,
                         size_t *ne) {
  E13Node *hn = NULL;
  int32_t *ord = NULL;
  double *est = NULL;
  int rc = frontier_sorted(c, ps, mode, d, n, &hn, &ord, &est);
  if (rc) return rc;
  const size_t keep = (n + 1) / 2;
  E13Node *tmp = (E13Node *)malloc(n * sizeof(E13Node));
  if (!tmp) rc = E13_ENOMEM;
  if (!rc) {
    for (size_t i = 0; i < n; i++) tmp[i] = hn[ord[i]];
    if (hipMemcpyAsync(d, tmp, keep * sizeof(E13Node), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        (n > keep &&
         hipMemcpyAsync(seeds, tmp + keep, (n - keep) * sizeof(E13Node), hipMemcpyHostToDevice, c->stream) !=
             hipSuccess) ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = E13_EHIP;
  }
  free(tmp);
  free(hn);
  free(ord);
  free(est);
  *ne = keep;
  return rc;
}

/* Order step: the whole frontier costliest first, so the waves that are dealt
   seeds round-robin each start with one of the largest subtrees; the estimates
   are kept in the new order for the cost log. */
static int learned_order(e13_ctx *c, const e13_problem *ps, int mode, E13Node *d, size_t n) {
  E13Node *hn = NULL;
  int32_t *ord = NULL;
  double *est = NULL;
  int rc = frontier_sorted(c, ps, mode, d, n, &hn, &ord, &est);
  if (rc) return rc;
  E13Node *tmp = (E13Node *)malloc(n * sizeof(E13Node));
  if (!tmp) rc = E13_ENOMEM;
  if (!rc && n > c->seed_est_cap) {
    double *se = (double *)realloc(c->seed_est, n * sizeof(double));
    if (!se) rc = E13_ENOMEM;
    else {
      c->seed_est = se;
      c->seed_est_cap = n;
    }
  }
  if (!rc) {
    for (size_t i = 0; i < n; i++) {
      tmp[i] = hn[ord[i]];
      c->seed_est[i] = est[ord[i]];
    }
    c->seed_est_n = n;
    if (hipMemcpyAsync(d, tmp, n * sizeof(E13Node), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = E13_EHIP;
  }
  free(tmp);
  free(hn);
  free(ord);
  free(est);
  return rc;
}

/* A seeded random order of the frontier (tuning.learned = 2; a check that the
   order of the frontier never changes the results). */
static int random_order(e13_ctx *c, E13Node *d, size_t n) {
  E13Node *hn = (E13Node *)malloc(n * sizeof(E13Node));
  int rc = hn ? E13_OK : E13_ENOMEM;
  if (!rc && (hipMemcpyAsync(hn, d, n * sizeof(E13Node), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipStreamSynchronize(c->stream) != hipSuccess))
    rc = E13_EHIP;
  if (!rc) {
    uint64_t st = 0x0D0E5EEDull + n;
    for (size_t i = n - 1; i > 0; i--) {
      const size_t j = (size_t)(e13i_splitmix(&st) % (uint64_t)(i + 1));
      E13Node t = hn[i];
      hn[i] = hn[j];
      hn[j] = t;
    }
    if (hipMemcpyAsync(d, hn, n * sizeof(E13Node), hipMemcpyHostToDevice, c->stream) != hipSuccess ||
        hipStreamSynchronize(c->stream) != hipSuccess)
      rc = E13_EHIP;
  }
  free(hn);
  return rc;
}

/* The cost log of a finished search: each seed's features and subtree counts. */
static int collect_cost_log(e13_ctx *c, const e13_problem *ps, size_t n) {
  c->cost_n = 0;
  if (!n) return E13_OK;
  if ((int64_t)n > c->cost_cap) {
    e13_cost_rec *r = (e13_cost_rec *)realloc(c->cost_log, n * sizeof(e13_cost_rec));
    if (!r) return E13_ENOMEM;
    c->cost_log = r;
    c->cost_cap = (int64_t)n;
  }
  E13Node *hn = (E13Node *)malloc(n * sizeof(E13Node));
  uint64_t *cnt = (uint64_t *)malloc(2 * n * sizeof(uint64_t));
  int rc = E13_OK;
  if (!hn || !cnt) rc = E13_ENOMEM;
  if (!rc && (hipMemcpyAsync(hn, c->d_seeds, n * sizeof(E13Node), hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
              hipMemcpyAsync(cnt, c->d_seed_cnt, 2 * n * sizeof(uint64_t), hipMemcpyDeviceToHost, c->stream) !=
                  hipSuccess ||
              hipStreamSynchronize(c->stream) != hipSuccess))
    rc = E13_EHIP;
  if (!rc) {
    for (size_t i = 0; i < n; i++) {
      e13_cost_rec *r = &c->cost_log[i];
      e13i_node_features(&ps[hn[i].prob], &hn[i], r->x);
      r->leaves = cnt[2 * i];
      r->steps = cnt[2 * i + 1];
      r->prob = hn[i].prob;
      r->depth = hn[i].depth;
      r->est_steps = c->seed_est_n == n ? c->seed_est[i] : 0.0;
    }
    c->cost_n = (int64_t)n;
  }
  free(hn);
  free(cnt);
  return rc;
}

static double host_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

/* The persistent form: every problem in one launch (see k_search). */
/* Workgroups of the persistent kernel that fit a CU: the runtime's figure, and
   no more than the CU's LDS holds with each workgroup's share rounded up to
   the allocation unit, taken as 1/128 of the CU's LDS (1280 bytes on gfx950:
   measured, the runtime's figure ignores it; 512 bytes on 64 KB parts).  The
   kernel also checks at start that every workgroup is resident. */
static int search_fit(const e13_ctx *c, int wg, int gate_hash, int np, int tw) {
  const int f = e13k_search_max_wgs(wg, gate_hash, np, tw);
  const size_t unit = c->lds_cu >= 128 * 512 ? c->lds_cu / 128 : 512;
  const size_t lds = (e13k_search_lds_bytes(wg, gate_hash, np, tw) + unit - 1) / unit * unit;
  const int f_lds = (int)(c->lds_cu / lds);
  return f < f_lds ? f : f_lds;
}

int e13_search_many(e13_ctx *c, const e13_problem *ps, int np, e13_counts *out, e13_survivor *surv,
                    uint64_t surv_cap, uint64_t *n_surv) {
  if (!c || !ps || !out || np < 1 || np > E13_MAXPROB) return E13_EINVAL;
  for (int k = 0; k < np; k++)
    if (!check_problem(&ps[k])) return E13_EINVAL;
  const int timing = getenv("E13_TIMING") && *getenv("E13_TIMING") != '0'; /* host phase times on stderr */
  double tm[5] = {0, 0, 0, 0, 0};
  tm[0] = timing ? host_ms() : 0;
  HIPCHK(hipSetDevice(c->device));
  const int profile = e13_search_profile(c, ps, np, NULL);
  const e13_tuning *tu = profile == E13_PROFILE_SMALL ? &c->tune_small : &c->tune;
  c->stats.profile = profile;
  const int gh = tu->gate == 2;
  int wg = tu->wg_threads;
  while (wg > 64 && e13k_search_lds_bytes(wg, gh, np, E13_TW) > c->lds_max) wg -= 64;
  if (e13k_search_lds_bytes(wg, gh, np, E13_TW) > c->lds_max) return E13_EINVAL;
  /* every workgroup of the persistent kernel must be resident at once: a wave
     that could start only after others exit would hold its frontier share and
     its place in the busy count while they wait for it */
  const int fit_wide = search_fit(c, wg, gh, np, E13_TW);
  if (fit_wide < 1) return E13_EINVAL;
  /* The narrow form: with a shorter lane tape in LDS more workgroups fit a CU.
     tape_words 0: when the whole tape does not fit wgs_per_cu workgroups, the
     widest tape with which the most of them (up to wgs_per_cu) fit; 18..42:
     that tape; 43: the whole tape.  Machines that go past the short tape
     finish in a launch with the whole tape (the wide launch below). */
  int tw = E13_TW, fit = fit_wide < tu->wgs_per_cu ? fit_wide : tu->wgs_per_cu;
  if (c->narrow_off) {
    /* a short-tape launch of this context found workgroups not resident: whole tape only */
  } else if (!gh && tu->tape_words >= E13_TW_MIN && tu->tape_words < E13_TW) {
    const int f = search_fit(c, wg, 0, np, tu->tape_words);
    if (f < 1) return E13_EINVAL;
    tw = tu->tape_words;
    fit = f < tu->wgs_per_cu ? f : tu->wgs_per_cu;
  } else if (!gh && tu->tape_words == 0) {
    for (int t = E13_TW - 1; t >= E13_TW_MIN && fit < tu->wgs_per_cu; t--) {
      int f = search_fit(c, wg, 0, np, t);
      if (f > tu->wgs_per_cu) f = tu->wgs_per_cu;
      if (f > fit) {
        fit = f;
        tw = t;
      }
    }
  }
  /* launch shapes: [0] the first launch (narrow when tw < E13_TW), [1] the whole tape */
  int grid_k[2], waves_k[2], stk_k[2];
  grid_k[0] = c->ncu * fit;
  grid_k[1] = c->ncu * (fit_wide < tu->wgs_per_cu ? fit_wide : tu->wgs_per_cu);
  size_t need_stacks = 0;
  for (int k = 0; k < 2; k++) {
    waves_k[k] = grid_k[k] * (wg / 64);
    /* each wave's stack is a ring of a power-of-two number of nodes, at least its
       share of stack_nodes (a mask, not a 64-bit remainder, maps the ring index
       to a slot) */
    stk_k[k] = 64;
    while ((int64_t)stk_k[k] * waves_k[k] < tu->stack_nodes && stk_k[k] < (1 << 24)) stk_k[k] *= 2;
    if ((size_t)waves_k[k] * (size_t)stk_k[k] > need_stacks) need_stacks = (size_t)waves_k[k] * (size_t)stk_k[k];
    if (waves_k[k] > 64 * 64) return E13_EINVAL; /* the hungry map is read as at most 64 words */
  }
  const int max_waves = waves_k[0] > waves_k[1] ? waves_k[0] : waves_k[1];
  int rc;
  if ((rc = e13i_ensure_buf((void **)&c->d_stacks, &c->stacks_cap, need_stacks, sizeof(E13Node)))) return rc;
  if (c->mail_cap < (size_t)max_waves) {
    hipFree(c->d_mail);
    hipFree(c->d_hungry);
    c->d_mail = NULL;
    c->d_hungry = NULL;
    c->mail_cap = 0;
    if (hipMalloc((void **)&c->d_mail, (size_t)max_waves * sizeof(E13Mail)) != hipSuccess ||
        hipMalloc((void **)&c->d_hungry, 64 * sizeof(uint64_t)) != hipSuccess)
      return E13_ENOMEM;
    c->mail_cap = (size_t)max_waves;
  }
  const size_t wide_cap = tu->wide_nodes > 0 ? (size_t)tu->wide_nodes : ((size_t)1 << 19);
  if (tw < E13_TW) { /* the wide list, and the wave records a stop of the narrow launch leaves */
    if ((rc = e13i_ensure_buf((void **)&c->d_wide, &c->wide_cap, wide_cap, sizeof(E13Node)))) return rc;
    if ((rc = e13i_ensure_buf((void **)&c->d_wave_state, &c->wave_state_cap, 3 * (size_t)max_waves, sizeof(int64_t))))
      return rc;
  }
  if (!c->d_sch && hipMalloc((void **)&c->d_sch, sizeof(E13Sched)) != hipSuccess) return E13_ENOMEM;
  if (!c->d_pbs && hipMalloc((void **)&c->d_pbs, E13_MAXPROB * sizeof(E13Problem)) != hipSuccess) return E13_ENOMEM;
  if (!c->d_pcnt && hipMalloc((void **)&c->d_pcnt, E13_MAXPROB * E13_PCNT * sizeof(uint64_t)) != hipSuccess)
    return E13_ENOMEM;
  size_t want_surv = surv_cap > 4096 ? (size_t)surv_cap : 4096;
  if ((rc = e13i_ensure_buf((void **)&c->d_surv, &c->surv_cap, want_surv, sizeof(E13SurvRec)))) return rc;

  /* Breadth-first seeding: a few k_expand launches per problem grow a frontier
     of about four nodes per wave; leaves met on the way are counted here.  The
     persistent launch then starts with the frontier dealt to the waves. */
  const size_t seed_target = (size_t)tu->seed_per_wave * (size_t)waves_k[0];
  /* frontier order: 1 by the installed network, 3 by probes from each node (non-learned) */
  const int order_mode = tu->learned == 1 && c->have_model ? 1 : (tu->learned == 3 ? 3 : 0);
  c->seed_est_n = 0;
  if ((rc = e13i_ensure_buf((void **)&c->d_seeds, &c->seeds_cap, 64 * (size_t)max_waves + (size_t)np, sizeof(E13Node))))
    return rc;
  if ((rc = e13i_ensure_buf((void **)&c->d_stack, &c->stack_cap, c->seeds_cap, sizeof(E13Node)))) return rc;
  E13Problem pbs[E13_MAXPROB];
  for (int k = 0; k < np; k++) {
    e13i_fill_problem(&ps[k], &pbs[k]);
    memset(&out[k], 0, sizeof out[k]);
  }
  HIPCHK(hipMemcpyAsync(c->d_pbs, pbs, (size_t)np * sizeof(E13Problem), hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->d_ctr, 0, sizeof(E13Counters), c->stream));
  size_t nseeds = 0;
  uint64_t prev[3 + E13_NS + 1] = {0};
  /* checkpoints: a search with a checkpoint present resumes from its nodes and counts */
  const int ckpt = c->ckpt_path[0] != 0;
  int resumed = 0;
  if (ckpt) {
    if ((rc = e13i_ckpt3_load(c, ps, np, out, &nseeds)) < 0) return rc;
    resumed = rc == 1;
    e13i_ckpt_stop_clear(c); /* left over from the last clean stop */
  }
  for (int k = 0; k < np && !resumed; k++) {
    const int nslot = 4 * ps[k].nq + 1;
    const size_t target_k = (seed_target + (size_t)np - 1) / (size_t)np;
    E13Node root;
    e13i_root_node(&ps[k], &root);
    root.prob = k;
    HIPCHK(hipMemcpyAsync(c->d_stack, &root, sizeof root, hipMemcpyHostToDevice, c->stream));
    size_t top = 1;
    E13ExpandArgs ea;
    memset(&ea, 0, sizeof ea);
    ea.pb = c->d_pbs + k;
    ea.nslot = (uint32_t)nslot;
    ea.nq = ps[k].nq;
    ea.ctr = c->d_ctr;
    ea.surv = c->d_surv;
    ea.key = c->d_key;
    ea.tgt = c->d_tgt;
    ea.surv_cap = (uint32_t)c->surv_cap;
    ea.slice = tu->slice;
    ea.cap = tu->cap;
    ea.refill_tries = tu->refill_tries;
    ea.chunk = 64;
    for (int level = 0; level < 12 && top > 0 && top < target_k && seed_target > 0; level++) {
      if (order_mode && top > 1) { /* split by estimated cost: expand the costlier half, keep the rest as seeds */
        size_t ne = top;
        if ((rc = learned_split(c, ps, order_mode, c->d_stack, top, c->d_seeds + nseeds, &ne))) return rc;
        nseeds += top - ne;
        top = ne;
      }
      const size_t room = c->seeds_cap - nseeds;
      if (top * (size_t)nslot > c->stack_cap || top * (size_t)nslot > room) break;
      if ((rc = e13i_ensure_buf((void **)&c->d_out, &c->out_cap, top * (size_t)nslot, sizeof(E13Node)))) return rc;
      ea.in = c->d_stack;
      ea.nin = (uint32_t)top;
      ea.out = c->d_out;
      ea.out_cap = (uint32_t)(top * (size_t)nslot);
      HIPCHK(hipMemsetAsync(&c->d_ctr->item, 0, 2 * sizeof(uint32_t), c->stream));
      HIPCHK(hipEventRecord(c->ev0, c->stream));
      int lr = e13k_launch_expand(&ea, c->ncu * tu->wgs_per_cu, 256, tu->gate == 2, 1, c->stream);
      if (lr != 0) {
        fprintf(stderr, "libe13: expand launch failed: %s\n", hipGetErrorString((hipError_t)lr));
        return E13_EHIP;
      }
      HIPCHK(hipEventRecord(c->ev1, c->stream));
      uint32_t io[4];
      HIPCHK(hipMemcpyAsync(io, &c->d_ctr->item, sizeof io, hipMemcpyDeviceToHost, c->stream));
      HIPCHK(hipStreamSynchronize(c->stream));
      float ems = 0;
      if (hipEventElapsedTime(&ems, c->ev0, c->ev1) == hipSuccess) c->stats.kernel_ms += ems;
      c->stats.launches++;
      if (io[3]) return io[3] & 4u ? E13_ELDS : E13_ESTACK;
      if (io[1])
        HIPCHK(hipMemcpyAsync(c->d_stack, c->d_out, (size_t)io[1] * sizeof(E13Node), hipMemcpyDeviceToDevice,
                              c->stream));
      top = io[1];
    }
    if (top)
      HIPCHK(hipMemcpyAsync(c->d_seeds + nseeds, c->d_stack, top * sizeof(E13Node), hipMemcpyDeviceToDevice,
                            c->stream));
    nseeds += top;
    /* this problem's share of the counters so far */
    E13Counters h1;
    HIPCHK(hipMemcpyAsync(&h1, c->d_ctr, sizeof h1, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    out[k].leaves += h1.leaves - prev[0];
    out[k].steps += h1.steps - prev[1];
    out[k].survivors += h1.survivors - prev[2];
    for (int j = 0; j <= E13_NS; j++) out[k].dead_at_stage[j] += h1.hist[j] - prev[3 + j];
    prev[0] = h1.leaves;
    prev[1] = h1.steps;
    prev[2] = h1.survivors;
    for (int j = 0; j <= E13_NS; j++) prev[3 + j] = h1.hist[j];
  }
  if (!resumed && order_mode && nseeds > 1 && (rc = learned_order(c, ps, order_mode, c->d_seeds, nseeds))) return rc;
  if (!resumed && tu->learned == 2 && nseeds > 1 && (rc = random_order(c, c->d_seeds, nseeds))) return rc;
  const int logging = tu->cost_log == 1 && !ckpt; /* subtree costs need the frontier of one launch */
  if (logging) { /* number the frontier and count each subtree */
    if ((rc = e13i_ensure_buf((void **)&c->d_seed_cnt, &c->seed_cnt_cap, nseeds > 0 ? 2 * nseeds : 2,
                              sizeof(uint64_t))))
      return rc;
    HIPCHK(hipMemsetAsync(c->d_seed_cnt, 0, 2 * nseeds * sizeof(uint64_t), c->stream));
    if (e13k_launch_seed_ids(c->d_seeds, (uint32_t)nseeds, c->stream)) return E13_EHIP;
  }
  E13Sched sch;
  if (ckpt && (rc = e13i_ckpt3_prepare(c, max_waves))) return rc;
  int64_t call_launches = 0;
  double t_ck = host_ms(), last_stop_check = 0;
  E13Counters h;
  float ms = 0;
  if (!resumed) c->stats.frontier = (int64_t)nseeds;
  c->stats.tape_words = tw;
  const size_t n_logged = nseeds;
  /* One launch at a time: the narrow form (or the whole tape when tw is
     E13_TW), then, when machines went past the narrow tape, a launch with the
     whole tape from the list of them; with checkpoints, again from the nodes a
     clean stop left. */
  const E13Node *seeds_dev = c->d_seeds;
  int wide_next = tw == E13_TW;
  for (;;) {
  if (nseeds == 0) {
    /* the breadth-first seeding finished every tree (or a resumed checkpoint
       holds no nodes): the counts are complete, and a launch with no work
       would have no wave to end it */
    HIPCHK(hipMemcpyAsync(&h, c->d_ctr, sizeof h, hipMemcpyDeviceToHost, c->stream));
    HIPCHK(hipStreamSynchronize(c->stream));
    break;
  }
  const int narrow = !wide_next;
  const int kk = narrow ? 0 : 1;
  const int grid = grid_k[kk], total_waves = waves_k[kk], stk_cap = stk_k[kk];
  const int hungry_words = (total_waves + 63) / 64;
  if (nseeds > (size_t)total_waves * (size_t)stk_cap) return E13_ESTACK;
  if (call_launches > 0) { /* per-launch figures restart (the survivor count goes on) */
    HIPCHK(hipMemsetAsync(&c->d_ctr->steps, 0, sizeof(uint64_t), c->stream));
    HIPCHK(hipMemsetAsync(&c->d_ctr->lane_slots, 0, 2 * sizeof(uint64_t), c->stream));
    HIPCHK(hipMemsetAsync(c->d_ctr->cyc, 0, sizeof c->d_ctr->cyc, c->stream));
  }
  memset(&sch, 0, sizeof sch);
  sch.busy = nseeds < (size_t)total_waves ? (int32_t)nseeds : total_waves; /* waves dealt a seed */
  if (narrow && tw < E13_TW) {
    sch.wide = c->d_wide;
    sch.wide_cap = (uint32_t)wide_cap;
  }
  HIPCHK(hipMemcpyAsync(c->d_sch, &sch, sizeof sch, hipMemcpyHostToDevice, c->stream));
  HIPCHK(hipMemsetAsync(c->d_mail, 0, (size_t)total_waves * sizeof(E13Mail), c->stream));
  HIPCHK(hipMemsetAsync(c->d_hungry, 0, 64 * sizeof(uint64_t), c->stream));
  HIPCHK(hipMemsetAsync(c->d_pcnt, 0, (size_t)np * E13_PCNT * sizeof(uint64_t), c->stream));

  E13SearchArgs a;
  memset(&a, 0, sizeof a);
  a.stacks = c->d_stacks;
  a.mail = c->d_mail;
  a.hungry = c->d_hungry;
  a.hungry_words = hungry_words;
  a.sch = c->d_sch;
  a.pbs = c->d_pbs;
  a.pcnt = c->d_pcnt;
  a.ctr = c->d_ctr;
  a.surv = c->d_surv;
  a.key = c->d_key;
  a.tgt = c->d_tgt;
  a.np = np;
  a.slice = tu->slice;
  a.stk_cap = stk_cap;
  a.refill_tries = tu->refill_tries;
  a.donate_every = tu->donate_max;
  a.sched_every = (uint32_t)tu->sched_every;
  a.idle_naps = tu->idle_naps;
  a.mail_every = tu->mail_every;
  a.solo = tu->solo_steps;
  a.evict_after = tu->evict_after;
  a.cut_live = tu->cut_live;
  a.phase_cycles = tu->phase_cycles;
  if (tu->nt1_skip > 0) {
    a.skip_settled = tu->nt1_skip == 1;
  } else { /* automatic: delete-mode searches, whose edits break the start, have many such machines */
    for (int k = 0; k < np; k++) a.skip_settled |= ps[k].budget > 0;
  }
  a.pool_min_idle = 16;
  a.total_waves = total_waves;
  a.surv_cap = (uint32_t)c->surv_cap;
  a.spin_limit = 1u << 20; /* mailbox re-reads before a delivery counts as lost */
  a.idle_limit = 3000000000u; /* idle watchdog: 30 s of the 100 MHz constant clock */
  a.seeds = seeds_dev;
  a.nseeds = (uint32_t)nseeds;
  a.tw = narrow ? tw : E13_TW;
  if (narrow && tw < E13_TW) a.wave_state = c->d_wave_state; /* a full wide list stops the launch cleanly */
  if (logging) {
    a.seed_cnt = c->d_seed_cnt;
    a.nseed_cnt = (uint32_t)n_logged;
  }
  if (ckpt) {
    *c->h_stop = 0u;
    __atomic_thread_fence(__ATOMIC_SEQ_CST);
    a.stop = c->d_stop;
    a.wave_state = c->d_wave_state;
  }
  /* the trace is of the call's first launch (with a short lane tape the main
     one; the launches after it run its leftovers) */
  if (call_launches == 0) c->trace_waves = c->trace_per_wave = 0;
  if (tu->trace_events > 0 && call_launches == 0) {
    const size_t per = (size_t)tu->trace_events + 1, words = per * (size_t)total_waves;
    if ((rc = e13i_ensure_buf((void **)&c->d_trace, &c->trace_words, words, sizeof(uint64_t)))) return rc;
    HIPCHK(hipMemsetAsync(c->d_trace, 0, words * sizeof(uint64_t), c->stream));
    a.trace = c->d_trace;
    a.trace_cap = (int32_t)per;
    c->trace_waves = total_waves;
    c->trace_per_wave = (int)per;
  }
  if (timing) {
    HIPCHK(hipStreamSynchronize(c->stream));
    tm[1] = host_ms();
  }
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  int lr = e13k_launch_search(&a, grid, wg, gh, c->stream);
  if (lr != 0) {
    fprintf(stderr, "libe13: search launch failed: %s\n", hipGetErrorString((hipError_t)lr));
    return E13_EHIP;
  }
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  int stop_file = 0;
  if (ckpt) { /* wait, raising the stop flag for the stop file or the checkpoint interval */
    int raised = 0;
    for (;;) {
      const hipError_t q = hipStreamQuery(c->stream);
      if (q == hipSuccess) break;
      if (q != hipErrorNotReady) return E13_EHIP;
      const double now = host_ms();
      if (!raised && now - last_stop_check > 100.0) {
        last_stop_check = now;
        stop_file = e13i_ckpt_stop_requested(c);
      }
      if (!raised && (stop_file || now - t_ck > 1000.0 * c->ckpt_every_s)) {
        __atomic_store_n(c->h_stop, 1u, __ATOMIC_SEQ_CST);
        raised = 1;
      }
      usleep(1000);
    }
  }
  uint64_t pcnt[E13_MAXPROB * E13_PCNT];
  HIPCHK(hipMemcpyAsync(pcnt, c->d_pcnt, (size_t)np * E13_PCNT * sizeof(uint64_t), hipMemcpyDeviceToHost,
                        c->stream));
  HIPCHK(hipMemcpyAsync(&h, c->d_ctr, sizeof h, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipMemcpyAsync(&sch, c->d_sch, sizeof sch, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (timing) tm[2] = host_ms();
  HIPCHK(hipEventElapsedTime(&ms, c->ev0, c->ev1));
  c->stats.kernel_ms += ms;
  c->stats.launches++;
  c->stats.rounds += (int64_t)h.rounds;
  c->stats.lane_slots += (int64_t)h.lane_slots;
  for (int j = 0; j < 4; j++) c->stats.cycles[j] += (int64_t)h.cyc[j];
  c->stats.slice = tu->slice;
  c->stats.wg_threads = wg;
  if (narrow || tw == E13_TW) c->stats.grid = grid; /* the search's main launch, not its whole-tape rest */
  c->stats.gate = tu->gate;
  c->stats.form = 3;
  c->stats.donations += sch.donations;
  c->stats.pool_takes += sch.pool_takes;
  if (sch.error == 5 && narrow && tw < E13_TW) {
    /* not every workgroup became resident (every wave left before any work):
       this context runs the whole tape from now on, and the launch is redone */
    fprintf(stderr, "libe13: short-tape launch of %d workgroups not all resident: whole tape from now on\n", grid);
    c->narrow_off = 1;
    tw = E13_TW;
    c->stats.tape_words = tw;
    wide_next = 1;
    continue;
  }
  if (sch.error) {
    fprintf(stderr, "libe13: persistent search aborted (%s)\n",
            sch.error == 1   ? "a private node stack is full"
            : sch.error == 2 ? "a mailbox delivery never arrived complete"
            : sch.error == 3 ? "idle watchdog"
            : sch.error == 5 ? "not every workgroup resident at once"
                             : "dynamic LDS does not start at LDS address 0");
    return sch.error == 1   ? E13_ESTACK
           : sch.error == 2 ? E13_ELOST
           : sch.error == 3 ? E13_EWATCHDOG
           : sch.error == 5 ? E13_ERESIDENT
                            : E13_ELDS;
  }
  c->stats.search_ms += ms;
  c->stats.search_launches++;
  if (!narrow && tw < E13_TW) {
    c->stats.wide_launches++;
    c->stats.wide_ms += ms;
  }
  for (int k = 0; k < np; k++) {
    const uint64_t *q = pcnt + k * E13_PCNT;
    c->stats.search_steps += (int64_t)q[1];
    if (!narrow && tw < E13_TW) c->stats.wide_steps += (int64_t)q[1];
    out[k].leaves += q[0];
    out[k].steps += q[1];
    out[k].survivors += q[2];
    for (int j = 0; j <= E13_NS; j++) out[k].dead_at_stage[j] += q[3 + j];
  }
  call_launches++;
  /* machines a narrow launch left for the whole tape (a full list stopped it) */
  const size_t n_wide = narrow && tw < E13_TW ? (sch.wide_n < wide_cap ? (size_t)sch.wide_n : wide_cap) : 0;
  const int halted = narrow && tw < E13_TW && sch.halt != 0;
  c->stats.wide_nodes += (int64_t)n_wide;
  if (!ckpt && !halted) {
    if (n_wide == 0) break;
    seeds_dev = c->d_wide; /* the next launch runs them, on the whole tape */
    nseeds = n_wide;
    wide_next = 1;
    continue;
  }
  /* what a clean stop left: every wave's stack, the deliveries still in
     mailboxes, and the wide list */
  size_t n_left = 0;
  if ((rc = e13i_ckpt3_gather(c, total_waves, stk_cap, c->d_wide, n_wide, &n_left, &nseeds))) return rc;
  seeds_dev = c->d_seeds;
  /* the next launch has the whole tape when only wide nodes are left, or when
     the full list stopped this one (a narrow launch would stop again at once);
     else it is narrow, and its refill sends wide nodes to the list again */
  wide_next = tw == E13_TW || (narrow && (halted || (n_left == 0 && n_wide > 0)));
  if (nseeds == 0) break; /* the search ended before the flag was seen */
  if (!ckpt) continue;
  if ((rc = e13i_ckpt3_save(c, ps, np, out, nseeds))) return rc;
  t_ck = host_ms();
  if (stop_file) {
    e13i_ckpt_stop_clear(c);
    return E13_ESTOPPED;
  }
  if (c->ckpt_max_launches > 0 && call_launches >= c->ckpt_max_launches) return E13_ESTOPPED;
  } /* next launch, from the gathered nodes */
  if (logging && (rc = collect_cost_log(c, ps, n_logged))) return rc;
  uint64_t ns = h.nsurv;
  if (n_surv) *n_surv = ns;
  uint64_t keep = ns < c->surv_cap ? ns : c->surv_cap;
  if (surv && surv_cap && keep) {
    uint64_t nk = keep < surv_cap ? keep : surv_cap;
    E13SurvRec *rec = (E13SurvRec *)malloc(nk * sizeof(E13SurvRec));
    if (!rec) return E13_ENOMEM;
    HIPCHK(hipMemcpy(rec, c->d_surv, nk * sizeof(E13SurvRec), hipMemcpyDeviceToHost));
    finish_all(ps, np, rec, surv, nk);
    free(rec);
  }
  if (timing) {
    tm[3] = host_ms();
    fprintf(stderr, "libe13: search phases: frontier %.2f ms, persistent kernel %.2f ms (events %.2f), results %.2f ms\n",
            tm[1] - tm[0], tm[2] - tm[1], ms, tm[3] - tm[2]);
  }
  return E13_OK;
}

int e13_get_trace(e13_ctx *c, uint64_t *buf, uint64_t cap_words, int32_t *waves, int32_t *per_wave) {
  if (!c) return E13_EINVAL;
  if (waves) *waves = c->trace_waves;
  if (per_wave) *per_wave = c->trace_per_wave;
  const uint64_t words = (uint64_t)c->trace_waves * (uint64_t)c->trace_per_wave;
  if (!buf || !words) return E13_OK;
  if (cap_words < words) return E13_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  HIPCHK(hipMemcpy(buf, c->d_trace, words * sizeof(uint64_t), hipMemcpyDeviceToHost));
  return E13_OK;
}

/* ------------------------------------------------------------------------ */
/* Complete tables.                                                          */
/* ------------------------------------------------------------------------ */
static int batch_launch(e13_ctx *c, const void *tabs_dev, int nq, int n, const void *caps_dev, void *out_dev) {
  E13BatchArgs a;
  memset(&a, 0, sizeof a);
  a.tabs = (const uint8_t *)tabs_dev;
  a.caps = (const int32_t *)caps_dev;
  a.out = (E13LeafRec *)out_dev;
  a.ctr = c->d_ctr;
  a.key = c->d_key;
  a.tgt = c->d_tgt;
  a.n = n;
  a.nq = nq;
  a.slice = c->tune.slice;
  a.dl = c->d_dl;
  a.solo = c->tune.solo_steps;
  HIPCHK(hipMemsetAsync(c->d_ctr, 0, sizeof(E13Counters), c->stream));
  const int wg = fit_wg(c, c->tune.wg_threads, c->tune.gate == 2);
  if (e13k_lds_bytes(wg, c->tune.gate == 2) > c->lds_max) return E13_EINVAL;
  /* form 1: one table per lane, a grid covering the batch; otherwise a
     persistent grid whose lanes take the next table as soon as theirs ends */
  const int refill = c->tune.form != 1;
  long need = ((long)n + wg - 1) / wg;
  int grid = refill ? c->ncu * c->tune.wgs_per_cu : (int)need;
  if (need < grid) grid = (int)(need > 0 ? need : 1);
  /* claims of 128 tables per atomic once the batch is many times the lanes;
     smaller batches claim one table per idle lane, so every wave gets work */
  a.chunk = (long)n >= 16L * grid * wg ? 128u : 0u;
  HIPCHK(hipEventRecord(c->ev0, c->stream));
  int lr = e13k_launch_batch(&a, grid, wg, c->tune.gate == 2, refill, c->stream);
  if (lr != 0) {
    fprintf(stderr, "libe13: batch launch failed: %s\n", hipGetErrorString((hipError_t)lr));
    return E13_EHIP;
  }
  HIPCHK(hipEventRecord(c->ev1, c->stream));
  c->stats.launches++;
  c->stats.grid = grid;
  c->stats.wg_threads = wg;
  c->stats.slice = c->tune.slice;
  c->stats.gate = c->tune.gate;
  return E13_OK;
}

int e13_simulate_batch_device(e13_ctx *c, const void *tabs_dev, int nq, int n, const void *caps_dev, void *out_dev) {
  if (!c || !tabs_dev || !out_dev || n < 1 || nq < 1 || nq > E13_MAXNQ) return E13_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  return batch_launch(c, tabs_dev, nq, n, caps_dev, out_dev);
}

int e13_simulate_batch(e13_ctx *c, const int8_t *tabs, int ncells, int n, const int32_t *caps, e13_leaf *out) {
  if (!c || !tabs || !out || n < 1 || ncells < 2 || ncells % 2 || ncells > 2 * E13_MAXNQ) return E13_EINVAL;
  const int nq = ncells / 2;
  HIPCHK(hipSetDevice(c->device));
  uint8_t *h = (uint8_t *)malloc((size_t)n * E13_MAXCELLS);
  if (!h) return E13_ENOMEM;
  for (int k = 0; k < n; k++) {
    for (int i = 0; i < E13_MAXCELLS; i++) {
      int v = i < ncells ? tabs[(size_t)k * ncells + i] : E13_HALT;
      if (v < -1 || v >= 4 * nq) {
        free(h);
        return E13_EINVAL;
      }
      h[(size_t)k * E13_MAXCELLS + i] = v < 0 ? E13_C_HALT : (uint8_t)v;
    }
  }
  int rc;
  size_t cap_before = c->batch_cap;
  if ((rc = e13i_ensure_buf(&c->d_tabs, &c->batch_cap, (size_t)n, E13_MAXCELLS))) {
    free(h);
    return rc;
  }
  if (c->batch_cap != cap_before || !c->d_caps || !c->d_leaves) {
    hipFree(c->d_caps);
    hipFree(c->d_leaves);
    c->d_caps = c->d_leaves = NULL;
    if (hipMalloc(&c->d_caps, c->batch_cap * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&c->d_leaves, c->batch_cap * sizeof(E13LeafRec)) != hipSuccess) {
      free(h);
      return E13_ENOMEM;
    }
  }
  HIPCHK(hipMemcpyAsync(c->d_tabs, h, (size_t)n * E13_MAXCELLS, hipMemcpyHostToDevice, c->stream));
  if (caps) HIPCHK(hipMemcpyAsync(c->d_caps, caps, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream));
  rc = batch_launch(c, c->d_tabs, nq, n, caps ? c->d_caps : NULL, c->d_leaves);
  if (rc) {
    free(h);
    return rc;
  }
  E13LeafRec *lr = (E13LeafRec *)malloc((size_t)n * sizeof(E13LeafRec));
  if (!lr) {
    free(h);
    return E13_ENOMEM;
  }
  HIPCHK(hipMemcpyAsync(lr, c->d_leaves, (size_t)n * sizeof(E13LeafRec), hipMemcpyDeviceToHost, c->stream));
  E13Counters hc;
  HIPCHK(hipMemcpyAsync(&hc, c->d_ctr, sizeof hc, hipMemcpyDeviceToHost, c->stream));
  HIPCHK(hipStreamSynchronize(c->stream));
  if (hc.overflow & 4u) {
    free(lr);
    free(h);
    fprintf(stderr, "libe13: dynamic LDS does not start at LDS address 0\n");
    return E13_ELDS;
  }
  float ms = 0;
  if (hipEventElapsedTime(&ms, c->ev0, c->ev1) == hipSuccess) c->stats.kernel_ms += ms;
  for (int k = 0; k < n; k++) {
    out[k].result = lr[k].result;
    out[k].stage = lr[k].stage;
    out[k].steps = lr[k].steps;
    out[k].st0 = lr[k].st0;
    out[k].head0 = lr[k].head0;
    out[k].abs_ok = lr[k].abs_ok;
    out[k].rel_ok = lr[k].rel_ok;
    out[k].event = lr[k].pad;
    for (int j = 0; j < 3; j++) out[k].ch[j] = lr[k].ch[j];
  }
  free(lr);
  free(h);
  return E13_OK;
}

/* ------------------------------------------------------------------------ */
/* Plain re-check (after recheck.c:57-92, stages 1..9, search deadlines).     */
/* ------------------------------------------------------------------------ */
int e13_recheck(const int8_t *tab, int nq, int st0, int head0, int abs_ok, int rel_ok, int64_t steps,
                uint32_t *read_mask) {
  enum { L = 64, R = 1300, T = L + R };
  uint8_t tape[T];
  uint8_t ctr[E13_NS + 1][L];
  memset(tape, 0, sizeof tape);
  const int rel = !abs_ok && rel_ok;
  int head = 0, st = 0, ones = 0;
  int64_t t = 0;
  uint32_t mask = 0;
  if (read_mask) *read_mask = 0;
  if (!tab || nq < 1 || nq > 16) return 0;
  for (int n = 1; n <= E13_NS; n++) {
    const int blk = 1 << n, want = rel ? head0 + blk - 2 : head0;
    const int64_t limit = 4 * (int64_t)T14S[n] + 64;
    for (;;) {
      if (st == st0 && head == want && ones == blk) {
        int all = 1;
        for (int k = -1; k <= blk - 2; k++)
          if (!tape[k + L]) {
            all = 0;
            break;
          }
        if (all) break;
      }
      if (t > limit) return 0;
      const int sym = tape[head + L], cell = 2 * st + sym;
      mask |= 1u << cell;
      const int code = tab[cell];
      if (code < 0) return 0; /* halt, or a never-chosen cell */
      const int w = code & 1;
      if (w != sym) {
        tape[head + L] = (uint8_t)w;
        if (head >= -2) ones += w ? 1 : -1;
      }
      head += (code & 2) ? 1 : -1;
      if (head <= -L || head >= R - 1) return 0;
      st = code >> 2;
      if (st >= nq) return 0;
      t++;
    }
    for (int k = 0; k < L; k++) ctr[n][k] = k < L - 2 ? tape[k] : 0; /* cells -64..-3 */
    const int cls = n <= 2 ? 0 : (n <= 8 ? 1 : 2), first = cls == 0 ? 1 : (cls == 1 ? 3 : 9);
    if (n != first && memcmp(ctr[n], ctr[first], L)) return 0;
  }
  if (!memcmp(ctr[1], ctr[3], L) || !memcmp(ctr[3], ctr[9], L) || !memcmp(ctr[1], ctr[9], L)) return 0;
  if (t != steps) return 0;
  if (read_mask) *read_mask = mask;
  return 1;
}

/* ------------------------------------------------------------------------ */
/* Output lines.                                                             */
/* ------------------------------------------------------------------------ */
int e13_format_survivor(const e13_problem *p, const e13_survivor *s, char *buf, int buflen) {
  if (!p || !s || !buf || buflen < 256) return E13_EINVAL;
  int n = snprintf(buf, (size_t)buflen, "SURVIVOR X=%d st0=%d head0=%d abs=%d rel=%d steps=%lld table=", p->X, s->st0,
                   s->head0, s->abs_ok, s->rel_ok, (long long)s->steps);
  for (int q = 0; q < p->nq && n < buflen - 16; q++) {
    for (int a = 0; a < 2; a++) {
      int i = 2 * q + a, code = s->tab[i];
      if (!((s->read_mask >> i) & 1)) buf[n++] = '?';
      if (code < 0)
        n += snprintf(buf + n, (size_t)(buflen - n), "---");
      else
        n += snprintf(buf + n, (size_t)(buflen - n), "%d%c%d", code & 1, (code & 2) ? 'R' : 'L', code >> 2);
      if (a == 0) buf[n++] = ' ';
    }
    if (q + 1 < p->nq) buf[n++] = '|';
  }
  buf[n] = 0;
  return n;
}

int e13_format_hash(const e13_problem *p, const e13_survivor *s, char *buf, int buflen) {
  if (!p || !s || !buf || buflen < 96) return E13_EINVAL;
  return snprintf(buf, (size_t)buflen, "HASH X=%d ch0=%016llx ch1=%016llx ch2=%016llx", p->X,
                  (unsigned long long)s->ch[0], (unsigned long long)s->ch[1], (unsigned long long)s->ch[2]);
}

void e13_synth_table(uint64_t seed0, uint64_t k, int8_t *tab28, int32_t *cap) {
  uint64_t st = seed0 + k;
  for (int i = 0; i < 28; i++) tab28[i] = (int8_t)(e13i_splitmix(&st) % 56);
  tab28[e13i_splitmix(&st) % 28] = E13_HALT;
  double u = (double)(e13i_splitmix(&st) >> 11) * (1.0 / 9007199254740992.0);
  *cap = (int32_t)floor(10.0 * pow(2e4, u));
}

/* ------------------------------------------------------------------------ */
/* Synthetic stress batch.                                                   */
/* ------------------------------------------------------------------------ */
static uint64_t rec_mix(uint64_t x) { return e13i_splitmix(&x); }

static uint64_t record_hash(uint64_t k, const e13_leaf *l) {
  uint64_t h = rec_mix(k);
  h = rec_mix(h ^ (uint64_t)(uint32_t)l->result);
  h = rec_mix(h ^ (uint64_t)(uint32_t)l->stage);
  h = rec_mix(h ^ (uint64_t)(uint32_t)l->steps);
  h = rec_mix(h ^ (uint64_t)(uint32_t)l->st0);
  h = rec_mix(h ^ (uint64_t)(uint32_t)l->head0);
  h = rec_mix(h ^ (uint64_t)(uint32_t)(l->abs_ok | (l->rel_ok << 1)));
  h = rec_mix(h ^ l->ch[0]);
  h = rec_mix(h ^ l->ch[1]);
  h = rec_mix(h ^ l->ch[2]);
  return h;
}

int e13_batch_digest(const e13_leaf *lv, int64_t n, uint64_t k0, e13_batch_sum *out) {
  if (!lv || !out || n < 0) return E13_EINVAL;
  memset(out, 0, sizeof *out);
  out->n = (uint64_t)n;
  for (int64_t j = 0; j < n; j++) {
    out->digest += record_hash(k0 + (uint64_t)j, &lv[j]);
    out->steps += (uint64_t)lv[j].steps;
    if (lv[j].result == E13_SURV) out->survivors++;
    else if (lv[j].stage >= 0 && lv[j].stage <= E13_NS) out->dead_at_stage[lv[j].stage]++;
  }
  return E13_OK;
}

int e13_synth_run(e13_ctx *c, uint64_t seed0, uint64_t k0, int64_t n, e13_batch_sum *out) {
  if (!c || !out || n < 1 || n > (1LL << 30)) return E13_EINVAL;
  HIPCHK(hipSetDevice(c->device));
  uint8_t *h = (uint8_t *)malloc((size_t)n * E13_MAXCELLS);
  int32_t *caps = (int32_t *)malloc((size_t)n * sizeof(int32_t));
  E13LeafRec *lr = (E13LeafRec *)malloc((size_t)n * sizeof(E13LeafRec));
  e13_leaf *lv = (e13_leaf *)malloc((size_t)n * sizeof(e13_leaf));
  int rc = E13_OK;
  if (!h || !caps || !lr || !lv) {
    rc = E13_ENOMEM;
    goto done;
  }
  for (int64_t j = 0; j < n; j++) {
    int8_t t[28];
    e13_synth_table(seed0, k0 + (uint64_t)j, t, &caps[j]);
    uint8_t *row = h + (size_t)j * E13_MAXCELLS;
    for (int i = 0; i < E13_MAXCELLS; i++) row[i] = i < 28 && t[i] >= 0 ? (uint8_t)t[i] : E13_C_HALT;
  }
  size_t cap_before = c->batch_cap;
  if ((rc = e13i_ensure_buf(&c->d_tabs, &c->batch_cap, (size_t)n, E13_MAXCELLS))) goto done;
  if (c->batch_cap != cap_before || !c->d_caps || !c->d_leaves) {
    hipFree(c->d_caps);
    hipFree(c->d_leaves);
    c->d_caps = c->d_leaves = NULL;
    if (hipMalloc(&c->d_caps, c->batch_cap * sizeof(int32_t)) != hipSuccess ||
        hipMalloc(&c->d_leaves, c->batch_cap * sizeof(E13LeafRec)) != hipSuccess) {
      rc = E13_ENOMEM;
      goto done;
    }
  }
  if (hipMemcpyAsync(c->d_tabs, h, (size_t)n * E13_MAXCELLS, hipMemcpyHostToDevice, c->stream) != hipSuccess ||
      hipMemcpyAsync(c->d_caps, caps, (size_t)n * sizeof(int32_t), hipMemcpyHostToDevice, c->stream) != hipSuccess) {
    rc = E13_EHIP;
    goto done;
  }
  if ((rc = batch_launch(c, c->d_tabs, 14, (int)n, c->d_caps, c->d_leaves))) goto done;
  E13Counters hc;
  float ms = 0;
  if (hipMemcpyAsync(lr, c->d_leaves, (size_t)n * sizeof(E13LeafRec), hipMemcpyDeviceToHost, c->stream) !=
          hipSuccess ||
      hipMemcpyAsync(&hc, c->d_ctr, sizeof hc, hipMemcpyDeviceToHost, c->stream) != hipSuccess ||
      hipStreamSynchronize(c->stream) != hipSuccess || hipEventElapsedTime(&ms, c->ev0, c->ev1) != hipSuccess) {
    rc = E13_EHIP;
    goto done;
  }
  if (hc.overflow & 4u) {
    fprintf(stderr, "libe13: dynamic LDS does not start at LDS address 0\n");
    rc = E13_ELDS;
    goto done;
  }
  c->stats.kernel_ms += ms;
  c->stats.rounds += (int64_t)hc.rounds;
  c->stats.lane_slots += (int64_t)hc.lane_slots;
  c->stats.form = c->tune.form == 1 ? 1 : 2;
  for (int64_t j = 0; j < n; j++) {
    lv[j].result = lr[j].result;
    lv[j].stage = lr[j].stage;
    lv[j].steps = lr[j].steps;
    lv[j].st0 = lr[j].st0;
    lv[j].head0 = lr[j].head0;
    lv[j].abs_ok = lr[j].abs_ok;
    lv[j].rel_ok = lr[j].rel_ok;
    lv[j].event = lr[j].pad;
    memcpy(lv[j].ch, lr[j].ch, sizeof lv[j].ch);
  }
  rc = e13_batch_digest(lv, n, k0, out);
  out->kernel_ms = ms;
done:
  free(h);
  free(caps);
  free(lr);
  free(lv);
  return rc;
}

