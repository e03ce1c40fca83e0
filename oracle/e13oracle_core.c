/* e13oracle_core.c -- CPU oracle (test infrastructure only; see e13oracle.h).
 *
 * A restatement of the reference CPU search src/e13search.c: same constants,
 * same Zobrist keys, same run()/explore()/setup()/sym_next()/run_delete()
 * semantics.  Written as a separate program: the simulation state is kept in
 * an eo_sim, the stage logic is its own function, and the recursion carries
 * an explicit frame.  File:line citations are to the reference tree.
 */
#include "e13oracle.h"

#include <math.h>
#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

uint64_t eo_key[EO_TAPE];
uint64_t eo_tgt[EO_NS + 1];
const long eo_t14s[EO_NS + 1] = {0, 2, 22, 90, 268, 944, 3434, 13450, 52068, 208864}; /* e13search.c:70 */
const int eo_cval[EO_NS + 1] = {0, 0, 0, 1, 1, 1, 1, 1, 1, 2};                       /* e13search.c:71 */

/* splitmix64 (Steele, Lea & Flood 2014), constants as at e13search.c:98-103. */
uint64_t eo_splitmix(uint64_t *state) {
  *state += 0x9E3779B97F4A7C15ULL;
  uint64_t z = *state;
  z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ULL;
  z = (z ^ (z >> 27)) * 0x94D049BB133111EBULL;
  return z ^ (z >> 31);
}

/* KEY[i] = i-th splitmix64 output from seed 20261028; TGT[n] = XOR of the keys
   of cells -1 .. 2^n-2 (e13search.c:437-444). */
void eo_init_keys(void) {
  uint64_t st = 20261028ULL;
  for (int i = 0; i < EO_TAPE; i++) eo_key[i] = eo_splitmix(&st);
  eo_tgt[0] = 0;
  for (int n = 1; n <= EO_NS; n++) {
    uint64_t h = 0;
    for (int cell = -1; cell <= (1 << n) - 2; cell++) h ^= eo_key[cell + EO_LO];
    eo_tgt[n] = h;
  }
}

static long long stage_deadline(const int32_t *dl, int n) {
  return dl ? (long long)dl[n] : 4LL * eo_t14s[n] + 64; /* BUDGET_MULT*T14S[n]+64, e13search.c:128,159 */
}

/* e13search.c:123-129 */
void eo_sim_init(eo_sim *s, const int32_t *dl) {
  memset(s, 0, sizeof *s);
  s->st0 = -1;
  s->abs_ok = 1;
  s->rel_ok = 1;
  s->nt = 1;
  s->deadline = stage_deadline(dl, 1);
}

/* The stage-start rule applied when the block hash equals TGT[nt]
   (e13search.c:136-161).  Returns 0 to go on, EO_DEAD or EO_SURV. */
static int stage_match(eo_sim *s, const int32_t *dl) {
  int n = s->nt, c = eo_cval[n];
  int a_ok = s->abs_ok, r_ok = s->rel_ok, ok = 1;
  if (n > 1) {
    if (s->st != s->st0) ok = 0;
    if (s->head != s->head0) a_ok = 0;
    if (s->head - (1 << n) != s->head0 - 2) r_ok = 0;
    if (!a_ok && !r_ok) ok = 0;
  }
  if (!ok) return 0; /* not a stage start: ignored */
  if (s->cs[c] && s->ch[c] != s->hc) return EO_DEAD;
  for (int j = 0; j < 3; j++)
    if (j != c && s->cs[j] && s->ch[j] == s->hc) return EO_DEAD;
  s->cs[c] = 1;
  s->ch[c] = s->hc;
  if (n == 1) {
    s->st0 = s->st;
    s->head0 = s->head;
  }
  s->abs_ok = a_ok;
  s->rel_ok = r_ok;
  s->stage = n;
  if (n == EO_NS) return EO_SURV;
  s->nt = n + 1;
  s->deadline = stage_deadline(dl, n + 1);
  return 0;
}

/* run(): e13search.c:131-177.  Stops before reading a cell with fixed[i]==0
   (returns i), or returns EO_DEAD / EO_SURV. */
int eo_run(eo_sim *s, const int *tab, const uint8_t *fixed, const int32_t *dl) {
  for (;;) {
    if (s->hb == eo_tgt[s->nt]) {
      int r = stage_match(s, dl);
      if (r) return r;
    }
    if (s->t > s->deadline) return EO_DEAD;
    int pos = s->head + EO_LO;
    int sym = s->tape[pos];
    int cell = 2 * s->st + sym;
    if (!fixed[cell]) return cell;
    int code = tab[cell];
    if (code < 0) return EO_DEAD; /* halt */
    int w = code & 1;
    if (w != sym) {
      s->tape[pos] = (uint8_t)w;
      if (s->head >= -2)
        s->hb ^= eo_key[pos];
      else
        s->hc ^= eo_key[pos];
    }
    s->head += (code & 2) ? 1 : -1;
    if (s->head <= -EO_LO || s->head >= EO_HI - 1) return EO_DEAD; /* the step is not counted */
    s->st = code >> 2;
    s->t++;
  }
}

/* e13search.c:105-108 */
int eo_parse_cell(const char *s) {
  if (s[0] == '-') return EO_HALT;
  return (atoi(s + 2) << 2) | ((s[1] == 'R') << 1) | (s[0] - '0');
}

/* e13search.c:110-121 */
int eo_load(const char *path, int *base) {
  FILE *f = fopen(path, "r");
  if (!f) return -1;
  char a[16], b[16];
  int nq = 0;
  while (nq < EO_MAXQ && fscanf(f, "%15s %15s", a, b) == 2) {
    base[2 * nq] = eo_parse_cell(a);
    base[2 * nq + 1] = eo_parse_cell(b);
    nq++;
  }
  fclose(f);
  return nq;
}

/* set_sym(): e13search.c:423-426 */
void eo_set_sym(eo_problem *p, int lo, int hi) {
  p->sym_lo = lo - (lo > p->X);
  p->sym_hi = hi - (hi >= p->X);
}

/* setup(): e13search.c:317-355 */
int eo_setup(eo_problem *p) {
  int nforced = 0, row = 0;
  for (int q = 0; q < p->nq_base; q++) {
    if (q == p->X) continue;
    for (int a = 0; a < 2; a++) {
      int code = p->base[2 * q + a];
      if (p->freemode && p->freebase[2 * q + a]) {
        code = EO_FREE;
      } else if (code >= 0 && (code >> 2) == p->X) {
        code = EO_FREE;
        nforced++;
      } else if (code >= 0) {
        int nx = code >> 2;
        if (p->X >= 0 && nx > p->X) nx--;
        code = (nx << 2) | (code & 3);
      }
      p->tab[2 * row + a] = code;
    }
    row++;
  }
  p->nq = row;
  p->sym_n0 = 0;
  if (p->sym_lo >= 0) {
    uint8_t named[EO_MAXQ] = {0};
    /* e13search.c:341 scans 2*(NQ-1) cells, NQ the base states: with no state
       deleted (X = -1) the last row is left out */
    for (int i = 0; i < 2 * (p->nq_base - 1); i++) {
      int c = p->tab[i];
      if (c >= 0 && (c >> 2) >= p->sym_lo && (c >> 2) <= p->sym_hi) named[(c >> 2) - p->sym_lo] = 1;
    }
    int n0 = 0;
    while (n0 <= p->sym_hi - p->sym_lo && named[n0]) n0++;
    for (int q = p->sym_lo + n0; q <= p->sym_hi; q++)
      if (named[q - p->sym_lo] || p->tab[2 * q] != EO_FREE || p->tab[2 * q + 1] != EO_FREE) return -2;
    if (p->sym_lo == 0) return -3;
    p->sym_n0 = n0;
  }
  int fmax = p->freemode ? nforced : p->fmax;
  if (nforced > fmax) {
    p->budget = -1;
    return -1;
  }
  p->budget = fmax - nforced;
  return 0;
}

/* --------------------------------------------------------------------------
   The tree walk (explore, e13search.c:214-254) and run_delete
   (e13search.c:357-420).
   -------------------------------------------------------------------------- */

typedef struct {
  long long t;
  int tab[2 * EO_MAXQ];
  uint8_t fixed[2 * EO_MAXQ];
  int budget, nintro;
} eo_node;

typedef struct {
  const eo_problem *p;
  int split, collecting;
  eo_counts c;
  eo_node *front;
  long nfront, capfront;
  eo_surv_fn fn;
  void *user;
  eo_sim frames[2 * EO_MAXQ + 2];
} eo_walk;

/* sym_next(): e13search.c:206-212 */
static int sym_next(const eo_problem *p, int code, int nintro) {
  if (p->sym_lo < 0 || code < 0) return nintro;
  int q = code >> 2;
  if (q < p->sym_lo || q > p->sym_hi) return nintro;
  int k = q - p->sym_lo;
  if (k < nintro) return nintro;
  return k == nintro ? nintro + 1 : -1;
}

static void walk(eo_walk *w, int depth, int splits, int *tab, uint8_t *fixed, int budget, int nintro) {
  eo_sim *s = &w->frames[depth];
  long long t0 = s->t;
  int r = eo_run(s, tab, fixed, NULL);
  w->c.steps += s->t - t0;
  if (r == EO_DEAD) {
    w->c.hist[s->stage]++;
    w->c.leaves++;
    return;
  }
  if (r == EO_SURV) {
    w->c.survivors++;
    w->c.leaves++;
    if (w->fn) {
#pragma omp critical(eo_report)
      w->fn(w->user, s, tab, fixed, w->p->X, w->p->nq);
    }
    return;
  }
  int cell = r, orig = tab[cell];
  int branching = orig == EO_FREE || budget > 0;
  if (w->collecting && splits == w->split) {
    if (w->nfront == w->capfront) {
      w->capfront = w->capfront ? 2 * w->capfront : 1024;
      w->front = (eo_node *)realloc(w->front, sizeof(eo_node) * (size_t)w->capfront);
      if (!w->front) {
        fprintf(stderr, "oracle: out of memory at front=%ld\n", w->nfront);
        exit(1);
      }
    }
    eo_node *nd = &w->front[w->nfront++];
    nd->t = s->t;
    memcpy(nd->tab, tab, sizeof nd->tab);
    memcpy(nd->fixed, fixed, sizeof nd->fixed);
    nd->budget = budget;
    nd->nintro = nintro;
    return;
  }
  eo_sim *child = &w->frames[depth + 1];
  fixed[cell] = 1;
  if (orig != EO_FREE) { /* keep the base value: same budget */
    *child = *s;
    walk(w, depth + 1, splits + branching, tab, fixed, budget, nintro);
  }
  if (branching) {
    int nb = orig == EO_FREE ? budget : budget - 1;
    for (int v = 0; v < 4 * w->p->nq; v++) {
      int ni = sym_next(w->p, v, nintro);
      if (v == orig || ni < 0) continue;
      tab[cell] = v;
      *child = *s;
      walk(w, depth + 1, splits + 1, tab, fixed, nb, ni);
    }
  }
  tab[cell] = orig;
  fixed[cell] = 0;
}

int eo_search(eo_problem *p, int split, int threads, eo_counts *out, eo_surv_fn fn, void *user) {
  memset(out, 0, sizeof *out);
  if (p->budget < 0) return -1;
  eo_walk *top = (eo_walk *)calloc(1, sizeof(eo_walk));
  int tab[2 * EO_MAXQ];
  uint8_t fixed[2 * EO_MAXQ] = {0};
  memcpy(tab, p->tab, sizeof tab);
  top->p = p;
  top->split = split;
  top->collecting = 1;
  top->fn = fn;
  top->user = user;
  eo_sim_init(&top->frames[0], NULL);
  walk(top, 0, 0, tab, fixed, p->budget, p->sym_n0);
  eo_counts tot = top->c;
  long nfront = top->nfront;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    eo_walk *w = (eo_walk *)calloc(1, sizeof(eo_walk));
    w->p = p;
    w->split = split;
    w->fn = fn;
    w->user = user;
    long k;
#pragma omp for schedule(dynamic, 1)
    for (k = 0; k < nfront; k++) {
      eo_node *nd = &top->front[k];
      int t2[2 * EO_MAXQ];
      uint8_t f2[2 * EO_MAXQ];
      memcpy(t2, nd->tab, sizeof t2);
      memcpy(f2, nd->fixed, sizeof f2);
      eo_sim_init(&w->frames[0], NULL);
      walk(w, 0, split, t2, f2, nd->budget, nd->nintro);
      w->c.steps -= nd->t; /* the replay up to the node was counted when it was collected */
    }
#pragma omp critical(eo_merge)
    {
      for (int j = 0; j <= EO_NS; j++) tot.hist[j] += w->c.hist[j];
      tot.leaves += w->c.leaves;
      tot.survivors += w->c.survivors;
      tot.steps += w->c.steps;
    }
    free(w);
  }
  tot.nfront = nfront;
  *out = tot;
  free(top->front);
  free(top);
  return 0;
}

/* run_full(): e13search.c:428-433 */
int eo_run_full(const int *tab, eo_sim *s, const int32_t *dl) {
  uint8_t fixed[2 * EO_MAXQ];
  memset(fixed, 1, sizeof fixed);
  eo_sim_init(s, dl);
  return eo_run(s, tab, fixed, dl);
}

/* --------------------------------------------------------------------------
   Plain re-check of a survivor (after recheck.c:57-92, stages 1..9 only, with
   the search's deadlines and window).  No hashing: the block and the counter
   are compared as exact cell sets.
   -------------------------------------------------------------------------- */
int eo_recheck(const int *tab, int nq, int st0, int head0, int abs_ok, int rel_ok, long long steps,
               uint32_t *read_mask) {
  uint8_t tape[EO_TAPE];
  memset(tape, 0, sizeof tape);
  uint8_t counters[EO_NS + 1][EO_LO];
  int rel = !abs_ok && rel_ok;
  int head = 0, st = 0, ones = 0;
  long long t = 0;
  uint32_t mask = 0;
  for (int n = 1; n <= EO_NS; n++) {
    int blk = 1 << n;
    int want = rel ? head0 + blk - 2 : head0;
    long long limit = 4LL * eo_t14s[n] + 64;
    for (;;) {
      if (st == st0 && head == want && ones == blk) {
        int exact = 1;
        for (int c = -1; c <= blk - 2; c++)
          if (!tape[c + EO_LO]) { exact = 0; break; }
        if (exact) break;
      }
      if (t > limit) return 0;
      int sym = tape[head + EO_LO], cell = 2 * st + sym;
      if (cell >= 2 * nq) return 0;
      mask |= 1u << cell;
      int code = tab[cell];
      if (code < 0) return 0; /* halt or never-chosen cell read */
      int w = code & 1;
      if (w != sym) {
        tape[head + EO_LO] = (uint8_t)w;
        if (head >= -2) ones += w ? 1 : -1;
      }
      head += (code & 2) ? 1 : -1;
      if (head <= -EO_LO || head >= EO_HI - 1) return 0;
      st = code >> 2;
      t++;
    }
    for (int c = 0; c < EO_LO; c++) counters[n][c] = c < EO_LO - 2 ? tape[c] : 0; /* cells -64..-3 */
    int cls = eo_cval[n], first = cls == 0 ? 1 : (cls == 1 ? 3 : 9);
    if (n != first && memcmp(counters[n], counters[first], EO_LO)) return 0;
  }
  if (!memcmp(counters[1], counters[3], EO_LO) || !memcmp(counters[3], counters[9], EO_LO) ||
      !memcmp(counters[1], counters[9], EO_LO))
    return 0;
  if (t != steps) return 0;
  if (read_mask) *read_mask = mask;
  return 1;
}

/* --------------------------------------------------------------------------
   Synthetic stress tables: table k draws from splitmix64 started at
   seed0 + k.  28 cells each get v = next() % 56 (write v&1, direction
   (v>>1)&1, next state v>>2 in 0..13); then the cell next() % 28 becomes the
   halt; then cap = floor(10 * 2e4^u) with u = (next() >> 11) * 2^-53, so the
   cap is log-uniform in [10, 2e5].  The cap replaces every stage deadline.
   -------------------------------------------------------------------------- */
void eo_synth_table(uint64_t seed0, uint64_t k, int *tab28, int32_t *cap) {
  uint64_t st = seed0 + k;
  for (int i = 0; i < 28; i++) tab28[i] = (int)(eo_splitmix(&st) % 56);
  tab28[eo_splitmix(&st) % 28] = EO_HALT;
  double u = (double)(eo_splitmix(&st) >> 11) * (1.0 / 9007199254740992.0);
  *cap = (int32_t)floor(10.0 * pow(2e4, u));
}

/* --------------------------------------------------------------------------
   Order-independent digest of a synthetic batch (tables k0 .. k0+n-1).  Per
   table: h = mix(k); then h = mix(h ^ x) for x = result, stage, t, st0, head0,
   abs_ok | rel_ok << 1 (each as a 32-bit two's-complement value), ch0, ch1,
   ch2; the digest is the sum of h mod 2^64, with mix(x) the splitmix64
   output for state x.  Totals: steps, survivors, dead-at-stage histogram.
   -------------------------------------------------------------------------- */
static uint64_t sd_mix(uint64_t x) { return eo_splitmix(&x); }

uint64_t eo_synth_record_hash(uint64_t k, int result, int stage, long long t, int st0, int head0, int abs_ok,
                              int rel_ok, const uint64_t ch[3]) {
  uint64_t h = sd_mix(k);
  h = sd_mix(h ^ (uint64_t)(uint32_t)result);
  h = sd_mix(h ^ (uint64_t)(uint32_t)stage);
  h = sd_mix(h ^ (uint64_t)(uint32_t)t);
  h = sd_mix(h ^ (uint64_t)(uint32_t)st0);
  h = sd_mix(h ^ (uint64_t)(uint32_t)head0);
  h = sd_mix(h ^ (uint64_t)(uint32_t)(abs_ok | (rel_ok << 1)));
  h = sd_mix(h ^ ch[0]);
  h = sd_mix(h ^ ch[1]);
  h = sd_mix(h ^ ch[2]);
  return h;
}

void eo_synth_digest(uint64_t seed0, uint64_t k0, long long n, int threads, eo_synth_sum *out) {
  eo_synth_sum tot;
  memset(&tot, 0, sizeof tot);
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    eo_synth_sum me;
    memset(&me, 0, sizeof me);
#pragma omp for schedule(dynamic, 256)
    for (long long j = 0; j < n; j++) {
      int tab[2 * EO_MAXQ];
      int32_t cap, dl[EO_NS + 1];
      eo_synth_table(seed0, k0 + (uint64_t)j, tab, &cap);
      for (int s = 0; s <= EO_NS; s++) dl[s] = cap;
      eo_sim *sim = (eo_sim *)malloc(sizeof(eo_sim));
      int r = eo_run_full(tab, sim, dl);
      me.digest += eo_synth_record_hash(k0 + (uint64_t)j, r, sim->stage, sim->t, sim->st0, sim->head0, sim->abs_ok,
                                        sim->rel_ok, sim->ch);
      me.steps += (uint64_t)sim->t;
      if (r == EO_SURV) me.survivors++;
      else me.hist[sim->stage]++;
      free(sim);
    }
#pragma omp critical
    {
      tot.digest += me.digest;
      tot.steps += me.steps;
      tot.survivors += me.survivors;
      for (int s = 0; s <= EO_NS; s++) tot.hist[s] += me.hist[s];
    }
  }
  tot.n = (uint64_t)n;
  *out = tot;
}


/* --------------------------------------------------------------------------
   Probe estimator: probe() (e13search.c:256-313) and the estimate mode
   (e13search.c:473-510).
   -------------------------------------------------------------------------- */

/* Probe p's own splitmix64 stream, started at a hashed point.  The reference
   starts one stream per OpenMP thread, at splitmix(0x5EED0000 + thread +
   (seed << 20)) (e13search.c:487-491), and a thread's probes share it, so a
   probe's draws depend on the schedule; one stream per probe makes every
   probe's result a function of (problem, seed, p) alone. */
uint64_t eo_probe_stream(uint64_t seed, uint64_t p) {
  uint64_t s = 0x5EED0000ULL + (seed << 40) + p;
  return eo_splitmix(&s);
}

typedef struct {
  eo_sim s, kids[4 * EO_MAXQ + 1];
} eo_probe_ws;

/* probe(): e13search.c:256-313.  One random root-to-leaf path; every child
   of a node on the path is run and counted with the path weight w, the path
   continues through a uniformly drawn live child and w grows by the number
   of live children.  Adds to out[0..3] = leaves, nodes, steps, survivors in
   the reference's order of additions, so the doubles agree bit for bit. */
static void probe_ws(const eo_problem *p, uint64_t *rng, double *out, eo_probe_ws *ws) {
  int tab[2 * EO_MAXQ], nintro = p->sym_n0, budget = p->budget;
  uint8_t fixed[2 * EO_MAXQ] = {0};
  double w = 1;
  eo_sim *s = &ws->s;
  memcpy(tab, p->tab, sizeof tab);
  eo_sim_init(s, NULL);
  int i = eo_run(s, tab, fixed, NULL);
  out[1] += 1;
  out[2] += (double)s->t;
  if (i < 0) {
    out[0] += 1;
    if (i == EO_SURV) out[3] += 1;
    return;
  }
  for (;;) {
    int orig = tab[i], vals[4 * EO_MAXQ + 1], nbs[4 * EO_MAXQ + 1], nis[4 * EO_MAXQ + 1], k = 0;
    if (orig != EO_FREE) { /* the keep child first (e13search.c:279) */
      vals[k] = orig;
      nbs[k] = budget;
      nis[k] = nintro;
      k++;
    }
    if (orig == EO_FREE || budget > 0) {
      for (int v = 0; v < 4 * p->nq; v++) {
        int ni = sym_next(p, v, nintro);
        if (v == orig || ni < 0) continue;
        vals[k] = v;
        nbs[k] = orig == EO_FREE ? budget : budget - 1;
        nis[k] = ni;
        k++;
      }
    }
    int live[4 * EO_MAXQ + 1], nlive = 0;
    fixed[i] = 1;
    for (int c = 0; c < k; c++) {
      eo_sim *ch = &ws->kids[c];
      *ch = *s;
      tab[i] = vals[c];
      long long t0 = ch->t;
      int r = eo_run(ch, tab, fixed, NULL);
      out[1] += w;
      out[2] += w * (double)(ch->t - t0);
      if (r < 0) {
        out[0] += w;
        if (r == EO_SURV) out[3] += w;
      } else {
        live[nlive++] = c;
      }
    }
    if (!nlive) return;
    int c = live[eo_splitmix(rng) % (uint64_t)nlive];
    w *= nlive;
    tab[i] = vals[c];
    budget = nbs[c];
    nintro = nis[c];
    *s = ws->kids[c];
    i = eo_run(s, tab, fixed, NULL); /* stops again at once at the same unbound cell */
  }
}

void eo_probe(const eo_problem *p, uint64_t *rng, double out[4]) {
  eo_probe_ws *ws = (eo_probe_ws *)malloc(sizeof(eo_probe_ws));
  probe_ws(p, rng, out, ws);
  free(ws);
}

/* The estimate mode (e13search.c:473-510) with one stream per probe: probes
   0..n-1 on `threads` OpenMP threads (<= 0: the default), the per-probe sums
   into per[4*k..4*k+3] when per is given, then the totals added in probe order
   and the mean and standard error as the reference prints them
   (e13search.c:503-506). */
int eo_estimate(const eo_problem *p, uint64_t seed, long long probes, int threads, double *per,
                eo_estimate_sum *out) {
  memset(out, 0, sizeof *out);
  if (p->budget < 0 || probes <= 0) return -1;
  double *buf = per ? per : (double *)malloc(sizeof(double) * 4 * (size_t)probes);
  if (!buf) return -1;
  if (threads > 0) omp_set_num_threads(threads);
#pragma omp parallel
  {
    eo_probe_ws *ws = (eo_probe_ws *)malloc(sizeof(eo_probe_ws));
    long long k;
#pragma omp for schedule(dynamic, 256)
    for (k = 0; k < probes; k++) {
      double o[4] = {0, 0, 0, 0};
      uint64_t rng = eo_probe_stream(seed, (uint64_t)k);
      probe_ws(p, &rng, o, ws);
      memcpy(buf + 4 * k, o, sizeof o);
    }
    free(ws);
  }
  out->probes = probes;
  for (long long k = 0; k < probes; k++)
    for (int j = 0; j < 4; j++) {
      out->sum[j] += buf[4 * k + j];
      out->sq[j] += buf[4 * k + j] * buf[4 * k + j];
    }
  for (int j = 0; j < 4; j++) {
    double m = out->sum[j] / (double)probes, var = out->sq[j] / (double)probes - m * m;
    out->mean[j] = m;
    out->se[j] = var > 0 ? sqrt(var / (double)probes) : 0.0;
  }
  if (!per) free(buf);
  return 0;
}
