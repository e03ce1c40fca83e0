/* e13oracle.c -- command line of the CPU oracle (test infrastructure only).
 *
 * Modes and output lines follow the reference program src/e13search.c
 * (main at :435-472, report at :184-200, summary at :413-416):
 *
 *   e13oracle base    machine.txt
 *   e13oracle mutant  machine.txt
 *   e13oracle delete  machine.txt FMAX X...
 *   e13oracle free    machine.txt X CELL...
 *   e13oracle freesym machine.txt X LO HI CELL...
 *   e13oracle estimate machine.txt PROBES X FMAX LO HI [CELL...]
 *                     (probe i draws from its own stream, see e13oracle_core.c;
 *                     seed from E13_SEED)
 *
 * Additions that the reference does not print: after every SURVIVOR line a
 *   HASH X=.. ch0=%016llx ch1=.. ch2=..
 * line with the counter hashes, and after every summary line a
 *   X=.. digest=%016llx
 * line, an order-independent digest of the (SURVIVOR, HASH) pairs.
 * Extra modes for the tests:
 *   e13oracle synth SEED0 K0 N        synthetic stress tables K0..K0+N-1: one
 *                                     line "k result stage t st0 head0 ch0 ch1 ch2"
 *   e13oracle synthsum SEED0 K0 N     the same tables: one line with the
 *                                     order-independent digest and totals"
 *   e13oracle hash  machine-lines <   SURVIVOR lines on stdin -> HASH lines
 *                                     (all cells fixed; '?---' read as halt)
 *
 * Environment: E13_SPLIT (default 2), OMP_NUM_THREADS.
 */
#include "e13oracle.h"

#include <omp.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

static void print_cell(FILE *o, int code) { /* e13search.c:179-182 */
  if (code < 0)
    fputs("---", o);
  else
    fprintf(o, "%d%c%d", code & 1, (code & 2) ? 'R' : 'L', code >> 2);
}

static uint64_t mix64(uint64_t x) {
  uint64_t s = x;
  return eo_splitmix(&s);
}

static uint64_t text_hash(const char *s) { /* FNV-1a folded through splitmix */
  uint64_t h = 0xcbf29ce484222325ULL;
  for (; *s; s++) h = (h ^ (uint8_t)*s) * 0x100000001b3ULL;
  return mix64(h);
}

typedef struct {
  uint64_t digest;
} report_ctx;

static void report(void *user, const eo_sim *s, const int *tab, const uint8_t *fixed, int X, int nq) {
  report_ctx *rc = (report_ctx *)user;
  char line[1024];
  int n = snprintf(line, sizeof line, "SURVIVOR X=%d st0=%d head0=%d abs=%d rel=%d steps=%lld table=", X, s->st0,
                   s->head0, s->abs_ok, s->rel_ok, s->t);
  for (int q = 0; q < nq; q++) {
    for (int a = 0; a < 2; a++) {
      int i = 2 * q + a, code = tab[i];
      if (!fixed[i]) line[n++] = '?';
      if (code < 0)
        n += snprintf(line + n, sizeof line - n, "---");
      else
        n += snprintf(line + n, sizeof line - n, "%d%c%d", code & 1, (code & 2) ? 'R' : 'L', code >> 2);
      if (a == 0) line[n++] = ' ';
    }
    line[n++] = q + 1 < nq ? '|' : '\n';
  }
  line[n] = 0;
  char hl[160];
  snprintf(hl, sizeof hl, "HASH X=%d ch0=%016llx ch1=%016llx ch2=%016llx\n", X, (unsigned long long)s->ch[0],
           (unsigned long long)s->ch[1], (unsigned long long)s->ch[2]);
  fputs(line, stdout);
  fputs(hl, stdout);
  fflush(stdout);
  rc->digest += text_hash(line) ^ mix64(text_hash(hl));
}

static int run_search(const char *path, int X, int fmax, int freemode, const uint8_t *freebase, int lo, int hi,
                      int split) {
  eo_problem p;
  memset(&p, 0, sizeof p);
  p.nq_base = eo_load(path, p.base);
  if (p.nq_base < 0) {
    fprintf(stderr, "cannot open %s\n", path);
    exit(1);
  }
  p.X = X;
  p.fmax = fmax;
  p.freemode = freemode;
  if (freebase) memcpy(p.freebase, freebase, sizeof p.freebase);
  p.sym_lo = p.sym_hi = -1;
  if (lo >= 0) eo_set_sym(&p, lo, hi);
  int rc = eo_setup(&p);
  if (rc == -2 || rc == -3) {
    fprintf(stderr, rc == -2 ? "symmetric state is named by a fixed cell or has one\n"
                             : "start state in the symmetric states\n");
    exit(1);
  }
  if (rc == -1) {
    int nf = 0;
    for (int i = 0; i < 2 * p.nq; i++) nf += p.tab[i] == EO_FREE;
    fprintf(stderr, "X=%d forced=%d > fmax\n", X, nf);
    return 0;
  }
  report_ctx ctx = {0};
  eo_counts c;
  double w0 = omp_get_wtime();
  eo_search(&p, split, 0, &c, report, &ctx);
  fprintf(stderr,
          "X=%d budget=%d split=%d front=%ld leaves=%lld survivors=%lld steps=%lld wall=%.1fs dead-at-stage(0..%d):",
          X, p.budget, split, c.nfront, c.leaves, c.survivors, c.steps, omp_get_wtime() - w0, EO_NS);
  for (int j = 0; j <= EO_NS; j++) fprintf(stderr, " %lld", c.hist[j]);
  fprintf(stderr, "\nX=%d digest=%016llx\n", X, (unsigned long long)ctx.digest);
  fflush(stderr);
  return 0;
}

/* Parse the table= part of a SURVIVOR line into codes; '?---' and '---' -> halt. */
static int parse_table(const char *tb, int *tab) {
  int n = 0;
  const char *p = tb;
  while (*p && *p != '\n' && n < 2 * EO_MAXQ) {
    while (*p == ' ' || *p == '|') p++;
    if (!*p || *p == '\n') break;
    if (*p == '?') p++;
    char cell[16];
    int k = 0;
    while (*p && *p != ' ' && *p != '|' && *p != '\n' && k < 15) cell[k++] = *p++;
    cell[k] = 0;
    tab[n++] = eo_parse_cell(cell);
  }
  return n;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: e13oracle base|mutant|delete|free|freesym|synth|hash machine.txt [...]\n");
    return 1;
  }
  int split = 2;
  const char *se = getenv("E13_SPLIT");
  if (se) split = atoi(se);
  eo_init_keys();
  const char *mode = argv[1];
  if (!strcmp(mode, "synth")) {
    uint64_t seed0 = strtoull(argv[2], NULL, 0), k0 = argc > 3 ? strtoull(argv[3], NULL, 0) : 0;
    long n = argc > 4 ? atol(argv[4]) : 16;
    for (long j = 0; j < n; j++) {
      int tab[2 * EO_MAXQ];
      int32_t cap, dl[EO_NS + 1];
      eo_synth_table(seed0, k0 + (uint64_t)j, tab, &cap);
      for (int s = 0; s <= EO_NS; s++) dl[s] = cap;
      eo_sim sim;
      int r = eo_run_full(tab, &sim, dl);
      printf("%llu %d %d %lld %d %d %016llx %016llx %016llx\n", (unsigned long long)(k0 + j), r, sim.stage, sim.t,
             sim.st0, sim.head0, (unsigned long long)sim.ch[0], (unsigned long long)sim.ch[1],
             (unsigned long long)sim.ch[2]);
    }
    return 0;
  }
  if (!strcmp(mode, "synthsum")) {
    uint64_t seed0 = strtoull(argv[2], NULL, 0), k0 = argc > 3 ? strtoull(argv[3], NULL, 0) : 0;
    long long n = argc > 4 ? atoll(argv[4]) : 16;
    eo_synth_sum sm;
    eo_synth_digest(seed0, k0, n, 0, &sm);
    printf("synth seed0=0x%llx k0=%llu n=%llu digest=%016llx steps=%llu survivors=%llu dead-at-stage(0..9):",
           (unsigned long long)seed0, (unsigned long long)k0, (unsigned long long)sm.n,
           (unsigned long long)sm.digest, (unsigned long long)sm.steps, (unsigned long long)sm.survivors);
    for (int j = 0; j <= EO_NS; j++) printf(" %llu", (unsigned long long)sm.hist[j]);
    printf("\n");
    return 0;
  }
  if (!strcmp(mode, "hash")) {
    char line[8192];
    while (fgets(line, sizeof line, stdin)) {
      char *tb = strstr(line, "table=");
      if (!tb) continue;
      int tab[2 * EO_MAXQ], X = 0;
      char *xp = strstr(line, "X=");
      if (xp) X = atoi(xp + 2);
      parse_table(tb + 6, tab);
      eo_sim s;
      int r = eo_run_full(tab, &s, NULL);
      printf("HASH X=%d ch0=%016llx ch1=%016llx ch2=%016llx result=%d steps=%lld\n", X, (unsigned long long)s.ch[0],
             (unsigned long long)s.ch[1], (unsigned long long)s.ch[2], r, s.t);
    }
    return 0;
  }
  int base[2 * EO_MAXQ];
  int nq = eo_load(argv[2], base);
  if (nq < 0) {
    fprintf(stderr, "cannot open %s\n", argv[2]);
    return 1;
  }
  if (!strcmp(mode, "base")) {
    eo_sim s;
    int r = eo_run_full(base, &s, NULL);
    printf("base states=%d result=%s stage=%d st0=%d head0=%d abs=%d rel=%d steps=%lld\n", nq,
           r == EO_SURV ? "SURV" : "DEAD", s.stage, s.st0, s.head0, s.abs_ok, s.rel_ok, s.t);
  } else if (!strcmp(mode, "mutant")) {
    for (int i = 0; i < 2 * nq; i++)
      for (int v = -1; v < 4 * nq; v++) {
        if (v == base[i]) continue;
        int tab[2 * EO_MAXQ];
        memcpy(tab, base, sizeof tab);
        tab[i] = v;
        eo_sim s;
        int r = eo_run_full(tab, &s, NULL);
        printf("%d %d %d %d\n", i, v, s.stage, r == EO_SURV);
      }
  } else if (!strcmp(mode, "delete")) {
    int fmax = atoi(argv[3]);
    for (int i = 4; i < argc; i++) run_search(argv[2], atoi(argv[i]), fmax, 0, NULL, -1, -1, split);
  } else if (!strcmp(mode, "free") || !strcmp(mode, "freesym")) {
    uint8_t fb[2 * EO_MAXQ] = {0};
    int X = atoi(argv[3]), a0 = 4, lo = -1, hi = -1;
    if (!strcmp(mode, "freesym")) {
      lo = atoi(argv[4]);
      hi = atoi(argv[5]);
      a0 = 6;
    }
    for (int i = a0; i < argc; i++) fb[2 * (argv[i][0] - 'A') + (argv[i][1] - '0')] = 1;
    run_search(argv[2], X, 0, 1, fb, lo, hi, split);
  } else if (!strcmp(mode, "estimate")) {
    /* e13search.c:473-510: estimate machine PROBES X FMAX LO HI [CELL...];
       LO < 0: no symmetry; cells: free mode.  One stream per probe. */
    if (argc < 8) {
      fprintf(stderr, "usage: e13oracle estimate machine.txt PROBES X FMAX LO HI [CELL...]\n");
      return 1;
    }
    eo_problem p;
    memset(&p, 0, sizeof p);
    p.nq_base = nq;
    memcpy(p.base, base, sizeof p.base);
    long long probes = atoll(argv[3]);
    p.X = atoi(argv[4]);
    p.fmax = atoi(argv[5]);
    p.sym_lo = p.sym_hi = -1;
    if (atoi(argv[6]) >= 0) eo_set_sym(&p, atoi(argv[6]), atoi(argv[7]));
    for (int i = 8; i < argc; i++) p.freebase[2 * (argv[i][0] - 'A') + (argv[i][1] - '0')] = 1;
    p.freemode = argc > 8;
    int rc = eo_setup(&p);
    if (rc == -2 || rc == -3) {
      fprintf(stderr, rc == -2 ? "symmetric state is named by a fixed cell or has one\n"
                               : "start state in the symmetric states\n");
      return 1;
    }
    if (rc == -1) {
      fprintf(stderr, "X=%d forced > fmax\n", p.X);
      return 1;
    }
    const char *seed_env = getenv("E13_SEED");
    uint64_t seed = seed_env ? strtoull(seed_env, NULL, 10) : 0;
    eo_estimate_sum es;
    double w0 = omp_get_wtime();
    eo_estimate(&p, seed, probes, 0, NULL, &es);
    printf("estimate X=%d budget=%d probes=%lld wall=%.1fs\n", p.X, p.budget, probes, omp_get_wtime() - w0);
    const char *name[4] = {"leaves", "nodes", "steps", "survivors"};
    for (int j = 0; j < 4; j++) printf("  %-9s %.3e  (+- %.1e standard error)\n", name[j], es.mean[j], es.se[j]);
  } else {
    fprintf(stderr, "unknown mode %s\n", mode);
    return 1;
  }
  (void)print_cell;
  return 0;
}
