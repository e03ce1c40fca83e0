/* e13_cli.c -- e13search_gpu: the reference's command line over libe13.so.
 *
 *   e13search_gpu base    machine.txt
 *   e13search_gpu mutant  machine.txt
 *   e13search_gpu delete  machine.txt FMAX X...
 *   e13search_gpu free    machine.txt X CELL...
 *   e13search_gpu freesym machine.txt X LO HI CELL...
 *   e13search_gpu gpu     machine.txt X LO HI CELL...   (the GPU driver's argv, src/e13gpu.py:293-301;
 *                                       LO = -1: no symmetry)
 *   e13search_gpu estimate machine.txt PROBES X FMAX LO HI [CELL...]   (probe estimator; LO < 0: no
 *                                       symmetry; cells: free mode; probe i draws from its own stream of
 *                                       seed $E13_SEED, default 0)
 *   e13search_gpu synth   SEED0 K0 N      (the synthetic stress batch; one summary line)
 *   e13search_gpu deep    NMAX < survivors.txt   (src/deep.c's stage check to NMAX, host threads)
 *   e13search_gpu deep_big NMAX < survivors.txt  (src/deep_big.c: 4M cells each side)
 *   e13search_gpu recheck NMAX < survivors.txt   (src/recheck.c's counter re-check, host threads)
 *   e13search_gpu deep_gpu NMAX < survivors.txt, recheck_gpu NMAX < survivors.txt   (the same on the
 *                                       GPU, one wave per survivor; NMAX <= 18)
 *   e13search_gpu check   survivors.txt          (src/check.py's crafted-start check, host only)
 *   e13search_gpu classify survivors.txt deep.txt...   (src/classify_deep.py, host only)
 *   e13search_gpu tune    machine.txt FMAX X...   (autotune on that search; saves the settings
 *   e13search_gpu tune    machine.txt freesym X LO HI CELL...   to $E13_TUNING_OUT or to the
 *                                       device's tuning file of the profile the search runs with)
 *
 * Argument meaning, stdout/stderr split and line formats follow
 * src/e13search.c main() (:435-472), report() (:184-200) and the summary at
 * :413-416; the GPU summary follows src/e13gpu.py:456-459.  After each
 * SURVIVOR line comes a HASH line with the counter hashes, and after each
 * summary a digest line (order-independent over the survivor and hash lines).
 * Survivors whose plain re-check disagrees are flagged on stderr.
 *
 * Environment: E13_FORM (1 plain, 2 batched), E13_SLICE, E13_CAP, E13_BATCH,
 * E13_GATE (1 count, 2 hash), E13_WG, E13_WGS_PER_CU, E13_DEVICE,
 * E13_REFILL_TRIES, E13_SEED_PER_WAVE, E13_SCHED_EVERY, E13_DONATE_MAX,
 * E13_IDLE_NAPS, E13_MAIL_EVERY, E13_GRAPH_CHAIN, E13_SOLO_STEPS, E13_EVICT_AFTER, E13_CUT_LIVE (negative: off),
 * E13_PROBE_SLICE, E13_NT1_SKIP (1 on, 2 off), E13_PHASE_CYCLES=1 (the persistent search's per-phase wave cycles),
 * E13_TAPE_WORDS (lane tape words of the persistent search: 18..42 short, 43 whole, negative automatic),
 * E13_WIDE_NODES (capacity of the list of machines past a short tape);
 * E13_TUNING_FILE (stored settings; <name>_small.<ext> beside it for small searches); E13_TRACE=N with
 * E13_TRACE_FILE=path writes the persistent search's wave event log (N events per wave) to path; the
 * library prints its host phase times with E13_TIMING=1 and a line per autotuner trial with E13_TUNE_LOG=1.
 * Checkpoints (e13gpu.py:19-23): E13_CKPT=file saves the search every E13_CKPT_SEC seconds (default
 * 1800) and after E13_CKPT_LAUNCHES launches if set; creating <file>.stop saves and exits with code 3;
 * the same command resumes.  The persistent form keeps one file for all X of a delete search; the host
 * loop (E13_FORM=1, 2) one per X (<file>.X<n>).
 */
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>
#include <unistd.h>

#include "e13.h"

static double now_s(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec + 1e-9 * (double)ts.tv_nsec;
}

static int env_int(const char *name) {
  const char *v = getenv(name);
  return v ? atoi(v) : 0;
}

static uint64_t mix64(uint64_t x) {
  x += 0x9E3779B97F4A7C15ULL;
  x = (x ^ (x >> 30)) * 0xBF58476D1CE4E5B9ULL;
  x = (x ^ (x >> 27)) * 0x94D049BB133111EBULL;
  return x ^ (x >> 31);
}

static uint64_t text_hash(const char *s) {
  uint64_t h = 0xcbf29ce484222325ULL;
  for (; *s; s++) h = (h ^ (uint8_t)*s) * 0x100000001b3ULL;
  return mix64(h);
}

static void print_cycles(const e13_stats *a, const e13_stats *b) {
  double c[4], t = 0;
  for (int j = 0; j < 4; j++) {
    c[j] = (double)(b->cycles[j] - a->cycles[j]);
    t += c[j];
  }
  if (t > 0)
    fprintf(stderr, "  wave cycles: refill %.1f%% steps %.1f%% finish %.1f%% pool/idle %.1f%% (%.3e per round)\n",
            100 * c[0] / t, 100 * c[1] / t, 100 * c[2] / t, 100 * c[3] / t,
            (b->rounds - a->rounds) > 0 ? t / (double)(b->rounds - a->rounds) : 0.0);
}

/* The wave event log of the last search, as int32 waves, int32 words per wave,
   then the words (e13_get_trace). */
static void dump_trace(e13_ctx *ctx) {
  const char *path = getenv("E13_TRACE_FILE");
  int32_t waves = 0, per = 0;
  if (!path || e13_get_trace(ctx, NULL, 0, &waves, &per) || !waves) return;
  uint64_t n = (uint64_t)waves * (uint64_t)per;
  uint64_t *buf = (uint64_t *)malloc(n * sizeof *buf);
  FILE *f = fopen(path, "wb");
  if (buf && f && !e13_get_trace(ctx, buf, n, &waves, &per)) {
    fwrite(&waves, sizeof waves, 1, f);
    fwrite(&per, sizeof per, 1, f);
    fwrite(buf, sizeof *buf, n, f);
  }
  if (f) fclose(f);
  free(buf);
}

static char *read_file(const char *path) {
  FILE *f = fopen(path, "rb");
  if (!f) return NULL;
  fseek(f, 0, SEEK_END);
  long n = ftell(f);
  fseek(f, 0, SEEK_SET);
  char *b = (char *)malloc((size_t)n + 1);
  if (!b) {
    fclose(f);
    return NULL;
  }
  size_t got = fread(b, 1, (size_t)n, f);
  b[got] = 0;
  fclose(f);
  return b;
}

/* Checkpoint file of the host-loop search ($E13_CKPT, e13gpu.py:19-23); a
   delete search over several X keeps one file per X (<file>.X<n>). */
static const char *g_ckpt;
static int g_ckpt_per_x;

static void set_checkpoint(e13_ctx *ctx, int X) {
  char path[1024];
  if (g_ckpt_per_x) snprintf(path, sizeof path, "%s.X%d", g_ckpt, X);
  else snprintf(path, sizeof path, "%s", g_ckpt);
  const char *sec = getenv("E13_CKPT_SEC");
  const char *nl = getenv("E13_CKPT_LAUNCHES");
  if (e13_set_checkpoint(ctx, path, sec ? atof(sec) : 0.0, nl ? atoll(nl) : 0)) {
    fprintf(stderr, "bad checkpoint path %s\n", path);
    exit(1);
  }
}

/* The summary line of the C program (e13search.c:413-416): split is
   $E13_SPLIT (default 2) as there, front the nodes the parallel phase started
   from (this library's breadth-first frontier); split, front and wall depend
   on the schedule, the other fields do not. */
static void c_summary(int X, int budget, const e13_stats *st, const e13_counts *cnt, double wall) {
  const char *sp = getenv("E13_SPLIT");
  fprintf(stderr, "X=%d budget=%d split=%d front=%lld leaves=%llu survivors=%llu steps=%llu wall=%.1fs "
                  "dead-at-stage(0..%d):",
          X, budget, sp ? atoi(sp) : 2, (long long)st->frontier, (unsigned long long)cnt->leaves,
          (unsigned long long)cnt->survivors, (unsigned long long)cnt->steps, wall, E13_NS);
  for (int j = 0; j <= E13_NS; j++) fprintf(stderr, " %llu", (unsigned long long)cnt->dead_at_stage[j]);
  fprintf(stderr, "\n");
}

/* gpu_line = 1: the GPU driver's summary line (e13gpu.py:456-459) instead of
   the C program's. */
static int run_one(e13_ctx *ctx, const int32_t *base, int nq_base, int X, int fmax, const uint8_t *freebase,
                   int lo, int hi, int gpu_line) {
  e13_problem p;
  int rc = e13_setup(base, nq_base, X, fmax, freebase, lo, hi, &p);
  if (rc == E13_EFORCED) {
    fprintf(stderr, "X=%d forced=%d > fmax\n", X, p.nforced); /* e13search.c:353 */
    return 0;
  }
  if (rc == E13_ESYM) {
    fprintf(stderr, "symmetric state is named by a fixed cell or has one\n");
    exit(1);
  }
  if (rc == E13_ESYMSTART) {
    fprintf(stderr, "start state in the symmetric states\n");
    exit(1);
  }
  if (rc) {
    fprintf(stderr, "setup failed: %s\n", e13_strerror(rc));
    exit(1);
  }
  const uint64_t cap = 200000;
  e13_survivor *sv = (e13_survivor *)calloc(cap, sizeof *sv);
  e13_counts cnt;
  uint64_t ns = 0;
  e13_stats st0, st1;
  e13_get_stats(ctx, &st0);
  if (g_ckpt) set_checkpoint(ctx, X);
  double w0 = now_s();
  rc = e13_search(ctx, &p, &cnt, sv, cap, &ns);
  double wall = now_s() - w0;
  if (rc == E13_ESTOPPED) { /* e13gpu.py:398-409 */
    e13_stats sst;
    e13_get_stats(ctx, &sst);
    fprintf(stderr, "STOPPED cleanly at wall=%.0fs X=%d peak stack=%lld; run the same command to resume\n", wall, X,
            (long long)sst.peak_stack);
    exit(3);
  }
  if (rc) {
    fprintf(stderr, "search failed: %s\n", e13_strerror(rc));
    exit(1);
  }
  e13_get_stats(ctx, &st1);
  uint64_t digest = 0;
  char line[1024], hl[256];
  for (uint64_t k = 0; k < ns && k < cap; k++) {
    e13_format_survivor(&p, &sv[k], line, sizeof line);
    e13_format_hash(&p, &sv[k], hl, sizeof hl);
    printf("%s\n%s\n", line, hl);
    if (!sv[k].recheck_ok) fprintf(stderr, "RECHECK FAILED: %s\n", line);
    char l2[1100], h2[300];
    snprintf(l2, sizeof l2, "%s\n", line);
    snprintf(h2, sizeof h2, "%s\n", hl);
    digest += text_hash(l2) ^ mix64(text_hash(h2));
  }
  fflush(stdout);
  if (gpu_line) {
    fprintf(stderr, "X=%d budget=%d gpu leaves=%llu survivors=%llu steps=%llu wall=%.1fs dead-at-stage(0..%d):", X,
            p.budget, (unsigned long long)cnt.leaves, (unsigned long long)cnt.survivors,
            (unsigned long long)cnt.steps, wall, E13_NS);
    for (int j = 0; j <= E13_NS; j++) fprintf(stderr, " %llu", (unsigned long long)cnt.dead_at_stage[j]);
    fprintf(stderr, "\n");
  } else {
    c_summary(X, p.budget, &st1, &cnt, wall);
  }
  fprintf(stderr,
          "  launches=%lld peak stack=%lld nodes kernel=%.1fms %.2e steps/s rounds=%lld lane-slots=%lld "
          "donations=%lld pool-takes=%lld settings=%s\nX=%d digest=%016llx\n",
          (long long)(st1.launches - st0.launches), (long long)st1.peak_stack, st1.kernel_ms - st0.kernel_ms,
          wall > 0 ? (double)cnt.steps / wall : 0.0, (long long)(st1.rounds - st0.rounds),
          (long long)(st1.lane_slots - st0.lane_slots), (long long)(st1.donations - st0.donations),
          (long long)(st1.pool_takes - st0.pool_takes), st1.profile == E13_PROFILE_SMALL ? "small-search" : "main", X,
          (unsigned long long)digest);
  print_cycles(&st0, &st1);
  dump_trace(ctx);
  if (ns > cap) fprintf(stderr, "WARNING: %llu survivors, only %llu printed\n", (unsigned long long)ns,
                        (unsigned long long)cap);
  fflush(stderr);
  free(sv);
  return 0;
}

/* Every X of a delete search in one persistent launch; the same lines as
   run_one, per X, in argument order. */
static int run_many(e13_ctx *ctx, const int32_t *base, int nq_base, int nx, char **xs, int fmax) {
  e13_problem ps[16];
  int xv[16], np = 0;
  for (int i = 0; i < nx; i++) {
    int X = atoi(xs[i]);
    e13_problem p;
    int rc = e13_setup(base, nq_base, X, fmax, NULL, -1, -1, &p);
    if (rc == E13_EFORCED) {
      fprintf(stderr, "X=%d forced=%d > fmax\n", X, p.nforced);
      continue;
    }
    if (rc || np == 16) {
      fprintf(stderr, "setup failed: %s\n", e13_strerror(rc ? rc : E13_EINVAL));
      exit(1);
    }
    xv[np] = X;
    ps[np++] = p;
  }
  if (!np) return 0;
  const uint64_t cap = 200000;
  e13_survivor *sv = (e13_survivor *)calloc(cap, sizeof *sv);
  e13_counts cnt[16];
  uint64_t ns = 0;
  e13_stats st0, st1;
  e13_get_stats(ctx, &st0);
  if (g_ckpt) set_checkpoint(ctx, 0); /* one file for the whole persistent search */
  double w0 = now_s();
  int rc = e13_search_many(ctx, ps, np, cnt, sv, cap, &ns);
  double wall = now_s() - w0;
  if (rc == E13_ESTOPPED) { /* e13gpu.py:398-409 */
    fprintf(stderr, "STOPPED cleanly at wall=%.0fs; run the same command to resume\n", wall);
    exit(3);
  }
  if (rc) {
    fprintf(stderr, "search failed: %s\n", e13_strerror(rc));
    exit(1);
  }
  e13_get_stats(ctx, &st1);
  uint64_t digest[16] = {0}, tot = 0;
  char line[1024], hl[256];
  for (int k = 0; k < np; k++)
    for (uint64_t j = 0; j < ns && j < cap; j++) {
      if (sv[j].prob != k) continue;
      e13_format_survivor(&ps[k], &sv[j], line, sizeof line);
      e13_format_hash(&ps[k], &sv[j], hl, sizeof hl);
      printf("%s\n%s\n", line, hl);
      if (!sv[j].recheck_ok) fprintf(stderr, "RECHECK FAILED: %s\n", line);
      char l2[1100], h2[300];
      snprintf(l2, sizeof l2, "%s\n", line);
      snprintf(h2, sizeof h2, "%s\n", hl);
      digest[k] += text_hash(l2) ^ mix64(text_hash(h2));
    }
  fflush(stdout);
  for (int k = 0; k < np; k++) {
    c_summary(xv[k], ps[k].budget, &st1, &cnt[k], wall);
    fprintf(stderr, "X=%d digest=%016llx\n", xv[k], (unsigned long long)digest[k]);
    tot += cnt[k].steps;
  }
  fprintf(stderr, "  one launch for %d problems (%s settings): kernel=%.1fms %.2e steps/s, rounds=%lld lane-slots=%lld "
                  "donations=%lld pool-takes=%lld\n",
          np, st1.profile == E13_PROFILE_SMALL ? "small-search" : "main", st1.kernel_ms - st0.kernel_ms, wall > 0 ? (double)tot / wall : 0.0, (long long)(st1.rounds - st0.rounds),
          (long long)(st1.lane_slots - st0.lane_slots), (long long)(st1.donations - st0.donations),
          (long long)(st1.pool_takes - st0.pool_takes));
  print_cycles(&st0, &st1);
  dump_trace(ctx);
  free(sv);
  return 0;
}

int main(int argc, char **argv) {
  if (argc < 3) {
    fprintf(stderr, "usage: e13search_gpu base|mutant|delete|free|freesym|gpu|estimate|tune machine.txt [...]\n"
                    "       e13search_gpu synth SEED0 K0 N | deep|deep_big|recheck|deep_gpu|recheck_gpu NMAX < survivors.txt |\n"
                    "       e13search_gpu check survivors.txt | classify survivors.txt deep.txt...\n");
    return 1;
  }
  if (!strcmp(argv[1], "deep") || !strcmp(argv[1], "deep_big") || !strcmp(argv[1], "recheck") ||
      !strcmp(argv[1], "deep_gpu") || !strcmp(argv[1], "recheck_gpu")) {
    /* deep NMAX < survivors.txt (src/deep.c; deep_big: src/deep_big.c), recheck NMAX < survivors.txt
       (src/recheck.c): host threads; deep_gpu / recheck_gpu: the same on the GPU */
    const int rech = !strcmp(argv[1], "recheck") || !strcmp(argv[1], "recheck_gpu"),
              big = !strcmp(argv[1], "deep_big"), gpu = strstr(argv[1], "_gpu") != NULL;
    const int nmax = atoi(argv[2]);
    size_t cap = 1 << 20, len = 0;
    char *text = (char *)malloc(cap);
    for (size_t k; text && (k = fread(text + len, 1, cap - len - 1, stdin)) > 0;) {
      len += k;
      if (len + 1 == cap) {
        char *t2 = (char *)realloc(text, cap *= 2);
        if (!t2) free(text);
        text = t2;
      }
    }
    if (!text) return 1;
    text[len] = 0;
    const char *th = getenv("OMP_NUM_THREADS");
    const int threads = th && atoi(th) > 0 ? atoi(th) : (int)sysconf(_SC_NPROCESSORS_ONLN);
    const int64_t outlen = (rech ? 20480 : 2048) * (int64_t)(len / 64 + 16);
    char *out = (char *)malloc((size_t)outlen);
    e13_ctx *dctx = NULL;
    int rc = !out ? E13_ENOMEM : E13_OK;
    if (!rc && gpu) rc = e13_open(env_int("E13_DEVICE"), NULL, &dctx);
    if (!rc)
      rc = gpu ? (rech ? e13_recheck_counter_gpu(dctx, text, nmax, out, outlen)
                       : e13_deep_check_gpu(dctx, text, nmax, out, outlen))
               : (rech ? e13_recheck_counter(text, nmax, threads, out, outlen)
                       : e13_deep_check(text, nmax, big, threads, out, outlen));
    if (dctx) e13_close(dctx);
    if (rc < 0) {
      fprintf(stderr, "%s: %s\n", argv[1], e13_strerror(rc));
      return 1;
    }
    fputs(out, stdout);
    if (rech) { /* recheck.c:100 */
      int good = 0;
      for (const char *l = out; (l = strstr(l, ": CONSISTENT")) != NULL; l++) good++;
      fprintf(stderr, "machines %d consistent %d failed %d\n", rc, good, rc - good);
    }
    free(out);
    free(text);
    return 0;
  }
  if (!strcmp(argv[1], "classify") && argc >= 4) { /* classify SURVIVORS DEEPOUT... (src/classify_deep.py) */
    char *txt[64] = {0};
    size_t len[64] = {0};
    for (int i = 2; i < argc && i < 64; i++) { /* the files, read whole */
      FILE *f = fopen(argv[i], "rb");
      if (!f) {
        fprintf(stderr, "cannot open %s\n", argv[i]);
        return 1;
      }
      fseek(f, 0, SEEK_END);
      len[i] = (size_t)ftell(f);
      fseek(f, 0, SEEK_SET);
      txt[i] = (char *)calloc(len[i] + 2, 1);
      if (!txt[i] || fread(txt[i], 1, len[i], f) != len[i]) return 1;
      if (len[i] && txt[i][len[i] - 1] != '\n') txt[i][len[i]++] = '\n';
      fclose(f);
    }
    size_t dl = 0;
    for (int i = 3; i < argc && i < 64; i++) dl += len[i];
    char *deep = (char *)calloc(dl + 1, 1), *out = (char *)malloc(1 << 20);
    if (!deep || !out) return 1;
    for (int i = 3, o = 0; i < argc && i < 64; i++) {
      memcpy(deep + o, txt[i], len[i]);
      o += (int)len[i];
    }
    const int rc = e13_classify_deep(txt[2], deep, out, 1 << 20);
    if (rc < 0) {
      fprintf(stderr, "classify: %s\n", e13_strerror(rc));
      return 1;
    }
    fputs(out, stdout);
    return 0;
  }
  if (!strcmp(argv[1], "check")) { /* check survivors.txt (src/check.py): host only, no GPU */
    FILE *f = fopen(argv[2], "r");
    if (!f) {
      fprintf(stderr, "cannot open %s\n", argv[2]);
      return 1;
    }
    char line[8192];
    char *out = (char *)malloc(1 << 16);
    while (out && fgets(line, sizeof line, f)) {
      if (!strstr(line, "table=")) continue;
      const int rc = e13_check_survivor(line, out, 1 << 16);
      if (rc < 0) {
        fprintf(stderr, "check: %s\n", e13_strerror(rc));
        return 1;
      }
      fputs(out, stdout);
      fflush(stdout);
    }
    fclose(f);
    free(out);
    return 0;
  }
  if (!strcmp(argv[1], "synth")) { /* synth SEED0 K0 N: the synthetic stress batch */
    e13_tuning st;
    memset(&st, 0, sizeof st);
    st.form = env_int("E13_FORM");
    st.slice = env_int("E13_SLICE");
    st.wg_threads = env_int("E13_WG");
    st.wgs_per_cu = env_int("E13_WGS_PER_CU");
    st.gate = env_int("E13_GATE");
    st.solo_steps = env_int("E13_SOLO_STEPS");
    e13_ctx *sc = NULL;
    int src = e13_open(env_int("E13_DEVICE"), &st, &sc);
    if (src) {
      fprintf(stderr, "e13_open: %s\n", e13_strerror(src));
      return 1;
    }
    uint64_t seed0 = strtoull(argv[2], NULL, 0), k0 = argc > 3 ? strtoull(argv[3], NULL, 0) : 0;
    long long n = argc > 4 ? atoll(argv[4]) : 16;
    e13_batch_sum sm;
    double w0 = now_s();
    src = e13_synth_run(sc, seed0, k0, n, &sm);
    double wall = now_s() - w0;
    if (src) {
      fprintf(stderr, "synth: %s\n", e13_strerror(src));
      return 1;
    }
    printf("synth seed0=0x%llx k0=%llu n=%llu digest=%016llx steps=%llu survivors=%llu dead-at-stage(0..9):",
           (unsigned long long)seed0, (unsigned long long)k0, (unsigned long long)sm.n,
           (unsigned long long)sm.digest, (unsigned long long)sm.steps, (unsigned long long)sm.survivors);
    for (int j = 0; j < 10; j++) printf(" %llu", (unsigned long long)sm.dead_at_stage[j]);
    printf("\n");
    fprintf(stderr, "  kernel=%.2fms %.3e steps/s %.3e tables/s (wall %.2fs incl. generation and copies)\n",
            sm.kernel_ms, (double)sm.steps / (sm.kernel_ms * 1e-3), (double)sm.n / (sm.kernel_ms * 1e-3), wall);
    e13_close(sc);
    return 0;
  }
  char *text = read_file(argv[2]);
  if (!text) {
    fprintf(stderr, "cannot open %s\n", argv[2]);
    return 1;
  }
  int32_t base[32];
  int nq = e13_parse_machine(text, base, 16);
  free(text);
  if (nq < 1) {
    fprintf(stderr, "cannot parse %s\n", argv[2]);
    return 1;
  }
  e13_tuning tu;
  memset(&tu, 0, sizeof tu);
  tu.form = env_int("E13_FORM");
  tu.slice = env_int("E13_SLICE");
  tu.cap = env_int("E13_CAP");
  tu.batch = env_int("E13_BATCH");
  tu.gate = env_int("E13_GATE");
  tu.wg_threads = env_int("E13_WG");
  tu.wgs_per_cu = env_int("E13_WGS_PER_CU");
  tu.refill_tries = env_int("E13_REFILL_TRIES");
  tu.seed_per_wave = env_int("E13_SEED_PER_WAVE");
  tu.sched_every = env_int("E13_SCHED_EVERY");
  tu.donate_max = env_int("E13_DONATE_MAX");
  tu.idle_naps = env_int("E13_IDLE_NAPS");
  tu.mail_every = env_int("E13_MAIL_EVERY");
  tu.trace_events = env_int("E13_TRACE");
  tu.graph_chain = env_int("E13_GRAPH_CHAIN");
  tu.solo_steps = env_int("E13_SOLO_STEPS");
  tu.evict_after = env_int("E13_EVICT_AFTER");
  tu.cut_live = env_int("E13_CUT_LIVE");
  tu.phase_cycles = env_int("E13_PHASE_CYCLES");
  tu.nt1_skip = env_int("E13_NT1_SKIP");
  tu.probe_slice = env_int("E13_PROBE_SLICE");
  tu.tape_words = env_int("E13_TAPE_WORDS");
  tu.wide_nodes = env_int("E13_WIDE_NODES");
  e13_ctx *ctx = NULL;
  int rc = e13_open(env_int("E13_DEVICE"), &tu, &ctx);
  if (rc) {
    fprintf(stderr, "e13_open: %s\n", e13_strerror(rc));
    return 1;
  }
  const char *mode = argv[1];
  g_ckpt = getenv("E13_CKPT");
  if (g_ckpt && !*g_ckpt) g_ckpt = NULL;
  if (!strcmp(mode, "base") || !strcmp(mode, "mutant")) {
    int ncells = 2 * nq, n = 0;
    int8_t *tabs = (int8_t *)malloc((size_t)ncells * (size_t)(ncells * (4 * nq + 1) + 1));
    int *ci = (int *)malloc(sizeof(int) * (size_t)(ncells * (4 * nq + 1) + 1));
    int *cv = (int *)malloc(sizeof(int) * (size_t)(ncells * (4 * nq + 1) + 1));
    if (!strcmp(mode, "base")) {
      for (int i = 0; i < ncells; i++) tabs[i] = (int8_t)base[i];
      n = 1;
    } else { /* every single-cell change (e13search.c:451-460) */
      for (int i = 0; i < ncells; i++)
        for (int v = -1; v < 4 * nq; v++) {
          if (v == base[i]) continue;
          for (int j = 0; j < ncells; j++) tabs[(size_t)n * ncells + j] = (int8_t)base[j];
          tabs[(size_t)n * ncells + i] = (int8_t)v;
          ci[n] = i;
          cv[n] = v;
          n++;
        }
    }
    e13_leaf *lv = (e13_leaf *)calloc((size_t)n, sizeof *lv);
    rc = e13_simulate_batch(ctx, tabs, ncells, n, NULL, lv);
    if (rc) {
      fprintf(stderr, "simulate: %s\n", e13_strerror(rc));
      return 1;
    }
    if (!strcmp(mode, "base"))
      printf("base states=%d result=%s stage=%d st0=%d head0=%d abs=%d rel=%d steps=%d\n", nq,
             lv[0].result == E13_SURV ? "SURV" : "DEAD", lv[0].stage, lv[0].st0, lv[0].head0, lv[0].abs_ok,
             lv[0].rel_ok, lv[0].steps);
    else
      for (int k = 0; k < n; k++) printf("%d %d %d %d\n", ci[k], cv[k], lv[k].stage, lv[k].result == E13_SURV);
    free(lv);
    free(tabs);
    free(ci);
    free(cv);
  } else if (!strcmp(mode, "delete")) {
    if (argc < 4) return 1;
    int fmax = atoi(argv[3]);
    e13_tuning cur;
    e13_get_tuning(ctx, &cur);
    g_ckpt_per_x = cur.form != 3 && argc - 4 > 1; /* the host loop keeps one file per X */
    if (cur.form == 3 && argc - 4 > 1)
      run_many(ctx, base, nq, argc - 4, argv + 4, fmax);
    else
      for (int i = 4; i < argc; i++) run_one(ctx, base, nq, atoi(argv[i]), fmax, NULL, -1, -1, 0);
  } else if (!strcmp(mode, "tune")) { /* tune machine.txt FMAX X... | tune machine.txt freesym X LO HI CELL... */
    if (argc < 5) return 1;
    int np = 0;
    e13_problem ps[16];
    if (!strcmp(argv[3], "freesym")) { /* autotune on that free-cell search */
      if (argc < 8) return 1;
      uint8_t fb[32] = {0};
      for (int i = 7; i < argc; i++) {
        int q = argv[i][0] - 'A', s = argv[i][1] - '0';
        if (q < 0 || q >= nq || (s != 0 && s != 1)) {
          fprintf(stderr, "bad cell %s\n", argv[i]);
          return 1;
        }
        fb[2 * q + s] = 1;
      }
      if (e13_setup(base, nq, atoi(argv[4]), 0, fb, atoi(argv[5]), atoi(argv[6]), &ps[0]) == E13_OK) np = 1;
    } else { /* autotune on that delete search */
      int fmax = atoi(argv[3]);
      for (int i = 4; i < argc && np < 16; i++) {
        e13_problem p;
        if (e13_setup(base, nq, atoi(argv[i]), fmax, NULL, -1, -1, &p) == E13_OK) ps[np++] = p;
      }
    }
    /* the stored profile these problems run with is the one the result replaces */
    const int profile = np ? e13_search_profile(ctx, ps, np, NULL) : E13_PROFILE_MAIN;
    e13_tuning best;
    e13_tune_report rep;
    rc = np ? e13_autotune(ctx, ps, np, 3, &best, &rep) : E13_EINVAL;
    if (rc) {
      fprintf(stderr, "autotune: %s\n", e13_strerror(rc));
      return 1;
    }
    char path[1024];
    const char *out = getenv("E13_TUNING_OUT");
    if (out && *out) snprintf(path, sizeof path, "%s", out);
    else if ((profile == E13_PROFILE_SMALL ? e13_tuning_file_small : e13_tuning_file)(ctx, path, sizeof path))
      path[0] = 0;
    fprintf(stderr, "autotune (%s profile): %d trials, %.1f ms -> %.1f ms; wgs_per_cu=%d slice=%d refill_tries=%d "
                    "solo_steps=%d seed_per_wave=%d donate_max=%d sched_every=%d wg_threads=%d idle_naps=%d "
                    "evict_after=%d cut_live=%d\n",
            profile == E13_PROFILE_SMALL ? "small-search" : "main", rep.trials, rep.first_ms, rep.best_ms,
            best.wgs_per_cu, best.slice, best.refill_tries, best.solo_steps, best.seed_per_wave, best.donate_max,
            best.sched_every, best.wg_threads, best.idle_naps, best.evict_after, best.cut_live);
    if (path[0] && e13_save_tuning(&best, path) == E13_OK) fprintf(stderr, "autotune: saved to %s\n", path);
  } else if (!strcmp(mode, "estimate")) { /* estimate machine.txt PROBES X FMAX LO HI [CELL...] (e13search.c:473-510) */
    if (argc < 8) {
      fprintf(stderr, "usage: e13search_gpu estimate machine.txt PROBES X FMAX LO HI [CELL...]\n");
      return 1;
    }
    const long long probes = atoll(argv[3]);
    const int X = atoi(argv[4]), fmax = atoi(argv[5]), lo = atoi(argv[6]), hi = atoi(argv[7]);
    uint8_t fb[32] = {0};
    for (int i = 8; i < argc; i++) {
      int q = argv[i][0] - 'A', s = argv[i][1] - '0';
      if (q < 0 || q >= nq || (s != 0 && s != 1)) {
        fprintf(stderr, "bad cell %s\n", argv[i]);
        return 1;
      }
      fb[2 * q + s] = 1;
    }
    e13_problem p;
    rc = e13_setup(base, nq, X, fmax, argc > 8 ? fb : NULL, lo, lo >= 0 ? hi : -1, &p);
    if (rc == E13_EFORCED) {
      fprintf(stderr, "X=%d forced=%d > fmax\n", X, p.nforced);
      return 1;
    }
    if (rc) {
      fprintf(stderr, "setup: %s\n", e13_strerror(rc));
      return 1;
    }
    const char *seed_env = getenv("E13_SEED");
    const uint64_t seed = seed_env ? strtoull(seed_env, NULL, 10) : 0;
    e13_estimate_sum es;
    const double w0 = now_s();
    rc = probes > 0 ? e13_estimate(ctx, &p, seed, probes, NULL, &es) : E13_EINVAL;
    if (rc) {
      fprintf(stderr, "estimate: %s\n", e13_strerror(rc));
      return 1;
    }
    printf("estimate X=%d budget=%d probes=%lld wall=%.1fs\n", X, p.budget, probes, now_s() - w0);
    const char *name[4] = {"leaves", "nodes", "steps", "survivors"};
    for (int j = 0; j < 4; j++) printf("  %-9s %.3e  (+- %.1e standard error)\n", name[j], es.mean[j], es.se[j]);
    fprintf(stderr, "  gpu: kernel=%.1fms %.3e probes/s %.3e steps/s (%llu steps), launches=%lld "
                    "lanes running per round=%.1f\n",
            es.kernel_ms, es.kernel_ms > 0 ? 1e3 * (double)probes / es.kernel_ms : 0.0,
            es.kernel_ms > 0 ? 1e3 * (double)es.sim_steps / es.kernel_ms : 0.0, (unsigned long long)es.sim_steps,
            (long long)es.launches, es.rounds ? (double)es.lane_slots / (double)es.rounds : 0.0);
  } else if (!strcmp(mode, "gpu")) { /* the GPU driver's argv: machine.txt X LO HI CELL... (e13gpu.py:293-301) */
    if (argc < 6) {
      fprintf(stderr, "usage: e13search_gpu gpu machine.txt X LO HI CELL...   (LO = -1: no symmetry)\n");
      return 1;
    }
    uint8_t fb[32] = {0};
    const int X = atoi(argv[3]), lo = atoi(argv[4]), hi = atoi(argv[5]);
    for (int i = 6; i < argc; i++) {
      int q = argv[i][0] - 'A', s = argv[i][1] - '0';
      if (q < 0 || q >= nq || (s != 0 && s != 1)) {
        fprintf(stderr, "bad cell %s\n", argv[i]);
        return 1;
      }
      fb[2 * q + s] = 1;
    }
    e13_problem p;
    if (e13_setup(base, nq, X, 0, fb, lo, lo >= 0 ? hi : -1, &p) == E13_OK) { /* the driver's first line (:318) */
      e13_tuning cur;
      e13_get_tuning(ctx, &cur);
      fprintf(stderr, "X=%d nq=%d free=%d sym=%d..%d n0=%d node=%dB form=%d\n", X, p.nq, argc - 6, p.sym_lo, p.sym_hi,
              p.sym_n0, 304, cur.form);
    }
    run_one(ctx, base, nq, X, 0, fb, lo, lo >= 0 ? hi : -1, 1);
  } else if (!strcmp(mode, "free") || !strcmp(mode, "freesym")) {
    if (argc < 4) return 1;
    uint8_t fb[32] = {0};
    int X = atoi(argv[3]), a0 = 4, lo = -1, hi = -1;
    if (!strcmp(mode, "freesym")) {
      if (argc < 6) return 1;
      lo = atoi(argv[4]);
      hi = atoi(argv[5]);
      a0 = 6;
    }
    for (int i = a0; i < argc; i++) {
      int q = argv[i][0] - 'A', s = argv[i][1] - '0';
      if (q < 0 || q >= nq || (s != 0 && s != 1)) {
        fprintf(stderr, "bad cell %s\n", argv[i]);
        return 1;
      }
      fb[2 * q + s] = 1;
    }
    run_one(ctx, base, nq, X, 0, fb, lo, hi, 0);
  } else {
    fprintf(stderr, "unknown mode %s\n", mode);
    return 1;
  }
  e13_close(ctx);
  return 0;
}
