/* e13_tune.c -- the autotuner: picks launch and schedule settings for this
 * device by measuring them on a real workload, and stores them per device.
 *
 * No reference counterpart (the reference adapts only its GPU batch size,
 * src/e13gpu.py:427-430, and takes E13_SLICE from the environment, :346).
 *
 * Coordinate descent over the knobs that move the persistent search: resident
 * waves per CU (with one lane per leaf, the occupancy knob: a step is a serial
 * dependency chain, so a leaf has no work for a second lane), the refill
 * granularity (steps per round, refill passes), the tape placement of a lone
 * running lane (LDS loop only, or the register loop with its round length),
 * the work-queue depth (frontier per wave, nodes per delivery pass, rounds
 * between hungry-map reads) and the step count after which a machine is handed
 * to an idle wave to run alone.  Every trial's counts must equal the first
 * trial's; the settings change the schedule, never the result.
 */
#define _GNU_SOURCE 1 /* dladdr */
#include <dlfcn.h>
#include <time.h>
#include <libgen.h>
#include <stdlib.h>
#include <string.h>

#include "e13_ctx.h"

#define NKNOB 11
static const char *knob_name[NKNOB] = {"wgs_per_cu", "slice",       "refill_tries", "solo_steps", "seed_per_wave",
                                       "donate_max", "sched_every", "wg_threads",   "idle_naps",  "evict_after",
                                       "cut_live"};
/* knobs whose "off" reads back as 0 and is set (and stored) as -1 */
static int off_knob(int k) { return k == 3 || k == 4 || k == 9 || k == 10; }

static int32_t *knob(e13_tuning *t, int k) {
  switch (k) {
  case 0: return &t->wgs_per_cu;
  case 1: return &t->slice;
  case 2: return &t->refill_tries;
  case 3: return &t->solo_steps;
  case 4: return &t->seed_per_wave;
  case 5: return &t->donate_max;
  case 6: return &t->sched_every;
  case 7: return &t->wg_threads;
  case 8: return &t->idle_naps;
  case 9: return &t->evict_after;
  default: return &t->cut_live;
  }
}

/* Candidate values (-1: off / none, as tuning_merge reads it). */
static const int32_t cand[NKNOB][6] = {
    {1, 2, 3, 4, 0},  {16, 32, 64, 128, 256, 0}, {1, 2, 4, 0, 0},     {-1, 1024, 4096, 16384, 65536, 0}, {-1, 2, 8, 0, 0},
    {2, 4, 8, 16, 0}, {8, 32, 128, 0, 0},   {128, 256, 0, 0, 0}, {1, 4, 16, 0, 0},                  {-1, 4096, 16384, 65536, 0},
    {-1, 16, 32, 48, 0}};

static int same_counts(const e13_counts *a, const e13_counts *b, int np) {
  for (int k = 0; k < np; k++) {
    if (a[k].leaves != b[k].leaves || a[k].steps != b[k].steps || a[k].survivors != b[k].survivors) return 0;
    for (int j = 0; j <= E13_NS; j++)
      if (a[k].dead_at_stage[j] != b[k].dead_at_stage[j]) return 0;
  }
  return 1;
}

static double now_ms(void) {
  struct timespec ts;
  clock_gettime(CLOCK_MONOTONIC, &ts);
  return (double)ts.tv_sec * 1e3 + (double)ts.tv_nsec * 1e-6;
}

/* Wall time of the search call under tuning t (the median of reps: the best
   of a few runs would favour settings whose times spread widely): the
   persistent kernel, the frontier launches before it and their host round
   trips; counts checked against ref. */
static int trial(e13_ctx *c, const e13_tuning *t, const e13_problem *ps, int np, int reps, e13_counts *ref,
                 int *have_ref, double *ms) {
  e13_tuning t2 = *t; /* "off" is 0 in a read-back tuning, -1 when set */
  for (int k = 0; k < NKNOB; k++)
    if (off_knob(k) && *knob(&t2, k) == 0) *knob(&t2, k) = -1;
  int rc = e13_set_tuning(c, &t2);
  if (rc) return rc;
  e13_counts cnt[E13_MAXPROB];
  double times[64];
  if (reps > 64) reps = 64;
  for (int r = 0; r < reps; r++) {
    uint64_t ns = 0;
    const double t0 = now_ms();
    if ((rc = e13_search_many(c, ps, np, cnt, NULL, 0, &ns))) return rc;
    times[r] = now_ms() - t0;
    if (!*have_ref) {
      memcpy(ref, cnt, (size_t)np * sizeof *cnt);
      *have_ref = 1;
    } else if (!same_counts(ref, cnt, np)) {
      fprintf(stderr, "libe13: autotune: counts changed under a tuning setting\n");
      return E13_EINVAL;
    }
  }
  for (int i = 1; i < reps; i++) /* insertion sort, then the median */
    for (int j = i; j > 0 && times[j] < times[j - 1]; j--) {
      const double x = times[j];
      times[j] = times[j - 1];
      times[j - 1] = x;
    }
  *ms = reps % 2 ? times[reps / 2] : 0.5 * (times[reps / 2 - 1] + times[reps / 2]);
  return E13_OK;
}

int e13_autotune(e13_ctx *c, const e13_problem *ps, int np, int reps, e13_tuning *best, e13_tune_report *rep) {
  if (!c || !ps || np < 1 || np > E13_MAXPROB || reps < 1) return E13_EINVAL;
  e13_tuning cur, start, installed;
  e13_get_tuning(c, &installed);
  e13_search_profile(c, ps, np, &start); /* the profile these problems would run with */
  cur = start;
  cur.form = 3;
  e13_counts ref[E13_MAXPROB];
  int have_ref = 0, trials = 0, rc;
  double best_ms, ms;
  const int log = getenv("E13_TUNE_LOG") && *getenv("E13_TUNE_LOG") != '0'; /* a line per trial on stderr */
  if ((rc = trial(c, &cur, ps, np, reps, ref, &have_ref, &best_ms))) goto fail;
  trials++;
  if (log) fprintf(stderr, "autotune: start -> %.1f ms\n", best_ms);
  const double first_ms = best_ms;
  for (int pass = 0; pass < 2; pass++) {
    int changed = 0;
    for (int k = 0; k < NKNOB; k++) {
      const int32_t keep = *knob(&cur, k);
      int32_t pick = keep;
      for (int i = 0; i < 6 && cand[k][i] != 0; i++) {
        const int32_t v = cand[k][i];
        /* the current value (as stored: off is 0 after merging) is already measured */
        if (v == keep || (v < 0 && keep == 0)) continue;
        e13_tuning t = cur;
        *knob(&t, k) = v;
        if (log) fprintf(stderr, "autotune: %s=%d ...", knob_name[k], (int)v);
        if ((rc = trial(c, &t, ps, np, reps, ref, &have_ref, &ms))) goto fail;
        trials++;
        if (log) fprintf(stderr, " %.1f ms\n", ms);
        if (ms < 0.98 * best_ms) { /* a clear gain only: timing noise does not move the choice */
          best_ms = ms;
          pick = v;
        }
      }
      if (pick != keep) changed = 1;
      *knob(&cur, k) = pick;
    }
    if (!changed) break;
  }
  for (int k = 0; k < NKNOB; k++)
    if (off_knob(k) && *knob(&cur, k) == 0) *knob(&cur, k) = -1;
  if ((rc = e13_set_tuning(c, &cur))) goto fail;
  e13_get_tuning(c, &cur);
  if (best) *best = cur;
  if (rep) {
    memset(rep, 0, sizeof *rep);
    rep->trials = trials;
    rep->first_ms = first_ms;
    rep->best_ms = best_ms;
  }
  return E13_OK;
fail:
  e13_set_tuning(c, &installed);
  return rc;
}

int e13_save_tuning(const e13_tuning *t, const char *path) {
  if (!t || !path) return E13_EINVAL;
  FILE *f = fopen(path, "w");
  if (!f) return E13_EINVAL;
  e13_tuning tmp = *t;
  fprintf(f, "# e13 tuning (autotuned): name value\n");
  for (int k = 0; k < NKNOB; k++) {
    int32_t v = *knob(&tmp, k);
    if (off_knob(k)) v = v > 0 ? v : -1; /* off is written as -1, which tuning_merge reads as off */
    fprintf(f, "%s %d\n", knob_name[k], (int)v);
  }
  fprintf(f, "gate %d\n", (int)t->gate);
  fclose(f);
  return E13_OK;
}

int e13_load_tuning(const char *path, e13_tuning *t) {
  if (!path || !t) return E13_EINVAL;
  FILE *f = fopen(path, "r");
  if (!f) return E13_EINVAL;
  memset(t, 0, sizeof *t);
  char name[64];
  int v;
  char line[256];
  while (fgets(line, sizeof line, f)) {
    if (line[0] == '#' || sscanf(line, "%63s %d", name, &v) != 2) continue;
    for (int k = 0; k < NKNOB; k++)
      if (!strcmp(name, knob_name[k])) *knob(t, k) = v;
    if (!strcmp(name, "gate")) t->gate = v;
    if (!strcmp(name, "tape_words")) t->tape_words = v;
    if (!strcmp(name, "wide_nodes")) t->wide_nodes = v;
  }
  fclose(f);
  return E13_OK;
}

/* Where this device's stored tuning lives: $E13_TUNING_FILE, else
   <library dir>/../tuning/<arch>_<cus>cu.txt; the small-search profile at the
   same path with _small before the extension. */
static void small_name(char *buf, int len) {
  char tmp[1024];
  snprintf(tmp, sizeof tmp, "%s", buf);
  char *dot = strrchr(tmp, '.'), *slash = strrchr(tmp, '/');
  if (dot && (!slash || dot > slash)) {
    *dot = 0;
    snprintf(buf, (size_t)len, "%s_small.%s", tmp, dot + 1);
  } else {
    snprintf(buf, (size_t)len, "%s_small", tmp);
  }
}

int e13i_tuning_path(const char *arch, int ncu, int small, char *buf, int len) {
  const char *env = getenv("E13_TUNING_FILE");
  if (env && *env) {
    snprintf(buf, (size_t)len, "%s", env);
    if (small) small_name(buf, len);
    return 1;
  }
  Dl_info info;
  if (!dladdr((void *)&e13_autotune, &info) || !info.dli_fname) return 0;
  char tmp[1024];
  snprintf(tmp, sizeof tmp, "%s", info.dli_fname);
  char a[64];
  snprintf(a, sizeof a, "%s", arch);
  char *colon = strchr(a, ':');
  if (colon) *colon = 0;
  snprintf(buf, (size_t)len, "%s/../tuning/%s_%dcu.txt", dirname(tmp), a, ncu);
  if (small) small_name(buf, len);
  return 1;
}

int e13_tuning_file(e13_ctx *c, char *buf, int len) {
  if (!c || !buf || len < 2) return E13_EINVAL;
  return e13i_tuning_path(c->arch, c->ncu, 0, buf, len) ? E13_OK : E13_EINVAL;
}

int e13_tuning_file_small(e13_ctx *c, char *buf, int len) {
  if (!c || !buf || len < 2) return E13_EINVAL;
  return e13i_tuning_path(c->arch, c->ncu, 1, buf, len) ? E13_OK : E13_EINVAL;
}
