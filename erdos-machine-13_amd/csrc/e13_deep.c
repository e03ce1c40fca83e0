/* e13_deep.c -- deep stage check of survivors (the reference's src/deep.c):
 * each SURVIVOR line's table is run from the blank tape on a tape long enough
 * for stage NMAX, looking for the stage starts n = 1..NMAX (state st0, head at
 * head0, or at head0 - 2 + 2^n for a relative start, and cells -1 .. 2^n - 2
 * exactly the ones at cells >= -2), and at each start the counter's ones left
 * of cell -2 are counted.  It stops at the first stage not reached within its
 * step limit, at a halt, at a free cell the search never read, or off the
 * tape.  The output lines are the reference's, survivor by survivor.
 *
 * These are a few long sequential runs (a survivor to stage 15 is ~1e9
 * steps), so they run on host threads, one survivor at a time per thread.
 */
#include <pthread.h>
#include <stdarg.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "e13.h"

#define DEEP_LEFT 4096 /* cells left of cell 0 (deep.c:13) */

typedef struct {
  int st0, rel, nq;
  long long head0;
  int8_t wr[32][2], dir[32][2], nx[32][2];
  uint8_t stop[32][2]; /* 1 halt, 2 a free cell the search never read ('?---') */
} deep_machine;

static int parse_table(const char *p, deep_machine *m);

/* The fields of one SURVIVOR line (st0=, head0=, abs=/rel=, table=); 0 if the
   line has no table or no st0 (deep.c:20-47). */
static int deep_parse(const char *line, deep_machine *m) {
  const char *tp = strstr(line, "table="), *sp = strstr(line, "st0=");
  if (!tp || !sp) return 0;
  memset(m, 0, sizeof *m);
  m->st0 = atoi(sp + 4);
  const char *hp = strstr(line, "head0="), *ap = strstr(line, "abs="), *rp = strstr(line, "rel=");
  m->head0 = hp ? atoll(hp + 6) : -2;
  m->rel = ap && rp && !atoi(ap + 4) && atoi(rp + 4);
  return parse_table(tp + 6, m);
}

/* The same for the counter re-check, whose defaults differ: st0 0, head0 -2,
   abs 1, rel 0, and only table= is required (recheck.c:31-38). */
static int deep_parse_recheck(const char *line, deep_machine *m) {
  const char *tp = strstr(line, "table=");
  if (!tp) return 0;
  memset(m, 0, sizeof *m);
  const char *sp = strstr(line, "st0="), *hp = strstr(line, "head0="), *ap = strstr(line, "abs="),
             *rp = strstr(line, "rel=");
  m->st0 = sp ? atoi(sp + 4) : 0;
  m->head0 = hp ? atoi(hp + 6) : -2;
  const int abs_ok = ap ? atoi(ap + 4) : 1, rel_ok = rp ? atoi(rp + 4) : 0;
  m->rel = !abs_ok && rel_ok;
  return parse_table(tp + 6, m);
}

/* Table cells: rows separated by '|', the two cells of a row by spaces; '?'
   marks a cell never read, '---' a halt. */
static int parse_table(const char *p, deep_machine *m) {
  int q = 0;
  for (;;) { /* rows separated by '|', the two cells of a row by spaces */
    if (q >= 32) return 0;
    for (int a = 0; a < 2; a++) {
      while (*p == ' ') p++;
      const int unread = *p == '?';
      if (unread) p++;
      if (p[0] == '-') {
        m->stop[q][a] = unread ? 2 : 1;
        p += 3;
      } else {
        m->wr[q][a] = (int8_t)(p[0] - '0');
        m->dir[q][a] = p[1] == 'R' ? 1 : -1;
        char *e;
        m->nx[q][a] = (int8_t)strtol(p + 2, &e, 10);
        p = e;
      }
    }
    q++;
    while (*p == ' ') p++;
    if (*p != '|') break;
    p++;
  }
  m->nq = q;
  return 1;
}

static int appendf(char *out, int len, int *pos, const char *fmt, ...) {
  va_list ap;
  va_start(ap, fmt);
  const int k = vsnprintf(out + *pos, (size_t)(len > *pos ? len - *pos : 0), fmt, ap);
  va_end(ap);
  if (k < 0 || *pos + k >= len) return -1;
  *pos += k;
  return 0;
}

/* One survivor (number id) to stage nmax on tape[0 .. DEEP_LEFT + 2^nmax + 4096). */
static int deep_run(const deep_machine *m, int id, int nmax, uint8_t *tape, char *out, int len) {
  const long long size = DEEP_LEFT + (1LL << nmax) + 4096;
  memset(tape, 0, (size_t)size);
  uint8_t *t0 = tape + DEEP_LEFT; /* cell k at t0[k] */
  long long head = 0, t = 0, ones = 0; /* ones at cells >= -2 */
  int st = 0, pos = 0, failed = 0;
  if (snprintf(out, (size_t)len, "survivor %d (nq=%d st0=%d%s):", id, m->nq, m->st0, m->rel ? " rel" : "") >= len)
    return -1;
  pos = (int)strlen(out);
  for (int n = 1; n <= nmax && !failed; n++) {
    const long long blk = 1LL << n, limit = t + 64LL * (1LL << (2 * n)) + 100000;
    const long long at = m->rel ? blk + m->head0 - 2 : m->head0;
    for (;;) {
      if (st == m->st0 && head == at && ones == blk) { /* the block -1 .. 2^n-2 all ones */
        long long k = -1;
        while (k < blk - 1 && t0[k]) k++;
        if (k == blk - 1) break;
      }
      if (t >= limit) {
        if (appendf(out, len, &pos, " | stage %d NOT reached (limit, t=%lld)", n, t)) return -1;
        failed = 1;
        break;
      }
      if (head < -DEEP_LEFT + 2 || head > blk + 2000) {
        if (appendf(out, len, &pos, " | stage %d: off tape at %lld", n, head)) return -1;
        failed = 1;
        break;
      }
      const int s = t0[head];
      if (m->stop[st][s] == 2) {
        if (appendf(out, len, &pos, " | FREE cell %c%d read before stage %d (t=%lld)", 'A' + st, s, n, t)) return -1;
        failed = 1;
        break;
      }
      if (m->stop[st][s]) {
        if (appendf(out, len, &pos, " | HALT before stage %d (t=%lld)", n, t)) return -1;
        failed = 1;
        break;
      }
      const int w = m->wr[st][s];
      if (w != s) {
        t0[head] = (uint8_t)w;
        if (head >= -2) ones += w ? 1 : -1;
      }
      head += m->dir[st][s];
      st = m->nx[st][s];
      t++;
    }
    if (failed) break;
    long long cnt = 0, lo = 0; /* the counter: ones left of cell -2 */
    for (long long k = -3; k >= -DEEP_LEFT + 1; k--)
      if (t0[k]) {
        cnt++;
        lo = k;
      }
    if (appendf(out, len, &pos, " n%d:%lld", n, cnt)) return -1;
    if (cnt && appendf(out, len, &pos, "@%lld", lo)) return -1;
  }
  if (appendf(out, len, &pos, "\n")) return -1;
  return pos;
}

int e13_deep_check_line(const char *line, int id, int nmax, char *out, int outlen) {
  if (!line || !out || outlen < 2 || nmax < 1 || nmax > 24) return E13_EINVAL;
  deep_machine m;
  if (!deep_parse(line, &m)) return E13_EINVAL;
  uint8_t *tape = (uint8_t *)malloc((size_t)(DEEP_LEFT + (1LL << nmax) + 4096));
  if (!tape) return E13_ENOMEM;
  const int r = deep_run(&m, id, nmax, tape, out, outlen);
  free(tape);
  return r < 0 ? E13_EINVAL : r;
}

/* Many survivors on host threads: every line of `text` with table= and st0=
   is one survivor, numbered from 1 in input order; the result lines go to out
   in the same order. */
typedef struct {
  const char **lines;
  int n, nmax, recheck; /* recheck: the counter re-check instead of the deep check */
  char *res;            /* n x DEEP_OUT */
  int next, err;
  pthread_mutex_t mu;
} deep_job;

#define DEEP_OUT 20480 /* a result line (a counter re-check line holds two sets of up to 8000 bytes) */

static int rc_run(const deep_machine *m, int id, int nmax, uint8_t *tape, char (*ctr)[8192], char *out, int len);
static int deep_parse_recheck(const char *line, deep_machine *m);

static void *deep_worker(void *arg) {
  deep_job *j = (deep_job *)arg;
  uint8_t *tape = (uint8_t *)malloc(j->recheck ? (size_t)(2 * 4000000LL + 1) :
                                                 (size_t)(DEEP_LEFT + (1LL << j->nmax) + 4096));
  char (*ctr)[8192] = j->recheck ? (char (*)[8192])calloc(20, 8192) : NULL;
  for (;;) {
    pthread_mutex_lock(&j->mu);
    const int i = j->next++;
    pthread_mutex_unlock(&j->mu);
    if (i >= j->n) break;
    deep_machine m;
    char *o = j->res + (size_t)i * DEEP_OUT;
    const int ok = tape && (j->recheck ? ctr && deep_parse_recheck(j->lines[i], &m) &&
                                             rc_run(&m, i + 1, j->nmax, tape, ctr, o, DEEP_OUT) >= 0
                                       : deep_parse(j->lines[i], &m) &&
                                             deep_run(&m, i + 1, j->nmax, tape, o, DEEP_OUT) >= 0);
    if (!ok) {
      pthread_mutex_lock(&j->mu);
      j->err = 1;
      pthread_mutex_unlock(&j->mu);
    }
  }
  free(tape);
  free(ctr);
  return NULL;
}

static int deep_many(const char *text, int nmax, int threads, char *out, int64_t outlen, int recheck) {
  /* the survivor lines, in order */
  char *copy = strdup(text);
  if (!copy) return E13_ENOMEM;
  int cap = 1024, n = 0;
  const char **lines = (const char **)malloc(sizeof(char *) * (size_t)cap);
  for (char *save = NULL, *l = strtok_r(copy, "\n", &save); lines && l; l = strtok_r(NULL, "\n", &save)) {
    if (!strstr(l, "table=") || (!recheck && !strstr(l, "st0="))) continue;
    if (n == cap) {
      cap *= 2;
      const char **nl = (const char **)realloc(lines, sizeof(char *) * (size_t)cap);
      if (!nl) {
        free(lines);
        lines = NULL;
        break;
      }
      lines = nl;
    }
    lines[n++] = l;
  }
  deep_job j;
  memset(&j, 0, sizeof j);
  j.lines = lines;
  j.n = n;
  j.nmax = nmax;
  j.recheck = recheck;
  j.res = n ? (char *)calloc((size_t)n, DEEP_OUT) : NULL;
  int rc = lines && (!n || j.res) ? E13_OK : E13_ENOMEM;
  if (!rc && n) {
    pthread_mutex_init(&j.mu, NULL);
    int nt = threads > 0 ? threads : 1;
    if (nt > n) nt = n;
    if (nt > 256) nt = 256;
    pthread_t tid[256];
    int started = 0;
    for (; started < nt; started++)
      if (pthread_create(&tid[started], NULL, deep_worker, &j)) break;
    if (!started) deep_worker(&j);
    for (int i = 0; i < started; i++) pthread_join(tid[i], NULL);
    pthread_mutex_destroy(&j.mu);
    if (j.err) rc = E13_EINVAL;
  }
  int64_t pos = 0;
  for (int i = 0; !rc && i < n; i++) {
    const size_t k = strlen(j.res + (size_t)i * DEEP_OUT);
    if (pos + (int64_t)k >= outlen) {
      rc = E13_EINVAL;
      break;
    }
    memcpy(out + pos, j.res + (size_t)i * DEEP_OUT, k);
    pos += (int64_t)k;
  }
  out[pos < outlen ? pos : outlen - 1] = 0;
  free(j.res);
  free(lines);
  free(copy);
  return rc ? rc : n;
}

/* ------------------------------------------------------------------------ */
/* Counter re-check (the reference's src/recheck.c): the same stage starts   */
/* from the blank tape on a tape of 4e6 cells each side, with the EXACT set  */
/* of counter ones (cells <= -3) at each start: the set must be the same at */
/* every stage of one counter class c(n) (0 for n <= 2, 1 for 3..8, 2 from 9 */
/* on) and differ between classes.  One verdict line per survivor.          */
/* ------------------------------------------------------------------------ */
#define RC_SIDE 4000000LL /* cells each side of cell 0 (recheck.c:17-18) */
#define RC_SET 8192       /* bytes of one counter set as text (the worker's ctr rows) */

/* The counter set at a stage start: "k," for every one at cells -3 down to
   minhead - 1, while the text is under 8000 bytes (recheck.c:81-83). */
static void rc_counter(const uint8_t *t0, long long minhead, char *q) {
  int len = 0;
  q[0] = 0;
  for (long long k = -3; k >= minhead - 1 && len < 8000; k--)
    if (t0[k]) len += snprintf(q + len, (size_t)(RC_SET - len), "%lld,", k);
}

static int rc_run(const deep_machine *m, int id, int nmax, uint8_t *tape, char (*ctr)[8192], char *out, int len) {
  memset(tape, 0, (size_t)(2 * RC_SIDE + 1));
  uint8_t *t0 = tape + RC_SIDE;
  long long head = 0, t = 0, minhead = 0, ones = 0;
  int st = 0, failed = 0, pos = 0;
  char verdict[2 * RC_SET + 128];
  verdict[0] = 0;
  for (int n = 1; n <= nmax && !failed; n++) {
    const long long blk = 1LL << n, limit = t + 8LL * (1LL << (2 * n)) + 100000;
    const long long at = m->rel ? m->head0 + blk - 2 : m->head0;
    for (;;) {
      if (st == m->st0 && head == at && ones == blk) {
        long long k = -1;
        while (k < blk - 1 && t0[k]) k++;
        if (k == blk - 1) break;
      }
      if (t >= limit) {
        snprintf(verdict, sizeof verdict, "stage %d not reached (limit)", n);
        failed = 1;
        break;
      }
      if (head <= -RC_SIDE + 2 || head >= RC_SIDE - 2) {
        snprintf(verdict, sizeof verdict, "stage %d: off tape", n);
        failed = 1;
        break;
      }
      const int s = t0[head];
      if (m->stop[st][s] == 2) {
        snprintf(verdict, sizeof verdict, "free cell %c%d read before stage %d", 'A' + st, s, n);
        failed = 1;
        break;
      }
      if (m->stop[st][s]) {
        snprintf(verdict, sizeof verdict, "halt before stage %d", n);
        failed = 1;
        break;
      }
      const int w = m->wr[st][s];
      if (w != s) {
        t0[head] = (uint8_t)w;
        if (head >= -2) ones += w ? 1 : -1;
      }
      head += m->dir[st][s];
      if (head < minhead) minhead = head;
      st = m->nx[st][s];
      t++;
    }
    if (failed) break;
    rc_counter(t0, minhead, ctr[n]);
    const int c = n <= 2 ? 0 : (n <= 8 ? 1 : 2), first = c == 0 ? 1 : (c == 1 ? 3 : 9);
    if (n != first && strcmp(ctr[n], ctr[first])) {
      snprintf(verdict, sizeof verdict, "counter changes at n=%d (%s -> %s)", n, ctr[first][0] ? ctr[first] : "none",
               ctr[n][0] ? ctr[n] : "none");
      failed = 1;
    }
  }
  if (!failed) {
    if (nmax >= 9 && (!strcmp(ctr[1], ctr[3]) || !strcmp(ctr[3], ctr[9]) || !strcmp(ctr[1], ctr[9])))
      snprintf(verdict, sizeof verdict, "counter equal for different c");
    else
      snprintf(verdict, sizeof verdict, "CONSISTENT to n=%d (c2 counter %s)", nmax, nmax >= 9 ? ctr[9] : "");
  }
  if (appendf(out, len, &pos, "%d: %s\n", id, verdict)) return -1;
  return pos;
}

int e13_recheck_counter_line(const char *line, int id, int nmax, char *out, int outlen) {
  if (!line || !out || outlen < 2 || nmax < 1 || nmax > 19) return E13_EINVAL;
  deep_machine m;
  if (!deep_parse_recheck(line, &m)) return E13_EINVAL;
  uint8_t *tape = (uint8_t *)malloc((size_t)(2 * RC_SIDE + 1));
  char (*ctr)[RC_SET] = (char (*)[RC_SET])calloc(20, RC_SET);
  int r = E13_ENOMEM;
  if (tape && ctr) {
    r = rc_run(&m, id, nmax, tape, ctr, out, outlen);
    if (r < 0) r = E13_EINVAL;
  }
  free(tape);
  free(ctr);
  return r;
}

int e13_deep_check(const char *text, int nmax, int threads, char *out, int64_t outlen) {
  if (!text || !out || outlen < 1 || nmax < 1 || nmax > 24) return E13_EINVAL;
  return deep_many(text, nmax, threads, out, outlen, 0);
}

int e13_recheck_counter(const char *text, int nmax, int threads, char *out, int64_t outlen) {
  if (!text || !out || outlen < 1 || nmax < 1 || nmax > 19) return E13_EINVAL;
  return deep_many(text, nmax, threads, out, outlen, 1);
}
